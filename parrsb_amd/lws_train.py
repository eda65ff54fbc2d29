"""Training and evaluation of the learned warm start, on the GPU, with the
library's own forward / backward / Adam kernels (rsbgpu_lws_*).

Training data are the Fiedler vectors the plain FP64 solver computes while
partitioning the seeded meshes of meshgen.warm_start_mesh: at every
bisection level, after the plain solve, a few Adam steps fit the network's
prediction to that level's Fiedler vectors (loss 1 - cos^2 per segment,
averaged over the level's segments), then the level is bisected as
parrsb_part_mesh does and the next level follows.

    python -m parrsb_amd.lws_train --meshes 0:600 --out parrsb_amd/lws_weights.bin

Parameter file: b"LWS1", int32 count, count float32 values (little endian).
"""
from __future__ import annotations

import argparse
import json
import math
import struct
import sys
import time

import numpy as np

from . import Device, lib, meshgen, warm_start_nparams

MAGIC = b"LWS1"
H, NIN, NL = 64, 8, 4


def save_params(path, P):
    P = np.ascontiguousarray(P, dtype="<f4")
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<i", P.size))
        f.write(P.tobytes())


def load_params(path):
    with open(path, "rb") as f:
        if f.read(4) != MAGIC:
            raise ValueError(f"{path}: not a warm-start parameter file")
        (n,) = struct.unpack("<i", f.read(4))
        P = np.frombuffer(f.read(4 * n), dtype="<f4").astype(np.float32)
    if P.size != n or n != warm_start_nparams():
        raise ValueError(f"{path}: {P.size} parameters, the network has {warm_start_nparams()}")
    return P


def init_params(seed=0):
    """He-style initialisation in the layout of lws.hip:
    W0 [64 x 8] | b0 | 4 x (W_l [64 x 128] | b_l) | w_o | b_o."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal(H * NIN) * math.sqrt(2.0 / 6), np.zeros(H)]
    for _ in range(NL):
        parts += [rng.standard_normal(H * 2 * H) * math.sqrt(1.0 / (2 * H)), np.zeros(H)]
    parts += [rng.standard_normal(H) * math.sqrt(1.0 / H), np.zeros(1)]
    P = np.concatenate(parts).astype(np.float32)
    assert P.size == warm_start_nparams()
    return P


# split arithmetic of the level loop (the same as the C driver's, rsb.c:
# first p - (n mod p) ranks hold floor(n/p) elements, the others one more)
def lower_count(n, p):
    q, rem, hl = n // p, n % p, (p + 1) // 2
    return q * hl + max(hl - (p - rem), 0)


def balance_expected(n, p, b):
    q, rem = n // p, n % p
    nph = (p + 1) // 2 if b == 0 else p // 2
    return q * nph + rem // 2 + (rem % 2) * (1 - b)


def run_levels(dev, nel, nparts, on_level=None, precision=64, max_iter=50, max_passes=50,
               tol=1e-5):
    """The RSB level loop through the step-by-step device API; on_level(cut,
    sizes, out) runs after each level's Fiedler solve, before the sort
    (sizes: elements of the level's segments that are bisected)."""
    segs = [(0, nel, nparts, 0)]
    ncuts = meshgen.ncuts(nparts)
    for cut in range(ncuts):
        cs, n, p, r0 = (np.array(x, dtype=np.int32) for x in zip(*segs))
        dev.set_segments(cs, n, p, r0)
        dev.prepartition(1)
        dev.level_laplacian()
        out = dev.fiedler(len(segs), precision, max_iter, max_passes, tol)
        if on_level is not None:
            on_level(cut, n[p > 1], out)
        dev.fiedler_sort()
        nl = n.copy()
        snd = np.full(len(segs), -1, np.int32)
        cnt = np.zeros(len(segs), np.int32)
        for s, (c0, ns, ps, _) in enumerate(segs):
            if ps <= 1:
                continue
            l = lower_count(ns, ps)
            e0, e1 = balance_expected(ns, ps, 0), balance_expected(ns, ps, 1)
            nl[s] = l
            if l > e0:
                snd[s], cnt[s] = 0, l - e0
            elif ns - l > e1:
                snd[s], cnt[s] = 1, ns - l - e1
        nl, _, _ = dev.bisect(nl, snd, cnt, 1)
        nxt = []
        for s, (c0, ns, ps, rs) in enumerate(segs):
            if ps <= 1:
                nxt.append((c0, ns, ps, rs))
                continue
            pl, pu = (ps + 1) // 2, ps // 2
            nxt += [(c0, int(nl[s]), pl, rs), (c0 + int(nl[s]), ns - int(nl[s]), pu, rs + pl)]
        segs = nxt


def parts_for(nel, per_part=1000):
    """Power-of-two part count giving about per_part elements per part."""
    return 1 << max(1, int(round(math.log2(max(nel / per_part, 2)))))


def train(meshes, steps_per_level=2, lr=1e-3, seed=0, init=None, out=None, log=print,
          per_part=1000, max_n=0):
    """Adam on the levels of meshgen.warm_start_mesh(i) for i in meshes; with
    max_n > 0 only levels whose segments have at most max_n elements (where
    the start is used) are trained on."""
    P0 = np.asarray(init if init is not None else init_params(seed), dtype=np.float32)
    holder = Device()  # owns the parameter, gradient and moment buffers
    z = np.zeros_like(P0)
    dP, dG, dm, dv = holder.upload(P0), holder.upload(z), holder.upload(z), holder.upload(z)
    step = [0]
    hist = []
    t0 = time.time()
    for idx, i in enumerate(meshes):
        vtx, xyz, nv, info = meshgen.warm_start_mesh(i)
        nparts = parts_for(info["nel"], per_part)
        losses = []

        def on_level(cut, sizes, o):
            if max_n > 0 and (sizes.size == 0 or sizes.max() > max_n):
                return
            for _ in range(steps_per_level):
                loss = dev.lws_train_step(dP, dG, float(cut))
                step[0] += 1
                rc = lib().rsbgpu_lws_adam(dP, dG, dm, dv, step[0], lr)
                if rc:
                    raise RuntimeError("rsbgpu_lws_adam failed")
                losses.append(loss)

        with Device() as dev:
            dev.set_mesh(vtx, xyz, nv)
            run_levels(dev, info["nel"], nparts, on_level)
        if losses:
            hist.append(dict(mesh=i, nel=info["nel"], kind=info["kind"], nparts=nparts,
                             loss_first=float(np.mean(losses[::steps_per_level])),
                             loss_last=float(np.mean(losses[steps_per_level - 1::steps_per_level]))))
        if ((idx + 1) % 20 == 0 or idx + 1 == len(meshes)) and hist:
            recent = hist[-20:]
            log(f"lws_train: {idx + 1}/{len(meshes)} meshes, {step[0]} steps, "
                f"{time.time() - t0:.0f} s, loss {np.mean([h['loss_first'] for h in recent]):.4f}")
            if out:
                save_params(out, holder.download(dP, P0))
    P = holder.download(dP, P0)
    for d in (dP, dG, dm, dv):
        holder.free(d)
    holder.close()
    if out:
        save_params(out, P)
    return P, hist


def evaluate(params, mesh, nparts, passes=1, precision=32, max_n=0):
    """Partition `mesh` = (vtx, xyz, nv) with the plain and with the learned
    start through parrsb_part_mesh; returns the reports' iteration counts,
    Lanczos time, fallbacks and partition quality of both."""
    import parrsb_amd as pa
    vtx, xyz, nv = mesh
    res = {}
    for name, P in (("plain", None), ("learned", params)):
        pa.set_precision(precision)
        pa.set_warm_start(P, passes, max_n)
        part, _ = pa.part_mesh(vtx, xyz, nv, nparts)
        r = pa.last_report()
        q = pa.quality(part, vtx, nv, nparts)
        res[name] = dict(work=float(np.sum(r["niter_mean"])), seconds=r["seconds"],
                         lanczos_ms=r["phase_ms"]["lanczos"], niter_max=r["niter_max"],
                         niter_mean=[round(x, 1) for x in r["niter_mean"]],
                         nfallback=r["nfallback"], nsegs=r["nsegs"],
                         fails=int(np.sum(r["nconverged_fail"])), edge_cut=q["edge_cut"],
                         max_part=q["max_part"], min_part=q["min_part"],
                         components=q["components"])
    pa.set_warm_start(None)
    res["work_ratio"] = res["learned"]["work"] / max(res["plain"]["work"], 1e-30)
    res["edge_cut_ratio"] = res["learned"]["edge_cut"] / max(res["plain"]["edge_cut"], 1)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--meshes", default="0:600", help="training mesh range a:b")
    ap.add_argument("--steps", type=int, default=2, help="Adam steps per level")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="parameter file to continue from")
    ap.add_argument("--out", default=None)
    ap.add_argument("--eval", default=None,
                    help="held-out evaluation: 'pipe:nr,nt,nz,nparts' after training")
    ap.add_argument("--passes", type=int, default=1)
    ap.add_argument("--max-n", type=int, default=2048,
                    help="train on and evaluate with segments of at most this many elements")
    ap.add_argument("--history", default=None, help="write the per-mesh losses (JSON)")
    a = ap.parse_args(argv)
    lo, hi = (int(x) for x in a.meshes.split(":"))
    log = lambda m: (print(m), sys.stdout.flush())
    init = load_params(a.init) if a.init else None
    if hi > lo:
        P, hist = train(range(lo, hi), a.steps, a.lr, a.seed, init, a.out, log, max_n=a.max_n)
        if a.history:
            with open(a.history, "w") as f:
                json.dump(hist, f)
    else:
        P = init
    if a.eval:
        kind, dims = a.eval.split(":")
        nr, nt, nz, nparts = (int(x) for x in dims.split(","))
        mesh = meshgen.pipe_hex(nr, nt, nz, seed=7)
        res = evaluate(P, mesh, nparts, a.passes, max_n=a.max_n)
        log(json.dumps(res))


if __name__ == "__main__":
    main()
