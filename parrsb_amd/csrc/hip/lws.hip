// Learned warm start: a message-passing network over the level graph that
// predicts a start vector for each segment's Lanczos solve.
//
//   x_i   = (u1, u2, u3, degree/26, level/16, 1, 0, 0): the centroid of row i
//           relative to its segment's bounding-box centre, divided by the
//           largest extent, axes ordered by decreasing extent
//   h0    = relu(W0 x + b0)                                   8 -> 64
//   a_l   = sum_j w_ij h_l,j / sum_j w_ij   (shared-vertex weights, same
//           segment)
//   h_l+1 = relu([h_l | a_l] W_l^T + b_l)                   128 -> 64, l = 0..3
//   y     = w_o . h4 + b_o
// h and a are BF16; the 128 x 64 products run on MFMA (BF16 in, FP32
// accumulate).  Training (forward, backward, Adam) uses the same kernels.
//
// Parameters (FP32 master copy, offsets in floats):
//   W0 [64 x 8] | b0 [64] | 4 x (W_l [64 x 128] | b_l [64]) | w_o [64] | b_o
#include <hip/hip_bf16.h>

#include "state.hpp"

#define LWS_H 64
#define LWS_IN 8
#define LWS_L 4
#define LWS_P0 0
#define LWS_B0 (LWS_H * LWS_IN)
#define LWS_PL(l) (LWS_B0 + LWS_H + (l) * (LWS_H * 2 * LWS_H + LWS_H))
#define LWS_BL(l) (LWS_PL(l) + LWS_H * 2 * LWS_H)
#define LWS_WO LWS_PL(LWS_L)
#define LWS_BO (LWS_WO + LWS_H)
#define LWS_NPARAM (LWS_BO + 1)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float f) {
  // round to nearest even
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// ---------------------------------------------------------------------------
// C[M x N] = A[M x K] B[K x N] (+ bias, relu): A and B BF16 row-major, FP32
// accumulate on MFMA 16x16x16.  A wavefront computes 16 x N strips (N <= 128,
// K a multiple of 16); B is staged once per workgroup in LDS, transposed
// (N x K) so that a lane's four k values of one column are one 8-byte read,
// and the workgroup then walks strips.  Lane layout of
// v_mfma_f32_16x16x16_bf16: A: lane l holds row l%16, k = 4*(l/16)+0..3;
// B: lane l holds column l%16, same k; C: lane l holds column l%16, rows
// 4*(l/16)+0..3.  Epilogue: +bias[n], optional relu, BF16 or FP32 out.
#define LWS_GEMM_BLOCKS 2048
template <int N, int K, bool RELU, bool OUT_BF16>
__global__ void __launch_bounds__(256) k_gemm_rows(int M, const unsigned short *__restrict__ A0,
                                                   const unsigned short *__restrict__ A1,
                                                   int lda, const unsigned short *__restrict__ Bg,
                                                   const float *__restrict__ bias,
                                                   void *__restrict__ C, int ldc) {
  // A = [A0 | A1] when A1 is given (each lda wide), else A0 (K wide)
  __shared__ __align__(16) unsigned short Bt[N * K];
  for (int i = threadIdx.x; i < K * N; i += blockDim.x) {
    const int k = i / N, n = i % N;
    Bt[n * K + k] = Bg[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kq = (lane >> 4) * 4;
  float bc[N / 16];
#pragma unroll
  for (int nb = 0; nb < N / 16; nb++) bc[nb] = bias ? bias[nb * 16 + (lane & 15)] : 0.f;
  const int nstrips = (M + 15) / 16;
  for (int strip = blockIdx.x * 4 + wave; strip < nstrips; strip += gridDim.x * 4) {
    const int row0 = strip * 16, r = row0 + (lane & 15);
    f32x4 acc[N / 16];
#pragma unroll
    for (int nb = 0; nb < N / 16; nb++) acc[nb] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
      bf16x4 a = (bf16x4){0, 0, 0, 0};
      if (r < M) {
        const unsigned short *ap = (A1 && k0 >= lda) ? A1 + (size_t)r * lda + (k0 - lda) + kq
                                                     : A0 + (size_t)r * lda + k0 + kq;
        a = *(const bf16x4 *)ap;
      }
#pragma unroll
      for (int nb = 0; nb < N / 16; nb++) {
        const int c = nb * 16 + (lane & 15);
        const bf16x4 b = *(const bf16x4 *)&Bt[c * K + k0 + kq];
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nb], 0, 0, 0);
      }
    }
#pragma unroll
    for (int nb = 0; nb < N / 16; nb++) {
      const int c = nb * 16 + (lane & 15);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int rr = row0 + kq + i;
        if (rr >= M) continue;
        float v = acc[nb][i] + bc[nb];
        if (RELU) v = v > 0.f ? v : 0.f;
        if (OUT_BF16) ((unsigned short *)C)[(size_t)rr * ldc + c] = f2bf(v);
        else ((float *)C)[(size_t)rr * ldc + c] = v;
      }
    }
  }
}

// dW[N x K] += D[M x N]^T X[M x K] over a chunk of rows per workgroup
// (MFMA with the row index as the reduction dimension), FP32 atomics.
template <int N, int K>
__global__ void __launch_bounds__(256) k_gemm_wgrad(int M, int rows_per_block,
                                                    const unsigned short *__restrict__ D, int ldd,
                                                    const unsigned short *__restrict__ X0,
                                                    const unsigned short *__restrict__ X1,
                                                    int ldx, float *__restrict__ dW) {
  // X = [X0 | X1] (each ldx wide) when X1 is given
  // each wave owns a 16 x 16 block of dW at a time: (N/16) x (K/16) blocks
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = blockIdx.x * rows_per_block, r1 = min(M, r0 + rows_per_block);
  const int kq = (lane >> 4) * 4;
  for (int blk = wave; blk < (N / 16) * (K / 16); blk += 4) {
    const int nb = blk / (K / 16), kb = blk % (K / 16);
    f32x4 acc = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int m0 = r0; m0 < r1; m0 += 16) {
      // A = D^T: row n (lane%16), reduction index m = m0 + kq + i
      // B = X: column k (lane%16), reduction index m = m0 + kq + i
      bf16x4 a, b;
      for (int i = 0; i < 4; i++) {
        const int m = m0 + kq + i;
        short av = 0, bv = 0;
        if (m < r1) {
          const int kk = kb * 16 + (lane & 15);
          av = (short)D[(size_t)m * ldd + nb * 16 + (lane & 15)];
          bv = (short)((X1 && kk >= ldx) ? X1[(size_t)m * ldx + kk - ldx] : X0[(size_t)m * ldx + kk]);
        }
        a[i] = av;
        b[i] = bv;
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    for (int i = 0; i < 4; i++) {
      const int n = nb * 16 + (lane >> 4) * 4 + i, k = kb * 16 + (lane & 15);
      atomicAdd(&dW[(size_t)n * K + k], acc[i]);
    }
  }
}

// ---------------------------------------------------------------------------
// Features of every real row of the level.  Bounding box per segment: the 64
// rows of a wavefront belong to one segment (segments start 64-aligned), so
// each wavefront reduces its rows first and issues one atomic per bound.
__global__ void k_lws_bbox(const int *__restrict__ row_elem, const int *__restrict__ segof,
                           const double *__restrict__ coord, int nel, int64_t mpad, int ndim,
                           unsigned long long *__restrict__ box) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int e = row < mpad ? row_elem[row] : -1;
  // segment of the wavefront: its first row is a real row unless the whole
  // wavefront is padding past the last segment
  const int e0 = __shfl(e, 0, RSB_WAVE);
  if (e0 < 0) return;
  const int s = segof[e0];
  const int lane = threadIdx.x & (RSB_WAVE - 1);
  for (int d = 0; d < ndim; d++) {
    uint64_t mn = ~0ull, mx = 0ull;
    if (e >= 0) mn = mx = dkey(coord[(size_t)d * nel + e]);
    for (int o = RSB_WAVE / 2; o > 0; o >>= 1) {
      const uint64_t a = __shfl_xor(mn, o, RSB_WAVE), b = __shfl_xor(mx, o, RSB_WAVE);
      mn = a < mn ? a : mn;
      mx = b > mx ? b : mx;
    }
    if (lane == 0) {
      atomicMin(&box[(size_t)s * 6 + 2 * d], (unsigned long long)mn);
      atomicMax(&box[(size_t)s * 6 + 2 * d + 1], (unsigned long long)mx);
    }
  }
}

__global__ void k_lws_box_init(int nseg, unsigned long long *__restrict__ box) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  for (int d = 0; d < 3; d++) box[s * 6 + 2 * d] = ~0ull, box[s * 6 + 2 * d + 1] = 0ull;
}

__global__ void k_lws_features(const int *__restrict__ row_elem, const int *__restrict__ segof,
                               const double *__restrict__ coord, int nel, int64_t mpad,
                               int ndim, const unsigned long long *__restrict__ box,
                               const int *__restrict__ rowlen, float level,
                               float *__restrict__ x) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= mpad) return;
  float *o = x + (size_t)row * LWS_IN;
  const int e = row_elem[row];
  if (e < 0) {
    for (int k = 0; k < LWS_IN; k++) o[k] = 0.f;
    return;
  }
  const int s = segof[e];
  double lo[3] = {0, 0, 0}, len[3] = {0, 0, 0}, c[3] = {0, 0, 0};
  for (int d = 0; d < ndim; d++) {
    lo[d] = dkey_inv(box[(size_t)s * 6 + 2 * d]);
    len[d] = dkey_inv(box[(size_t)s * 6 + 2 * d + 1]) - lo[d];
    c[d] = coord[(size_t)d * nel + e];
  }
  // axes by decreasing extent (ties: lower axis first)
  int ax[3] = {0, 1, 2};
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++)
      if (len[ax[j]] > len[ax[i]]) {
        int t = ax[i];
        ax[i] = ax[j];
        ax[j] = t;
      }
  const double L = len[ax[0]] > 0 ? len[ax[0]] : 1.0;
  for (int k = 0; k < 3; k++) {
    const int d = ax[k];
    o[k] = (float)((c[d] - lo[d] - 0.5 * len[d]) / L);
  }
  o[3] = rowlen[row] / 26.f;
  o[4] = level / 16.f;
  o[5] = 1.f;
  o[6] = 0.f;
  o[7] = 0.f;
}

// h0 = relu(W0 x + b0), FP32 math, BF16 out: one thread per row, its 8
// inputs and 64 outputs moved as 16-byte words (each output summed over k
// in order).
__global__ void k_lws_in(int64_t M, const float *__restrict__ x, const float *__restrict__ P,
                         unsigned short *__restrict__ h) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= M) return;
  static_assert(LWS_IN == 8 && LWS_H % 8 == 0, "16-byte words of inputs and outputs");
  const float4 x0 = *(const float4 *)(x + r * LWS_IN), x1 = *(const float4 *)(x + r * LWS_IN + 4);
  const float xv[LWS_IN] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
  uint4 *o = (uint4 *)(h + r * LWS_H);
  for (int c0 = 0; c0 < LWS_H; c0 += 8) {
    unsigned pk[4];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int c = c0 + q;
      float v = P[LWS_B0 + c];
#pragma unroll
      for (int k = 0; k < LWS_IN; k++) v += P[LWS_P0 + c * LWS_IN + k] * xv[k];
      const unsigned b = f2bf(v > 0.f ? v : 0.f);
      if (q & 1) pk[q >> 1] |= b << 16;
      else pk[q >> 1] = b;
    }
    o[c0 / 8] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
  }
}

// Weighted neighbour aggregation over the level Laplacian (SELL-64 global
// form): out_i = (norm ? 1/d_i : 1) sum_j w_ij (pre ? in_j / d_j : in_j),
// d_i = sum_j w_ij.  One thread per (row, 8 channels).
__global__ void k_lws_agg(int64_t mpad, const int64_t *__restrict__ slice_off,
                          const int *__restrict__ slice_w, const uint32_t *__restrict__ ell,
                          const unsigned short *__restrict__ in, int ldin, int norm, int pre,
                          const float *__restrict__ dinv, const unsigned short *__restrict__ add,
                          int ldadd, unsigned short *__restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = t >> 3;
  const int cg = (int)(t & 7) * 8;
  if (row >= mpad) return;
  const int64_t sl = row >> 6;
  const uint32_t *e = ell + slice_off[sl] + (row & 63);
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  float dsum = 0.f;
  const int wd = slice_w[sl];
  // four entries and their 16-byte gathers in flight, then summed in order
  for (int j0 = 0; j0 < wd; j0 += 4) {
    uint32_t q[4];
    uint4 v[4];
    float dv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) q[u] = j0 + u < wd ? e[(size_t)(j0 + u) * RSB_WAVE] : 0u;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t c = (q[u] >> 28) ? (q[u] & 0x0fffffffu) : (uint32_t)row;
      v[u] = *(const uint4 *)(in + (size_t)c * ldin + cg);
      dv[u] = pre ? dinv[c] : 1.f;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const float w = (float)(q[u] >> 28);
      if (w == 0.f) continue;
      const float sc = pre ? w * dv[u] : w;
      const unsigned vv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
      for (int k = 0; k < 4; k++) {
        acc[2 * k] += sc * bf2f((unsigned short)(vv[k] & 0xffff));
        acc[2 * k + 1] += sc * bf2f((unsigned short)(vv[k] >> 16));
      }
      dsum += w;
    }
  }
  const float s = norm ? (dsum > 0.f ? 1.f / dsum : 0.f) : 1.f;
  for (int k = 0; k < 8; k++) {
    float v = acc[k] * s;
    if (add) v += bf2f(add[(size_t)row * ldadd + cg + k]);
    out[(size_t)row * LWS_H + cg + k] = f2bf(v);
  }
}

__global__ void k_lws_dinv(int64_t mpad, const int64_t *__restrict__ slice_off,
                           const int *__restrict__ slice_w, const uint32_t *__restrict__ ell,
                           float *__restrict__ dinv) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= mpad) return;
  const int64_t sl = row >> 6;
  const uint32_t *e = ell + slice_off[sl] + (row & 63);
  float d = 0.f;
  for (int j = 0; j < slice_w[sl]; j++) d += (float)(e[(size_t)j * RSB_WAVE] >> 28);
  dinv[row] = d > 0.f ? 1.f / d : 0.f;
}

// y = w_o . h + b_o (FP32): one thread per row, its 64 activations read as
// 16-byte words, summed over c in order.
__global__ void k_lws_out(int64_t M, const unsigned short *__restrict__ h,
                          const float *__restrict__ P, float *__restrict__ y) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= M) return;
  const uint4 *hv = (const uint4 *)(h + r * LWS_H);
  uint4 w[LWS_H / 8];
#pragma unroll
  for (int q = 0; q < LWS_H / 8; q++) w[q] = hv[q];
  float v = P[LWS_BO];
#pragma unroll
  for (int q = 0; q < LWS_H / 8; q++) {
    const unsigned u[4] = {w[q].x, w[q].y, w[q].z, w[q].w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      v += P[LWS_WO + 8 * q + 2 * k] * bf2f((unsigned short)(u[k] & 0xffffu));
      v += P[LWS_WO + 8 * q + 2 * k + 1] * bf2f((unsigned short)(u[k] >> 16));
    }
  }
  y[r] = v;
}

// BF16 copies of the layer weights, transposed to K x N (row-major) for the
// forward product ([h|a] (M x 128) times W_l^T (128 x 64)).
__global__ void k_lws_wt(const float *__restrict__ P, unsigned short *__restrict__ wt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= LWS_L * 2 * LWS_H * LWS_H) return;
  const int l = i / (2 * LWS_H * LWS_H), j = i % (2 * LWS_H * LWS_H);
  const int k = j / LWS_H, n = j % LWS_H;  // wt[l][k][n] = W_l[n][k]
  wt[i] = f2bf(P[LWS_PL(l) + n * 2 * LWS_H + k]);
}

// Row-major BF16 copies of W_l (64 x 128) for the backward product
// dz (M x 64) times W_l (64 x 128).
__global__ void k_lws_w(const float *__restrict__ P, unsigned short *__restrict__ w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= LWS_L * 2 * LWS_H * LWS_H) return;
  const int l = i / (2 * LWS_H * LWS_H), j = i % (2 * LWS_H * LWS_H);
  w[i] = f2bf(P[LWS_PL(l) + j]);
}

// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Training loss of a segment: its Rayleigh quotient relative to the Fiedler
// value the plain solver found, R(y) / lambda - 1 with R(y) = y.Ly / yc.yc
// (yc = y - mean; L annihilates constants, so y.Ly = yc.L yc).  It is
// smallest on the Fiedler eigenspace (any vector of a degenerate one) and
// weighs each eigencomponent of the error by lambda_k / lambda_2: the
// high-frequency error that restarted CG-Lanczos is slowest to remove costs
// the most.  Averaged over the level's segments.

// ly = L y on every row of the level (difference form, SELL-64 global entries)
__global__ void k_lws_lap(int64_t mpad, const int64_t *__restrict__ slice_off,
                          const int *__restrict__ slice_w, const uint32_t *__restrict__ ell,
                          const float *__restrict__ y, float *__restrict__ ly) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= mpad) return;
  const int64_t sl = row >> 6;
  const uint32_t *e = ell + slice_off[sl] + (row & 63);
  const float yi = y[row];
  float acc = 0.f;
  for (int j = 0; j < slice_w[sl]; j++) {
    const uint32_t q = e[(size_t)j * RSB_WAVE];
    acc += (float)(q >> 28) * (yi - y[q & 0x0fffffffu]);
  }
  ly[row] = acc;
}

// sums per segment: phase 0: S[8s] = sum y; phase 1: S[8s+2] = y.Ly,
// S[8s+3] = yc.yc
__global__ void k_lws_sums(const int *__restrict__ tile, const int *__restrict__ seg, int nseg,
                           const float *__restrict__ y, const float *__restrict__ ly, int phase,
                           double *__restrict__ S) {
  const int ti = blockIdx.x, s = tile[3 * ti];
  const int row0 = tile[3 * ti + 1], row1 = tile[3 * ti + 2];
  const int n = seg[nseg + s], pstart = seg[4 * nseg + s];
  const double my = phase ? S[s * 8 + 0] / n : 0.0;
  double v[2] = {0, 0};
  for (int row = row0 + threadIdx.x; row < row1; row += blockDim.x) {
    if (row - pstart >= n) continue;
    if (phase == 0) {
      v[0] += y[row];
    } else {
      const double a = y[row] - my;
      v[0] += (double)y[row] * ly[row], v[1] += a * a;
    }
  }
  __shared__ double red[(RSB_BLOCK / RSB_WAVE) * 2];
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    if (phase == 0) atomicAdd(&S[s * 8 + 0], v[0]);
    else atomicAdd(&S[s * 8 + 2], v[0]), atomicAdd(&S[s * 8 + 3], v[1]);
  }
}

// dL/dy = 2 (Ly - R yc) / (Q lambda), times `scale` (1 / segments); the loss
// of segment s goes to S[8s+5].  lam: the segment's Fiedler value (lz sc).
__global__ void k_lws_dy(const int *__restrict__ tile, const int *__restrict__ seg, int nseg,
                         const float *__restrict__ y, const float *__restrict__ ly,
                         const double *__restrict__ lam, int lam_stride,
                         double *__restrict__ S, float scale, float *__restrict__ dy) {
  const int ti = blockIdx.x, s = tile[3 * ti];
  const int row0 = tile[3 * ti + 1], row1 = tile[3 * ti + 2];
  const int n = seg[nseg + s], pstart = seg[4 * nseg + s];
  const double my = S[s * 8 + 0] / n, Ry = S[s * 8 + 2], Q = S[s * 8 + 3];
  const double lm = lam[(size_t)s * lam_stride];
  // lambda ~ 0: disconnected segment, no usable scale
  const bool ok = Q > 0 && lm > 1e-12;
  const double R = ok ? Ry / Q : 0.0;
  for (int row = row0 + threadIdx.x; row < row1; row += blockDim.x) {
    float g = 0.f;
    if (ok && row - pstart < n) g = (float)(scale * 2.0 * (ly[row] - R * (y[row] - my)) / (Q * lm));
    dy[row] = g;
  }
  if (threadIdx.x == 0 && row0 == pstart) S[s * 8 + 5] = ok ? R / lm - 1.0 : 0.0;
}

// dh4 = dy w_o (BF16); dw_o += sum dy h4, db_o += sum dy
__global__ void k_lws_out_bwd(int64_t M, const float *__restrict__ dy,
                              const unsigned short *__restrict__ h, const float *__restrict__ P,
                              unsigned short *__restrict__ dh, float *__restrict__ G) {
  const int c = threadIdx.x & 63;
  float gw = 0.f, gb = 0.f;
  for (int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < M; r += (int64_t)gridDim.x * 4) {
    const float d = dy[r];
    dh[r * LWS_H + c] = f2bf(d * P[LWS_WO + c]);
    gw += d * bf2f(h[r * LWS_H + c]);
    if (c == 0) gb += d;
  }
  atomicAdd(&G[LWS_WO + c], gw);
  if (c == 0) atomicAdd(&G[LWS_BO], gb);
}

// dz = dh * (h > 0) in place; db += sum dz
__global__ void k_lws_relu_bwd(int64_t M, const unsigned short *__restrict__ hout,
                               unsigned short *__restrict__ dz, float *__restrict__ db) {
  const int c = threadIdx.x & 63;
  float g = 0.f;
  for (int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < M; r += (int64_t)gridDim.x * 4) {
    const size_t i = (size_t)r * LWS_H + c;
    const float v = hout[i] ? bf2f(dz[i]) : 0.f;  // relu output 0 -> no gradient
    dz[i] = f2bf(v);
    g += v;
  }
  atomicAdd(&db[c], g);
}

// dW0 += dz0^T x (64 x 8)
__global__ void k_lws_in_bwd(int64_t M, const unsigned short *__restrict__ dz,
                             const float *__restrict__ x, float *__restrict__ G) {
  const int c = threadIdx.x & 63, kg = threadIdx.x >> 6;  // 4 groups of 2 inputs
  float g0 = 0.f, g1 = 0.f;
  for (int64_t r = blockIdx.x; r < M; r += gridDim.x) {
    const float d = bf2f(dz[r * LWS_H + c]);
    g0 += d * x[r * LWS_IN + 2 * kg];
    g1 += d * x[r * LWS_IN + 2 * kg + 1];
  }
  atomicAdd(&G[LWS_P0 + c * LWS_IN + 2 * kg], g0);
  atomicAdd(&G[LWS_P0 + c * LWS_IN + 2 * kg + 1], g1);
}

__global__ void k_lws_adam(float *__restrict__ P, const float *__restrict__ G,
                           float *__restrict__ m, float *__restrict__ v, int n, float lr,
                           float b1, float b2, float eps, float c1, float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float mi = b1 * m[i] + (1.f - b1) * g, vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi, v[i] = vi;
  P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
struct lws_ws {
  float *x, *y, *t, *dy, *dinv;
  unsigned short *h[LWS_L + 1], *a[LWS_L], *wt, *w, *g0, *g1, *dha;
  double *S;
};

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

static size_t lws_ws_bytes(int64_t M, int nseg, bool train) {
  const size_t hb = align256(2 * (size_t)M * LWS_H);
  size_t b = align256(4 * (size_t)M * LWS_IN) + 4 * align256(4 * (size_t)M) +
             2 * align256(2 * (size_t)LWS_L * 2 * LWS_H * LWS_H) + align256(8 * (size_t)nseg * 8);
  b += train ? hb * (2 * LWS_L + 1) + 2 * hb + 2 * hb : hb * 3;
  return b + 4096;
}

static lws_ws lws_carve(void *base, int64_t M, int nseg, bool train) {
  lws_ws w;
  char *p = (char *)base;
  auto take = [&](size_t bytes) {
    char *q = p;
    p += align256(bytes);
    return (void *)q;
  };
  const size_t hb = 2 * (size_t)M * LWS_H;
  w.x = (float *)take(4 * (size_t)M * LWS_IN);
  w.y = (float *)take(4 * (size_t)M);
  w.t = (float *)take(4 * (size_t)M);
  w.dy = (float *)take(4 * (size_t)M);
  w.dinv = (float *)take(4 * (size_t)M);
  w.wt = (unsigned short *)take(2 * (size_t)LWS_L * 2 * LWS_H * LWS_H);
  w.w = (unsigned short *)take(2 * (size_t)LWS_L * 2 * LWS_H * LWS_H);
  w.S = (double *)take(8 * (size_t)nseg * 8);
  if (train) {
    for (int l = 0; l <= LWS_L; l++) w.h[l] = (unsigned short *)take(hb);
    for (int l = 0; l < LWS_L; l++) w.a[l] = (unsigned short *)take(hb);
    w.g0 = (unsigned short *)take(hb);
    w.dha = (unsigned short *)take(2 * hb);
    w.g1 = w.g0;
  } else {
    unsigned short *h0 = (unsigned short *)take(hb), *h1 = (unsigned short *)take(hb);
    unsigned short *a = (unsigned short *)take(hb);
    for (int l = 0; l <= LWS_L; l++) w.h[l] = (l & 1) ? h1 : h0;
    for (int l = 0; l < LWS_L; l++) w.a[l] = a;
    w.g0 = w.g1 = w.dha = nullptr;
  }
  return w;
}

static int lws_forward(rsbgpu *g, const float *P, lws_ws &w, float level) {
  const int64_t M = g->mpad;
  hipStream_t st = g->stream;
  unsigned long long *box = (unsigned long long *)w.S;  // 6 per segment, reused
  k_lws_box_init<<<nblocks(g->nseg, 256), 256, 0, st>>>(g->nseg, box);
  k_lws_bbox<<<nblocks(M, 256), 256, 0, st>>>(g->d_row_elem, g->d_segof, g->d_coord, g->nel, M,
                                              g->ndim, box);
  k_lws_features<<<nblocks(M, 256), 256, 0, st>>>(g->d_row_elem, g->d_segof, g->d_coord, g->nel,
                                                  M, g->ndim, box, g->d_rowlen, level, w.x);
  k_lws_dinv<<<nblocks(M, 256), 256, 0, st>>>(M, g->d_slice_off, g->d_slice_w, g->d_ell, w.dinv);
  k_lws_wt<<<nblocks(LWS_L * 2 * LWS_H * LWS_H, 256), 256, 0, st>>>(P, w.wt);
  k_lws_w<<<nblocks(LWS_L * 2 * LWS_H * LWS_H, 256), 256, 0, st>>>(P, w.w);
  k_lws_in<<<nblocks(M, 256), 256, 0, st>>>(M, w.x, P, w.h[0]);
  RSB_LAUNCH_CHECK();
  for (int l = 0; l < LWS_L; l++) {
    k_lws_agg<<<nblocks(M * 8, 256), 256, 0, st>>>(M, g->d_slice_off, g->d_slice_w, g->d_ell,
                                                   w.h[l], LWS_H, 1, 0, w.dinv, nullptr, 0,
                                                   w.a[l]);
    k_gemm_rows<LWS_H, 2 * LWS_H, true, true><<<std::min<unsigned>(nblocks(M, 64), LWS_GEMM_BLOCKS), 256, 0, st>>>(
        (int)M, w.h[l], w.a[l], LWS_H, w.wt + (size_t)l * 2 * LWS_H * LWS_H, P + LWS_BL(l),
        w.h[l + 1], LWS_H);
    RSB_LAUNCH_CHECK();
  }
  k_lws_out<<<nblocks(M, 256), 256, 0, st>>>(M, w.h[LWS_L], P, w.y);
  RSB_LAUNCH_CHECK();
  return 0;
}

static int lws_ws_get(rsbgpu *g, bool train, lws_ws &w) {
  size_t need = lws_ws_bytes(g->mpad, g->nseg, train);
  if (need > g->lws_ws_bytes) {
    if (g->d_lws_ws) RSB_TRY(rsb_dev_free(g->d_lws_ws));
    g->d_lws_ws = nullptr;
    RSB_TRY(rsb_dev_malloc(&g->d_lws_ws, need));
    g->lws_ws_bytes = need;
  }
  w = lws_carve(g->d_lws_ws, g->mpad, g->nseg, train);
  return 0;
}

// Start vector of every row of the level from the network (used by
// rsbgpu_fiedler with init_mode 2).
int rsb_lws_predict(rsbgpu *g, const float *P, float level, float *d_rows) {
  lws_ws w;
  int rc = lws_ws_get(g, false, w);
  if (rc) return rc;
  if ((rc = lws_forward(g, P, w, level))) return rc;
  RSB_TRY(hipMemcpyAsync(d_rows, w.y, sizeof(float) * g->mpad, hipMemcpyDeviceToDevice,
                         g->stream));
  return 0;
}

extern "C" int rsbgpu_lws_nparams(void) { return LWS_NPARAM; }

extern "C" int rsbgpu_lws_set_params(rsbgpu *g, const float *d_params, float level,
                                     int passes, int max_n) {
  if (passes < 1) RSB_FAIL(3, "learned start needs at least one pass");
  if (max_n < 0) RSB_FAIL(3, "negative segment-size limit");
  g->lws_params = d_params;
  g->lws_level = level;
  g->lws_passes = passes;
  g->lws_max_n = max_n;
  return 0;
}

extern "C" int rsbgpu_lws_predict(rsbgpu *g, const float *d_params, float level, float *d_rows) {
  RSB_TRY(hipSetDevice(g->device));
  int rc = rsb_lws_predict(g, d_params, level, d_rows);
  if (!rc) RSB_TRY(hipStreamSynchronize(g->stream));
  return rc;
}

extern "C" int rsbgpu_lws_train_step(rsbgpu *g, const float *d_params, float *d_grad,
                                     float level, double *loss_out) {
  RSB_TRY(hipSetDevice(g->device));
  const int64_t M = g->mpad;
  const int nt = g->ntiles, nseg = g->nseg;
  *loss_out = 0.0;
  hipStream_t st = g->stream;
  RSB_TRY(hipMemsetAsync(d_grad, 0, sizeof(float) * LWS_NPARAM, st));
  if (nt == 0) return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
  lws_ws w;
  int rc = lws_ws_get(g, true, w);
  if (rc) return rc;
  if ((rc = lws_forward(g, d_params, w, level))) return rc;
  // loss and dL/dy (w.t holds L y)
  if (!g->lz.sc || g->lz_nseg_cap < nseg) RSB_FAIL(3, "training needs the level's Fiedler values");
  RSB_TRY(hipMemsetAsync(w.S, 0, sizeof(double) * nseg * 8, st));
  k_lws_lap<<<nblocks(M, 256), 256, 0, st>>>(M, g->d_slice_off, g->d_slice_w, g->d_ell, w.y, w.t);
  k_lws_sums<<<nt, RSB_BLOCK, 0, st>>>(g->d_tile, g->d_seg, nseg, w.y, w.t, 0, w.S);
  k_lws_sums<<<nt, RSB_BLOCK, 0, st>>>(g->d_tile, g->d_seg, nseg, w.y, w.t, 1, w.S);
  int nact = 0;
  for (int s = 0; s < nseg; s++) nact += (g->h_p[s] > 1 && g->h_n[s] >= 2);
  const float scale = nact ? 1.f / nact : 0.f;
  k_lws_dy<<<nt, 256, 0, st>>>(g->d_tile, g->d_seg, nseg, w.y, w.t, g->lz.sc + SC_LAMBDA, LZ_NSC,
                               w.S, scale, w.dy);
  RSB_LAUNCH_CHECK();
  // backward
  k_lws_out_bwd<<<256, 256, 0, st>>>(M, w.dy, w.h[LWS_L], d_params, w.g0, d_grad);
  for (int l = LWS_L - 1; l >= 0; l--) {
    // g0 holds dL/dh_{l+1}: dz = g0 * relu'(h_{l+1})
    k_lws_relu_bwd<<<256, 256, 0, st>>>(M, w.h[l + 1], w.g0, d_grad + LWS_BL(l));
    k_gemm_wgrad<LWS_H, 2 * LWS_H><<<nblocks(M, 1024), 256, 0, st>>>(
        (int)M, 1024, w.g0, LWS_H, w.h[l], w.a[l], LWS_H, d_grad + LWS_PL(l));
    // d[h|a] = dz W_l  (M x 128)
    k_gemm_rows<2 * LWS_H, LWS_H, false, true><<<std::min<unsigned>(nblocks(M, 64), LWS_GEMM_BLOCKS), 256, 0, st>>>(
        (int)M, w.g0, nullptr, LWS_H, w.w + (size_t)l * 2 * LWS_H * LWS_H, nullptr, w.dha,
        2 * LWS_H);
    // dh_l = dha[:, :64] + sum_i w_ij dha[i, 64:] / d_i
    k_lws_agg<<<nblocks(M * 8, 256), 256, 0, st>>>(M, g->d_slice_off, g->d_slice_w, g->d_ell,
                                                   w.dha + LWS_H, 2 * LWS_H, 0, 1, w.dinv, w.dha,
                                                   2 * LWS_H, w.g0);
    RSB_LAUNCH_CHECK();
  }
  k_lws_relu_bwd<<<256, 256, 0, st>>>(M, w.h[0], w.g0, d_grad + LWS_B0);
  k_lws_in_bwd<<<1024, 256, 0, st>>>(M, w.g0, w.x, d_grad);
  RSB_LAUNCH_CHECK();
  std::vector<double> S((size_t)nseg * 8);
  RSB_TRY(hipMemcpyAsync(S.data(), w.S, sizeof(double) * nseg * 8, hipMemcpyDeviceToHost, st));
  RSB_TRY(hipStreamSynchronize(st));
  double L = 0.0;
  for (int s = 0; s < nseg; s++)
    if (g->h_p[s] > 1 && g->h_n[s] >= 2) L += S[(size_t)s * 8 + 5];
  *loss_out = nact ? L / nact : 0.0;
  return 0;
}

extern "C" int rsbgpu_lws_adam(float *d_params, const float *d_grad, float *d_m, float *d_v,
                               int step, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  k_lws_adam<<<nblocks(LWS_NPARAM, 256), 256>>>(d_params, d_grad, d_m, d_v, LWS_NPARAM, lr, b1,
                                                b2, eps, c1, c2);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
}

// Parity hook: forward and backward of the network on one segment given on
// the host: coord [3 x M] (SoA), level, its graph as CSR (rowptr[M+1], col,
// w), its Fiedler value lam, params; writes y[M], grad and the loss.
extern "C" int rsbgpu_lws_test(int M, const double *coord_in, float level, const int *rowptr,
                               const int *col, const int *wgt, double lam,
                               const float *params, float *y, float *grad, double *loss) {
  rsbgpu *g = nullptr;
  if (rsbgpu_open(&g, 0)) return 2;
  int rc = 0;
  // Build a one-segment level: rows = 0..M-1 padded to 64, SELL from the CSR.
  {
    const int64_t mp = ((int64_t)M + 63) / 64 * 64;
    const int64_t ns = mp / 64;
    std::vector<int> sw(ns + 1, 0), rl(mp, 0), re(mp, -1), seg(5), tiles, segt(2);
    std::vector<int64_t> so(ns + 1, 0);
    for (int r = 0; r < M; r++) rl[r] = rowptr[r + 1] - rowptr[r], re[r] = r;
    for (int64_t s = 0; s < ns; s++)
      for (int k = 0; k < 64; k++) sw[s] = std::max(sw[s], rl[s * 64 + k]);
    for (int64_t s = 0; s < ns; s++) so[s + 1] = so[s] + (int64_t)sw[s] * 64;
    std::vector<uint32_t> ell(so[ns] + 1);
    for (int64_t r = 0; r < mp; r++) {
      int64_t s = r / 64, ln = r % 64;
      for (int j = 0; j < sw[s]; j++) {
        uint32_t v = (uint32_t)r;
        if (r < M && j < rl[r]) v = (uint32_t)col[rowptr[r] + j] | ((uint32_t)wgt[rowptr[r] + j] << 28);
        ell[so[s] + (int64_t)j * 64 + ln] = v;
      }
    }
    seg = {0, M, 2, 0, 0};
    for (int64_t r = 0; r < mp; r += RSB_TILE) tiles.insert(tiles.end(), {0, (int)r, (int)std::min<int64_t>(mp, r + RSB_TILE)});
    segt = {0, (int)(tiles.size() / 3)};
    g->nel = M, g->ndim = 3, g->nseg = 1, g->mpad = mp, g->nslices = ns;
    g->ntiles = (int)(tiles.size() / 3);
    g->h_p = {2}, g->h_n = {M}, g->h_cstart = {0}, g->h_r0 = {0}, g->h_pstart = {0};
    std::vector<double> coord(coord_in, coord_in + 3 * (size_t)M), sc(LZ_NSC, 0.0);
    sc[SC_LAMBDA] = lam;
    std::vector<int> segof(M, 0);
#define UP(dst, src, n, T)                                                               \
  if (!rc && dev_alloc(g, dst, (size_t)(n)) != hipSuccess) rc = 2;                         \
  if (!rc && hipMemcpy(dst, (src), sizeof(T) * (size_t)(n), hipMemcpyHostToDevice) != hipSuccess) rc = 2;
    UP(g->d_slice_w, sw.data(), ns + 1, int);
    UP(g->d_slice_off, so.data(), ns + 1, int64_t);
    UP(g->d_ell, ell.data(), so[ns] + 1, uint32_t);
    UP(g->d_rowlen, rl.data(), mp, int);
    UP(g->d_row_elem, re.data(), mp, int);
    UP(g->d_seg, seg.data(), 5, int);
    UP(g->d_tile, tiles.data(), tiles.size(), int);
    UP(g->d_seg_tiles, segt.data(), 2, int);
    UP(g->d_coord, coord.data(), 3 * (size_t)M, double);
    UP(g->lz.sc, sc.data(), LZ_NSC, double);
    g->lz_nseg_cap = 1;
    UP(g->d_segof, segof.data(), M, int);
#undef UP
  }
  float *dP = nullptr, *dG = nullptr, *dY = nullptr;
  double L = 0;
  if (!rc && rsb_dev_malloc((void **)&dP, sizeof(float) * LWS_NPARAM) != hipSuccess) rc = 2;
  if (!rc && rsb_dev_malloc((void **)&dG, sizeof(float) * LWS_NPARAM) != hipSuccess) rc = 2;
  if (!rc && rsb_dev_malloc((void **)&dY, sizeof(float) * g->mpad) != hipSuccess) rc = 2;
  if (!rc && hipMemcpy(dP, params, sizeof(float) * LWS_NPARAM, hipMemcpyHostToDevice) != hipSuccess) rc = 2;
  if (!rc && hipMemset(dG, 0, sizeof(float) * LWS_NPARAM) != hipSuccess) rc = 2;
  if (!rc) rc = rsbgpu_lws_train_step(g, dP, dG, level, &L);
  if (!rc) rc = rsbgpu_lws_predict(g, dP, level, dY);
  if (!rc && hipMemcpy(y, dY, sizeof(float) * M, hipMemcpyDeviceToHost) != hipSuccess) rc = 2;
  if (!rc && hipMemcpy(grad, dG, sizeof(float) * LWS_NPARAM, hipMemcpyDeviceToHost) != hipSuccess) rc = 2;
  *loss = L;
  if (dP) rsb_dev_free(dP);
  if (dG) rsb_dev_free(dG);
  if (dY) rsb_dev_free(dY);
  rsbgpu_close(g);
  return rc;
}
