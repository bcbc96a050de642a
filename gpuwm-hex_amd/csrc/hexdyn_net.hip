/*
 * Learned large-step corrector: a message-passing graph network on the
 * Voronoi mesh that maps the state after one coarse large step (doubled dt,
 * fewer acoustic substeps) to the correction that brings it back onto the
 * fine-dt trajectory.  No reference counterpart: the reference dycore has no
 * learned component; the corrector sits after the reference's step boundary
 * (cuda_driver.py:5751 step_device) and the plain step is its fallback.
 *
 * Network (hidden width H = 128, NL message-passing layers):
 *   features  x_c = normalised column of 6 channels x 64 rows: the cell wind
 *             (U_east, U_north) by a least-squares fit of the normal
 *             velocities on edgesOnCell, w, theta_m, rho_zz, qv, plus the
 *             Coriolis parameter in a padding row                 [C][384]
 *   encoder   h0 = relu(x W_in + b_in)                           [C][128]
 *   layer l   a_c = mean_{j in cellsOnCell(c)} h_j
 *             h_{l+1} = h_l + relu([h_l, a] W_l + b_l)            [C][128]
 *   decoder   y = h_NL W_out + b_out                              [C][320]
 *             y = (dU_east, dU_north, dtheta_m, drho_zz, dw) x 64 rows
 *   edges     du_e = cos(a_e) (dU_east(c0)+dU_east(c1))/2
 *                  + sin(a_e) (dU_north(c0)+dU_north(c1))/2  (cellsOnEdge)
 * Every GEMM is a bf16 MFMA (v_mfma_f32_16x16x32_bf16) with fp32
 * accumulation; weights and activations are bf16, the Adam master weights,
 * moments and gradients fp32.  Weight gradients are reduced over cells in
 * per-block partials summed in a fixed order (repeatable bits).
 *
 * Loss: mean squared error of the normalised corrections of u (edges),
 * theta_m and rho_zz (cells) and w (interior interfaces), one mean per
 * variable, summed.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hexdyn.h"
#include "hexdyn_internal.h"

namespace hexdyn {
namespace net {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int HID = 128;
constexpr int KIN = 6 * HX_ROW;    // 384 input features per cell
constexpr int NOUT = 5 * HX_ROW;   // 320 outputs per cell
enum { CH_UE = 0, CH_UN = 1, CH_W = 2, CH_TH = 3, CH_RHO = 4, CH_QV = 5 };
enum { O_UE = 0, O_UN = 1, O_TH = 2, O_RHO = 3, O_W = 4 };
constexpr int GPART = 128;         // blocks of the weight-gradient reductions

__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------ GEMM
// out[M][N] = epi(A[M][K] . B), A = [A1 (K1 columns) | A2 (K - K1 columns)]
// row-major bf16, B given transposed (Bt[N][K], bf16).  A block of four
// wavefronts stages Bt in LDS once and walks 16-row tiles (one tile per
// wavefront at a time, all N columns: N/16 MFMA accumulators).
// MFMA operand layout (16x16x32): lane l supplies row (or column) l % 16
// and the 8 consecutive k values 8 * (l / 16) .. +7; the accumulator lane l
// holds column l % 16, rows 4 * (l / 16) + i.
enum { EPI_RELU = 0, EPI_RESID_RELU = 1, EPI_F32 = 2 };
constexpr int GEMM_WAVES = 8;  // wavefronts per block (2 per SIMD beside the block's LDS copy of B)
constexpr int GEMM_TR = 2;     // 16-row tiles per wavefront at a time (each B fragment feeds TR MFMAs)

template <int K, int N, int K1, int EPI>
__global__ void __launch_bounds__(64 * GEMM_WAVES) k_gemm(const bf16 *__restrict__ A1, const bf16 *__restrict__ A2,
                                                          const bf16 *__restrict__ Bt, const float *__restrict__ bias,
                                                          const bf16 *__restrict__ resid, bf16 *__restrict__ out_b,
                                                          float *__restrict__ out_f, bf16 *__restrict__ relu_out,
                                                          int M) {
  static_assert(K % 32 == 0 && N % 16 == 0 && K1 % 32 == 0 && K1 <= K, "tile shape");
  constexpr int LDK = K + 8;  // row stride in LDS (16-byte pad against bank conflicts)
  constexpr int KA1 = K1;
  constexpr int KA2 = K - K1;
  constexpr int NT = N / 16;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  bf16 *sB = (bf16 *)lds_raw;
  for (int i = (int)threadIdx.x; i < N * (K / 8); i += 64 * GEMM_WAVES) {
    const int n = i / (K / 8), kc = i % (K / 8);
    *(bf16x8 *)(sB + n * LDK + kc * 8) = *(const bf16x8 *)(Bt + (size_t)n * K + kc * 8);
  }
  __syncthreads();
  const int w = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
  const int lr = l & 15, lg = l >> 4;
  const int ngroups = (M + 16 * GEMM_TR - 1) / (16 * GEMM_TR);
  for (int gt = (int)blockIdx.x * GEMM_WAVES + w; gt < ngroups; gt += (int)gridDim.x * GEMM_WAVES) {
    int row[GEMM_TR];
#pragma unroll
    for (int r = 0; r < GEMM_TR; ++r) row[r] = min((gt * GEMM_TR + r) * 16 + lr, M - 1);
    f32x4 acc[GEMM_TR][NT];
#pragma unroll
    for (int r = 0; r < GEMM_TR; ++r)
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[r][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k0 = 0; k0 < K; k0 += 32) {
      const int kk = k0 + lg * 8;
      bf16x8 a[GEMM_TR];
#pragma unroll
      for (int r = 0; r < GEMM_TR; ++r) {
        if (KA2 == 0 || k0 < KA1)
          a[r] = *(const bf16x8 *)(A1 + (size_t)row[r] * KA1 + kk);
        else
          a[r] = *(const bf16x8 *)(A2 + (size_t)row[r] * (KA2 > 0 ? KA2 : 1) + (kk - KA1));
      }
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const bf16x8 b = *(const bf16x8 *)(sB + (j * 16 + lr) * LDK + kk);
#pragma unroll
        for (int r = 0; r < GEMM_TR; ++r) acc[r][j] = mfma(a[r], b, acc[r][j]);
      }
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int col = j * 16 + lr;
      const float bj = bias ? bias[col] : 0.0f;
#pragma unroll
      for (int r = 0; r < GEMM_TR; ++r) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rr = (gt * GEMM_TR + r) * 16 + lg * 4 + i;
          if (rr >= M) continue;
          const float v = acc[r][j][i] + bj;
          const size_t o = (size_t)rr * N + col;
          if (EPI == EPI_RELU) {
            out_b[o] = (bf16)fmaxf(v, 0.0f);
          } else if (EPI == EPI_RESID_RELU) {
            const float rl = fmaxf(v, 0.0f);
            if (relu_out) relu_out[o] = (bf16)rl;
            out_b[o] = (bf16)((float)resid[o] + rl);
          } else {
            out_f[o] = v;
          }
        }
      }
    }
  }
}

// partial[blockIdx][KA][NB] = sum over this block's row chunks of A^T B,
// A = [A1 (K1 cols) | A2] [M][KA], B [M][NB] (bf16).  Chunks of 64 rows are
// staged transposed in LDS (At[k][m], Bt[n][m]) so each MFMA operand is one
// 16-byte LDS read; eight wavefronts split the (KA/16) x (NB/16) tiles.
template <int KA, int K1, int NB>
__global__ void __launch_bounds__(512) k_gemm_tn(const bf16 *__restrict__ A1, const bf16 *__restrict__ A2,
                                                 const bf16 *__restrict__ Bm, float *__restrict__ partial, int M) {
  constexpr int MC = 64, LDM = MC + 8;
  constexpr int TILES = (KA / 16) * (NB / 16);
  static_assert(TILES % 8 == 0, "tiles split over 8 wavefronts");
  constexpr int TPW = TILES / 8;
  constexpr int KA2 = KA - K1;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  bf16 *sA = (bf16 *)lds_raw;      // [KA][LDM]
  bf16 *sBt = sA + KA * LDM;       // [NB][LDM]
  const int w = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
  const int lr = l & 15, lg = l >> 4;
  f32x4 acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nchunks = (M + MC - 1) / MC;
  for (int ch = (int)blockIdx.x; ch < nchunks; ch += (int)gridDim.x) {
    __syncthreads();
    for (int i = (int)threadIdx.x; i < MC * (KA / 8); i += 512) {
      const int m = i / (KA / 8), kc = (i % (KA / 8)) * 8;
      const int gm = ch * MC + m;
      bf16x8 v;
      if (gm < M) {
        if (KA2 == 0 || kc < K1)
          v = *(const bf16x8 *)(A1 + (size_t)gm * K1 + kc);
        else
          v = *(const bf16x8 *)(A2 + (size_t)gm * (KA2 > 0 ? KA2 : 1) + (kc - K1));
      } else {
        for (int j = 0; j < 8; ++j) v[j] = (bf16)0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) sA[(kc + j) * LDM + m] = v[j];
    }
    for (int i = (int)threadIdx.x; i < MC * (NB / 8); i += 512) {
      const int m = i / (NB / 8), nc = (i % (NB / 8)) * 8;
      const int gm = ch * MC + m;
      bf16x8 v;
      if (gm < M) {
        v = *(const bf16x8 *)(Bm + (size_t)gm * NB + nc);
      } else {
        for (int j = 0; j < 8; ++j) v[j] = (bf16)0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) sBt[(nc + j) * LDM + m] = v[j];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int tile = w * TPW + q;
      const int ti = tile / (NB / 16), tj = tile % (NB / 16);
#pragma unroll
      for (int m0 = 0; m0 < MC; m0 += 32) {
        const bf16x8 a = *(const bf16x8 *)(sA + (ti * 16 + lr) * LDM + m0 + lg * 8);
        const bf16x8 b = *(const bf16x8 *)(sBt + (tj * 16 + lr) * LDM + m0 + lg * 8);
        acc[q] = mfma(a, b, acc[q]);
      }
    }
  }
  float *out = partial + (size_t)blockIdx.x * KA * NB;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int tile = w * TPW + q;
    const int ti = tile / (NB / 16), tj = tile % (NB / 16);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(size_t)(ti * 16 + lg * 4 + i) * NB + tj * 16 + lr] = acc[q][i];
  }
}

// partial[blockIdx][NB] = column sums of B over this block's row chunks
template <int NB>
__global__ void __launch_bounds__(512) k_colsum(const bf16 *__restrict__ Bm, float *__restrict__ partial, int M) {
  constexpr int MC = 64;
  const int n = (int)threadIdx.x;
  if (n >= NB) return;
  float s = 0.0f;
  const int nchunks = (M + MC - 1) / MC;
  for (int ch = (int)blockIdx.x; ch < nchunks; ch += (int)gridDim.x)
    for (int m = ch * MC; m < min(M, ch * MC + MC); ++m) s += (float)Bm[(size_t)m * NB + n];
  partial[(size_t)blockIdx.x * NB + n] = s;
}

// out[i] = sum_b partial[b][i], b ascending
__global__ void __launch_bounds__(256) k_reduce(const float *__restrict__ partial, int nb, int n,
                                                float *__restrict__ out) {
  const int i = (int)(blockIdx.x * 256 + threadIdx.x);
  if (i >= n) return;
  float s = 0.0f;
  for (int b = 0; b < nb; ++b) s += partial[(size_t)b * n + i];
  out[i] = s;
}

// ---------------------------------------------------------- mesh kernels
// Lane k of wavefront (blockIdx*4 + w) owns row k of one cell or edge.
__device__ __forceinline__ int ent4() { return (int)(blockIdx.x * 4 + (threadIdx.x >> 6)); }
__device__ __forceinline__ int lane() { return (int)(threadIdx.x & 63); }
// element k of row e of a dycore array (rows rs floats apart; lanes k >= rs
// read the row's last element, as the dycore kernels do)
__device__ __forceinline__ size_t rowi(int e, int k, int rs) { return (size_t)e * rs + (k < rs ? k : rs - 1); }
__device__ __forceinline__ float rowv(const float *p, int e, int k, int rs) { return p[rowi(e, k, rs)]; }

// Least-squares cell wind from the normal velocities of the cell's edges:
// minimise sum_s (n_s . U - u_s)^2, n_s = (cos a_s, sin a_s).
template <int NE>
__device__ __forceinline__ void lsq_frame(const Instance &I, int c, float &i11, float &i12, float &i22, int *eo,
                                          float *cs, float *sn, int &n) {
  n = I.m.neoc[c];
  float m11 = 0.f, m12 = 0.f, m22 = 0.f;
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    eo[s] = I.m.eoc[c * NE + s];
    cs[s] = 0.f;
    sn[s] = 0.f;
    if (s < n) {
      cs[s] = I.m.cosang[eo[s]];
      sn[s] = I.m.sinang[eo[s]];
      m11 += cs[s] * cs[s];
      m12 += cs[s] * sn[s];
      m22 += sn[s] * sn[s];
    }
  }
  const float det = m11 * m22 - m12 * m12;
  i11 = m22 / det;
  i12 = -m12 / det;
  i22 = m11 / det;
}

// Feature columns of the published state (raw, before normalisation).
template <int NE>
__device__ __forceinline__ void raw_features(const Instance &I, int c, int k, float raw[6]) {
  const Work &B = I.b;
  const int n = I.m.neoc[c];
  int eo[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) eo[s] = I.m.eoc[c * NE + s];
  // every load first (slot tables are padded, so each slot is readable); the
  // least-squares frame and sums below are lsq_frame's, in the same order
  float u[NE], ca[NE], sa[NE], fe[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    u[s] = rowv(B.p_u, eo[s], k, I.n.RS);
    ca[s] = I.m.cosang[eo[s]];
    sa[s] = I.m.sinang[eo[s]];
    fe[s] = I.m.fedge[eo[s]];
  }
  const float wk = rowv(B.p_w, c, k, I.n.RS), thk = rowv(B.p_theta, c, k, I.n.RS), rk = rowv(B.rho, c, k, I.n.RS);
  const float qk = I.n.NT > 0 ? rowv(B.q, c, k, I.n.RS) : 0.0f;
  float m11 = 0.f, m12 = 0.f, m22 = 0.f;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) {
      m11 += ca[s] * ca[s];
      m12 += ca[s] * sa[s];
      m22 += sa[s] * sa[s];
    }
  const float det = m11 * m22 - m12 * m12;
  const float i11 = m22 / det, i12 = -m12 / det, i22 = m11 / det;
  float b1 = 0.f, b2 = 0.f, fs = 0.f;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) {
      b1 += ca[s] * u[s];
      b2 += sa[s] * u[s];
      fs += fe[s];
    }
  const int L = I.n.L;
  raw[CH_UE] = (k < L) ? i11 * b1 + i12 * b2 : (k == L ? fs / (float)n : 0.0f);
  raw[CH_UN] = (k < L) ? i12 * b1 + i22 * b2 : 0.0f;
  raw[CH_W] = (k <= L) ? wk : 0.0f;
  raw[CH_TH] = (k < L) ? thk : 0.0f;
  raw[CH_RHO] = (k < L) ? rk : 0.0f;
  raw[CH_QV] = (k < L && I.n.NT > 0) ? qk : 0.0f;
}

template <int NE>
__global__ void __launch_bounds__(256) k_features(const Instance I, const float *__restrict__ mean,
                                                  const float *__restrict__ scale, bf16 *__restrict__ X,
                                                  float *__restrict__ raw_out) {
  const int c = ent4();
  if (c >= I.n.C) return;
  const int k = lane();
  float raw[6];
  raw_features<NE>(I, c, k, raw);
#pragma unroll
  for (int ch = 0; ch < 6; ++ch) {
    const int j = ch * HX_ROW + k;
    X[(size_t)c * KIN + j] = (bf16)((raw[ch] - mean[j]) * scale[j]);
    if (raw_out) raw_out[(size_t)c * KIN + j] = raw[ch];
  }
}

// a_c = mean over cellsOnCell of h (128 channels: lane and lane + 64)
template <int NE>
__global__ void __launch_bounds__(256) k_aggregate(const Instance I, const bf16 *__restrict__ Hin,
                                                   bf16 *__restrict__ Aout) {
  const int c = ent4();
  if (c >= I.n.C) return;
  const int k = lane();
  const int n = I.m.neoc[c];
  // every neighbour row first (the slot table is padded), then the sums in slot order
  float h0[NE], h1[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    const int j = I.m.coc[c * NE + s];
    h0[s] = (float)Hin[(size_t)j * HID + k];
    h1[s] = (float)Hin[(size_t)j * HID + 64 + k];
  }
  float s0 = 0.f, s1 = 0.f;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) {
      s0 += h0[s];
      s1 += h1[s];
    }
  Aout[(size_t)c * HID + k] = (bf16)(s0 / (float)n);
  Aout[(size_t)c * HID + 64 + k] = (bf16)(s1 / (float)n);
}

// dH_l = dH_{l+1} + dZ[:, :128] + sum over cellsOnCell j of dZ_j[128:] / n_j
// (the mean's adjoint; cellsOnCell is symmetric on a Voronoi mesh)
template <int NE>
__global__ void __launch_bounds__(256) k_aggregate_back(const Instance I, const float *__restrict__ dHnext,
                                                        const float *__restrict__ dZ, float *__restrict__ dH) {
  const int c = ent4();
  if (c >= I.n.C) return;
  const int k = lane();
  const int n = I.m.neoc[c];
  float z0[NE], z1[NE], nj[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    const int j = I.m.coc[c * NE + s];
    nj[s] = (float)I.m.neoc[j];
    z0[s] = dZ[(size_t)j * 2 * HID + HID + k];
    z1[s] = dZ[(size_t)j * 2 * HID + HID + 64 + k];
  }
  float a0 = 0.f, a1 = 0.f;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) {
      const float inv = 1.0f / nj[s];
      a0 += z0[s] * inv;
      a1 += z1[s] * inv;
    }
  const size_t o = (size_t)c * HID;
  dH[o + k] = (dHnext[o + k] + dZ[(size_t)c * 2 * HID + k]) + a0;
  dH[o + 64 + k] = (dHnext[o + 64 + k] + dZ[(size_t)c * 2 * HID + 64 + k]) + a1;
}

// dP = dH * [mask > 0] (bf16 for the next MFMA)
__global__ void __launch_bounds__(256) k_relu_back(const float *__restrict__ dH, const bf16 *__restrict__ mask,
                                                   bf16 *__restrict__ dP, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  dP[i] = (bf16)((float)mask[i] > 0.0f ? dH[i] : 0.0f);
}

__global__ void __launch_bounds__(256) k_f32_to_bf16(const float *__restrict__ x, bf16 *__restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  y[i] = (bf16)x[i];
}

// Published fields of the coarse and fine handles that the target reads.
struct TargetArgs {
  int C, E, L, RS;
  const float *th_c, *rho_c, *w_c, *u_c;
  const float *th_f, *rho_f, *w_f, *u_f;
};

// normalised targets: (fine - coarse) / scale on the published states
__global__ void __launch_bounds__(256) k_target_cells(const TargetArgs a, const float *__restrict__ oscale,
                                                      float *__restrict__ Tc) {
  const int c = ent4();
  if (c >= a.C) return;
  const int k = lane();
  const int L = a.L;
  const float th = k < L ? (rowv(a.th_f, c, k, a.RS) - rowv(a.th_c, c, k, a.RS)) / oscale[O_TH * 64 + k] : 0.f;
  const float rh = k < L ? (rowv(a.rho_f, c, k, a.RS) - rowv(a.rho_c, c, k, a.RS)) / oscale[O_RHO * 64 + k] : 0.f;
  const float ww = (k >= 1 && k < L) ? (rowv(a.w_f, c, k, a.RS) - rowv(a.w_c, c, k, a.RS)) / oscale[O_W * 64 + k] : 0.f;
  float *t = Tc + (size_t)c * 3 * HX_ROW;
  t[k] = th;
  t[HX_ROW + k] = rh;
  t[2 * HX_ROW + k] = ww;
}

__global__ void __launch_bounds__(256) k_target_edges(const TargetArgs a, const float *__restrict__ oscale,
                                                      float *__restrict__ Te) {
  const int e = ent4();
  if (e >= a.E) return;
  const int k = lane();
  Te[(size_t)e * HX_ROW + k] = k < a.L ? (rowv(a.u_f, e, k, a.RS) - rowv(a.u_c, e, k, a.RS)) / oscale[O_UE * 64 + k] : 0.f;
}

__device__ __forceinline__ float wave_sum_f(float x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_down(x, off, 64);
  return x;
}

// normal-wind prediction of an edge from its two cells' outputs (normalised units)
__device__ __forceinline__ float edge_pred(const Instance &I, const float *Y, int e, int k) {
  const int c0 = I.m.coe[2 * e], c1 = I.m.coe[2 * e + 1];
  const float ue = 0.5f * (Y[(size_t)c0 * NOUT + O_UE * 64 + k] + Y[(size_t)c1 * NOUT + O_UE * 64 + k]);
  const float un = 0.5f * (Y[(size_t)c0 * NOUT + O_UN * 64 + k] + Y[(size_t)c1 * NOUT + O_UN * 64 + k]);
  return I.m.cosang[e] * ue + I.m.sinang[e] * un;
}

// g_e = dLoss/d(du_e) and the per-block partial sum of squared errors
__global__ void __launch_bounds__(256) k_loss_edges(const Instance I, const float *__restrict__ Y,
                                                    const float *__restrict__ Te, float *__restrict__ g,
                                                    float *__restrict__ lpart, float inv_n) {
  __shared__ float s_part[4];
  const int e = ent4();
  const int k = lane();
  float d2 = 0.f;
  if (e < I.n.E) {
    float gv = 0.f;
    if (k < I.n.L) {
      const float d = edge_pred(I, Y, e, k) - Te[(size_t)e * HX_ROW + k];
      gv = 2.0f * d * inv_n;
      d2 = d * d;
    }
    g[(size_t)e * HX_ROW + k] = gv;
  }
  d2 = wave_sum_f(d2);
  if (k == 0) s_part[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) lpart[blockIdx.x] = ((s_part[0] + s_part[1]) + s_part[2]) + s_part[3];
}

// dY of every output channel (bf16) and the per-block partial sums of squared
// errors of theta, rho and w; the wind channels gather the edge gradients
template <int NE>
__global__ void __launch_bounds__(256) k_loss_cells(const Instance I, const float *__restrict__ Y,
                                                    const float *__restrict__ Tc, const float *__restrict__ g,
                                                    bf16 *__restrict__ dY, float *__restrict__ lpart,
                                                    float inv_nc, float inv_nw) {
  __shared__ float s_part[3][4];
  const int c = ent4();
  const int k = lane();
  const int L = I.n.L;
  float d2[3] = {0.f, 0.f, 0.f};
  if (c < I.n.C) {
    const float *y = Y + (size_t)c * NOUT;
    const float *t = Tc + (size_t)c * 3 * HX_ROW;
    bf16 *dy = dY + (size_t)c * NOUT;
    const int n = I.m.neoc[c];
    float ge = 0.f, gn = 0.f;
#pragma unroll
    for (int s = 0; s < NE; ++s) {
      if (s < n) {
        const int e = I.m.eoc[c * NE + s];
        const float gv = 0.5f * g[(size_t)e * HX_ROW + k];
        ge += I.m.cosang[e] * gv;
        gn += I.m.sinang[e] * gv;
      }
    }
    dy[O_UE * 64 + k] = (bf16)(k < L ? ge : 0.f);
    dy[O_UN * 64 + k] = (bf16)(k < L ? gn : 0.f);
    const bool lay = k < L, intr = k >= 1 && k < L;
    const float dth = lay ? y[O_TH * 64 + k] - t[k] : 0.f;
    const float drh = lay ? y[O_RHO * 64 + k] - t[HX_ROW + k] : 0.f;
    const float dw = intr ? y[O_W * 64 + k] - t[2 * HX_ROW + k] : 0.f;
    dy[O_TH * 64 + k] = (bf16)(2.0f * dth * inv_nc);
    dy[O_RHO * 64 + k] = (bf16)(2.0f * drh * inv_nc);
    dy[O_W * 64 + k] = (bf16)(2.0f * dw * inv_nw);
    d2[0] = dth * dth;
    d2[1] = drh * drh;
    d2[2] = dw * dw;
  }
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const float x = wave_sum_f(d2[v]);
    if (k == 0) s_part[v][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int v = (int)threadIdx.x;
    lpart[(size_t)v * gridDim.x + blockIdx.x] = ((s_part[v][0] + s_part[v][1]) + s_part[v][2]) + s_part[v][3];
  }
}

// sum of nb floats in ascending order, in binary64 (one thread)
__global__ void k_sum_ordered(const float *__restrict__ p, int nb, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += (double)p[b];
  out[0] = s;
}

// ---------------------------------------------------------- application
// Per-block partials of sum_c,k area * dzw * drho (the raw density
// correction) and of the volume sum_c,k area * dzw.
__global__ void __launch_bounds__(256) k_mass_partial(const Instance I, const float *__restrict__ Y,
                                                      const float *__restrict__ oscale, double *__restrict__ part) {
  __shared__ double s_part[2][4];
  const int c = ent4();
  const int k = lane();
  double v = 0.0, vol = 0.0;
  if (c < I.n.C && k < I.n.L) {
    const double dr = (double)(Y[(size_t)c * NOUT + O_RHO * 64 + k] * oscale[O_RHO * 64 + k]);
    vol = (1.0 / (double)I.m.invarea[c]) * (1.0 / (double)I.z.rdzw[k]);
    v = vol * dr;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    v += __shfl_down(v, off, 64);
    vol += __shfl_down(vol, off, 64);
  }
  if (k == 0) {
    s_part[0][threadIdx.x >> 6] = v;
    s_part[1][threadIdx.x >> 6] = vol;
  }
  __syncthreads();
  if (threadIdx.x < 2)
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] =
        ((s_part[threadIdx.x][0] + s_part[threadIdx.x][1]) + s_part[threadIdx.x][2]) + s_part[threadIdx.x][3];
}

// shift = (sum of the correction partials) / (sum of the volume partials),
// one block, fixed order
__global__ void __launch_bounds__(256) k_mass_shift(const double *__restrict__ part, int nb,
                                                    float *__restrict__ shift) {
  __shared__ double s[2][256];
  const int t = (int)threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int i = t; i < nb; i += 256) {
    a += part[i];
    b += part[(size_t)nb + i];
  }
  s[0][t] = a;
  s[1][t] = b;
  __syncthreads();
  for (int off = 128; off >= 1; off >>= 1) {
    if (t < off) {
      s[0][t] = s[0][t] + s[0][t + off];
      s[1][t] = s[1][t] + s[1][t + off];
    }
    __syncthreads();
  }
  if (t == 0) shift[0] = (float)(s[0][0] / s[1][0]);
}

// cells: rho' = rho + drho - shift, theta' = theta + dtheta, rtheta' = rho' theta',
// rw' = rw + dw * (fzm rho'_k + fzp rho'_{k-1}) (fzm zz_k + fzp zz_{k-1}) inside the column
__global__ void __launch_bounds__(256) k_apply_cells(const Instance I, const float *__restrict__ Y,
                                                     const float *__restrict__ oscale, const float *__restrict__ shift,
                                                     float gain) {
  const int c = ent4();
  if (c >= I.n.C) return;
  const int k = lane();
  const int L = I.n.L;
  const Work &B = I.b;
  const float *y = Y + (size_t)c * NOUT;
  const size_t o = rowi(c, k, I.n.RS);
  float rho = B.rho[o];
  if (k < L) {
    const float drho = gain * (y[O_RHO * 64 + k] * oscale[O_RHO * 64 + k] - shift[0]);
    const float th = B.p_theta[o] + gain * (y[O_TH * 64 + k] * oscale[O_TH * 64 + k]);
    rho = rho + drho;
    B.rho[o] = rho;
    B.rt[o] = rho * th;
  }
  const float rho_m = __shfl_up(rho, 1, 64);
  const float zk = I.z.zz[o];
  const float zkm = __shfl_up(zk, 1, 64);
  if (k >= 1 && k < L) {
    const float fzm = I.z.fzm[k], fzp = I.z.fzp[k];
    const float dw = gain * (y[O_W * 64 + k] * oscale[O_W * 64 + k]);
    B.rw[o] = B.rw[o] + dw * ((fzm * rho + fzp * rho_m) * (fzm * zk + fzp * zkm));
  }
}

// edges (after the cells): u' = u + du, ru' = (rho'_0 + rho'_1)/2 u', and
// the carried u recomputed from ru' as the diagnostic rebuild forms it,
// 2 ru' / (rho'_0 + rho'_1)
__global__ void __launch_bounds__(256) k_apply_edges(const Instance I, const float *__restrict__ Y,
                                                     const float *__restrict__ oscale, float gain) {
  const int e = ent4();
  if (e >= I.n.E) return;
  const int k = lane();
  if (k >= I.n.L) return;
  const int c0 = I.m.coe[2 * e], c1 = I.m.coe[2 * e + 1];
  const size_t o = rowi(e, k, I.n.RS);
  const float u = I.b.p_u[o] + gain * (edge_pred(I, Y, e, k) * oscale[O_UE * 64 + k]);
  const float r0 = rowv(I.b.rho, c0, k, I.n.RS), r1 = rowv(I.b.rho, c1, k, I.n.RS);
  const float ru = (0.5f * (r0 + r1)) * u;
  I.b.ru[o] = ru;
  I.b.p_u[o] = (2.0f * ru) / (r0 + r1);
}

// ---------------------------------------------------------- checks
__device__ __forceinline__ float nanmax(float a, float b) { return (b != b || b > a) ? b : a; }

// Per-block partials of the checks: max |u| dt / dc (horizontal CFL),
// max |w| dt rdzw / zz (vertical CFL), max |divergence| dt (a NaN anywhere
// wins), and the dry mass sum areaCell * dzw * rho_zz in binary64.
template <int NE>
__global__ void __launch_bounds__(256) k_checks_cells(const Instance I, float dt, float *__restrict__ fpart,
                                                      double *__restrict__ mpart) {
  __shared__ float s_f[3][4];
  __shared__ double s_m[4];
  const int c = ent4();
  const int k = lane();
  const int L = I.n.L;
  const Work &B = I.b;
  float div = 0.f, cfl = 0.f, vc = 0.f;
  double mass = 0.0;
  if (c < I.n.C) {
    const int n = I.m.neoc[c];
    // every load first (slot tables are padded, so each slot is readable),
    // then the sums in slot order
    int e[NE];
#pragma unroll
    for (int s = 0; s < NE; ++s) e[s] = I.m.eoc[c * NE + s];
    float u[NE], sd[NE], idc[NE];
#pragma unroll
    for (int s = 0; s < NE; ++s) {
      u[s] = rowv(B.p_u, e[s], k, I.n.RS);
      sd[s] = I.m.csdv[c * NE + s];
      idc[s] = I.m.invdc[e[s]];
    }
    const float wk = rowv(B.p_w, c, k, I.n.RS), zk = rowv(I.z.zz, c, k, I.n.RS), rk = rowv(B.rho, c, k, I.n.RS);
    const float rdzw = I.z.rdzw[k];
#pragma unroll
    for (int s = 0; s < NE; ++s) {
      if (s < n) {
        div += sd[s] * u[s];
        cfl = nanmax(cfl, fabsf(u[s]) * dt * idc[s]);
      }
    }
    div = fabsf(div * I.m.invarea[c]) * dt;
    if (k >= 1 && k < L) vc = fabsf(wk) * dt * rdzw / zk;
    if (k < L) mass = (1.0 / (double)I.m.invarea[c]) * (1.0 / (double)rdzw) * (double)rk;
    if (k >= L) {
      div = 0.f;
      cfl = 0.f;
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    cfl = nanmax(cfl, __shfl_down(cfl, off, 64));
    vc = nanmax(vc, __shfl_down(vc, off, 64));
    div = nanmax(div, __shfl_down(div, off, 64));
    mass += __shfl_down(mass, off, 64);
  }
  const int w = (int)(threadIdx.x >> 6);
  if (k == 0) {
    s_f[0][w] = cfl;
    s_f[1][w] = vc;
    s_f[2][w] = div;
    s_m[w] = mass;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int v = (int)threadIdx.x;
    fpart[(size_t)v * gridDim.x + blockIdx.x] = nanmax(nanmax(nanmax(s_f[v][0], s_f[v][1]), s_f[v][2]), s_f[v][3]);
  } else if (threadIdx.x == 3) {
    mpart[blockIdx.x] = ((s_m[0] + s_m[1]) + s_m[2]) + s_m[3];
  }
}

// one block: out = {cfl_h, cfl_v, div (floats), dry mass (double)} from the
// partials in a fixed order
__global__ void __launch_bounds__(256) k_checks_final(const float *__restrict__ fpart, const double *__restrict__ mpart,
                                                      int nb, float *__restrict__ fout, double *__restrict__ mout) {
  __shared__ float s_f[3][256];
  __shared__ double s_m[256];
  const int t = (int)threadIdx.x;
  float f[3] = {0.f, 0.f, 0.f};
  double m = 0.0;
  for (int b = t; b < nb; b += 256) {
#pragma unroll
    for (int v = 0; v < 3; ++v) f[v] = nanmax(f[v], fpart[(size_t)v * nb + b]);
    m += mpart[b];
  }
#pragma unroll
  for (int v = 0; v < 3; ++v) s_f[v][t] = f[v];
  s_m[t] = m;
  __syncthreads();
  for (int off = 128; off >= 1; off >>= 1) {
    if (t < off) {
#pragma unroll
      for (int v = 0; v < 3; ++v) s_f[v][t] = nanmax(s_f[v][t], s_f[v][t + off]);
      s_m[t] = s_m[t] + s_m[t + off];
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int v = 0; v < 3; ++v) fout[v] = s_f[v][0];
    mout[0] = s_m[0];
  }
}

// ---------------------------------------------------------- optimiser
struct Seg {
  long long off;   // offset in the flat fp32 parameter vector
  int K, N;        // matrix shape (N = count, K = 1 for a bias)
  long long boff;  // offset in the bf16 weight buffers (matrices only, -1 for biases)
};

__global__ void __launch_bounds__(256) k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, size_t n, float lr,
                                              float b1, float b2, float eps, float bc1, float bc2) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / bc1, vh = vi / bc2;
  p[i] = p[i] - lr * (mh / (sqrtf(vh) + eps));
}

// bf16 copies of one weight matrix: W[k][n] (backward) and Wt[n][k] (forward)
__global__ void __launch_bounds__(256) k_to_bf16_pair(const float *__restrict__ p, int K, int N,
                                                      bf16 *__restrict__ W, bf16 *__restrict__ Wt) {
  const int i = (int)(blockIdx.x * 256 + threadIdx.x);
  if (i >= K * N) return;
  const int k = i / N, n = i % N;
  const bf16 b = (bf16)p[i];
  W[i] = b;
  Wt[(size_t)n * K + k] = b;
}

// logical (original) order export of per-cell rows of width w
__global__ void __launch_bounds__(256) k_export_rows(const float *__restrict__ src, const int *__restrict__ old_of_new,
                                                     int n, int w, float *__restrict__ dst) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)n * w) return;
  const int r = (int)(i / w), j = (int)(i % w);
  dst[(size_t)old_of_new[r] * w + j] = src[i];
}

__global__ void __launch_bounds__(256) k_export_rows_bf16(const bf16 *__restrict__ src,
                                                          const int *__restrict__ old_of_new, int n, int w,
                                                          float *__restrict__ dst) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)n * w) return;
  const int r = (int)(i / w), j = (int)(i % w);
  dst[(size_t)old_of_new[r] * w + j] = (float)src[i];
}

inline unsigned nblk(size_t n, int per = 256) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ====================================================================== Net
struct Net {
  int C = 0, E = 0, NL = 0, NE = 6;
  int dev_blocks = 256;  // persistent GEMM grid
  std::vector<void *> allocs;
  std::vector<Seg> segs;  // W_in, b_in, (W_l, b_l) x NL, W_out, b_out
  size_t n_params = 0, n_bf = 0;
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  bf16 *w_bf = nullptr, *wt_bf = nullptr;
  float *in_mean = nullptr, *in_scale = nullptr, *out_scale = nullptr;
  bf16 *X = nullptr;
  std::vector<bf16 *> H, A, R;  // H: NL+1, A and R: NL
  float *Y = nullptr;
  float *Tc = nullptr, *Te = nullptr, *g = nullptr;
  bf16 *dY = nullptr, *dP = nullptr;
  float *dH0 = nullptr, *dH1 = nullptr, *dZ = nullptr, *dH4 = nullptr;
  float *partial = nullptr, *lpart = nullptr;
  double *loss_out = nullptr, *mpart = nullptr;
  float *shift = nullptr, *checks = nullptr, *raw = nullptr, *cpart = nullptr;
  double *mass_out = nullptr;
  float check_host[6];  // three maxima, pad, then the binary64 mass
  int have_target = 0, have_forward = 0;
  std::string error;

  template <class T> int alloc(T **p, size_t count) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return -1;
    if (hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return -1;
    allocs.push_back(q);
    *p = (T *)q;
    return 0;
  }
  ~Net() {
    for (void *p : allocs) (void)hipFree(p);
  }
  const Seg &w_in() const { return segs[0]; }
  const Seg &b_in() const { return segs[1]; }
  const Seg &w_l(int l) const { return segs[2 + 2 * l]; }
  const Seg &b_l(int l) const { return segs[3 + 2 * l]; }
  const Seg &w_out() const { return segs[2 + 2 * NL]; }
  const Seg &b_out() const { return segs[3 + 2 * NL]; }
  const bf16 *Wt(const Seg &s) const { return wt_bf + s.boff; }
  const bf16 *W(const Seg &s) const { return w_bf + s.boff; }
  const float *P(const Seg &s) const { return params + s.off; }
  float *G(const Seg &s) const { return grads + s.off; }
};

namespace {

template <class KFn>
int set_lds(KFn fn, size_t bytes) {
  return hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
                 hipSuccess
             ? 0
             : -1;
}

template <int K, int N, int K1, int EPI>
int gemm(const Net &n, const bf16 *A1, const bf16 *A2, const bf16 *Bt, const float *bias, const bf16 *resid,
         bf16 *out_b, float *out_f, bf16 *relu_out, int M, hipStream_t s) {
  const size_t lds = (size_t)N * (K + 8) * sizeof(bf16);
  static int ready = 0;
  if (!ready) {
    if (set_lds(k_gemm<K, N, K1, EPI>, lds)) return -1;
    ready = 1;
  }
  const int groups = (M + 16 * GEMM_TR - 1) / (16 * GEMM_TR);
  const int grid = std::max(1, std::min(n.dev_blocks, (groups + GEMM_WAVES - 1) / GEMM_WAVES));
  prof_mark("k_net_gemm", s, 1);
  hipLaunchKernelGGL((k_gemm<K, N, K1, EPI>), dim3(grid), dim3(64 * GEMM_WAVES), lds, s, A1, A2, Bt, bias, resid,
                     out_b, out_f, relu_out, M);
  prof_mark("k_net_gemm", s, 0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int KA, int K1, int NB>
int gemm_tn(const Net &n, const bf16 *A1, const bf16 *A2, const bf16 *Bm, float *dW, int M, hipStream_t s) {
  const size_t lds = (size_t)(KA + NB) * (64 + 8) * sizeof(bf16);
  static int ready = 0;
  if (!ready) {
    if (set_lds(k_gemm_tn<KA, K1, NB>, lds)) return -1;
    ready = 1;
  }
  prof_mark("k_net_gemm_tn", s, 1);
  hipLaunchKernelGGL((k_gemm_tn<KA, K1, NB>), dim3(GPART), dim3(512), lds, s, A1, A2, Bm, n.partial, M);
  prof_mark("k_net_gemm_tn", s, 0);
  hipLaunchKernelGGL(k_reduce, dim3(nblk((size_t)KA * NB)), dim3(256), 0, s, (const float *)n.partial, GPART, KA * NB,
                     dW);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NB>
int colsum(const Net &n, const bf16 *Bm, float *db, int M, hipStream_t s) {
  hipLaunchKernelGGL(k_colsum<NB>, dim3(GPART), dim3(512), 0, s, Bm, n.partial, M);
  hipLaunchKernelGGL(k_reduce, dim3(nblk(NB)), dim3(256), 0, s, (const float *)n.partial, GPART, NB, db);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

#define NE_LAUNCH(NEV, KERN, grid, block, s, ...)                                   \
  do {                                                                              \
    if ((NEV) == 6)                                                                 \
      hipLaunchKernelGGL(KERN<6>, dim3(grid), dim3(block), 0, s, __VA_ARGS__);      \
    else                                                                            \
      hipLaunchKernelGGL(KERN<7>, dim3(grid), dim3(block), 0, s, __VA_ARGS__);      \
  } while (0)

}  // namespace

int features(Net &n, const Instance &I, hipStream_t s, float *raw = nullptr) {
  prof_mark("k_net_features", s, 1);
  NE_LAUNCH(n.NE, k_features, nblk(n.C, 4), 256, s, I, (const float *)n.in_mean, (const float *)n.in_scale, n.X,
            raw);
  prof_mark("k_net_features", s, 0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int forward(Net &n, const Instance &I, bool keep, hipStream_t s) {
  if (features(n, I, s)) return -1;
  if (gemm<KIN, HID, KIN, EPI_RELU>(n, n.X, nullptr, n.Wt(n.w_in()), n.P(n.b_in()), nullptr, n.H[0], nullptr,
                                    nullptr, n.C, s))
    return -1;
  for (int l = 0; l < n.NL; ++l) {
    prof_mark("k_net_aggregate", s, 1);
    NE_LAUNCH(n.NE, k_aggregate, nblk(n.C, 4), 256, s, I, (const bf16 *)n.H[l], n.A[l]);
    prof_mark("k_net_aggregate", s, 0);
    if (gemm<2 * HID, HID, HID, EPI_RESID_RELU>(n, n.H[l], n.A[l], n.Wt(n.w_l(l)), n.P(n.b_l(l)), n.H[l],
                                                n.H[l + 1], nullptr, keep ? n.R[l] : nullptr, n.C, s))
      return -1;
  }
  if (gemm<HID, NOUT, HID, EPI_F32>(n, n.H[n.NL], nullptr, n.Wt(n.w_out()), n.P(n.b_out()), nullptr, nullptr, n.Y,
                                    nullptr, n.C, s))
    return -1;
  n.have_forward = 1;
  return 0;
}

int loss(Net &n, const Instance &I, double *out4, hipStream_t s) {
  const int L = I.n.L;
  const unsigned ge = nblk(n.E, 4), gc = nblk(n.C, 4);
  hipLaunchKernelGGL(k_loss_edges, dim3(ge), dim3(256), 0, s, I, (const float *)n.Y, (const float *)n.Te, n.g,
                     n.lpart, 1.0f / (float)((double)n.E * L));
  hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(64), 0, s, (const float *)n.lpart, (int)ge, n.loss_out);
  float *lp = n.lpart + ge;
  NE_LAUNCH(n.NE, k_loss_cells, gc, 256, s, I, (const float *)n.Y, (const float *)n.Tc, (const float *)n.g, n.dY, lp,
            1.0f / (float)((double)n.C * L), 1.0f / (float)((double)n.C * (L - 1)));
  for (int v = 0; v < 3; ++v)
    hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(64), 0, s, (const float *)(lp + (size_t)v * gc), (int)gc,
                       n.loss_out + 1 + v);
  if (hipGetLastError() != hipSuccess) return -1;
  double raw[4];
  if (hipMemcpyAsync(raw, n.loss_out, sizeof raw, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -1;
  out4[1] = raw[0] / ((double)n.E * L);
  out4[2] = raw[1] / ((double)n.C * L);
  out4[3] = raw[2] / ((double)n.C * L);
  out4[4] = raw[3] / ((double)n.C * (L - 1));
  out4[0] = ((out4[1] + out4[2]) + out4[3]) + out4[4];
  return 0;
}

int backward(Net &n, const Instance &I, hipStream_t s) {
  const int C = n.C;
  const size_t ch = (size_t)C * HID;
  // decoder
  if (gemm_tn<HID, HID, NOUT>(n, n.H[n.NL], nullptr, n.dY, n.G(n.w_out()), C, s)) return -1;
  if (colsum<NOUT>(n, n.dY, n.G(n.b_out()), C, s)) return -1;
  if (gemm<NOUT, HID, NOUT, EPI_F32>(n, n.dY, nullptr, n.W(n.w_out()), nullptr, nullptr, nullptr, n.dH1, nullptr, C,
                                     s))
    return -1;
  float *dHnext = n.dH1, *dHcur = n.dH0;
  for (int l = n.NL - 1; l >= 0; --l) {
    hipLaunchKernelGGL(k_relu_back, dim3(nblk(ch)), dim3(256), 0, s, (const float *)dHnext, (const bf16 *)n.R[l], n.dP,
                       ch);
    if (gemm_tn<2 * HID, HID, HID>(n, n.H[l], n.A[l], n.dP, n.G(n.w_l(l)), C, s)) return -1;
    if (colsum<HID>(n, n.dP, n.G(n.b_l(l)), C, s)) return -1;
    if (gemm<HID, 2 * HID, HID, EPI_F32>(n, n.dP, nullptr, n.W(n.w_l(l)), nullptr, nullptr, nullptr, n.dZ, nullptr,
                                         C, s))
      return -1;
    NE_LAUNCH(n.NE, k_aggregate_back, nblk(C, 4), 256, s, I, (const float *)dHnext, (const float *)n.dZ, dHcur);
    std::swap(dHnext, dHcur);
  }
  // encoder (dHnext holds dL/dh0)
  hipLaunchKernelGGL(k_relu_back, dim3(nblk(ch)), dim3(256), 0, s, (const float *)dHnext, (const bf16 *)n.H[0], n.dP,
                     ch);
  if (gemm_tn<KIN, KIN, HID>(n, n.X, nullptr, n.dP, n.G(n.w_in()), C, s)) return -1;
  if (colsum<HID>(n, n.dP, n.G(n.b_in()), C, s)) return -1;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sync_bf16(Net &n, hipStream_t s) {
  for (const Seg &sg : n.segs) {
    if (sg.boff < 0) continue;
    hipLaunchKernelGGL(k_to_bf16_pair, dim3(nblk((size_t)sg.K * sg.N)), dim3(256), 0, s,
                       (const float *)(n.params + sg.off), sg.K, sg.N, n.w_bf + sg.boff, n.wt_bf + sg.boff);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int adam(Net &n, float lr, float b1, float b2, float eps, int t, hipStream_t s) {
  const float bc1 = (float)(1.0 - std::pow((double)b1, t)), bc2 = (float)(1.0 - std::pow((double)b2, t));
  hipLaunchKernelGGL(k_adam, dim3(nblk(n.n_params)), dim3(256), 0, s, n.params, (const float *)n.grads, n.adam_m,
                     n.adam_v, n.n_params, lr, b1, b2, eps, bc1, bc2);
  if (hipGetLastError() != hipSuccess) return -1;
  return sync_bf16(n, s);
}

int target(Net &n, const Instance &Ic, const Instance &If, hipStream_t s) {
  const TargetArgs a{n.C, n.E, Ic.n.L, Ic.n.RS, Ic.b.p_theta, Ic.b.rho, Ic.b.p_w, Ic.b.p_u,
                     If.b.p_theta, If.b.rho, If.b.p_w, If.b.p_u};
  hipLaunchKernelGGL(k_target_cells, dim3(nblk(n.C, 4)), dim3(256), 0, s, a, (const float *)n.out_scale, n.Tc);
  hipLaunchKernelGGL(k_target_edges, dim3(nblk(n.E, 4)), dim3(256), 0, s, a, (const float *)n.out_scale, n.Te);
  n.have_target = 1;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int apply(Net &n, const Instance &I, float gain, hipStream_t s) {
  const unsigned gc = nblk(n.C, 4);
  prof_mark("k_net_apply", s, 1);
  hipLaunchKernelGGL(k_mass_partial, dim3(gc), dim3(256), 0, s, I, (const float *)n.Y, (const float *)n.out_scale,
                     n.mpart);
  hipLaunchKernelGGL(k_mass_shift, dim3(1), dim3(256), 0, s, (const double *)n.mpart, (int)gc, n.shift);
  hipLaunchKernelGGL(k_apply_cells, dim3(gc), dim3(256), 0, s, I, (const float *)n.Y, (const float *)n.out_scale,
                     (const float *)n.shift, gain);
  hipLaunchKernelGGL(k_apply_edges, dim3(nblk(n.E, 4)), dim3(256), 0, s, I, (const float *)n.Y,
                     (const float *)n.out_scale, gain);
  prof_mark("k_net_apply", s, 0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int checks(Net &n, const Instance &I, float dt, float out[3], double *mass, hipStream_t s) {
  const unsigned gc = nblk(n.C, 4);
  prof_mark("k_net_checks", s, 1);
  NE_LAUNCH(n.NE, k_checks_cells, gc, 256, s, I, dt, n.cpart, n.mpart);
  hipLaunchKernelGGL(k_checks_final, dim3(1), dim3(256), 0, s, (const float *)n.cpart, (const double *)n.mpart,
                     (int)gc, n.checks, n.mass_out);
  prof_mark("k_net_checks", s, 0);
  if (hipGetLastError() != hipSuccess) return -1;
  // one host read: the three maxima and the binary64 mass side by side
  if (hipMemcpyAsync(n.check_host, n.checks, 4 * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&n.check_host[4], n.mass_out, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -1;
  for (int v = 0; v < 3; ++v) out[v] = n.check_host[v];
  std::memcpy(mass, &n.check_host[4], sizeof(double));
  return 0;
}

Net *create(const Instance &I, int layers, std::string *err) {
  Net *n = new Net();
  n->C = I.n.C;
  n->E = I.n.E;
  n->NL = layers;
  n->NE = I.n.ME;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
      cus = prop.multiProcessorCount;
  }
  n->dev_blocks = cus;
  // parameter segments
  size_t off = 0, boff = 0;
  auto mat = [&](int K, int N) {
    n->segs.push_back({(long long)off, K, N, (long long)boff});
    off += (size_t)K * N;
    boff += (size_t)K * N;
  };
  auto vec = [&](int N) {
    n->segs.push_back({(long long)off, 1, N, -1});
    off += (size_t)N;
  };
  mat(KIN, HID);
  vec(HID);
  for (int l = 0; l < layers; ++l) {
    mat(2 * HID, HID);
    vec(HID);
  }
  mat(HID, NOUT);
  vec(NOUT);
  n->n_params = off;
  n->n_bf = boff;
  const size_t C = (size_t)n->C, E = (size_t)n->E;
  int rc = 0;
  rc |= n->alloc(&n->params, off);
  rc |= n->alloc(&n->grads, off);
  rc |= n->alloc(&n->adam_m, off);
  rc |= n->alloc(&n->adam_v, off);
  rc |= n->alloc(&n->w_bf, boff);
  rc |= n->alloc(&n->wt_bf, boff);
  rc |= n->alloc(&n->in_mean, KIN);
  rc |= n->alloc(&n->in_scale, KIN);
  rc |= n->alloc(&n->out_scale, NOUT);
  rc |= n->alloc(&n->X, C * KIN);
  n->H.assign(layers + 1, nullptr);
  n->A.assign(layers, nullptr);
  n->R.assign(layers, nullptr);
  for (auto &p : n->H) rc |= n->alloc(&p, C * HID);
  for (auto &p : n->A) rc |= n->alloc(&p, C * HID);
  for (auto &p : n->R) rc |= n->alloc(&p, C * HID);
  rc |= n->alloc(&n->Y, C * NOUT);
  rc |= n->alloc(&n->Tc, C * 3 * HX_ROW);
  rc |= n->alloc(&n->Te, E * HX_ROW);
  rc |= n->alloc(&n->g, E * HX_ROW);
  rc |= n->alloc(&n->dY, C * NOUT);
  rc |= n->alloc(&n->dP, C * HID);
  rc |= n->alloc(&n->dH0, C * HID);
  rc |= n->alloc(&n->dH1, C * HID);
  rc |= n->alloc(&n->dZ, C * 2 * HID);
  rc |= n->alloc(&n->partial, (size_t)GPART * KIN * NOUT);
  rc |= n->alloc(&n->lpart, (size_t)nblk(E, 4) + 3 * (size_t)nblk(C, 4));
  rc |= n->alloc(&n->loss_out, 4);
  rc |= n->alloc(&n->mpart, 2 * (size_t)nblk(C, 4));
  rc |= n->alloc(&n->shift, 1);
  rc |= n->alloc(&n->checks, 4);
  rc |= n->alloc(&n->cpart, 3 * (size_t)nblk(C, 4));
  rc |= n->alloc(&n->mass_out, 1);
  if (rc) {
    *err = "device allocation for the corrector failed";
    delete n;
    return nullptr;
  }
  return n;
}

void destroy(Net *n) { delete n; }
size_t param_count(const Net *n) { return n->n_params; }

int set_params(Net &n, const float *host, hipStream_t s) {
  if (hipMemcpyAsync(n.params, host, n.n_params * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess) return -1;
  if (hipMemsetAsync(n.adam_m, 0, n.n_params * sizeof(float), s) != hipSuccess) return -1;
  if (hipMemsetAsync(n.adam_v, 0, n.n_params * sizeof(float), s) != hipSuccess) return -1;
  if (sync_bf16(n, s)) return -1;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

int get_vec(const Net &n, int which, float *host, hipStream_t s) {
  const float *src = which == 0 ? n.params : n.grads;
  if (hipMemcpyAsync(host, src, n.n_params * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

int set_scales(Net &n, const float *in_mean, const float *in_scale, const float *out_scale, hipStream_t s) {
  if (hipMemcpyAsync(n.in_mean, in_mean, KIN * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(n.in_scale, in_scale, KIN * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(n.out_scale, out_scale, NOUT * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess)
    return -1;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

int raw_features(Net &n, const Instance &I, const int *old_c, float *dst, hipStream_t s) {
  if (!n.raw && n.alloc(&n.raw, (size_t)n.C * KIN)) return -1;
  if (features(n, I, s, n.raw)) return -1;
  hipLaunchKernelGGL(k_export_rows, dim3(nblk((size_t)n.C * KIN)), dim3(256), 0, s, (const float *)n.raw, old_c, n.C,
                     KIN, dst);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int export_buf(Net &n, int what, const int *old_c, const int *old_e, float *dst, hipStream_t s) {
  switch (what) {
    case HEXDYN_NET_X:
      hipLaunchKernelGGL(k_export_rows_bf16, dim3(nblk((size_t)n.C * KIN)), dim3(256), 0, s, (const bf16 *)n.X, old_c,
                         n.C, KIN, dst);
      break;
    case HEXDYN_NET_Y:
      hipLaunchKernelGGL(k_export_rows, dim3(nblk((size_t)n.C * NOUT)), dim3(256), 0, s, (const float *)n.Y, old_c,
                         n.C, NOUT, dst);
      break;
    case HEXDYN_NET_TARGET_CELLS:
      hipLaunchKernelGGL(k_export_rows, dim3(nblk((size_t)n.C * 3 * HX_ROW)), dim3(256), 0, s, (const float *)n.Tc,
                         old_c, n.C, 3 * HX_ROW, dst);
      break;
    case HEXDYN_NET_TARGET_EDGES:
      hipLaunchKernelGGL(k_export_rows, dim3(nblk((size_t)n.E * HX_ROW)), dim3(256), 0, s, (const float *)n.Te, old_e,
                         n.E, HX_ROW, dst);
      break;
    default:
      return 1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace net
}  // namespace hexdyn
