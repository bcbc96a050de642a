/*
 * FP64 global integrals of the published state: dry mass, rho*theta_m,
 * total energy and the tracer masses, each a sum over cells and levels with
 * the cell-volume weight areaCell * dzw in binary64.
 *
 * The weights and products are formed in binary64 from the binary32 fields
 * (the reference's owned-union totals, cuda_fullstep_lts_active.py:1337-1360,
 * sum rho * (area * dzw) in float64 the same way).  Reduction order is fixed:
 * a lane tree inside each cell row, the block's four rows in order, then one
 * block sums the per-block partials in contiguous chunks with Neumaier
 * compensation and a fixed tree over the chunks.  Repeated calls on the same
 * state return identical bits.
 *
 * Total energy density: rho_zz * (c_v T + g z + K), T = theta_m * exner /
 * (1 + (R_v/R_d) q_v) with q_v the first scalar when the state is moist, z the
 * layer-mid height from zgrid (no potential term without zgrid), K the cell
 * kinetic energy (1/A) sum_e dc dv u_e^2 / 4.
 */
#include <hip/hip_runtime.h>

#include "hexdyn_internal.h"

namespace hexdyn {
namespace {

constexpr double kCv = 1004.5 - 287.0;
constexpr double kG = 9.80616;
constexpr double kRvRd = 461.6 / 287.0;

struct IntegralArgs {
  const float *rho, *rt, *q, *theta, *exner, *u;
  const float *zmid;   // [C][64] layer-mid heights (may be null)
  int NT, moist;
};

__device__ __forceinline__ double wave_sum(double x) {
  // fixed-order tree over the 64 lanes; lane 0 holds the total
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_down(x, off, 64);
  return x;
}

template <int NE>
__global__ void __launch_bounds__(256) k_integrals(const Instance I, const IntegralArgs a, double *partial) {
  extern __shared__ double s_part[];  // [4 waves][NQ]
  const int NQ = 3 + a.NT;
  const int w = (int)(threadIdx.x >> 6);
  const int c0 = blockIdx.x * 4 + w;
  const int k = (int)(threadIdx.x & 63);
  const bool live = c0 < I.n.C && k < I.n.L;
  const int c = c0 < I.n.C ? c0 : 0;
  const int rs = I.n.RS;
  const size_t row = (size_t)c * rs + (k < rs ? k : rs - 1);  // rows rs floats apart
  double wgt = 0.0;
  double rho = 0.0;
  if (live) {
    wgt = (1.0 / (double)I.m.invarea[c]) * (1.0 / (double)I.z.rdzw[k]);
    rho = (double)a.rho[row];
  }
  // kinetic energy at the cell centre
  double ke = 0.0;
  const int n = I.m.neoc[c];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    const int e = I.m.eoc[c * NE + s];
    const double ue = (double)a.u[(size_t)e * rs + (k < rs ? k : rs - 1)];
    if (s < n) ke += 0.25 * (double)I.m.dcdv[e] * ue * ue;
  }
  ke *= (double)I.m.invarea[c];
  double qv = 0.0;
  if (a.moist && a.NT > 0) qv = (double)a.q[row];
  const double temp = (double)a.theta[row] * (double)a.exner[row] / (1.0 + kRvRd * qv);
  const double z = a.zmid ? (double)a.zmid[row] : 0.0;
  double vals[3];
  vals[0] = wgt * rho;
  vals[1] = live ? wgt * (double)a.rt[row] : 0.0;
  vals[2] = live ? wgt * rho * (kCv * temp + kG * z + ke) : 0.0;
  for (int i = 0; i < 3; ++i) {
    const double t = wave_sum(vals[i]);
    if (k == 0) s_part[w * NQ + i] = t;
  }
  for (int sp = 0; sp < a.NT; ++sp) {
    const double x = live ? wgt * rho * (double)a.q[(size_t)sp * I.n.C * rs + row] : 0.0;
    const double t = wave_sum(x);
    if (k == 0) s_part[w * NQ + 3 + sp] = t;
  }
  __syncthreads();
  if ((int)threadIdx.x < NQ) {
    const int i = (int)threadIdx.x;
    partial[(size_t)blockIdx.x * NQ + i] = ((s_part[i] + s_part[NQ + i]) + s_part[2 * NQ + i]) + s_part[3 * NQ + i];
  }
}

// one block: quantity i of nb partials -> out[i]
__global__ void __launch_bounds__(256) k_integrals_final(const double *partial, int nb, int NQ, double *out) {
  __shared__ double s_sum[256];
  const int t = (int)threadIdx.x;
  const int chunk = (nb + 255) / 256;
  for (int i = 0; i < NQ; ++i) {
    double sum = 0.0, comp = 0.0;
    const int b0 = t * chunk, b1 = min(nb, b0 + chunk);
    for (int b = b0; b < b1; ++b) {
      const double x = partial[(size_t)b * NQ + i];
      const double tt = sum + x;
      comp += fabs(sum) >= fabs(x) ? (sum - tt) + x : (x - tt) + sum;  // Neumaier
      sum = tt;
    }
    s_sum[t] = sum + comp;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
      if (t < off) s_sum[t] = s_sum[t] + s_sum[t + off];
      __syncthreads();
    }
    if (t == 0) out[i] = s_sum[0];
    __syncthreads();
  }
}

}  // namespace

int integrals(const Instance &I, const float *rho, const float *rt, const float *q, const float *theta,
              const float *exner, const float *u, const float *zmid, double *partial, double *out,
              hipStream_t s) {
  IntegralArgs a{rho, rt, q, theta, exner, u, zmid, I.n.NT, I.n.NM > 0};
  const int nb = (I.n.C + 3) / 4;
  const int NQ = 3 + I.n.NT;
  const size_t shm = (size_t)4 * NQ * sizeof(double);
  if (I.n.ME == 6)
    hipLaunchKernelGGL(k_integrals<6>, dim3(nb), dim3(256), shm, s, I, a, partial);
  else
    hipLaunchKernelGGL(k_integrals<7>, dim3(nb), dim3(256), shm, s, I, a, partial);
  hipLaunchKernelGGL(k_integrals_final, dim3(1), dim3(256), 0, s, (const double *)partial, nb, NQ, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace hexdyn
