/*
 * hexdyn device kernels and the phase functions that launch them.
 *
 * This translation unit is compiled twice: with -ffp-contract=off into
 * namespace hexdyn::strict (every binary32 multiply and add rounded
 * separately, in the reference's operand order, so results equal the CPU
 * restatement bit for bit) and with -ffp-contract=fast into hexdyn::fast.
 *
 * Mapping: a block of 256 threads is 4 wavefronts; each wavefront owns one
 * entity (cell, edge or vertex) row of 64 floats and lane k owns level k.
 * Connectivity of the entity is wave-uniform (scalar loads); neighbour rows
 * are read as whole 256-byte lines, so every cellsOnEdge / edgesOnCell /
 * advCellsForEdge gather is one coalesced access per neighbour.  Blocks are
 * renumbered so each XCD (hardware assigns block b to XCD b % 8) works on a
 * contiguous range of the space-filling-curve order and its L2 sees the
 * neighbour rows it needs.  Vertical neighbours come from cross-lane moves;
 * the serial column recurrences (Thomas sweep, implicit coefficients) walk
 * the lanes with readlane so their arithmetic order is the reference's.
 *
 * Reference operators (operation order per element):
 *   solve diagnostics  cuda_horizontal_v841.py:67-200, diagnostics.py:95-306
 *   dyn tendencies     dynamics.py:86-268, dynamics_v841.py:128-185,
 *                      driver.py:933-1207, acoustic.py:246-281
 *   implicit coefs     acoustic_v841.py:81-233
 *   acoustic substep   acoustic_v841.py:296-565
 *   divergence damping mixing.py:851-945
 *   recovery           driver.py:3176-3290, driver.py:819-876, terrain.py:491-571
 *   scalar transport   transport.py:1649-1826, transport.py:1829-2175
 */
#include <hip/hip_runtime.h>

#include "hexdyn_internal.h"
#include "hexdyn_powf.h"

#ifndef HEXDYN_NS
#define HEXDYN_NS strict
#endif

namespace hexdyn {
namespace HEXDYN_NS {
namespace {

// Element k of row ent of p, addressed as a 32-bit byte offset from the
// wave-uniform base (every array is below 4 GiB), so loads and stores take
// the SGPR-base + VGPR-offset form instead of a per-lane 64-bit address.
// Rows are rs floats apart (the level count plus one, rounded up to 8:
// 224-byte rows at 55 levels, so a streamed array moves 12.5 % fewer bytes
// than with 64-float rows, while a gathered row still spans four 64-byte
// sectors).  Lanes k >= rs read the row's last element instead of the next
// row: for layer arrays that element is padding (zero), so every lane sees
// what a 64-float row would hold; for interface arrays it is the top value,
// which no lane at or below the top reads through a shuffle.  Stores are
// made for k <= L < rs only.
__device__ __forceinline__ uint32_t row_off(int rs, int ent, int k) {
  return (uint32_t)ent * (uint32_t)(rs * 4) + (uint32_t)(k < rs ? k : rs - 1) * 4u;
}
#define AT(p, ent, k) (*(float *)((char *)(p) + row_off(I.n.RS, (ent), (k))))
// gather kernels take up to 16 rows (wavefronts) per block: a per-kernel launch setting
#define KB __global__ void __launch_bounds__(1024)

// First entity of this block: blocks are renumbered so that XCD x (hardware
// block b runs on XCD b % 8) owns a contiguous range of the curve.  wpb is
// the number of wavefronts (entities) per block.
__device__ __forceinline__ int block_entity0(int wpb) {
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int x = b & 7, i = b >> 3, q = nb >> 3, r = nb & 7;
  const int lb = (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
  return lb * wpb;
}
// Column kernels: WPB rows per block, fixed at compile time.
template <int WPB>
__device__ __forceinline__ int wave_entity() {
  return __builtin_amdgcn_readfirstlane(block_entity0(WPB) + (int)(threadIdx.x >> 6));
}
// Gather kernels: the block size is a per-kernel launch setting, carried in
// the kernel's own Instance argument (I.n.WPB) rather than read from the
// dispatch packet, which would cost each wavefront a vector-memory round trip
// before its first row load.
__device__ __forceinline__ int wave_entity(const Instance &I) {
  return __builtin_amdgcn_readfirstlane(block_entity0(I.n.WPB) + (int)(threadIdx.x >> 6));
}
__device__ __forceinline__ int lane() { return (int)(threadIdx.x & 63); }
__device__ __forceinline__ float up1(float x) { return __shfl_up(x, 1, 64); }     // value of lane k-1
__device__ __forceinline__ float dn1(float x) { return __shfl_down(x, 1, 64); }   // value of lane k+1
__device__ __forceinline__ float rdl(float x, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
}
__device__ __forceinline__ float cs1(float x) { return __builtin_copysignf(1.0f, x); }
__device__ __forceinline__ float max0(float x) { return x > 0.0f ? x : 0.0f; }
__device__ __forceinline__ float min0(float x) { return x < 0.0f ? x : 0.0f; }
__device__ __forceinline__ float pmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float pmax(float a, float b) { return b > a ? b : a; }
__device__ __forceinline__ int cl(int k) { return k < 0 ? 0 : (k > 63 ? 63 : k); }
// Value of x held by lane cl(k + d): the same value a load of row[cl(k + d)]
// gives, taken from the row already in registers.  Whole-wave shuffles: call
// with every lane active (outside lane-dependent branches).
__device__ __forceinline__ float lane_off(float x, int k, int d) { return __shfl(x, cl(k + d), 64); }

// x / 12 with IEEE round-to-nearest.  For quotients in [2^-125, FLT_MAX] the
// product by RN(1/12) plus one FMA residual correction is the correctly
// rounded quotient (verified exhaustively over every binary32 x); zeros,
// tiny, infinite and NaN inputs take the division itself.
__device__ __forceinline__ float div12(float x) {
  constexpr float r12 = 0x1.555556p-4f;  // RN(1/12)
  const float q0 = x * r12;
  const float a = fabsf(q0);
  if (a >= 0x1p-125f && a <= 0x1.fffffep127f) return __builtin_fmaf(__builtin_fmaf(-q0, 12.0f, x), r12, q0);
  return x / 12.0f;
}

__device__ __forceinline__ float flux4(float qm2, float qm1, float q0, float qp1, float vel) {
  return div12(vel * (7.0f * (q0 + qm1) - (qp1 + qm2)));
}
__device__ __forceinline__ float flux3(float qm2, float qm1, float q0, float qp1, float vel, float coef) {
  return flux4(qm2, qm1, q0, qp1, vel) + div12((coef * fabsf(vel)) * ((qp1 - qm2) - 3.0f * (q0 - qm1)));
}

__device__ __forceinline__ void set_flag(int *flag) { atomicOr(flag, 1); }

// exner = arg^(R/cv) with the GNU powf bits the reference's v8.4.1 recovery
// uses (cuda_driver.py:1466-1472): the correctly rounded binary64 method,
// moved to the other neighbour where the host-built bitmap marks the argument.
__device__ __forceinline__ float exner_pow(const Instance &I, float arg) {
  const double d = hexdyn_pow_d(arg, I.k.rcv);
  return hexdyn_gnu_powf_fix(arg, (float)d, d, I.m.powf_ties, I.k.powf_near_tie);
}

// Per-slot coefficients of one entity (its row of N values in an [n][N]
// table), read by one vector load with lane s taking slot s and issued with
// the row loads; slot s is read from lane s where it is used.  The same
// values as scalar loads miss the scalar cache for every entity, and the
// compiler issues such loads next to their uses, after the rows have
// arrived, so each becomes a dependent round trip of the wavefront.
// Read a slot only in wave-uniform control flow: the compiler may move the
// load into the first branch that reads the row, and in a lane-divergent
// branch the lanes that hold the other slots would not execute it.  Code
// that needs slots under a lane condition takes them with slot_vals first.
template <int N>
struct SlotRow {
  float v;  // lane s: slot s
  __device__ __forceinline__ float operator[](int s) const { return rdl(v, s); }
};
template <int N>
__device__ __forceinline__ SlotRow<N> slot_row(const float *__restrict__ p) {
  const int l = lane();
  return {p[l < N ? l : N - 1]};
}
// every slot as a wave-uniform value (call with all lanes active)
template <int N>
__device__ __forceinline__ void slot_vals(SlotRow<N> &r, float (&o)[N]) {
#pragma unroll
  for (int s = 0; s < N; ++s) o[s] = r[s];
}
// Ends the load phase for x: the compiler may not sink the load of x into
// the first (uniform) branch that uses it, which would issue it, and the
// scalar index load it needs, one round trip after the others.
__device__ __forceinline__ void keep(float &x) { asm volatile("" : "+v"(x)); }
template <int N>
__device__ __forceinline__ void keep(float (&x)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) keep(x[i]);
}
template <int G, int N>
__device__ __forceinline__ void keep(float (&x)[G][N]) {
#pragma unroll
  for (int g = 0; g < G; ++g) keep(x[g]);
}
template <class... T>
__device__ __forceinline__ void keep(T &...x) {
  (keep(x), ...);
}
// Marks slot rows as read only after this point: called once the
// wavefront's row loads are issued, so the lane reads of the slots (which
// wait for the slot loads) are not scheduled ahead of the row loads.  Call
// with all lanes active.  Emits nothing.
template <class R>
__device__ __forceinline__ void settle1(R &row) {
  asm volatile("" : "+v"(row.v));
}
template <class... R>
__device__ __forceinline__ void settle(R &...rows) {
  (settle1(rows), ...);
}
// Ends the index phase of a wavefront: every scalar index load issued before
// the first fence completes in one wait, before any row load after it is
// issued.  Without it the compiler may issue one index table's loads only
// after waiting for another's, one more dependent round trip.
__device__ __forceinline__ void fence(int x) { asm volatile("" ::"s"(x)); }
// A per-entity value read by a vector load (every lane the same address) for
// the same reason; the zero offset is opaque, so the load stays a vector load.
__device__ __forceinline__ float uni_load(const float *__restrict__ p) {
  int z;
  asm("v_mov_b32 %0, 0" : "=v"(z));
  return p[z];
}

inline int nblocks(int n, int wpb) { return (n + wpb - 1) / wpb; }

// Column kernels: W cells (one wavefront each) per block.  The serial
// vertical recurrences of the block's cells run together on the lanes of
// one wavefront (lane = cell) through LDS rows padded to COLP floats, so
// W columns cost one column's dependent chain instead of W.  The implicit
// coefficients take 16 columns per block; the acoustic column update, whose
// two sweeps hold the block for longer, takes 8 (four blocks per CU
// overlap their loads with the others' sweeps: 57 -> 53 us; the
// coefficients were 1 us slower at 8).
constexpr int COLW_VC = 16;  // k_vcoef
constexpr int COLW_AC = 8;   // k_ac_cell (at least 7: its waves 0-6 stage the level rows)
constexpr int COLP = 65;
// The column blocks of a CU need every SIMD to hold 8 wavefronts: the
// register budget is capped for that (at 7 a CU would hold one block of 16).
#define KBCOL(W) __global__ void __launch_bounds__(64 * (W)) __attribute__((amdgpu_waves_per_eu(8, 8)))

// Per-level coefficient rows are the same for every column of a block: wave
// j < N of a column block copies row j into LDS once, instead of each of the
// 16 wavefronts reading every row.  Callers synchronise before reading.
template <int N>
__device__ __forceinline__ void stage_levels(float *s_lev, const float *const (&rows)[N], int w, int k) {
#pragma unroll
  for (int j = 0; j < N; ++j)
    if (w == j) s_lev[j * 64 + k] = rows[j][k];
}

// Gather helpers.  Connectivity rows are padded to fixed widths (NE = maxEdges
// slots per cell, NX = 2*maxEdges-2 slots per edge) with index 0, so every
// slot of an unrolled loop addresses a valid row.  All neighbour rows are
// loaded before any arithmetic, then combined in slot order for the valid
// slots only, which keeps the reference's association and lets the memory
// system overlap every gather of the wave.

// ------------------------------------------------------------------ diagnostics
KB k_diag_vertex(const Instance I, const float *__restrict__ u, int store_mix) {
  const int v = wave_entity(I);
  if (v >= I.n.V) return;
  const int k = lane();
  int e[3];
  float ue[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) e[s] = I.m.eov[3 * v + s];
  SlotRow<3> dc = slot_row<3>(I.m.vdc + 3 * v), dcdv = slot_row<3>(I.m.vdcdv + 3 * v);
  const float it = uni_load(I.m.invtri + v), fv = uni_load(I.m.fvert + v);
#pragma unroll
  for (int s = 0; s < 3; ++s) fence(e[s]);
#pragma unroll
  for (int s = 0; s < 3; ++s) ue[s] = AT(u, e[s], k);
  keep(ue, it, fv);
  settle(dc, dcdv);
  float vort = 0.0f, kev = 0.0f;
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    vort = vort + dc[s] * ue[s];
    kev = kev + dcdv[s] * (ue[s] * ue[s]);
  }
  vort = vort * it;
  kev = kev * (0.25f * it);
  const float pvv = fv + vort;
  if (k < I.n.L) {
    if (store_mix) AT(I.b.vorticity, v, k) = vort;
    AT(I.b.ke_vertex, v, k) = kev;
    AT(I.b.pv_vertex, v, k) = pvv;
  }
}

// divergence, kinetic energy (0.625/0.375 kite blend) and pv_cell of cell c
struct CellDiag {
  float div, ke, pv;
};
template <int NE>
struct CellDiagIn {
  float ue[NE], kv[NE], pv[NE];
  SlotRow<NE> sd, dd, kt;
  float ia;
  int n;
};
template <int NE>
__device__ __forceinline__ CellDiagIn<NE> diag_cell_load(const Instance &I, const float *__restrict__ u, int c,
                                                         int k) {
  CellDiagIn<NE> a;
  a.n = I.m.neoc[c];
  int e[NE], vv[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    e[s] = I.m.eoc[c * NE + s];
    vv[s] = I.m.voc[c * NE + s];
  }
  a.sd = slot_row<NE>(I.m.csdv + c * NE);
  a.dd = slot_row<NE>(I.m.cdcdv + c * NE);
  a.kt = slot_row<NE>(I.m.ckite + c * NE);
  a.ia = uni_load(I.m.invarea + c);
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    fence(e[s]);
    fence(vv[s]);
  }
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    a.ue[s] = AT(u, e[s], k);
    a.kv[s] = AT(I.b.ke_vertex, vv[s], k);
    a.pv[s] = AT(I.b.pv_vertex, vv[s], k);
  }
  keep(a.ue, a.kv, a.pv, a.ia);
  settle(a.sd, a.dd, a.kt);
  return a;
}
template <int NE>
__device__ __forceinline__ CellDiag diag_cell_compute(const Instance &I, const CellDiagIn<NE> &a) {
  const int n = a.n;
  const float ia = a.ia;
  float d = 0.0f, kc = 0.0f, pvc = 0.0f;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) {
      d = d + a.sd[s] * a.ue[s];
      kc = kc + 0.25f * (a.dd[s] * (a.ue[s] * a.ue[s]));
    }
  d = d * ia;
  kc = kc * ia;
  kc = kc * 0.625f;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) {
      const float kite = a.kt[s];
      float corr = 0.375f * kite;
      corr = corr * a.kv[s];
      corr = corr * ia;
      kc = kc + corr;
      pvc = pvc + (kite * a.pv[s]) * ia;
    }
  return {d, kc, pvc};
}
__device__ __forceinline__ void store_cell_diag(const Instance &I, int c, int k, CellDiag r, int store_mix) {
  if (k < I.n.L) {
    if (store_mix) AT(I.b.divergence, c, k) = r.div;
    AT(I.b.ke, c, k) = r.ke;
    if (I.k.apvm_on) AT(I.b.pv_cell, c, k) = r.pv;
  }
}

template <int NE>
KB k_diag_cell(const Instance I, const float *__restrict__ u, int store_mix) {
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  store_cell_diag(I, c, k, diag_cell_compute<NE>(I, diag_cell_load<NE>(I, u, c, k)), store_mix);
}

// pv_edge (with APVM) and, at rk step 3, the TRiSK tangential velocity v of
// edge e.  The bodies below only load and compute; kernels store afterwards,
// so no load of a pass waits behind one of its own stores.
struct EdgeDiag {
  float v, pv;
};
template <int NE, bool RECON>
struct EdgeDiagIn {
  float un[RECON ? 2 * NE - 2 : 1];
  SlotRow<2 * NE - 2> wt;
  float vrow, pv0, pv1, pc0, pc1, ue, idv, idc;
  int n;
  int v0, v1, c0, c1, nb[RECON ? 2 * NE - 2 : 1];  // scalar index loads
};
// first phase: indices, slot counts and per-edge coefficients (none of
// which needs another load first)
template <int NE, bool RECON>
__device__ __forceinline__ void diag_edge_index(const Instance &I, int e, EdgeDiagIn<NE, RECON> &a) {
  constexpr int NX = 2 * NE - 2;
  a.v0 = I.m.voe[2 * e];
  a.v1 = I.m.voe[2 * e + 1];
  a.c0 = I.m.coe[2 * e];
  a.c1 = I.m.coe[2 * e + 1];
  a.idv = uni_load(I.m.invdv + e);
  a.idc = uni_load(I.m.invdc + e);
  if (RECON) {
#pragma unroll
    for (int s = 0; s < NX; ++s) a.nb[s] = I.m.eoe[e * NX + s];
    a.n = I.m.neoe[e];
    a.wt = slot_row<NX>(I.m.woe + e * NX);
  } else {
    a.wt = {0.0f};
    a.n = 0;
  }
}
template <int NE, bool RECON>
__device__ __forceinline__ void diag_edge_fence(const EdgeDiagIn<NE, RECON> &a) {
  fence(a.v0);
  fence(a.v1);
  fence(a.c0);
  fence(a.c1);
  if (RECON) {
#pragma unroll
    for (int s = 0; s < 2 * NE - 2; ++s) fence(a.nb[s]);
  }
}
// second phase: the rows (pv_cell and u are read whether or not APVM is on:
// both buffers always exist)
template <int NE, bool RECON>
__device__ __forceinline__ void diag_edge_rows(const Instance &I, const float *__restrict__ u, int e, int k,
                                               EdgeDiagIn<NE, RECON> &a) {
  constexpr int NX = 2 * NE - 2;
  if (RECON) {
#pragma unroll
    for (int s = 0; s < NX; ++s) a.un[s] = AT(u, a.nb[s], k);
    a.vrow = 0.0f;
  } else {
    a.vrow = AT(I.b.v, e, k);
  }
  a.pv0 = AT(I.b.pv_vertex, a.v0, k);
  a.pv1 = AT(I.b.pv_vertex, a.v1, k);
  a.pc0 = AT(I.b.pv_cell, a.c0, k);
  a.pc1 = AT(I.b.pv_cell, a.c1, k);
  a.ue = AT(u, e, k);
}
template <int NE, bool RECON>
__device__ __forceinline__ EdgeDiagIn<NE, RECON> diag_edge_load(const Instance &I, const float *__restrict__ u, int e,
                                                                int k) {
  EdgeDiagIn<NE, RECON> a;
  diag_edge_index(I, e, a);
  diag_edge_fence(a);
  diag_edge_rows(I, u, e, k, a);
  settle(a.wt);
  return a;
}
template <int NE, bool RECON>
__device__ __forceinline__ EdgeDiag diag_edge_compute(const Instance &I, int e, const EdgeDiagIn<NE, RECON> &a) {
  constexpr int NX = 2 * NE - 2;
  float vv;
  if (RECON) {
    vv = 0.0f;
#pragma unroll
    for (int s = 0; s < NX; ++s)
      if (s < a.n) vv = vv + a.wt[s] * a.un[s];
  } else {
    vv = a.vrow;
  }
  const float pve = 0.5f * (a.pv0 + a.pv1);
  // computed either way (a select, not a branch, so the loads are not sunk)
  const float gt = (a.pv1 - a.pv0) * a.idv;
  const float gn = (a.pc1 - a.pc0) * a.idc;
  const float pva = pve - I.k.apvm_scale * (vv * gt + a.ue * gn);
  return {vv, I.k.apvm_on ? pva : pve};
}
template <int NE, bool RECON>
__device__ __forceinline__ EdgeDiag diag_edge_values(const Instance &I, const float *__restrict__ u, int e, int k) {
  return diag_edge_compute<NE, RECON>(I, e, diag_edge_load<NE, RECON>(I, u, e, k));
}
__device__ __forceinline__ void store_edge_diag(const Instance &I, int e, int k, bool recon, EdgeDiag d) {
  if (k < I.n.L) {
    if (recon) AT(I.b.v, e, k) = d.v;
    AT(I.b.pv_edge, e, k) = d.pv;
  }
}

// RECON: rk step 3, where v is reconstructed (cuda_horizontal_v841.py:1058-1097)
template <int NE, bool RECON>
KB k_diag_edge(const Instance I, const float *__restrict__ u) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  store_edge_diag(I, e, k, RECON, diag_edge_values<NE, RECON>(I, u, e, k));
}

// theta and w edge fluxes of a stage (defined with the tendencies)
struct EdgeFlux {
  float th, corr, w;
};
template <int NE>
struct EdgeFluxIn;
template <int NE, bool CORR>
__device__ __forceinline__ void tend_flux_index(const Instance &I, int e, int k, EdgeFluxIn<NE> &b);
template <int NE>
__device__ __forceinline__ void tend_flux_fence(const EdgeFluxIn<NE> &b);
template <int NE, bool CORR>
__device__ __forceinline__ void tend_flux_rows(const Instance &I, int e, int k, EdgeFluxIn<NE> &b);
template <int NE, bool CORR>
__device__ __forceinline__ EdgeFlux tend_flux_compute(const Instance &I, int e, const EdgeFluxIn<NE> &b);
template <bool CORR>
__device__ __forceinline__ void store_edge_flux(const Instance &I, int e, int k, EdgeFlux f);

// the edge diagnostics and, in the same pass, the next stage's theta/w edge
// fluxes: every input of those fluxes (theta, w, ru after recovery; the
// subcycle-start ru and theta) is final once recovery has run.  After rk
// step 3 (RECON) the next stage is a first stage (no RK correction flux);
// after rk steps 1 and 2 it is stage 2 or 3 (with the correction).
template <int NE, bool RECON>
KB k_diag_edge_flux(const Instance I, const float *__restrict__ u) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  EdgeDiagIn<NE, RECON> a;
  EdgeFluxIn<NE> b;
  diag_edge_index(I, e, a);
  tend_flux_index<NE, !RECON>(I, e, k, b);
  diag_edge_fence(a);
  tend_flux_fence(b);
  diag_edge_rows(I, u, e, k, a);
  tend_flux_rows<NE, !RECON>(I, e, k, b);
  keep(b.thn, b.wn, b.rue, b.sru, b.sth0, b.sth1, b.fzm, b.fzp, b.dv);
  keep(a.un, a.vrow, a.pv0, a.pv1, a.pc0, a.pc1, a.ue, a.idv, a.idc);
  settle(a.wt, b.av, b.a3v);
  const EdgeDiag d = diag_edge_compute<NE, RECON>(I, e, a);
  const EdgeFlux f = tend_flux_compute<NE, !RECON>(I, e, b);
  store_edge_diag(I, e, k, RECON, d);
  store_edge_flux<!RECON>(I, e, k, f);
}

// The same pass over a tile of HX_TILE_E edges, one wavefront each: the
// theta and w rows of the tile's distinct advection cells cross the texture
// path once per tile into LDS, and the ten-cell stencils read them there.
// The gathers are bound by the texture path's rate per row load (about ten
// cycles per 256-byte row per CU), and consecutive edges along the curve
// share most stencil cells (about 2.8 distinct cells per edge instead of 10).
template <int NE, bool RECON>
__global__ void __launch_bounds__(64 * HX_TILE_E) k_diag_edge_flux_t(const Instance I, const float *__restrict__ u) {
  constexpr int NX = 2 * NE - 2, J = (HX_TILE_U + HX_TILE_E - 1) / HX_TILE_E;
  constexpr bool CORR = !RECON;
  __shared__ float s_th[HX_TILE_U * 64], s_w[HX_TILE_U * 64];
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), k = lane();
  const int tile = block_entity0(1);
  const int e0 = tile * HX_TILE_E + w;
  const bool live = e0 < I.n.E;
  const int e = live ? e0 : I.n.E - 1;  // every wavefront stages its share and reaches the barrier
  const int ncell = I.m.ntcell[tile];
  EdgeDiagIn<NE, RECON> a;
  EdgeFluxIn<NE> b;
  diag_edge_index(I, e, a);
  b.c0 = I.m.coe[2 * e];
  b.c1 = I.m.coe[2 * e + 1];
#pragma unroll
  for (int sl = 0; sl < NX; ++sl) b.ac[sl] = I.m.advl[e * NX + sl];  // slot -> row of the tile's cell list
  b.na = I.m.nadv[e];
  b.av = slot_row<NX>(I.m.adv + e * NX);
  b.a3v = slot_row<NX>(I.m.adv3 + e * NX);
  b.dv = CORR ? uni_load(I.m.dv + e) : 0.0f;
  b.fzm = AT(I.z.fzm, 0, k);
  b.fzp = AT(I.z.fzp, 0, k);
  int cr[J];
#pragma unroll
  for (int j = 0; j < J; ++j) cr[j] = I.m.tcell[tile * HX_TILE_U + min(w + j * HX_TILE_E, HX_TILE_U - 1)];
  diag_edge_fence(a);
  fence(b.c0);
  fence(b.c1);
#pragma unroll
  for (int j = 0; j < J; ++j) fence(cr[j]);
  float st[J], sw[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    st[j] = sw[j] = 0.0f;
    if (w + j * HX_TILE_E < ncell) {  // this wavefront's rows of the tile's list
      st[j] = AT(I.b.theta, cr[j], k);
      sw[j] = AT(I.b.w, cr[j], k);
    }
  }
  diag_edge_rows(I, u, e, k, a);
  b.rue = AT(I.b.c_ru, e, k);
  b.sru = b.sth0 = b.sth1 = 0.0f;
  if (CORR) {
    b.sru = AT(I.b.s_ru, e, k);
    b.sth0 = AT(I.b.s_theta, b.c0, k);
    b.sth1 = AT(I.b.s_theta, b.c1, k);
  }
  keep(st, sw);
#pragma unroll
  for (int j = 0; j < J; ++j)
    if (w + j * HX_TILE_E < ncell) {
      s_th[(w + j * HX_TILE_E) * 64 + k] = st[j];
      s_w[(w + j * HX_TILE_E) * 64 + k] = sw[j];
    }
  __syncthreads();
#pragma unroll
  for (int sl = 0; sl < NX; ++sl) {
    b.thn[sl] = s_th[b.ac[sl] * 64 + k];
    b.wn[sl] = s_w[b.ac[sl] * 64 + k];
  }
  keep(b.thn, b.wn, b.rue, b.sru, b.sth0, b.sth1, b.fzm, b.fzp, b.dv);
  keep(a.un, a.vrow, a.pv0, a.pv1, a.pc0, a.pc1, a.ue, a.idv, a.idc);
  settle(a.wt, b.av, b.a3v);
  const EdgeDiag d = diag_edge_compute<NE, RECON>(I, e, a);
  const EdgeFlux f = tend_flux_compute<NE, CORR>(I, e, b);
  if (live) {
    store_edge_diag(I, e, k, RECON, d);
    store_edge_flux<CORR>(I, e, k, f);
  }
}

// ---------------------------------------------------------- moist coefficients
// the moist species are summed in order, their rows loaded HX_MOIST_G at a
// time before any of them is added (one round trip per batch, not per species)
constexpr int HX_MOIST_G = 8;
KB k_moist_cell(const Instance I) {
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int NM = I.n.NM;
  float t = 0.0f;
  for (int s0 = 0; s0 < NM; s0 += HX_MOIST_G) {
    float q[HX_MOIST_G];
#pragma unroll
    for (int g = 0; g < HX_MOIST_G; ++g) q[g] = s0 + g < NM ? AT(I.b.q, (s0 + g) * I.n.C + c, k) : 0.0f;
    keep(q);
#pragma unroll
    for (int g = 0; g < HX_MOIST_G; ++g)
      if (s0 + g < NM) t = t + q[g];
  }
  const float tm = up1(t);
  const float cqw = k >= 1 ? 1.0f / (1.0f + 0.5f * (t + tm)) : 1.0f;
  if (k < I.n.L) {
    AT(I.b.qtot, c, k) = t;
    AT(I.b.cqw, c, k) = cqw;
  }
}

KB k_moist_edge(const Instance I) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  const int c0 = I.m.coe[2 * e], c1 = I.m.coe[2 * e + 1];
  const int NM = I.n.NM;
  fence(c0);
  fence(c1);
  float t = 0.0f;
  for (int s0 = 0; s0 < NM; s0 += HX_MOIST_G) {
    float q0[HX_MOIST_G], q1[HX_MOIST_G];
#pragma unroll
    for (int g = 0; g < HX_MOIST_G; ++g) {
      q0[g] = s0 + g < NM ? AT(I.b.q, (s0 + g) * I.n.C + c0, k) : 0.0f;
      q1[g] = s0 + g < NM ? AT(I.b.q, (s0 + g) * I.n.C + c1, k) : 0.0f;
    }
    keep(q0, q1);
#pragma unroll
    for (int g = 0; g < HX_MOIST_G; ++g)
      if (s0 + g < NM) t = t + 0.5f * (q0[g] + q1[g]);
  }
  if (k < I.n.L) AT(I.b.cqu, e, k) = 1.0f / (1.0f + t);
}

// ------------------------------------------------------ implicit coefficients
KBCOL(COLW_VC) k_vcoef(const Instance I, float dts, const float *__restrict__ exn, const float *__restrict__ theta,
              const float *__restrict__ rtp) {
  constexpr int COLW = COLW_VC;
  __shared__ float s_a[COLW * COLP], s_b[COLW * COLP], s_c[COLW * COLP];
  __shared__ float s_lev[6 * 64];
  const int w = (int)(threadIdx.x >> 6);
  const int c0 = wave_entity<COLW>();
  const bool live = c0 < I.n.C;
  const int c = live ? c0 : 0;  // every wave reaches the barriers
  const int k = lane();
  const int L = I.n.L;
  const float g = I.k.gravity, rcv = I.k.rcv, c2 = I.k.c2;
  const float dts2 = dts * dts;
  // every load first; the level k-1 and k+1 values come from neighbouring lanes
  const float zk = AT(I.z.zz, c, k), exk = AT(exn, c, k), thk = AT(theta, c, k);
  const float cqw = AT(I.b.cqw, c, k), qt = AT(I.b.qtot, c, k), rtpk = AT(rtp, c, k);
  const float rbk = AT(I.z.rb, c, k), rtbk = AT(I.z.rtb, c, k), exbk = AT(I.z.exb, c, k);
  {
    const float *const lev[6] = {I.z.fzm, I.z.fzp, I.z.rdzu, I.z.rdzw, I.z.etp, I.z.ewp};
    stage_levels<6>(s_lev, lev, w, k);
  }
  __syncthreads();
  const float fzm = s_lev[0 * 64 + k], fzp = s_lev[1 * 64 + k], rdzu = s_lev[2 * 64 + k];
  const float rdzw = s_lev[3 * 64 + k], etp = s_lev[4 * 64 + k], ewp = s_lev[5 * 64 + k];
  const float zkm = lane_off(zk, k, -1), exkm = lane_off(exk, k, -1), thkm = lane_off(thk, k, -1);
  const float rdzwm = lane_off(rdzw, k, -1), etpm = lane_off(etp, k, -1);
  const float ewpm = lane_off(ewp, k, -1), ewpp = lane_off(ewp, k, 1);
  float cofwr = 0.0f, cofwz = 0.0f, coftz = 0.0f;
  if (k >= 1 && k < L) {
    const float iz = fzm * zk + fzp * zkm;
    cofwr = (0.5f * g) * iz;
    cofwz = (((c2 * iz) * rdzu) * cqw) * (fzm * exk + fzp * exkm);
    coftz = fzm * thk + fzp * thkm;
  }
  float cofwt = 0.0f;
  if (k < L) cofwt = (((((((0.5f * rcv) * zk) * g) * rbk) / (1.0f + qt)) * exk) / ((rtbk + rtpk) * exbk));
  const float coftz_m = up1(coftz), coftz_p = dn1(coftz), cofwt_m = up1(cofwt);
  float a = 0.0f, b = 1.0f, cc = 0.0f;
  if (k >= 1 && k < L) {
    a = (((-cofwz) * coftz_m) * rdzwm) * zkm + cofwr * rdzwm - (cofwt_m * coftz_m) * rdzwm;
    a = (a * etpm) * ewpm;
    b = (cofwz * coftz) * ((etp * rdzw) * zk + (etpm * rdzwm) * zkm) -
        coftz * ((etp * cofwt) * rdzw - (etpm * cofwt_m) * rdzwm) + cofwr * (etp * rdzw - etpm * rdzwm);
    b = b * ewp;
    cc = (((-cofwz) * coftz_p) * rdzw) * zk - cofwr * rdzw + (cofwt * coftz_p) * rdzw;
    cc = (cc * etp) * ewpp;
    if (k == L - 1) cc = 0.0f;
  }
  // alpha/gamma recurrence up the column (acoustic_v841.py:81-233), lane = cell
  s_a[w * COLP + k] = a;
  s_b[w * COLP + k] = b;
  s_c[w * COLP + k] = cc;
  __syncthreads();
  if (w == 0 && k < COLW) {
    float *pa = s_a + k * COLP, *pb = s_b + k * COLP;
    const float *pc = s_c + k * COLP;
    float gm = 0.0f;
    int singular = 0;
    for (int j = 1; j < L; ++j) {
      const float den = 1.0f + dts2 * (pb[j] - pa[j] * gm);
      if (den == 0.0f) singular = 1;
      const float al = 1.0f / den;
      gm = (dts2 * pc[j]) * al;
      pa[j] = al;
      pb[j] = gm;
    }
    if (singular && block_entity0(COLW) + k < I.n.C) set_flag(I.b.flag);
  }
  __syncthreads();
  const bool inner = k >= 1 && k < L;
  const float alpha = inner ? s_a[w * COLP + k] : 0.0f;
  const float gamma = inner ? s_b[w * COLP + k] : 0.0f;
  if (live && k <= L) {
    AT(I.b.coftz, c, k) = coftz;
    if (k < L) {
      AT(I.b.cofwz, c, k) = cofwz;
      AT(I.b.cofwt, c, k) = cofwt;
      if (I.k.store_flux) AT(I.b.atri, c, k) = a;  // export only: k_ac_cell forms a itself
      AT(I.b.alpha, c, k) = alpha;
      AT(I.b.gamma, c, k) = gamma;
    }
  }
}

// ------------------------------------------------------- horizontal mixing
// v8.4.1 2-D Smagorinsky mixing, once per dynamics subcycle from the
// subcycle-start state (cuda_horizontal_v841.py:205-583, host sequence
// :1270-1588; mpas_atm_dissipation_models.F smagorinsky_2d and the u/w/theta
// del2/del4 filters).  Edge density is 0.5*(rho[cell0]+rho[cell1]) of the
// current rho_zz, as the subcycle-start solve formed it.

// 2-D Smagorinsky eddy viscosity of level k (lane k) from the deformation
// sums of a cell's edge winds (smagorinsky_v841_f32)
__device__ __forceinline__ float kdiff_from_deformation(const Instance &I, int k, float dudx, float dudy, float dvdx,
                                                       float dvdy) {
  const float d11 = 2.0f * dudx, d22 = 2.0f * dvdy, d12 = dudy + dvdx;
  const float diff = d11 - d22;
  const float strain = __builtin_sqrtf(0.25f * (diff * diff) + d12 * d12);
  const float kd = fminf(I.k.strain_scale * strain, I.k.ceiling);
  // CAM floor over the top levels (native/upper_damping.rs:5-16,
  // cuda_upper_damping_v841.py:5-11): max(kdiff, base*(1 - (L-1-k)/n))
  const int L = I.n.L, ncam = I.k.cam_levels;
  if (k >= L - ncam && k < L) return fmaxf(kd, I.k.cam_base * (1.0f - (float)(L - 1 - k) / (float)ncam));
  return kd;
}




// divergence of delsq_u (laplacian_divergence_v841_f32)
template <int NE>
KB k_mix_lapdiv(const Instance I) {
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int n = I.m.neoc[c];
  int e[NE];
  float x[NE], f[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) e[s] = I.m.eoc[c * NE + s];
  SlotRow<NE> fr = slot_row<NE>(I.m.mx_fdiv + c * NE);
#pragma unroll
  for (int s = 0; s < NE; ++s) fence(e[s]);
#pragma unroll
  for (int s = 0; s < NE; ++s) x[s] = AT(I.b.delsq_u, e[s], k);
  keep(x);
  settle(fr);
  slot_vals(fr, f);
  float acc = 0.0f;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) acc = acc + f[s] * x[s];
  if (k < I.n.L) AT(I.b.delsq_div, c, k) = acc;
}

// vorticity of delsq_u (laplacian_vorticity_v841_f32)
KB k_mix_lapvort(const Instance I) {
  const int v = wave_entity(I);
  if (v >= I.n.V) return;
  const int k = lane();
  int e[3];
  float x[3], f[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) e[s] = I.m.eov[3 * v + s];
  SlotRow<3> fr = slot_row<3>(I.m.mx_fvort + 3 * v);
#pragma unroll
  for (int s = 0; s < 3; ++s) fence(e[s]);
#pragma unroll
  for (int s = 0; s < 3; ++s) x[s] = AT(I.b.delsq_u, e[s], k);
  keep(x);
  settle(fr);
  slot_vals(fr, f);
  float acc = 0.0f;
#pragma unroll
  for (int s = 0; s < 3; ++s) acc = acc + f[s] * x[s];
  if (k < I.n.L) AT(I.b.delsq_vort, v, k) = acc;
}

// del4 of u (momentum_filter_lap4_v841_f32): pool - rho_e*((ddiv*r_dc) - (dvort*r_dv))
KB k_mix_edge4(const Instance I) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  const int c0 = I.m.coe[2 * e], c1 = I.m.coe[2 * e + 1];
  const int v0 = I.m.voe[2 * e], v1 = I.m.voe[2 * e + 1];
  const float d0 = AT(I.b.delsq_div, c0, k), d1 = AT(I.b.delsq_div, c1, k);
  const float z0 = AT(I.b.delsq_vort, v0, k), z1 = AT(I.b.delsq_vort, v1, k);
  const float r0 = AT(I.b.c_rho, c0, k), r1 = AT(I.b.c_rho, c1, k);
  const float eru = AT(I.b.euler_ru, e, k);
  const float rdc = uni_load(I.m.mx_rdc + e), rdv = uni_load(I.m.mx_rdv + e);
  const float rho_e = 0.5f * (r0 + r1);
  const float diffusion = rho_e * ((d1 - d0) * rdc - (z1 - z0) * rdv);
  if (k < I.n.L) AT(I.b.euler_ru, e, k) = eru - diffusion;
}

// del2 of theta and of w (theta_filter_lap2_v841_f32, w_filter_lap2_v841_f32).
// Without the fourth-order part (MIX4 false) the results go straight to the
// Euler pools.
template <int NE, bool MIX4>
KB k_mix_cell(const Instance I) {
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int L = I.n.L;
  const int n = I.m.neoc[c];
  int e[NE], o[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    e[s] = I.m.eoc[c * NE + s];
    o[s] = I.m.ocell[c * NE + s];
  }
  SlotRow<NE> sgr = slot_row<NE>(I.m.csign + c * NE), f2r = slot_row<NE>(I.m.mx_fth2 + c * NE),
                    fwr = slot_row<NE>(I.m.mx_fw2 + c * NE);
  const float thc = AT(I.b.theta, c, k), rhc = AT(I.b.c_rho, c, k), wc = AT(I.b.w, c, k);
  const float kdc = AT(I.b.kdiff, c, k);
  const float erw = MIX4 ? 0.0f : AT(I.b.euler_rw, c, k);
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    fence(e[s]);
    fence(o[s]);
  }
  float tho[NE], rho[NE], wo[NE], kdo[NE], d2[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    tho[s] = AT(I.b.theta, o[s], k);
    rho[s] = AT(I.b.c_rho, o[s], k);
    wo[s] = AT(I.b.w, o[s], k);
    kdo[s] = AT(I.b.kdiff, o[s], k);
    d2[s] = uni_load(I.m.mx_del2 + e[s]);
  }
  keep(tho, rho, wo, kdo, d2);
  settle(sgr, f2r, fwr);
  float sg[NE], f2[NE], fw[NE];
  slot_vals(sgr, sg);
  slot_vals(f2r, f2);
  slot_vals(fwr, fw);
  float lap = 0.0f, tend = 0.0f, lapw = 0.0f, tendw = 0.0f;
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    // cell0/cell1 of edge s in cellsOnEdge order
    const bool first = sg[s] > 0.0f;
    const float th0 = first ? thc : tho[s], th1 = first ? tho[s] : thc;
    const float r0 = first ? rhc : rho[s], r1 = first ? rho[s] : rhc;
    const float w0 = first ? wc : wo[s], w1 = first ? wo[s] : wc;
    const float k0 = first ? kdc : kdo[s], k1 = first ? kdo[s] : kdc;
    const float rho_e = 0.5f * (r0 + r1);
    const float rho_em = up1(rho_e);
    const float k0m = up1(k0), k1m = up1(k1);
    if (s < n) {
      const float flux = (f2[s] * (th1 - th0)) * rho_e;
      lap = lap + flux;
      tend = tend + ((flux * 0.5f) * (k0 + k1)) * d2[s];
      const float fluxw = (fw[s] * (rho_e + rho_em)) * (w1 - w0);
      lapw = lapw + fluxw;
      const float sum_k = ((k0 + k1) + k0m) + k1m;
      tendw = tendw + ((fluxw * d2[s]) * 0.25f) * sum_k;
    }
  }
  if (k == 0 || k >= L) {
    lapw = 0.0f;
    tendw = 0.0f;
  }
  if (k < L) {
    AT(I.b.delsq_theta, c, k) = lap;
    AT(I.b.euler_theta, c, k) = tend;
    AT(I.b.delsq_w, c, k) = lapw;
  }
  if (MIX4) {
    if (k <= L) AT(I.b.mix_tw, c, k) = tendw;
  } else if (k <= L) {
    AT(I.b.euler_rw, c, k) = erw + tendw;
  }
}

// del4 of theta and of w (theta_filter_lap4_v841_f32, w_filter_lap4_v841_f32), then the w pool
template <int NE>
KB k_mix_cell4(const Instance I) {
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int L = I.n.L;
  const int n = I.m.neoc[c];
  int o[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) o[s] = I.m.ocell[c * NE + s];
  SlotRow<NE> sgr = slot_row<NE>(I.m.csign + c * NE), f4r = slot_row<NE>(I.m.mx_f4 + c * NE);
  const float dtc = AT(I.b.delsq_theta, c, k), dwc = AT(I.b.delsq_w, c, k);
  float tend = AT(I.b.euler_theta, c, k);
  float tendw = AT(I.b.mix_tw, c, k);
  const float erw = AT(I.b.euler_rw, c, k);
#pragma unroll
  for (int s = 0; s < NE; ++s) fence(o[s]);
  float dto[NE], dwo[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    dto[s] = AT(I.b.delsq_theta, o[s], k);
    dwo[s] = AT(I.b.delsq_w, o[s], k);
  }
  keep(dto, dwo);
  settle(sgr, f4r);
  float sg[NE], f4[NE];
  slot_vals(sgr, sg);
  slot_vals(f4r, f4);
  const bool wlev = k >= 1 && k < L;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) {
      const bool first = sg[s] > 0.0f;
      const float t0 = first ? dtc : dto[s], t1 = first ? dto[s] : dtc;
      const float w0 = first ? dwc : dwo[s], w1 = first ? dwo[s] : dwc;
      tend = tend - f4[s] * (t1 - t0);
      if (wlev) tendw = tendw - f4[s] * (w1 - w0);
    }
  if (k < L) AT(I.b.euler_theta, c, k) = tend;
  if (k <= L) AT(I.b.euler_rw, c, k) = erw + tendw;
}

// ------------------------------------------------------------ subcycle start
// MIX: also the 2-D Smagorinsky kdiff of the cell (mix_kdiff_value's
// arithmetic); MOIST: moisture species present
template <int NE, bool MIX, bool MOIST>
KB k_sub_cell(const Instance I) {
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int L = I.n.L;
  const int n = I.m.neoc[c];
  int e[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) e[s] = I.m.eoc[c * NE + s];
  SlotRow<NE> sd = slot_row<NE>(I.m.csdv + c * NE);
  SlotRow<NE> c2 = {0.0f}, s2 = {0.0f}, cs = {0.0f};
  if (MIX) {
    c2 = slot_row<NE>(I.m.defc2 + c * NE);
    s2 = slot_row<NE>(I.m.defs2 + c * NE);
    cs = slot_row<NE>(I.m.defcs + c * NE);
  }
  const float ia = uni_load(I.m.invarea + c);
  const float rdzw = AT(I.z.rdzw, 0, k), rdzu = AT(I.z.rdzu, 0, k);
  const float fzm = AT(I.z.fzm, 0, k), fzp = AT(I.z.fzp, 0, k);
  const float rho_p = AT(I.b.s_rho_p, c, k);
  const float qt = MOIST ? AT(I.b.qtot, c, k) : 0.0f;
  const float rb = MOIST ? AT(I.z.rb, c, k) : 0.0f;
  const float cqw = MOIST ? AT(I.b.cqw, c, k) : 0.0f;
  const float rw_k = AT(I.b.c_rw, c, k);
  const float pp_k = AT(I.b.pp, c, k);
#pragma unroll
  for (int s = 0; s < NE; ++s) fence(e[s]);
  float xv[NE], un[NE], vn[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    xv[s] = AT(I.b.c_ru, e[s], k);
    un[s] = MIX ? AT(I.b.u, e[s], k) : 0.0f;
    vn[s] = MIX ? AT(I.b.v, e[s], k) : 0.0f;
  }
  const float prho = I.k.physics ? AT(I.b.phys_rho, c, k) : 0.0f;  // the buffer exists with physics only
  keep(xv);
  if (MIX) keep(un, vn);
  settle(sd, c2, s2, cs);
  float kd = 0.0f;
  if (MIX) {
    float dudx = 0.0f, dudy = 0.0f, dvdx = 0.0f, dvdy = 0.0f;
#pragma unroll
    for (int s = 0; s < NE; ++s)
      if (s < n) {
        dudx = (dudx + c2[s] * un[s]) - cs[s] * vn[s];
        dudy = (dudy + cs[s] * un[s]) - s2[s] * vn[s];
        dvdx = (dvdx + cs[s] * un[s]) + c2[s] * vn[s];
        dvdy = (dvdy + s2[s] * un[s]) + cs[s] * vn[s];
      }
    kd = kdiff_from_deformation(I, k, dudx, dudy, dvdx, dvdy);
  }
  float acc = 0.0f;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) acc = acc + sd[s] * xv[s];
  acc = acc * ia;
  const float g = I.k.gravity;
  const float dpdz = MOIST ? (-g) * (rb * qt + rho_p * (1.0f + qt)) : (-g) * rho_p;
  // 0 - x, not -x: a zero divergence gives +0 (density_tendency_f32)
  const float trho = (0.0f - acc) - rdzw * (lane_off(rw_k, k, 1) - rw_k);
  const float pp_m1 = lane_off(pp_k, k, -1);
  const float dpdzm = up1(dpdz);
  float erw = 0.0f;
  if (k >= 1 && k < L) {
    const float inner = rdzu * (pp_k - pp_m1) - (fzm * dpdz + fzp * dpdzm);
    erw = MOIST ? 0.0f - cqw * inner : 0.0f - inner;  // euler_w(_moist_v841)_f32
  }
  if (k < L) {
    const float tr = I.k.physics ? trho + prho : trho;
    AT(I.b.dpdz, c, k) = dpdz;
    AT(I.b.tend_rho, c, k) = tr;
    AT(I.b.hdiv, c, k) = acc;  // stage-1 mass-flux divergence of the current ru
    if (MIX) AT(I.b.kdiff, c, k) = kd;
  }
  if (k <= L) AT(I.b.euler_rw, c, k) = erw;
}

// MIX: also add the del2 u increment of the 2-D Smagorinsky mixing to the
// pool (mixing.py del2 u term, in the same pass; kdiff is final by then)
template <bool MIX>
KB k_sub_edge(const Instance I) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  const int c0 = I.m.coe[2 * e], c1 = I.m.coe[2 * e + 1];
  int v0 = 0, v1 = 0;
  if (MIX) {
    v0 = I.m.voe[2 * e];
    v1 = I.m.voe[2 * e + 1];
  }
  fence(c0);
  fence(c1);
  fence(v0);
  fence(v1);
  // every row first: the pool's, then (with mixing) del2 u's: divergence,
  // vorticity, kdiff and rho of the edge's cells and vertices
  const float *pp = I.b.pp, *zz = I.z.zz, *dz = I.b.dpdz;
  const float idc = uni_load(I.m.invdc + e);
  const float pp0 = AT(pp, c0, k), pp1 = AT(pp, c1, k), zz0 = AT(zz, c0, k), zz1 = AT(zz, c1, k);
  const float dz0 = AT(dz, c0, k), dz1 = AT(dz, c1, k), zxu = AT(I.z.zxu, e, k), cqu = AT(I.b.cqu, e, k);
  float d0 = 0.0f, d1 = 0.0f, z0 = 0.0f, z1 = 0.0f, kd0 = 0.0f, kd1 = 0.0f, r0 = 0.0f, r1 = 0.0f;
  float midc = 0.0f, midv = 0.0f, del2 = 0.0f;
  if (MIX) {
    d0 = AT(I.b.divergence, c0, k);
    d1 = AT(I.b.divergence, c1, k);
    z0 = AT(I.b.vorticity, v0, k);
    z1 = AT(I.b.vorticity, v1, k);
    kd0 = AT(I.b.kdiff, c0, k);
    kd1 = AT(I.b.kdiff, c1, k);
    r0 = AT(I.b.c_rho, c0, k);
    r1 = AT(I.b.c_rho, c1, k);
    midc = uni_load(I.m.mx_idc + e);
    midv = uni_load(I.m.mx_idv + e);
    del2 = uni_load(I.m.mx_del2 + e);
  }
  keep(pp0, pp1, zz0, zz1, dz0, dz1, zxu, cqu);
  if (MIX) keep(d0, d1, z0, z1, kd0, kd1, r0, r1);
  const float a = ((pp1 - pp0) * idc) / (0.5f * (zz1 + zz0));
  const float bb = (0.5f * zxu) * (dz0 + dz1);
  const float pool = (0.0f - cqu) * (a - bb);
  if (MIX) {
    // del2 u at level k: the increment added to the u pool, and delsq_u
    const float lap = (d1 - d0) * midc - (z1 - z0) * midv;
    const float kedge = 0.5f * (kd0 + kd1);
    const float rho_e = 0.5f * (r0 + r1);
    const float inc = ((rho_e * kedge) * lap) * del2;
    if (k < I.n.L) {
      if (I.k.mix4) AT(I.b.delsq_u, e, k) = lap;
      AT(I.b.euler_ru, e, k) = pool + inc;
    }
  } else {
    if (k < I.n.L) AT(I.b.euler_ru, e, k) = pool;
  }
}

// ---------------------------------------------------------- stage tendencies

// momentum tendency of one stage (dynamics.py:223-268, dynamics_v841.py:128-185);
// returns lane k's tend_ru value, stored for levels k < L
template <int NE>
__device__ __forceinline__ float tend_mom_edge(const Instance &I, int e, int k) {
  constexpr int NX = 2 * NE - 2;
  const int L = I.n.L;
  const int c0 = I.m.coe[2 * e], c1 = I.m.coe[2 * e + 1];
  int nb[NX];
#pragma unroll
  for (int s = 0; s < NX; ++s) nb[s] = I.m.eoe[e * NX + s];
  const int nn = I.m.neoe[e];
  SlotRow<NX> wtr = slot_row<NX>(I.m.woe + e * NX), car = slot_row<NX>(I.m.ecos + e * NX),
                    sar = slot_row<NX>(I.m.esin + e * NX);
  const float fe = uni_load(I.m.fedge + e), idc = uni_load(I.m.invdc + e);
  const float fzm = AT(I.z.fzm, 0, k), fzp = AT(I.z.fzp, 0, k), rdzw = AT(I.z.rdzw, 0, k);
  const float ui = AT(I.z.uinit, 0, k), vi = AT(I.z.vinit, 0, k);
  const float u_k = AT(I.b.u, e, k);
  const float rw0 = AT(I.b.c_rw, c0, k), rw1 = AT(I.b.c_rw, c1, k);
  const float pve = AT(I.b.pv_edge, e, k);
  float un[NX], pvn[NX];
#pragma unroll
  for (int s = 0; s < NX; ++s) {
    un[s] = AT(I.b.u, nb[s], k);
    pvn[s] = AT(I.b.pv_edge, nb[s], k);
  }
  const float rho0 = AT(I.b.c_rho, c0, k), rho1 = AT(I.b.c_rho, c1, k);
  const float ke0 = AT(I.b.ke, c0, k), ke1 = AT(I.b.ke, c1, k);
  const float hd0 = AT(I.b.hdiv, c0, k), hd1 = AT(I.b.hdiv, c1, k);
  const float eru = AT(I.b.euler_ru, e, k);
  const float pru = I.k.physics ? AT(I.b.phys_ru, e, k) : 0.0f;  // the buffer exists with physics only
  keep(un, pvn);
  keep(u_k, rw0, rw1, pve, rho0, rho1, ke0, ke1, hd0, hd1, eru, fe, idc);
  SlotRow<NX> wt = wtr, ca = car, sa = sar;
  settle(wt, ca, sa);
  const float u_m1 = lane_off(u_k, k, -1), u_m2 = lane_off(u_k, k, -2), u_p1 = lane_off(u_k, k, 1);
  const float rwe = 0.5f * (rw0 + rw1);
  float wduz = 0.0f;
  if (k == 1 || k == L - 1) {
    wduz = rwe * (fzm * u_k + fzp * u_m1);
  } else if (k >= 2 && k < L - 1) {
    wduz = flux3(u_m2, u_m1, u_k, u_p1, rwe, 1.0f);
  }
  const float wduz_p = dn1(wduz);
  const float initial = (-rdzw) * (wduz_p - wduz);
  float qq = 0.0f;
#pragma unroll
  for (int s = 0; s < NX; ++s)
    if (s < nn) {
      const float work_pv = 0.5f * (pve + pvn[s]);
      qq = qq + (wt[s] * un[s]) * work_pv;
    }
#pragma unroll
  for (int s = 0; s < NX; ++s)
    if (s < nn) {
      const float ref_u = ui * ca[s] - vi * sa[s];
      qq = qq - (wt[s] * ref_u) * fe;
    }
  const float rho_edge = 0.5f * (rho0 + rho1);
  const float horiz = rho_edge * (qq - (ke1 - ke0) * idc);
  float t = (initial + horiz) - (u_k * 0.5f) * (hd0 + hd1);
  // upper-level Rayleigh damping of u, before the Euler pool is added
  // (native/upper_damping.rs:18-31; cuda_upper_damping_v841.py:12-18)
  const int nray = I.k.ray_levels;
  if (k >= L - nray && k < L) t = t - (rho_edge * u_k) * ((float)(k - (L - nray) + 1) * I.k.ray_inverse);
  const float tr = I.k.physics ? (t + eru) + pru : t + eru;
  if (k < L) AT(I.b.tend_ru, e, k) = tr;
  return tr;
}

template <int NE>
KB k_tend_mom(const Instance I) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  tend_mom_edge<NE>(I, e, lane());
}

// the momentum tendency and the first acoustic substep's edge update
// (ru_p = dts * tend_ru, acoustic_v841.py:296-330) in one pass.  The running
// sum ru_avg equals ru_p after one substep; it is stored only when a later
// substep reads it back (with two substeps the second forms p1 + p2 from
// ru_p, with one the recovery reads ru_p itself).  ru_p itself is stored
// only for a second substep's edge update or, at stage 3, for the step's
// output: the first substep's column update and a one-substep stage's
// recovery form dts * tend_ru from tend_ru.
template <int NE>
KB k_tend_mom_ac(const Instance I, float dts, int store_avg, int store_rup) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  const float tr = tend_mom_edge<NE>(I, e, k);
  if (k < I.n.L) {
    const float p = dts * tr;
    if (store_rup) AT(I.b.ru_p, e, k) = p;
    if (store_avg) AT(I.b.ru_avg, e, k) = p;
  }
}

// theta and w edge fluxes of edge e with the 3rd-order upwind-blended
// stencil (driver.py:1003-1034, 1140-1170) and, at stages 2-3, the theta RK
// correction flux (cuda_dynamics_v841.py:209-247)
// CORR: stages 2 and 3, which add the theta RK correction flux
template <int NE>
struct EdgeFluxIn {
  float thn[2 * NE - 2], wn[2 * NE - 2];
  SlotRow<2 * NE - 2> av, a3v;
  float rue, sru, sth0, sth1, fzm, fzp, dv;
  int na;
  int c0, c1, ac[2 * NE - 2];  // scalar index loads
};
template <int NE, bool CORR>
__device__ __forceinline__ void tend_flux_index(const Instance &I, int e, int k, EdgeFluxIn<NE> &b) {
  constexpr int NX = 2 * NE - 2;
  b.c0 = I.m.coe[2 * e];
  b.c1 = I.m.coe[2 * e + 1];
#pragma unroll
  for (int s = 0; s < NX; ++s) b.ac[s] = I.m.advc[e * NX + s];
  b.na = I.m.nadv[e];
  b.av = slot_row<NX>(I.m.adv + e * NX);
  b.a3v = slot_row<NX>(I.m.adv3 + e * NX);
  b.dv = CORR ? uni_load(I.m.dv + e) : 0.0f;
  b.fzm = AT(I.z.fzm, 0, k);
  b.fzp = AT(I.z.fzp, 0, k);
}
template <int NE>
__device__ __forceinline__ void tend_flux_fence(const EdgeFluxIn<NE> &b) {
  fence(b.c0);
  fence(b.c1);
#pragma unroll
  for (int s = 0; s < 2 * NE - 2; ++s) fence(b.ac[s]);
}
template <int NE, bool CORR>
__device__ __forceinline__ void tend_flux_rows(const Instance &I, int e, int k, EdgeFluxIn<NE> &b) {
  constexpr int NX = 2 * NE - 2;
#pragma unroll
  for (int s = 0; s < NX; ++s) {
    b.thn[s] = AT(I.b.theta, b.ac[s], k);
    b.wn[s] = AT(I.b.w, b.ac[s], k);
  }
  b.rue = AT(I.b.c_ru, e, k);
  b.sru = b.sth0 = b.sth1 = 0.0f;
  if (CORR) {
    b.sru = AT(I.b.s_ru, e, k);
    b.sth0 = AT(I.b.s_theta, b.c0, k);
    b.sth1 = AT(I.b.s_theta, b.c1, k);
  }
}
template <int NE, bool CORR>
__device__ __forceinline__ EdgeFluxIn<NE> tend_flux_load(const Instance &I, int e, int k) {
  EdgeFluxIn<NE> b;
  tend_flux_index<NE, CORR>(I, e, k, b);
  tend_flux_fence(b);
  tend_flux_rows<NE, CORR>(I, e, k, b);
  settle(b.av, b.a3v);
  return b;
}
template <int NE, bool CORR>
__device__ __forceinline__ EdgeFlux tend_flux_compute(const Instance &I, int e, const EdgeFluxIn<NE> &b) {
  constexpr int NX = 2 * NE - 2;
  const float rue_m = up1(b.rue);
  const float coef = I.k.coef3;
  const float rui = b.fzm * b.rue + b.fzp * rue_m;
  const float sv_t = cs1(b.rue), sv_w = cs1(rui);
  float et = 0.0f, ew = 0.0f;
#pragma unroll
  for (int s = 0; s < NX; ++s)
    if (s < b.na) {
      const float a = b.av[s], a3 = b.a3v[s];
      et = et + (a + (coef * sv_t) * a3) * b.thn[s];
      ew = ew + (a + (coef * sv_w) * a3) * b.wn[s];
    }
  EdgeFlux f;
  f.th = b.rue * et;
  f.corr = CORR ? ((b.dv * (b.sru - b.rue)) * 0.5f) * (b.sth1 + b.sth0) : 0.0f;
  f.w = rui * ew;
  return f;
}
template <int NE, bool CORR>
__device__ __forceinline__ EdgeFlux tend_flux_values(const Instance &I, int e, int k) {
  return tend_flux_compute<NE, CORR>(I, e, tend_flux_load<NE, CORR>(I, e, k));
}
template <bool CORR>
__device__ __forceinline__ void store_edge_flux(const Instance &I, int e, int k, EdgeFlux f) {
  if (k < I.n.L) {
    AT(I.b.fth, e, k) = f.th;
    if (CORR) AT(I.b.fcorr, e, k) = f.corr;
    if (k >= 1) AT(I.b.fw, e, k) = f.w;
  }
}

template <int NE, bool CORR>
KB k_tend_flux(const Instance I) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  store_edge_flux<CORR>(I, e, k, tend_flux_values<NE, CORR>(I, e, k));
}

// CORR: stages 2 and 3 (the theta RK correction flux); MIX: horizontal
// mixing on (its theta increment is added)
template <int NE, bool CORR, bool MIX>
KB k_tend_cell(const Instance I) {
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int L = I.n.L;
  const float coef = I.k.coef3;
  const int n = I.m.neoc[c];
  int e[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) e[s] = I.m.eoc[c * NE + s];
  SlotRow<NE> sgr = slot_row<NE>(I.m.csign + c * NE);
  const float ia = uni_load(I.m.invarea + c);
  const float fzm = AT(I.z.fzm, 0, k), fzp = AT(I.z.fzp, 0, k);
  const float rdzw = AT(I.z.rdzw, 0, k), rdzu = AT(I.z.rdzu, 0, k);
  // the column rows, read once; vertical neighbours by shuffles
  const float th_k = AT(I.b.theta, c, k), ths_k = AT(I.b.s_theta, c, k);
  const float rwc_k = AT(I.b.c_rw, c, k), rws_k = AT(I.b.s_rw, c, k), ww_k = AT(I.b.w, c, k);
  const float zz_k = AT(I.z.zz, c, k), erw_k = AT(I.b.euler_rw, c, k);
  const float eth_k = MIX ? AT(I.b.euler_theta, c, k) : 0.0f;
  float fth[NE], fco[NE], fwv[NE], tru[NE], zb[NE], zb3[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    zb[s] = AT(I.z.zb, c * NE + s, k);
    zb3[s] = AT(I.z.zb3, c * NE + s, k);
  }
#pragma unroll
  for (int s = 0; s < NE; ++s) fence(e[s]);
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    fth[s] = AT(I.b.fth, e[s], k);
    fwv[s] = AT(I.b.fw, e[s], k);
    tru[s] = AT(I.b.tend_ru, e[s], k);
    fco[s] = CORR ? AT(I.b.fcorr, e[s], k) : 0.0f;
  }
  const float prt_k = I.k.physics ? AT(I.b.phys_rt, c, k) : 0.0f;  // the buffer exists with physics only
  keep(fth, fwv, tru, zb, zb3);
  if (CORR) keep(fco);
  SlotRow<NE> sgs = sgr;
  settle(sgs);
  float sg[NE];
  slot_vals(sgs, sg);
  const float th_m1 = lane_off(th_k, k, -1), th_m2 = lane_off(th_k, k, -2), th_p1 = lane_off(th_k, k, 1);
  const float ths_m1 = lane_off(ths_k, k, -1), rwc_m1 = lane_off(rwc_k, k, -1);
  const float ww_m1 = lane_off(ww_k, k, -1), ww_m2 = lane_off(ww_k, k, -2), ww_p1 = lane_off(ww_k, k, 1);
  const float zz_m1 = lane_off(zz_k, k, -1);
  float vf = 0.0f;
  if (k == 1) {
    vf = rwc_k * (fzm * th_k + fzp * th_m1);
    vf = vf + (rws_k - rwc_k) * (fzm * ths_k + fzp * ths_m1);
  } else if (k >= 2 && k < L - 1) {
    vf = flux3(th_m2, th_m1, th_k, th_p1, rwc_k, coef);
    vf = vf + (rws_k - rwc_k) * (fzm * ths_k + fzp * ths_m1);
  } else if (k == L - 1) {
    vf = rws_k * (fzm * th_k + fzp * th_m1);
  }
  const float vf_p = dn1(vf);
  float wf = 0.0f;
  if (k == 1 || k == L - 1) {
    wf = (0.25f * (rwc_k + rwc_m1)) * (ww_k + ww_m1);
  } else if (k >= 2 && k < L - 1) {
    wf = flux3(ww_m2, ww_m1, ww_k, ww_p1, 0.5f * (rwc_k + rwc_m1), 1.0f);
  }
  const float wf_p = dn1(wf);
  float r = 0.0f;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) r = r - sg[s] * fth[s];
  if (CORR) {
#pragma unroll
    for (int s = 0; s < NE; ++s)
      if (s < n) r = r - sg[s] * fco[s];
  }
  r = r * ia;
  const float trt = r - rdzw * (vf_p - vf);
  float tru_m[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) tru_m[s] = up1(tru[s]);
  float rw_t = 0.0f;
  if (k >= 1 && k < L) {
    float q = 0.0f;
#pragma unroll
    for (int s = 0; s < NE; ++s)
      if (s < n) q = q - sg[s] * fwv[s];
    q = q * ia;
    q = q - rdzu * (wf_p - wf);
    q = q + erw_k;
#pragma unroll
    for (int s = 0; s < NE; ++s)
      if (s < n) {
        const float tk = tru[s];
        const float flux = sg[s] * (fzm * tk + fzp * tru_m[s]);
        q = q - (zb[s] + cs1(tk) * zb3[s]) * flux;
      }
    rw_t = q * (fzm * zz_k + fzp * zz_m1);
  }
  if (k < L) {
    const float tmix = MIX ? trt + eth_k : trt;
    AT(I.b.tend_rt, c, k) = I.k.physics ? tmix + prt_k : tmix;
  }
  if (k <= L) AT(I.b.tend_rw, c, k) = rw_t;
}

// ----------------------------------------------------------- acoustic substep
// ru_p + divergence-damping increment of edge e at level k (mixing.py:851-945)
// OLD0: the damped substep was a stage's first, whose rtheta_pp_old is zero
// (then neither stored nor read; x - 0 keeps the bits of x)
template <bool OLD0 = false>
__device__ __forceinline__ float damped_ru_p(const Instance &I, int e, int k, float rp, float coefficient) {
  const int c0 = I.m.coe[2 * e], c1 = I.m.coe[2 * e + 1];
  const float d0 = AT(I.b.rtheta_pp, c0, k) - (OLD0 ? 0.0f : AT(I.b.rtheta_pp_old, c0, k));
  const float d1 = AT(I.b.rtheta_pp, c1, k) - (OLD0 ? 0.0f : AT(I.b.rtheta_pp_old, c1, k));
  const float den = AT(I.b.s_theta, c0, k) + AT(I.b.s_theta, c1, k);
  return rp + ((coefficient * (d0 - d1)) * (1.0f - 0.0f)) / den;
}

// FIRST: the first substep (ru_p = dts * tend_ru).  DAMP: apply the previous
// substep's divergence damping to ru_p first (the same update k_divdamp
// makes, fused into this pass over the edges).  AVGRP: this is the second of
// two substeps and ru_avg was not stored at the first, where it equals the
// undamped ru_p read here (p1 + p2, the same operands in the same order as
// ru_avg + p).  Loads first, stores last.
template <bool FIRST, bool DAMP, bool AVGRP>
KB k_ac_edge(const Instance I, float dts, float damp_coef) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  if (k >= I.n.L) return;
  if (FIRST) {
    const float p = dts * AT(I.b.tend_ru, e, k);
    AT(I.b.ru_p, e, k) = p;
    AT(I.b.ru_avg, e, k) = p;
    return;
  }
  const int c0 = I.m.coe[2 * e], c1 = I.m.coe[2 * e + 1];
  const float idc = uni_load(I.m.invdc + e);
  const float tr = AT(I.b.tend_ru, e, k);
  const float rp_raw = AT(I.b.ru_p, e, k);
  const float ra_in = AVGRP ? 0.0f : AT(I.b.ru_avg, e, k);
  const float cqu = AT(I.b.cqu, e, k), zxu = AT(I.z.zxu, e, k);
  fence(c0);
  fence(c1);
  const float tp0 = AT(I.b.rtheta_pp, c0, k), tp1 = AT(I.b.rtheta_pp, c1, k);
  const float z0 = AT(I.z.zz, c0, k), z1 = AT(I.z.zz, c1, k);
  const float x0 = AT(I.b.s_exner, c0, k), x1 = AT(I.b.s_exner, c1, k);
  const float q0 = AT(I.b.rho_pp, c0, k), q1 = AT(I.b.rho_pp, c1, k);
  float to0 = 0.0f, to1 = 0.0f, st0 = 0.0f, st1 = 0.0f;
  if (DAMP) {
    if (!AVGRP) {  // after a stage's first substep (AVGRP: the second of two) rtheta_pp_old is zero
      to0 = AT(I.b.rtheta_pp_old, c0, k);
      to1 = AT(I.b.rtheta_pp_old, c1, k);
    }
    st0 = AT(I.b.s_theta, c0, k);
    st1 = AT(I.b.s_theta, c1, k);
  }
  const float ra_old = AVGRP ? rp_raw : ra_in;
  // ru_p + the divergence-damping increment (damped_ru_p's arithmetic)
  const float rp_in = DAMP ? rp_raw + ((damp_coef * ((tp0 - to0) - (tp1 - to1))) * (1.0f - 0.0f)) / (st0 + st1)
                           : rp_raw;
  float pg = ((tp1 - tp0) * idc) / (0.5f * (z1 + z0));
  pg = (((cqu * 0.5f) * I.k.c2) * (x0 + x1)) * pg;
  pg = pg + ((0.5f * zxu) * I.k.gravity) * (q0 + q1);
  const float p = rp_in + dts * (tr - pg);
  AT(I.b.ru_p, e, k) = p;
  AT(I.b.ru_avg, e, k) = ra_old + p;
}

// FIRST: the first substep of a stage (the perturbation fields start at
// zero).  Every global load is issued before the column sweep.
template <int NE, bool FIRST>
KBCOL(COLW_AC) k_ac_cell(const Instance I, float dts, int store_old) {
  constexpr int COLW = COLW_AC;
  static_assert(COLW >= 7, "waves 0-6 stage the level rows");
  __shared__ float s_x[COLW * COLP], s_d[COLW * COLP], s_l[COLW * COLP], s_g[COLW * COLP];
  __shared__ float s_lev[7 * 64];
  const int w = (int)(threadIdx.x >> 6);
  const int c0 = wave_entity<COLW>();
  const bool live = c0 < I.n.C;
  const int c = live ? c0 : 0;  // every wave reaches the barriers
  const int k = lane();
  const int L = I.n.L;
  const int n = I.m.neoc[c];
  int e[NE], o[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    e[s] = I.m.eoc[c * NE + s];
    o[s] = I.m.ocell[c * NE + s];  // the other cell of edge s
  }
  SlotRow<NE> sgr = slot_row<NE>(I.m.csign + c * NE), sdr = slot_row<NE>(I.m.csdv + c * NE);
  const float ia = uni_load(I.m.invarea + c);
  // wavefront w < 7 stages level row w for the block
  const float *const lev[7] = {I.z.rdzw, I.z.ewm, I.z.ewp, I.z.etp, I.z.etm, I.z.fzm, I.z.fzp};
  float lv = 0.0f;
  if (w < 7) lv = AT(lev[w], 0, k);
  const float sthc = AT(I.b.s_theta, c, k);
  float rpp = 0.0f, rtpp = 0.0f, rwp = 0.0f, wwa = 0.0f;
  if (!FIRST) {
    rpp = AT(I.b.rho_pp, c, k);
    rtpp = AT(I.b.rtheta_pp, c, k);
    rwp = AT(I.b.rw_p, c, k);
    wwa = AT(I.b.ww_avg, c, k);
  }
  const float coftz = AT(I.b.coftz, c, k), cofwt = AT(I.b.cofwt, c, k);
  const float cofwz = AT(I.b.cofwz, c, k);
  const float trho = AT(I.b.tend_rho, c, k), trt = AT(I.b.tend_rt, c, k), trw = AT(I.b.tend_rw, c, k);
  const float zk = AT(I.z.zz, c, k);
  const float al = AT(I.b.alpha, c, k), ga = AT(I.b.gamma, c, k);
  const float srw = AT(I.b.s_rw, c, k), crw = AT(I.b.c_rw, c, k), damp = AT(I.z.dss, c, k);
  const float crho = AT(I.b.c_rho, c, k), wk = AT(I.b.w, c, k);
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    fence(e[s]);
    fence(o[s]);
  }
  float rup[NE], sto[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    rup[s] = FIRST ? AT(I.b.tend_ru, e[s], k) : AT(I.b.ru_p, e[s], k);
    sto[s] = AT(I.b.s_theta, o[s], k);
  }
  keep(rup, sto);
  if (w < 7) s_lev[w * 64 + k] = lv;
  settle(sgr, sdr);
  float sg[NE], sd[NE];
  slot_vals(sgr, sg);
  slot_vals(sdr, sd);
  __syncthreads();
  const float rdzw = s_lev[0 * 64 + k], ewm = s_lev[1 * 64 + k], ewp = s_lev[2 * 64 + k];
  const float etp = s_lev[3 * 64 + k], etm = s_lev[4 * 64 + k], fzm = s_lev[5 * 64 + k], fzp = s_lev[6 * 64 + k];
  const float rtpp_old = rtpp;
  float r = 0.0f, t = 0.0f;
#pragma unroll
  for (int s = 0; s < NE; ++s)
    if (s < n) {
      // the first substep's ru_p is dts * tend_ru (the momentum pass's product)
      const float rp = FIRST ? dts * rup[s] : rup[s];
      const float flux = ((dts * sd[s]) * rp) * ia;  // (csign*dts)*dv, csign = +-1
      // theta at the edge: (theta[cell1] + theta[cell0]) in that order
      const bool first_cell = sg[s] > 0.0f;
      const float th0 = first_cell ? sthc : sto[s];
      const float th1 = first_cell ? sto[s] : sthc;
      r = r - flux;
      t = t - (flux * 0.5f) * (th1 + th0);
    }
  const float coftz_p = dn1(coftz);
  const float ewm_p = lane_off(ewm, k, 1), ewp_p = lane_off(ewp, k, 1);
  const float rwp_p = dn1(rwp);
  const float rs = ((rpp + dts * trho) + r) - (dts * rdzw) * (ewm_p * rwp_p - ewm * rwp);
  const float ts = ((rtpp + dts * trt) + t) - (dts * rdzw) * ((ewm_p * coftz_p) * rwp_p - (ewm * coftz) * rwp);
  if (k >= 1 && k < L) wwa = wwa + ewm * rwp;
  const float ts_m = up1(ts), rs_m = up1(rs), rtpp_m = up1(rtpp), rpp_m = up1(rpp), zk_m = up1(zk);
  const float cofwt_m = up1(cofwt);
  const float etp_m = lane_off(etp, k, -1), etm_m = lane_off(etm, k, -1);
  const float crho_m = up1(crho);
  // cofwr and the sub-diagonal a of the implicit w system at levels 1..L-1,
  // formed from zz, the level rows and the stored coefficients of levels k
  // and k-1 by the same operations as in k_vcoef (which stores cofwz, coftz
  // and cofwt for this kernel, and a only for export)
  const float rdzw_m = lane_off(rdzw, k, -1), ewp_m = lane_off(ewp, k, -1), coftz_m = up1(coftz);
  float cofwr = 0.0f, a = 0.0f;
  if (k >= 1 && k < L) {
    cofwr = (0.5f * I.k.gravity) * (fzm * zk + fzp * zk_m);
    a = (((-cofwz) * coftz_m) * rdzw_m) * zk_m + cofwr * rdzw_m - (cofwt_m * coftz_m) * rdzw_m;
    a = (a * etp_m) * ewp_m;
  }
  float col = rwp;
  if (k >= 1 && k < L) {
    float x = trw - cofwz * (((etp * zk) * ts - (etp_m * zk_m) * ts_m) + ((etm * zk) * rtpp - (etm_m * zk_m) * rtpp_m));
    x = x - cofwr * ((etp * rs + etp_m * rs_m) + (etm * rpp + etm_m * rpp_m));
    x = x + cofwt * (etp * ts + etm * rtpp);
    x = x + cofwt_m * (etp_m * ts_m + etm_m * rtpp_m);
    col = col + dts * x;
  }
  const float dts2 = dts * dts;
  const float ad = dts2 * a;
  // Thomas sweeps (acoustic_v841.py:296-565), lane = cell: forward for
  // levels 1..L-1, back substitution for levels L-1..0
  s_x[w * COLP + k] = col;
  s_d[w * COLP + k] = ad;
  s_l[w * COLP + k] = al;
  s_g[w * COLP + k] = ga;
  __syncthreads();
  if (w == 0 && k < COLW) {
    float *x = s_x + k * COLP;
    const float *pd = s_d + k * COLP, *pl = s_l + k * COLP, *pg = s_g + k * COLP;
    float p = x[0];
    for (int j = 1; j < L; ++j) {
      p = (x[j] - pd[j] * p) * pl[j];
      x[j] = p;
    }
    p = x[L];
    for (int j = L - 1; j >= 0; --j) {
      p = x[j] - pg[j] * p;
      x[j] = p;
    }
  }
  __syncthreads();
  col = s_x[w * COLP + k];
  if (k >= 1 && k < L) {
    const float delta = srw - crw;
    const float zi = fzm * zk + fzp * zk_m;
    const float ri = fzm * crho + fzp * crho_m;
    col = (col + delta - (((dts * damp) * zi) * ri) * wk) / (1.0f + dts * damp) - delta;
  }
  if (k >= 1 && k < L) wwa = wwa + ewp * col;
  const float col_p = dn1(col);
  const float rpp_new = rs - (dts * rdzw) * (ewp_p * col_p - ewp * col);
  const float rtpp_new = ts - (dts * rdzw) * ((ewp_p * coftz_p) * col_p - (ewp * coftz) * col);
  if (live && k <= L) {
    AT(I.b.rw_p, c, k) = col;
    AT(I.b.ww_avg, c, k) = wwa;
  }
  if (live && k < L) {
    AT(I.b.rho_pp, c, k) = rpp_new;
    AT(I.b.rtheta_pp, c, k) = rtpp_new;
    if (store_old) AT(I.b.rtheta_pp_old, c, k) = rtpp_old;
  }
}

KB k_divdamp(const Instance I, float coefficient) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  if (k >= I.n.L) return;
  AT(I.b.ru_p, e, k) = damped_ru_p(I, e, k, AT(I.b.ru_p, e, k), coefficient);
}

// --------------------------------------------------------------------- recovery
// w from rw and the terrain-following horizontal flux (terrain.py:491-571).
// ru_row(e) returns lane k's value of the edge momentum row of edge e.
template <int NE, class RU>
__device__ __forceinline__ float recover_w_column(const Instance &I, int c, int k, float rho_k, float rw_k, RU ru_row,
                                                  float *hdiv_out = nullptr) {
  const int L = I.n.L;
  const int n = I.m.neoc[c];
  int e[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) e[s] = I.m.eoc[c * NE + s];
  SlotRow<NE> sgr = slot_row<NE>(I.m.csign + c * NE);
  SlotRow<NE> sdr = slot_row<NE>(I.m.csdv + c * NE);
  const float ia = uni_load(I.m.invarea + c);
  const float zk = AT(I.z.zz, c, k);
  const float fzm = AT(I.z.fzm, 0, k), fzp = AT(I.z.fzp, 0, k);
  float ru[NE], zb[NE], zb3[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    zb[s] = AT(I.z.zb, c * NE + s, k);
    zb3[s] = AT(I.z.zb3, c * NE + s, k);
  }
#pragma unroll
  for (int s = 0; s < NE; ++s) fence(e[s]);
#pragma unroll
  for (int s = 0; s < NE; ++s) ru[s] = ru_row(e[s]);
  keep(ru, zb, zb3);
  SlotRow<NE> sgs = sgr, sd = sdr;
  settle(sgs, sd);
  float sg[NE];
  slot_vals(sgs, sg);
  const float zkm = lane_off(zk, k, -1);
  if (hdiv_out) {
    // mass-flux divergence of the recovered ru (dynamics.py:86-121), next stage's hdiv
    float acc = 0.0f;
#pragma unroll
    for (int s = 0; s < NE; ++s)
      if (s < n) acc = acc + sd[s] * ru[s];
    *hdiv_out = acc * ia;
  }
  float col = 0.0f;
  if (k >= 1 && k < L) col = rw_k / (fzm * zk + fzp * zkm);
  const float cf1 = I.z.cf1, cf2 = I.z.cf2, cf3 = I.z.cf3;
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    const float ru_m = up1(ru[s]);
    const float ru0 = rdl(ru[s], 0), ru1 = rdl(ru[s], 1), ru2 = rdl(ru[s], 2);
    if (s < n) {
      if (k == 0) {
        const float flux = (cf1 * ru0 + cf2 * ru1) + cf3 * ru2;
        col = col + (sg[s] * (zb[s] + cs1(flux) * zb3[s])) * flux;
      } else if (k < L) {
        const float flux = fzm * ru[s] + fzp * ru_m;
        col = col + (sg[s] * (zb[s] + cs1(flux) * zb3[s])) * flux;
      }
    }
  }
  const float r0 = rdl(rho_k, 0), r1 = rdl(rho_k, 1), r2 = rdl(rho_k, 2);
  const float rho_m = up1(rho_k);
  if (k == 0) {
    const float bottom = (cf1 * r0 + cf2 * r1) + cf3 * r2;
    col = col / bottom;
  } else if (k < L) {
    col = col / (fzm * rho_k + fzp * rho_m);
  } else {
    col = 0.0f;
  }
  return col;
}

// S3: the stage-3 recovery (exner, p' and the split-transport sums).  The
// loads and the terrain w recovery come first, branch-free, so every gather
// of the pass is in flight together; stores come last.
template <int NE, bool S3>
KB k_rec_cell(const Instance I, int sub, float inv_steps, int store_rtp) {
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int L = I.n.L;
  const bool lay = k < L, itf = k <= L, inner = k >= 1 && k < L;
  const float srp = AT(I.b.s_rho_p, c, k), rpp = AT(I.b.rho_pp, c, k);
  const float srtp = AT(I.b.s_rtheta_p, c, k), rtpp = AT(I.b.rtheta_pp, c, k);
  const float rb = AT(I.z.rb, c, k), rtb = AT(I.z.rtb, c, k);
  const float srw = AT(I.b.s_rw, c, k), rwp = AT(I.b.rw_p, c, k), wwa = AT(I.b.ww_avg, c, k);
  float zk = 0.0f, exb = 0.0f, spw = 0.0f;
  if (S3) {
    zk = AT(I.z.zz, c, k);
    exb = AT(I.z.exb, c, k);
    spw = AT(I.b.split_w, c, k);
  }
  const float dp = lay ? srp + rpp : 0.0f;
  const float rp = lay ? srtp + rtpp : 0.0f;
  const float rho = lay ? dp + rb : 1.0f;
  const float rtv = lay ? rp + rtb : 1.0f;
  const float rwv = inner ? srw + rwp : 0.0f;
  const float *cru = I.b.c_ru;
  float hd = 0.0f;
  // hdiv feeds the next stage's momentum tendency; after stage 3 the next
  // subcycle's opening pass forms it again
  const float wv = recover_w_column<NE>(I, c, k, rho, rwv, [&](int e) { return AT(cru, e, k); }, S3 ? nullptr : &hd);
  int bad = (lay && (!(rho > 0.0f) || !(rtv > 0.0f))) ? 1 : 0;
  const float th = rtv / rho;
  float ex = 0.0f, ppv = 0.0f;
  if (S3 && lay) {
    const float arg = (zk * I.k.cfac) * rtv;
    if (!(arg > 0.0f)) bad = 1;
    ex = exner_pow(I, arg);
    ppv = (zk * I.k.rgas) * (ex * rp + rtb * (ex - exb));
  }
  const float fw = srw + wwa * inv_steps;
  float split = 0.0f;
  if (S3) {
    // split-transport sums: copy, then accumulate; the last subcycle also
    // applies the 1/splits average (integration.py:27-82)
    const float acc = sub == 1 ? fw : fw + spw;
    split = sub == I.k.splits ? acc * I.k.split_inv : acc;
  }
  if (lay) {
    // rho_p and rtheta_m are read only as the next subcycle's start values
    // (stage 3) or by an export between phases
    if (S3 || I.k.store_flux) {
      AT(I.b.rho_p, c, k) = dp;
      AT(I.b.c_rt, c, k) = rtv;
    }
    if (store_rtp) AT(I.b.rtheta_p, c, k) = rp;
    AT(I.b.c_rho, c, k) = rho;
    AT(I.b.theta, c, k) = th;
    if (S3) {
      AT(I.b.exner, c, k) = ex;
      AT(I.b.pp, c, k) = ppv;
    }
    if (!S3) AT(I.b.hdiv, c, k) = hd;
  }
  if (itf) {
    AT(I.b.c_rw, c, k) = rwv;
    if (I.k.store_flux) AT(I.b.flux_w, c, k) = fw;
    if (S3) AT(I.b.split_w, c, k) = split;
    AT(I.b.w, c, k) = wv;
  }
  if (bad) set_flag(I.b.flag);
}

// DAMP: apply the last substep's divergence damping to ru_p first.
// AVGRP: one substep and ru_avg not stored, so the average is the undamped
// ru_p of that substep.  S3: accumulate the split-transport sums.  Loads
// first, stores last.
template <bool S3, bool DAMP, bool AVGRP>
KB k_rec_edge(const Instance I, int sub, float inv_steps, float damp_coef, float dts) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  if (k >= I.n.L) return;
  const int c0 = I.m.coe[2 * e], c1 = I.m.coe[2 * e + 1];
  const float sru = AT(I.b.s_ru, e, k);
  // one substep: its ru_p is dts * tend_ru, formed here as the momentum pass formed it
  const float rp_ld = AVGRP ? AT(I.b.tend_ru, e, k) : AT(I.b.ru_p, e, k);
  const float avg_ld = AVGRP ? 0.0f : AT(I.b.ru_avg, e, k);
  const float spu = S3 ? AT(I.b.split_u, e, k) : 0.0f;
  const float r0 = (AT(I.b.s_rho_p, c0, k) + AT(I.b.rho_pp, c0, k)) + AT(I.z.rb, c0, k);
  const float r1 = (AT(I.b.s_rho_p, c1, k) + AT(I.b.rho_pp, c1, k)) + AT(I.z.rb, c1, k);
  const float rp_raw = AVGRP ? dts * rp_ld : rp_ld;
  const float avg = AVGRP ? rp_raw : avg_ld;
  const float rp = DAMP ? damped_ru_p<AVGRP>(I, e, k, rp_raw, damp_coef) : rp_raw;
  const float ru = sru + rp;
  const float fu = sru + avg * inv_steps;
  float split = 0.0f;
  if (S3) {
    const float acc = sub == 1 ? fu : fu + spu;
    split = sub == I.k.splits ? acc * I.k.split_inv : acc;
  }
  const float uu = (2.0f * ru) / (r0 + r1);
  // the damped ru_p is read again only as the step's output (stage 3) or by
  // an export between phases; the next stage forms its own
  if (DAMP && (S3 || I.k.store_flux)) AT(I.b.ru_p, e, k) = rp;
  AT(I.b.c_ru, e, k) = ru;
  if (I.k.store_flux) AT(I.b.flux_u, e, k) = fu;
  if (S3) AT(I.b.split_u, e, k) = split;
  AT(I.b.u, e, k) = uu;
}

// rebuild of the saved diagnostics from a prognostic state (driver.py:2655-2700)
template <int NE>
KB k_rebuild_cell(const Instance I) {
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int L = I.n.L;
  // every load before any store (the terrain w recovery gathers among them)
  const float rho_k = AT(I.b.rho, c, k), rt = AT(I.b.rt, c, k), zk = AT(I.z.zz, c, k);
  const float rtb = AT(I.z.rtb, c, k), rb = AT(I.z.rb, c, k), exb = AT(I.z.exb, c, k);
  const float rwk = AT(I.b.rw, c, k);
  const float rho = k < L ? rho_k : 1.0f;
  const float *ru = I.b.ru;
  const float wv = recover_w_column<NE>(I, c, k, rho, rwk, [&](int e) { return AT(ru, e, k); });
  if (k < L) {
    const float ex = exner_pow(I, (zk * I.k.cfac) * rt);
    const float rtp = rt - rtb;
    AT(I.b.theta, c, k) = rt / rho;
    AT(I.b.exner, c, k) = ex;
    AT(I.b.rho_p, c, k) = rho - rb;
    AT(I.b.rtheta_p, c, k) = rtp;
    AT(I.b.pp, c, k) = (zk * I.k.rgas) * (ex * rtp + rtb * (ex - exb));
  }
  if (k <= L) AT(I.b.w, c, k) = wv;
}

KB k_rebuild_edge(const Instance I) {
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  if (k >= I.n.L) return;
  const int c0 = I.m.coe[2 * e], c1 = I.m.coe[2 * e + 1];
  AT(I.b.u, e, k) = (2.0f * AT(I.b.ru, e, k)) / (AT(I.b.rho, c0, k) + AT(I.b.rho, c1, k));
}

// ------------------------------------------------------------ scalar transport
// q0..q3: the column's values at levels k-2, k-1, k, k+1
__device__ __forceinline__ float vflux(float q0, float q1, float q2, float q3, float wk, int k, int L, float fzm,
                                       float fzp, float coef) {
  if (k == 1 || k == L - 1) return wk * (fzm * q2 + fzp * q1);
  const float base = div12(wk * (7.0f * (q2 + q1) - (q3 + q0)));
  const float corr = div12((coef * fabsf(wk)) * ((q3 - q0) - 3.0f * (q2 - q1)));
  return base + corr;
}

// Scalar kernels run one wavefront per entity over all species: the
// connectivity, stencil weights and the species-independent rows are read
// once, and the species are taken in batches of G whose rows are all
// requested before any of them is combined (each batch is one memory round
// trip; a batch's stores come after its loads, and the next batch's loads
// after those stores, since the kernels' input and output arrays may alias).
// G trades round trips against registers per wavefront.
#ifndef HEXDYN_G_SC_EDGE
#define HEXDYN_G_SC_EDGE 3
#endif
#ifndef HEXDYN_G_SC_CELL
#define HEXDYN_G_SC_CELL 3
#endif
#ifndef HEXDYN_G_FCT_EDGE1
#define HEXDYN_G_FCT_EDGE1 6  // every species in one batch: 152 us against 158 us in batches of 3 (x1.40962)
#endif
#ifndef HEXDYN_G_FCT_CELL1
#define HEXDYN_G_FCT_CELL1 2
#endif
#ifndef HEXDYN_G_FCT_EDGE2
#define HEXDYN_G_FCT_EDGE2 6
#endif
#ifndef HEXDYN_G_FCT_CELL2
#define HEXDYN_G_FCT_CELL2 3
#endif
__device__ __forceinline__ int clamp_sp(int sp, int nt) { return sp < nt ? sp : nt - 1; }

template <int NE>
KB k_sc_edge(const Instance I, const float *__restrict__ qs, float *__restrict__ ev) {
  constexpr int NX = 2 * NE - 2, G = HEXDYN_G_SC_EDGE;
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  const int NT = I.n.NT;
  int ac[NX];
#pragma unroll
  for (int s = 0; s < NX; ++s) ac[s] = I.m.advc[e * NX + s];
  const int na = I.m.nadv[e];
  SlotRow<NX> av = slot_row<NX>(I.m.adv + e * NX), a3v = slot_row<NX>(I.m.adv3 + e * NX);
  const float uu = AT(I.b.split_u, e, k);
#pragma unroll
  for (int s = 0; s < NX; ++s) fence(ac[s]);
  const float coef = I.k.coef3;
  for (int sp0 = 0; sp0 < NT; sp0 += G) {
    float v[G][NX];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int r0 = clamp_sp(sp0 + g, NT) * I.n.C;  // species rows start at s * C
#pragma unroll
      for (int s = 0; s < NX; ++s) v[g][s] = AT(qs, r0 + ac[s], k);
    }
    // the weights are formed after the batch's loads are issued
    float uk = uu;
    keep(uk);
    keep(v);
    SlotRow<NX> avs = av, a3s = a3v;
    settle(avs, a3s);
    const float sv = cs1(uk);
    float wt[NX];
#pragma unroll
    for (int s = 0; s < NX; ++s) wt[s] = avs[s] + (coef * sv) * a3s[s];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (sp0 + g >= NT) break;
      float r = 0.0f;
#pragma unroll
      for (int s = 0; s < NX; ++s)
        if (s < na) r = r + wt[s] * v[g][s];
      if (k < I.n.L) AT(ev, (sp0 + g) * I.n.E + e, k) = r;
    }
  }
}

template <int NE>
KB k_sc_cell(const Instance I, const float *__restrict__ qs, const float *__restrict__ ev, float *__restrict__ out,
             float dt, float wn, const float *__restrict__ src) {
  constexpr int G = HEXDYN_G_SC_CELL;
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int L = I.n.L, NT = I.n.NT;
  const int n = I.m.neoc[c];
  int e[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) e[s] = I.m.eoc[c * NE + s];
  SlotRow<NE> sgr = slot_row<NE>(I.m.csign + c * NE);
  const float ia = uni_load(I.m.invarea + c);
  const float sw = AT(I.b.split_w, c, k);
  const float orho = AT(I.b.rho, c, k), crho = AT(I.b.c_rho, c, k);
  const float fzm = AT(I.z.fzm, 0, k), fzp = AT(I.z.fzp, 0, k), rdzw = AT(I.z.rdzw, 0, k);
#pragma unroll
  for (int s = 0; s < NE; ++s) fence(e[s]);
  float uh[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) uh[s] = AT(I.b.split_u, e[s], k);
  for (int sp0 = 0; sp0 < NT; sp0 += G) {
    float qk[G], qo[G], sr[G], evv[G][NE];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int sq = clamp_sp(sp0 + g, NT);
      qk[g] = AT(qs, sq * I.n.C + c, k);
      qo[g] = AT(I.b.q, sq * I.n.C + c, k);
#pragma unroll
      for (int s = 0; s < NE; ++s) evv[g][s] = AT(ev, sq * I.n.E + e[s], k);
    }
    // held scalar tendency (transport.py:1800-1808)
#pragma unroll
    for (int g = 0; g < G; ++g) sr[g] = src ? AT(src, clamp_sp(sp0 + g, NT) * I.n.C + c, k) : 0.0f;
    keep(qk, qo, evv);
    keep(uh);
    SlotRow<NE> sgs = sgr;
    settle(sgs);
    float sg[NE];
    slot_vals(sgs, sg);
    const float target = (1.0f - wn) * orho + wn * crho;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (sp0 + g >= NT) break;
      const float q_m2 = lane_off(qk[g], k, -2), q_m1 = lane_off(qk[g], k, -1), q_p1 = lane_off(qk[g], k, 1);
      float vf = 0.0f;
      if (k >= 1 && k < L) vf = vflux(q_m2, q_m1, qk[g], q_p1, sw, k, L, fzm, fzp, I.k.coef3);
      const float vf_p = dn1(vf);
      float tc = 0.0f;
#pragma unroll
      for (int s = 0; s < NE; ++s)
        if (s < n) tc = tc + ((-sg[s]) * uh[s]) * evv[g][s];
      tc = tc * ia + sr[g];
      if (k < L) {
        const float num = qo[g] * orho + dt * (tc - rdzw * (vf_p - vf));
        AT(out, (sp0 + g) * I.n.C + c, k) = num * (1.0f / target);
      }
    }
  }
}

// old mixing ratio with the held scalar tendency, old + dt*source/rho_old
// (transport.py:1906), every species of cell c
KB k_fct_source(const Instance I, float dt) {
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  if (k >= I.n.L) return;
  const float rho = AT(I.b.rho, c, k);
  const int NT = I.n.NT;
  for (int s0 = 0; s0 < NT; s0 += HX_MOIST_G) {  // every row of a batch loaded before its stores
    float q[HX_MOIST_G], f[HX_MOIST_G];
#pragma unroll
    for (int g = 0; g < HX_MOIST_G; ++g) {
      const int r = (s0 + g < NT ? s0 + g : NT - 1) * I.n.C + c;
      q[g] = AT(I.b.q, r, k);
      f[g] = AT(I.b.phys_q, r, k);
    }
    keep(q, f);
#pragma unroll
    for (int g = 0; g < HX_MOIST_G; ++g)
      if (s0 + g < NT) AT(I.b.q_src, (s0 + g) * I.n.C + c, k) = q[g] + (dt * f[g]) / rho;
  }
}


template <int NE>
KB k_fct_edge1(const Instance I, const float *__restrict__ qs, float dt, const float *__restrict__ qold, float qadd) {
  constexpr int NX = 2 * NE - 2, G = HEXDYN_G_FCT_EDGE1;
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  const int NT = I.n.NT;
  int ac[NX];
#pragma unroll
  for (int s = 0; s < NX; ++s) ac[s] = I.m.advc[e * NX + s];
  const int c1 = I.m.coe[2 * e], c2 = I.m.coe[2 * e + 1];
  const int na = I.m.nadv[e];
  SlotRow<NX> av = slot_row<NX>(I.m.adv + e * NX), a3v = slot_row<NX>(I.m.adv3 + e * NX);
  const float dve = uni_load(I.m.dv + e);
  const float uu = AT(I.b.split_u, e, k);
#pragma unroll
  for (int s = 0; s < NX; ++s) fence(ac[s]);
  fence(c1);
  fence(c2);
  const float coef = I.k.coef3;
  const float dvdt = dve * dt;
  for (int sp0 = 0; sp0 < NT; sp0 += G) {
    float v[G][NX], q1[G], q2[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int p = clamp_sp(sp0 + g, NT) * I.n.C;
#pragma unroll
      for (int s = 0; s < NX; ++s) v[g][s] = AT(qs, p + ac[s], k);
      q1[g] = AT(qold, p + c1, k);
      q2[g] = AT(qold, p + c2, k);
    }
    float uk = uu;
    keep(uk);
    keep(v);
    SlotRow<NX> avs = av, a3s = a3v;
    settle(avs, a3s);
    const float sv = cs1(uk);
    float wt[NX];
#pragma unroll
    for (int s = 0; s < NX; ++s) wt[s] = avs[s] + (coef * sv) * a3s[s];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (sp0 + g >= NT) break;
      const float qa1 = q1[g] + qadd, qa2 = q2[g] + qadd;
      float ra;
      if (na == 10) {
        // ten-cell stencils associate from the first product (transport.py n_adv == 10 branch)
        float xa = wt[0] * v[g][0];
#pragma unroll
        for (int s = 1; s < NX; ++s)
          if (s < 10) xa = xa + wt[s] * v[g][s];
        ra = uk * xa;
      } else {
        ra = 0.0f;
#pragma unroll
        for (int s = 0; s < NX; ++s)
          if (s < na) ra = ra + (uk * wt[s]) * v[g][s];
      }
      const float upa = dvdt * (max0(uk) * qa1 + min0(uk) * qa2);  // formed again by k_fct_cell1
      if (k < I.n.L) AT(I.b.fct_hres, (sp0 + g) * I.n.E + e, k) = dt * ra - upa;
    }
  }
}

// The upwind flux of each edge, dv*dt*(max(u,0)*q[cell0] + min(u,0)*q[cell1])
// (k_fct_edge1 forms the same value for the high-order residual), is formed
// here from the edge's split u and the old mixing ratios of its two cells,
// which the bounds read anyway: one u row per edge instead of one flux row
// per edge and species.  PAIR: cellsOnCell slot s is the other cell of
// edgesOnCell slot s in every cell (as in MPAS meshes); otherwise that cell's
// row is read separately.
template <int NE, bool PAIR>
KB k_fct_cell1(const Instance I, const float *__restrict__ qs, float dt, const float *__restrict__ qold, float qadd) {
  constexpr int G = HEXDYN_G_FCT_CELL1, NP = PAIR ? 1 : NE;
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int L = I.n.L, NT = I.n.NT;
  const int n = I.m.neoc[c];
  int nb[NE], e[NE], oc[NP];
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    nb[s] = I.m.coc[c * NE + s];
    e[s] = I.m.eoc[c * NE + s];
    if (!PAIR) oc[s < NP ? s : 0] = I.m.ocell[c * NE + s];
  }
  SlotRow<NE> sgr = slot_row<NE>(I.m.csign + c * NE), sdr = slot_row<NE>(I.m.csdv + c * NE);
  const float ia = uni_load(I.m.invarea + c);
  const float orho = AT(I.b.rho, c, k);
  const float wk = AT(I.b.split_w, c, k);
  const float rdzw = AT(I.z.rdzw, 0, k), fzm = AT(I.z.fzm, 0, k), fzp = AT(I.z.fzp, 0, k);
#pragma unroll
  for (int s = 0; s < NE; ++s) {
    fence(nb[s]);
    fence(e[s]);
  }
  if (!PAIR)
#pragma unroll
    for (int s = 0; s < NP; ++s) fence(oc[s]);
  float uh[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) uh[s] = AT(I.b.split_u, e[s], k);
  for (int sp0 = 0; sp0 < NT; sp0 += G) {
    float qsk[G], qol[G], qn[G][NE], qp[G][NP], hres[G][NE];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int sq = clamp_sp(sp0 + g, NT);
      const int po = sq * I.n.C;
      qsk[g] = AT(qs, po + c, k);
#pragma unroll
      for (int s = 0; s < NE; ++s) {
        qn[g][s] = AT(qold, po + nb[s], k);
        hres[g][s] = AT(I.b.fct_hres, sq * I.n.E + e[s], k);
      }
      if (!PAIR)
#pragma unroll
        for (int s = 0; s < NP; ++s) qp[g][s] = AT(qold, po + oc[s], k);
      qol[g] = AT(qold, po + c, k);
    }
    keep(qsk, qol, qn, hres, uh);
    if (!PAIR) keep(qp);
    SlotRow<NE> sgs = sgr, sds = sdr;
    settle(sgs, sds);
    float sg[NE], sd[NE];
    slot_vals(sgs, sg);
    slot_vals(sds, sd);
    // the target density (transport.py:1964-1989): the old density plus dt
    // times the split mass-flux convergence, from rows this pass holds; the
    // first batch stores it for the second pass and checks it
    float hd = 0.0f;
#pragma unroll
    for (int s = 0; s < NE; ++s)
      if (s < n) hd = hd + (-(uh[s] * sd[s])) * ia;  // ((-csign)*uh)*dv, csign = +-1
    const float wk_p = lane_off(wk, k, 1);
    const float target = orho + dt * (hd - rdzw * (wk_p - wk));
    if (sp0 == 0 && k < L) {
      AT(I.b.target, c, k) = target;
      if (!(target > 0.0f)) set_flag(I.b.flag);
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (sp0 + g >= NT) break;
      const int sp = sp0 + g;
      const float qs_k = qsk[g];
      const float qs_m2 = lane_off(qs_k, k, -2), qs_m1 = lane_off(qs_k, k, -1), qs_p1 = lane_off(qs_k, k, 1);
      const float qo = qol[g] + qadd;
      const float qo_m = up1(qo), qo_p = dn1(qo);
      float a, b;
      if (k == 0) {
        a = qo; b = qo;
        a = pmin(a, qo_p); b = pmax(b, qo_p);
      } else if (k == L - 1) {
        a = qo_m; b = qo_m;
        a = pmin(a, qo); b = pmax(b, qo);
      } else {
        a = qo_m; b = qo_m;
        a = pmin(a, qo); b = pmax(b, qo);
        a = pmin(a, qo_p); b = pmax(b, qo_p);
      }
#pragma unroll
      for (int s = 0; s < NE; ++s)
        if (s < n) {
          const float q = qn[g][s] + qadd;
          a = pmin(a, q);
          b = pmax(b, q);
        }
      float mass = qo * orho;
      float vres = 0.0f;
      float fcol = 0.0f;
      if (k >= 1 && k < L) {
        vres = dt * vflux(qs_m2, qs_m1, qs_k, qs_p1, wk, k, L, fzm, fzp, I.k.coef3);
        fcol = dt * (max0(wk) * qo_m + min0(wk) * qo);
      } else {
        vres = dt * 0.0f;
      }
      const float fcol_p = dn1(fcol);
      if (k <= L - 2) mass = mass - fcol_p * rdzw;
      if (k >= 1 && k < L) {
        mass = mass + fcol * rdzw;
        vres = vres - fcol;
      }
      const float vres_p = dn1(vres);
      float sin_ = (-rdzw) * (min0(vres_p) - max0(vres));
      float sout = (-rdzw) * (max0(vres_p) - min0(vres));
#pragma unroll
      for (int s = 0; s < NE; ++s)
        if (s < n) {
          const float so = -sg[s];
          // edge s's cells in edge order: this one first where csign = +1
          const float qx = (PAIR ? qn[g][s] : qp[g][s < NP ? s : 0]) + qadd;
          const bool first = sg[s] > 0.0f;
          const float qa1 = first ? qo : qx, qa2 = first ? qx : qo;
          const float dvdt = __builtin_fabsf(sd[s]) * dt;  // |csign * dv| = dv
          const float upa = dvdt * (max0(uh[s]) * qa1 + min0(uh[s]) * qa2);
          mass = mass + (so * upa) * ia;
          const float sr = (-so) * hres[g][s];
          sout = sout - max0(sr) * ia;
          sin_ = sin_ - min0(sr) * ia;
        }
      float f = (b * target - mass) / (sin_ + 1.0e-20f);
      sin_ = pmin(1.0f, max0(f));
      f = (a * target - mass) / (sout - 1.0e-20f);
      sout = pmin(1.0f, max0(f));
      if (k < L) {
        AT(I.b.fct_mass, sp * I.n.C + c, k) = mass;
        AT(I.b.fct_sin, sp * I.n.C + c, k) = sin_;
        AT(I.b.fct_sout, sp * I.n.C + c, k) = sout;
      }
      if (k <= L) AT(I.b.fct_vres, sp * I.n.C + c, k) = vres;
    }
  }
}

KB k_fct_edge2(const Instance I) {
  constexpr int G = HEXDYN_G_FCT_EDGE2;
  const int e = wave_entity(I);
  if (e >= I.n.E) return;
  const int k = lane();
  if (k >= I.n.L) return;
  const int NT = I.n.NT;
  const int c1 = I.m.coe[2 * e], c2 = I.m.coe[2 * e + 1];
  for (int sp0 = 0; sp0 < NT; sp0 += G) {
    float f[G], so1[G], si2[G], si1[G], so2[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int sq = clamp_sp(sp0 + g, NT);
      const int ps = sq * I.n.C;
      f[g] = AT(I.b.fct_hres, sq * I.n.E + e, k);
      so1[g] = AT(I.b.fct_sout, ps + c1, k);
      si2[g] = AT(I.b.fct_sin, ps + c2, k);
      si1[g] = AT(I.b.fct_sin, ps + c1, k);
      so2[g] = AT(I.b.fct_sout, ps + c2, k);
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (sp0 + g >= NT) break;
      AT(I.b.fct_hres, (sp0 + g) * I.n.E + e, k) =
          max0(f[g]) * pmin(so1[g], si2[g]) + min0(f[g]) * pmin(si1[g], so2[g]);
    }
  }
}

template <int NE>
KB k_fct_cell2(const Instance I, float *__restrict__ out) {
  constexpr int G = HEXDYN_G_FCT_CELL2;
  const int c = wave_entity(I);
  if (c >= I.n.C) return;
  const int k = lane();
  const int L = I.n.L, NT = I.n.NT;
  const int n = I.m.neoc[c];
  int e[NE];
#pragma unroll
  for (int s = 0; s < NE; ++s) e[s] = I.m.eoc[c * NE + s];
  SlotRow<NE> sgr = slot_row<NE>(I.m.csign + c * NE);
  const float ia = uni_load(I.m.invarea + c);
  const float rdzw = AT(I.z.rdzw, 0, k);
  const float target = AT(I.b.target, c, k);
#pragma unroll
  for (int s = 0; s < NE; ++s) fence(e[s]);
  for (int sp0 = 0; sp0 < NT; sp0 += G) {
    float hres[G][NE], si[G], so[G], vr[G], ms[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int sq = clamp_sp(sp0 + g, NT);
#pragma unroll
      for (int s = 0; s < NE; ++s) hres[g][s] = AT(I.b.fct_hres, sq * I.n.E + e[s], k);
      si[g] = AT(I.b.fct_sin, sq * I.n.C + c, k);
      so[g] = AT(I.b.fct_sout, sq * I.n.C + c, k);
      vr[g] = AT(I.b.fct_vres, sq * I.n.C + c, k);
      ms[g] = AT(I.b.fct_mass, sq * I.n.C + c, k);
    }
    keep(hres, si, so, vr, ms);
    SlotRow<NE> sgs = sgr;
    settle(sgs);
    float sg[NE];
    slot_vals(sgs, sg);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (sp0 + g >= NT) break;
      const float si_m = up1(si[g]), so_m = up1(so[g]);
      float v = vr[g];
      if (k >= 1 && k < L) v = max0(v) * pmin(so_m, si[g]) + min0(v) * pmin(so[g], si_m);
      const float v_p = dn1(v);
      float mass = ms[g];
#pragma unroll
      for (int s = 0; s < NE; ++s)
        if (s < n) {
          const float sgn = -sg[s];
          mass = mass + (sgn * hres[g][s]) * ia;
        }
      mass = mass - rdzw * (v_p - v);
      if (k < L) AT(out, (sp0 + g) * I.n.C + c, k) = max0(mass / target);
    }
  }
}

// q < 0 -> 0 as a where-assignment (-0.0 and NaN keep their bits); src may be dst
KB k_clamp(const Instance I, const float *src, float *dst, int rows) {
  const int r = wave_entity(I);
  if (r >= rows) return;
  const int k = lane();
  if (k >= I.n.RS) return;
  const float x = AT(src, r, k);
  AT(dst, r, k) = x < 0.0f ? 0.0f : x;
}

// Wavefronts per block of a gather kernel: the handle's launch table when the
// first kernel argument is the Instance, else 4.
template <class T>
inline int wpb_from(const char *, const T &) {
  return 4;
}
inline int wpb_from(const char *name, const Instance &I) { return I.tune ? I.tune->waves_per_block(name) : 4; }
template <class A0, class... A>
inline int launch_wpb(const char *name, const A0 &a0, const A &...) {
  return wpb_from(name, a0);
}

template <class T>
inline void set_wpb(T &, int) {}
inline void set_wpb(Instance &I, int wpb) { I.n.WPB = wpb; }

template <typename K, typename... A>
inline void launch_named(const char *name, K kernel, int n, int gy, hipStream_t s, A... args) {
  if (n <= 0 || gy <= 0) return;
  const int wpb = launch_wpb(name, args...);
  (set_wpb(args, wpb), ...);  // the launch's own copy of the Instance carries its tile
  if (g_prof_on) prof_mark(name, s, 1);
  hipLaunchKernelGGL(kernel, dim3(nblocks(n, wpb), gy), dim3(64 * wpb), 0, s, args...);
  if (g_prof_on) prof_mark(name, s, 0);
}
#define launch(kernel, ...) launch_named(#kernel, kernel, __VA_ARGS__)

// column kernels: W cells per block of 64*W threads
template <int W, class K, class... A>
inline void launch_col_named(const char *name, K kernel, int n, hipStream_t s, A... args) {
  if (n <= 0) return;
  if (g_prof_on) prof_mark(name, s, 1);
  hipLaunchKernelGGL(kernel, dim3((n + W - 1) / W), dim3(64 * W), 0, s, args...);
  if (g_prof_on) prof_mark(name, s, 0);
}

inline int check() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : -1;
}

}  // namespace

// ------------------------------------------------------------------ phases
// NE: maxEdges of the mesh (6 or 7); kernels are instantiated for both.
#define NEDISPATCH(I, call6, call7) \
  do {                              \
    if ((I).n.ME == 6) {            \
      call6;                        \
    } else {                        \
      call7;                        \
    }                               \
  } while (0)
#define LAUNCH_NE(I, kernel, ...) \
  NEDISPATCH(I, launch_named(#kernel, kernel<6>, __VA_ARGS__), launch_named(#kernel, kernel<7>, __VA_ARGS__))

// The edge-tiled flux pass runs unless the mesh has a tile with more than
// HX_TILE_U distinct stencil cells or the caller set a launch shape for the
// gather form (k_diag_edge_flux) explicitly.
inline bool tiled_flux_on(const Instance &I) {
  return I.n.TU > 0 && !(I.tune && I.tune->find("k_diag_edge_flux") >= 0);
}

int solve_diagnostics(const Instance &I, int rk, hipStream_t s) {
  // divergence and vorticity are read only by the horizontal mixing, from
  // the diagnostics that open a subcycle (rk 3, and the step's first), or by
  // an export between phases
  const int store_mix = rk == 3 || I.k.store_flux ? 1 : 0;
  launch(k_diag_vertex, I.n.V, 1, s, I, (const float *)I.b.u, store_mix);
  LAUNCH_NE(I, k_diag_cell, I.n.C, 1, s, I, (const float *)I.b.u, store_mix);
  const float *u = I.b.u;
  const bool six = I.n.ME == 6;
  if (I.k.fused && tiled_flux_on(I)) {
    const int ntiles = (I.n.E + HX_TILE_E - 1) / HX_TILE_E;
    if (g_prof_on) prof_mark("k_diag_edge_flux_t", s, 1);
    auto kt = rk == 3 ? (six ? k_diag_edge_flux_t<6, true> : k_diag_edge_flux_t<7, true>)
                      : (six ? k_diag_edge_flux_t<6, false> : k_diag_edge_flux_t<7, false>);
    hipLaunchKernelGGL(kt, dim3(ntiles), dim3(64 * HX_TILE_E), 0, s, I, u);
    if (g_prof_on) prof_mark("k_diag_edge_flux_t", s, 0);
  } else if (I.k.fused) {  // diagnostics after stage rk feed stage rk % 3 + 1 (the next subcycle's first after rk 3)
    if (rk == 3)
      launch_named("k_diag_edge_flux", six ? k_diag_edge_flux<6, true> : k_diag_edge_flux<7, true>, I.n.E, 1, s, I,
                   u);
    else
      launch_named("k_diag_edge_flux", six ? k_diag_edge_flux<6, false> : k_diag_edge_flux<7, false>, I.n.E, 1, s,
                   I, u);
  } else {
    if (rk == 3)
      launch_named("k_diag_edge", six ? k_diag_edge<6, true> : k_diag_edge<7, true>, I.n.E, 1, s, I, u);
    else
      launch_named("k_diag_edge", six ? k_diag_edge<6, false> : k_diag_edge<7, false>, I.n.E, 1, s, I, u);
  }
  return check();
}

int begin_step(const Instance &I, hipStream_t s) {
  if (I.n.NM > 0) {
    launch(k_moist_cell, I.n.C, 1, s, I);
    launch(k_moist_edge, I.n.E, 1, s, I);
  }
  return solve_diagnostics(I, 3, s);
}

// The side stream (or s itself): ordered after everything s has queued.
inline hipStream_t fork_side(const Side *side, hipStream_t s, int &rc) {
  if (!side) return s;
  if (hipEventRecord(side->fork, s) != hipSuccess || hipStreamWaitEvent(side->s, side->fork, 0) != hipSuccess) rc = -1;
  return side->s;
}
// s waits for what the side stream has queued
inline void join_side(const Side *side, hipStream_t s, int &rc) {
  if (!side) return;
  if (hipEventRecord(side->join, side->s) != hipSuccess || hipStreamWaitEvent(s, side->join, 0) != hipSuccess) rc = -1;
}

int begin_subcycle(const Instance &I, int sub, hipStream_t s, const Side *side) {
  (void)sub;
  // the implicit coefficients read only the subcycle-start state, the
  // moisture and reference rows: beside the pools and the mixing
  int rc = 0;
  launch_col_named<COLW_VC>("k_vcoef", k_vcoef, I.n.C, fork_side(side, s, rc), I, I.k.dts[0],
                            (const float *)I.b.s_exner, (const float *)I.b.s_theta, (const float *)I.b.s_rtheta_p);
  // with mixing, kdiff rides on the cell pass and the del2 u increment on the
  // edge pass (the pool is written once)
  if (I.k.mixing) {
    if (I.n.NM > 0)
      NEDISPATCH(I, launch_named("k_sub_cell", k_sub_cell<6, true, true>, I.n.C, 1, s, I),
                 launch_named("k_sub_cell", k_sub_cell<7, true, true>, I.n.C, 1, s, I));
    else
      NEDISPATCH(I, launch_named("k_sub_cell", k_sub_cell<6, true, false>, I.n.C, 1, s, I),
                 launch_named("k_sub_cell", k_sub_cell<7, true, false>, I.n.C, 1, s, I));
    launch_named("k_sub_edge", k_sub_edge<true>, I.n.E, 1, s, I);
  } else {
    if (I.n.NM > 0)
      NEDISPATCH(I, launch_named("k_sub_cell", k_sub_cell<6, false, true>, I.n.C, 1, s, I),
                 launch_named("k_sub_cell", k_sub_cell<7, false, true>, I.n.C, 1, s, I));
    else
      NEDISPATCH(I, launch_named("k_sub_cell", k_sub_cell<6, false, false>, I.n.C, 1, s, I),
                 launch_named("k_sub_cell", k_sub_cell<7, false, false>, I.n.C, 1, s, I));
    launch_named("k_sub_edge", k_sub_edge<false>, I.n.E, 1, s, I);
  }
  if (I.k.mixing) {
    if (I.k.mix4) {
      LAUNCH_NE(I, k_mix_lapdiv, I.n.C, 1, s, I);
      launch(k_mix_lapvort, I.n.V, 1, s, I);
      launch(k_mix_edge4, I.n.E, 1, s, I);
    }
    if (I.k.mix4)
      NEDISPATCH(I, launch_named("k_mix_cell", k_mix_cell<6, true>, I.n.C, 1, s, I),
                 launch_named("k_mix_cell", k_mix_cell<7, true>, I.n.C, 1, s, I));
    else
      NEDISPATCH(I, launch_named("k_mix_cell", k_mix_cell<6, false>, I.n.C, 1, s, I),
                 launch_named("k_mix_cell", k_mix_cell<7, false>, I.n.C, 1, s, I));
    if (I.k.mix4) LAUNCH_NE(I, k_mix_cell4, I.n.C, 1, s, I);
  }
  join_side(side, s, rc);
  return rc ? rc : check();
}

int dyn_tend(const Instance &I, int stage, hipStream_t s, const Side *side) {
  // stage 2's implicit coefficients beside the stage's tendencies (which
  // neither read nor write them); joined before the acoustic substeps
  int rc = 0;
  const bool vc = stage == 2 && I.k.order == 3;
  if (vc)
    launch_col_named<COLW_VC>("k_vcoef", k_vcoef, I.n.C, fork_side(side, s, rc), I, I.k.dts[1],
                              (const float *)I.b.s_exner, (const float *)I.b.theta, (const float *)I.b.rtheta_p);
  if (I.k.fused) {
    // the edge fluxes came with the preceding diagnostics
    LAUNCH_NE(I, k_tend_mom_ac, I.n.E, 1, s, I, I.k.dts[stage - 1], I.k.steps[stage - 1] > 2 ? 1 : 0,
              I.k.steps[stage - 1] > 1 || stage == 3 ? 1 : 0);
  } else {
    LAUNCH_NE(I, k_tend_mom, I.n.E, 1, s, I);
    if (stage > 1)
      NEDISPATCH(I, launch_named("k_tend_flux", k_tend_flux<6, true>, I.n.E, 1, s, I),
                 launch_named("k_tend_flux", k_tend_flux<7, true>, I.n.E, 1, s, I));
    else
      NEDISPATCH(I, launch_named("k_tend_flux", k_tend_flux<6, false>, I.n.E, 1, s, I),
                 launch_named("k_tend_flux", k_tend_flux<7, false>, I.n.E, 1, s, I));
  }
  {
    const bool corr = stage > 1, mix = I.k.mixing != 0;
#define TEND_CELL(NE, CO, MX) launch_named("k_tend_cell", k_tend_cell<NE, CO, MX>, I.n.C, 1, s, I)
#define TEND_CELL_NE(NE)                           \
  (corr ? (mix ? TEND_CELL(NE, true, true) : TEND_CELL(NE, true, false)) \
        : (mix ? TEND_CELL(NE, false, true) : TEND_CELL(NE, false, false)))
    NEDISPATCH(I, TEND_CELL_NE(6), TEND_CELL_NE(7));
#undef TEND_CELL_NE
#undef TEND_CELL
  }
  if (vc) join_side(side, s, rc);
  return rc ? rc : check();
}

int acoustic_substep(const Instance &I, int stage, int ss, int damp_prev, hipStream_t s) {
  const float dts = I.k.dts[stage - 1];
  const int n = I.k.steps[stage - 1];
  if (!(I.k.fused && ss == 1))  // fused: the first edge update ran with the momentum tendency
  {
    const float coef = I.k.dts_inv_coef[stage - 1];
    const bool avgrp = I.k.fused && n == 2;
    if (ss == 1)
      launch_named("k_ac_edge", k_ac_edge<true, false, false>, I.n.E, 1, s, I, dts, coef);
    else if (damp_prev && avgrp)
      launch_named("k_ac_edge", k_ac_edge<false, true, true>, I.n.E, 1, s, I, dts, coef);
    else if (damp_prev)
      launch_named("k_ac_edge", k_ac_edge<false, true, false>, I.n.E, 1, s, I, dts, coef);
    else if (avgrp)
      launch_named("k_ac_edge", k_ac_edge<false, false, true>, I.n.E, 1, s, I, dts, coef);
    else
      launch_named("k_ac_edge", k_ac_edge<false, false, false>, I.n.E, 1, s, I, dts, coef);
  }
  // a stage's first substep leaves rtheta_pp_old zero; on the step path the
  // damping after it takes that zero without a store and a reload (the
  // second of two substeps' edge update, or a one-substep stage's recovery)
  const int store_old = I.k.fused && ss == 1 && n <= 2 ? 0 : 1;
  if (ss == 1)
    NEDISPATCH(I, launch_col_named<COLW_AC>("k_ac_cell", k_ac_cell<6, true>, I.n.C, s, I, dts, store_old),
               launch_col_named<COLW_AC>("k_ac_cell", k_ac_cell<7, true>, I.n.C, s, I, dts, store_old));
  else
    NEDISPATCH(I, launch_col_named<COLW_AC>("k_ac_cell", k_ac_cell<6, false>, I.n.C, s, I, dts, store_old),
               launch_col_named<COLW_AC>("k_ac_cell", k_ac_cell<7, false>, I.n.C, s, I, dts, store_old));
  return check();
}

int divergence_damping(const Instance &I, int stage, hipStream_t s) {
  launch(k_divdamp, I.n.E, 1, s, I, I.k.dts_inv_coef[stage - 1]);
  return check();
}

template <bool S3>
void recover_edges(const Instance &I, int sub, float inv, bool damp, bool avgrp, float coef, float dts,
                   hipStream_t s) {
  if (damp && avgrp)
    launch_named("k_rec_edge", k_rec_edge<S3, true, true>, I.n.E, 1, s, I, sub, inv, coef, dts);
  else if (damp)
    launch_named("k_rec_edge", k_rec_edge<S3, true, false>, I.n.E, 1, s, I, sub, inv, coef, dts);
  else if (avgrp)
    launch_named("k_rec_edge", k_rec_edge<S3, false, true>, I.n.E, 1, s, I, sub, inv, coef, dts);
  else
    launch_named("k_rec_edge", k_rec_edge<S3, false, false>, I.n.E, 1, s, I, sub, inv, coef, dts);
}

int recover(const Instance &I, int stage, int sub, int damp, hipStream_t s) {
  const float inv = I.k.inv_steps[stage - 1];
  const bool avgrp = I.k.fused && I.k.steps[stage - 1] == 1;
  const float coef = I.k.dts_inv_coef[stage - 1];
  // rtheta_p is read by the implicit coefficients of stage 2 (from stage 1)
  // and as the next subcycle's start (from stage 3)
  const int store_rtp = stage == 3 || (stage == 1 && I.k.order == 3) || I.k.store_flux ? 1 : 0;
  if (stage == 3) {
    recover_edges<true>(I, sub, inv, damp != 0, avgrp, coef, I.k.dts[stage - 1], s);
    NEDISPATCH(I, launch_named("k_rec_cell", k_rec_cell<6, true>, I.n.C, 1, s, I, sub, inv, store_rtp),
               launch_named("k_rec_cell", k_rec_cell<7, true>, I.n.C, 1, s, I, sub, inv, store_rtp));
  } else {
    recover_edges<false>(I, sub, inv, damp != 0, avgrp, coef, I.k.dts[stage - 1], s);
    NEDISPATCH(I, launch_named("k_rec_cell", k_rec_cell<6, false>, I.n.C, 1, s, I, sub, inv, store_rtp),
               launch_named("k_rec_cell", k_rec_cell<7, false>, I.n.C, 1, s, I, sub, inv, store_rtp));
  }
  return check();
}

int scalars_stage(const Instance &I, int stage, hipStream_t s) {
  const float dt = I.k.sca_dt[stage - 1];
  const float *qin = stage == 1 ? I.b.q : I.b.q_stage;  // stage 1 starts from the published scalars
  if (stage < 3 || !I.k.monotonic) {
    LAUNCH_NE(I, k_sc_edge, I.n.E, 1, s, I, qin, I.b.ev);
    LAUNCH_NE(I, k_sc_cell, I.n.C, 1, s, I, qin, (const float *)I.b.ev, I.b.q_next, dt,
              I.k.wn[stage - 1], (const float *)(I.k.physics_q ? I.b.phys_q : nullptr));
    return check();
  }
  // the limiter's old mixing ratio: the published one plus +0, or the one
  // the held scalar tendency moved (x + -0 == x keeps its bits)
  const float *qold = I.b.q;
  float qadd = 0.0f;
  if (I.k.physics_q) {
    launch(k_fct_source, I.n.C, 1, s, I, dt);
    qold = I.b.q_src;
    qadd = -0.0f;
  }
  LAUNCH_NE(I, k_fct_edge1, I.n.E, 1, s, I, qin, dt, qold, qadd);
  if (I.n.OCP)
    NEDISPATCH(I, launch_named("k_fct_cell1", k_fct_cell1<6, true>, I.n.C, 1, s, I, qin, dt, qold, qadd),
               launch_named("k_fct_cell1", k_fct_cell1<7, true>, I.n.C, 1, s, I, qin, dt, qold, qadd));
  else
    NEDISPATCH(I, launch_named("k_fct_cell1", k_fct_cell1<6, false>, I.n.C, 1, s, I, qin, dt, qold, qadd),
               launch_named("k_fct_cell1", k_fct_cell1<7, false>, I.n.C, 1, s, I, qin, dt, qold, qadd));
  launch(k_fct_edge2, I.n.E, 1, s, I);
  LAUNCH_NE(I, k_fct_cell2, I.n.C, 1, s, I, I.b.q_next);
  return check();
}

int rebuild_saved(const Instance &I, hipStream_t s) {
  LAUNCH_NE(I, k_rebuild_cell, I.n.C, 1, s, I);
  launch(k_rebuild_edge, I.n.E, 1, s, I);
  return check();
}

int rebuild_cells(const Instance &I, hipStream_t s) {
  LAUNCH_NE(I, k_rebuild_cell, I.n.C, 1, s, I);
  return check();
}

int clamp_scalars(const Instance &I, const float *src, float *dst, hipStream_t s) {
  launch(k_clamp, I.n.NT * I.n.C, 1, s, I, src, dst, I.n.NT * I.n.C);
  return check();
}

}  // namespace HEXDYN_NS
}  // namespace hexdyn
