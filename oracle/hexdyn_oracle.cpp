/*
 * TEST INFRASTRUCTURE — CPU oracle for the hexdyn dycore large step.
 *
 * A scalar C++ restatement of the reference's MPAS-A v8.4.2 (v8.4.1 lane)
 * dry/moist split-explicit step, in the reference's operation order, for
 * T = float (strict binary32: build with -ffp-contract=off, no fast-math) and
 * T = double (the FP64 reference the fast numerics are measured against).
 * Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product never does.
 *
 * Layout: MPAS logical order, level first, entity fastest ([L][C], [L][E],
 * [L+1][C], [T][L][C]).  Loops follow the reference text:
 *   step order              driver.py:3314-3605 (DryDycoreDriver.step)
 *   subcycle order          driver.py:2710-3313 (_advance_dynamics_subcycle)
 *   RK schedule             integration.py:110-160
 *   solve diagnostics       diagnostics.py:95-306, cuda_horizontal_v841.py:67-200
 *   mass flux divergence    dynamics.py:86-121
 *   density tendency        dynamics.py:124-148
 *   pressure gradient Euler dynamics.py:184-220
 *   Euler w                 driver.py:2790-2808
 *   dpdz (dry / moist)      driver.py:2777-2789
 *   vertical u transport    dynamics.py:223-268
 *   momentum (v8.4.1)       dynamics_v841.py:128-185
 *   theta tendency          driver.py:933-1079 (stored-inverse branch)
 *   w transport             driver.py:1081-1207 (stored-inverse branch)
 *   w -> omega tendency     acoustic.py:246-281
 *   vertical coefficients   acoustic_v841.py:81-233
 *   acoustic substep        acoustic_v841.py:296-565
 *   Thomas sweep            acoustic_v841.py:236-258
 *   divergence damping      mixing.py:840-945
 *   candidate / recovery    driver.py:3176-3290, driver.py:819-876 (exner, p'),
 *                           terrain.py:491-571 (u, w recovery)
 *   split flux              integration.py:27-82
 *   moist coefficients      regional_v841.py:1180-1219
 *   scalar transport        transport.py:1649-1826 (stages 1-2),
 *                           transport.py:1829-2175 (monotonic FCT, stage 3)
 *   post-transport clamp    regional_v841.py:1004 (putmask q<0 -> 0)
 * Parity status: see the test suite.  The exner power function is the host
 * C library's powf: the GNU powf, equal to the reference's own transcription
 * of it (native_powf.py) on every argument in [2^-12, 4) by its golden
 * vectors.  The mesh-scaling power is binary64 pow rounded once, the
 * reference's rkind_libm.powf_rkind.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../include/hexdyn.h"
#include "../gpuwm-hex_amd/csrc/hexdyn_powf.h"

namespace {

const double kGravity = 9.80616;
const double kRgas = 287.0;
const double kCp = 1004.5;
const double kP0 = 100000.0;

template <class T> static inline T pmax0(T x) { return x > T(0) ? x : T(0); }   // Python max(0, x)
template <class T> static inline T pmin0(T x) { return x < T(0) ? x : T(0); }   // Python min(0, x)
template <class T> static inline T pmin(T a, T b) { return b < a ? b : a; }     // Python min(a, b)
template <class T> static inline T pmax(T a, T b) { return b > a ? b : a; }     // Python max(a, b)
template <class T> static inline T copysign1(T x) { return std::signbit(x) ? T(-1) : T(1); }

template <class T> static inline T power(T x, T y);
// binary32: the host C library's powf, the GNU powf the reference's v8.4.1
// exner and native MPAS use (cuda_driver.py:1466-1472)
template <> inline float power<float>(float x, float y) { return ::powf(x, y); }
template <> inline double power<double>(double x, double y) { return std::pow(x, y); }

struct Stage {
  double large, acoustic;
  int steps;
};

static void schedule(double dt, int order, int substeps, int splits, Stage out[3]) {
  const double dt_dyn = dt / splits;
  if (order == 3) {
    out[0] = {dt_dyn / 3.0, dt_dyn / 3.0, 1};
    out[1] = {dt_dyn / 2.0, dt_dyn / substeps, std::max(1, substeps / 2)};
    out[2] = {dt_dyn, dt_dyn / substeps, substeps};
  } else {
    out[0] = {dt_dyn / 2.0, dt_dyn / substeps, std::max(1, substeps / 2)};
    out[1] = {dt_dyn / 2.0, dt_dyn / substeps, std::max(1, substeps / 2)};
    out[2] = {dt_dyn, dt_dyn / substeps, substeps};
  }
}

// Planted mistakes (test hook, Oracle::plant): each is a plausible error the
// tests must catch.
enum {
  PLANT_NONE = 0,
  PLANT_CORIOLIS_TARGET_ANGLE = 1,  // reference wind from the target edge's angle
  PLANT_CORIOLIS_NEIGHBOUR_F = 2,   // perturbation Coriolis with the neighbour's f
  PLANT_BLEND_SIGN = 3,             // third-order upwind blend with the wrong sign
  PLANT_ADV_SLOT = 4,               // advection weight s paired with cell s+1
  PLANT_CORRECTION_SIGN = 5,        // RK correction flux of theta with the wrong sign
  PLANT_VERTICAL_CORRECTION = 6,    // (rw_save - rw) reversed in the vertical theta flux
  PLANT_W_INTERP = 7,               // fzm/fzp swapped in the w-flux edge mass flux
  PLANT_OMEGA_SIGN = 8,             // omega conversion with the upwind sign reversed
};

template <class T> struct Oracle {
  // sizes
  int C, E, V, L, ME, ME2, MA, NT, NM;
  hexdyn_config cfg;
  // connectivity
  std::vector<int> coe, voe, eoc, coc, voc, neoc, kfc, eov, cov, eoe, neoe, nadv, advc;
  // geometry
  std::vector<T> woe, dv, invdc, invdv, invarea, invtri, kite, fedge, fvert, cosang, sinang, adv, adv3;
  std::vector<T> dcv, dcdv;       // dc, dc*dv per edge
  std::vector<T> csign;           // [C][ME] +1 first cell (diagnostics/dynamics sign)
  std::vector<T> vsign;           // [V][3]
  // vertical
  std::vector<T> zz, zxu, rdzw, rdzu, fzm, fzp, zbc, zb3c, dss, etp, etm, ewp, ewm, uinit, vinit, rb, rtb, exb;
  T cf1, cf2, cf3;
  // published state (outer)
  std::vector<T> rho, rt, ru, rw, q;
  std::vector<T> theta_m, exner, rho_p, rtheta_p, pp, u, w;
  double time_s = 0.0;
  // step working set
  std::vector<T> qtot, cqw, cqu;
  std::vector<T> o_rho, o_q;       // outer saved rho / scalars (transport t0)
  std::vector<T> c_rho, c_rt, c_ru, c_rw;          // current
  std::vector<T> s_rho, s_rt, s_ru, s_rw;          // subcycle start
  std::vector<T> c_theta, c_rho_p, c_rtheta_p;     // current diagnostics (u, w, exner, pp in place)
  std::vector<T> s_theta, s_rho_p, s_rtheta_p, s_exner;
  std::vector<T> cofwr, cofwz, cofwt, coftz, atri, btri, ctri, alpha, gamma;
  std::vector<T> dpdz, euler_ru, euler_rw, tend_rho;
  std::vector<T> tend_ru, tend_rt, tend_rw, hdiv;
  std::vector<T> ru_p, ru_avg, rw_p, ww_avg, rho_pp, rtheta_pp, rtheta_pp_old;
  std::vector<T> v, pv_edge, ke, divergence, vorticity, ke_vertex, pv_vertex, pv_cell;
  std::vector<T> flux_u, flux_w, split_u, split_w;
  std::vector<T> q_stage;
  // horizontal mixing
  int mixing = 0;
  std::vector<T> area, tri, density, defc2, defs2, defcs;  // areaCell, areaTriangle, meshDensity, weights
  std::vector<T> kdiff, euler_theta;
  // held physics tendencies (cuda_driver.py:4294-4302, 4443-4446, 4484-4494)
  int physics = 0;
  std::vector<T> phys_rho, phys_ru, phys_rt;
  // held scalar tendencies, the transport's scalar_tendency (cuda_driver.py:5162-5163)
  int physics_q = 0;
  std::vector<T> phys_q;
  // the stage's edge fluxes (kept for the tests: theta flux, RK correction flux, w flux)
  std::vector<T> fth_e, corr_e, fw_e;
  // Test hook: a deliberate mistake planted in one operator, so the tests can
  // show that their checks catch it (0 = none; see the PLANT_* codes).
  int plant = 0;
  Stage dyn[3], sca[3];
  int failed = 0;
  int in_step = 0;
  std::string error;

  size_t cL() const { return (size_t)L * C; }
  size_t eL() const { return (size_t)L * E; }
  size_t wL() const { return (size_t)(L + 1) * C; }

  template <class S> static std::vector<T> cast(const S *p, size_t n) {
    std::vector<T> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = (T)p[i];
    return out;
  }
  static std::vector<int> ints(const int32_t *p, size_t n) { return std::vector<int>(p, p + n); }

  int init(const hexdyn_mesh *m, const hexdyn_vertical *vt, const hexdyn_config *c) {
    C = m->n_cells; E = m->n_edges; V = m->n_vertices; ME = m->max_edges; ME2 = m->max_edges2; MA = m->max_adv;
    L = vt->n_levels; cfg = *c; NT = c->n_scalars; NM = c->n_moist;
    coe = ints(m->cells_on_edge, 2 * (size_t)E);
    voe = ints(m->vertices_on_edge, 2 * (size_t)E);
    eoc = ints(m->edges_on_cell, (size_t)C * ME);
    coc = ints(m->cells_on_cell, (size_t)C * ME);
    voc = ints(m->vertices_on_cell, (size_t)C * ME);
    neoc = ints(m->n_edges_on_cell, C);
    kfc = ints(m->kite_for_cell, (size_t)C * ME);
    eov = ints(m->edges_on_vertex, 3 * (size_t)V);
    cov = ints(m->cells_on_vertex, 3 * (size_t)V);
    eoe = ints(m->edges_on_edge, (size_t)E * ME2);
    neoe = ints(m->n_edges_on_edge, E);
    nadv = ints(m->n_adv_cells_for_edge, E);
    advc = ints(m->adv_cells_for_edge, (size_t)E * MA);
    woe = cast(m->weights_on_edge, (size_t)E * ME2);
    dcv = cast(m->dc_edge, E);
    dv = cast(m->dv_edge, E);
    invdc = cast(m->inv_dc_edge, E);
    invdv = cast(m->inv_dv_edge, E);
    invarea = cast(m->inv_area_cell, C);
    invtri = cast(m->inv_area_triangle, V);
    kite = cast(m->kite_areas_on_vertex, 3 * (size_t)V);
    fedge = cast(m->f_edge, E);
    fvert = cast(m->f_vertex, V);
    cosang = cast(m->cos_angle_edge, E);
    sinang = cast(m->sin_angle_edge, E);
    adv = cast(m->adv_coefs, (size_t)E * MA);
    adv3 = cast(m->adv_coefs_3rd, (size_t)E * MA);
    dcdv.resize(E);
    for (int e = 0; e < E; ++e) dcdv[e] = dcv[e] * dv[e];
    mixing = c->config_horiz_mixing == HEXDYN_MIXING_2D_SMAGORINSKY;
    if (mixing) {
      if (!m->deformation_coef_c2 || !m->deformation_coef_s2 || !m->deformation_coef_cs || !m->area_cell ||
          !m->area_triangle)
        return 1;
      area = cast(m->area_cell, C);
      tri = cast(m->area_triangle, V);
      if (m->mesh_density) density = cast(m->mesh_density, C);
      else density.assign(C, T(1));
      defc2 = cast(m->deformation_coef_c2, (size_t)C * ME);
      defs2 = cast(m->deformation_coef_s2, (size_t)C * ME);
      defcs = cast(m->deformation_coef_cs, (size_t)C * ME);
    }
    csign.assign((size_t)C * ME, T(0));
    for (int cc = 0; cc < C; ++cc)
      for (int s = 0; s < neoc[cc]; ++s) {
        const int e = eoc[(size_t)cc * ME + s];
        csign[(size_t)cc * ME + s] = coe[2 * (size_t)e] == cc ? T(1) : T(-1);
      }
    vsign.assign(3 * (size_t)V, T(0));
    for (int vv = 0; vv < V; ++vv)
      for (int s = 0; s < 3; ++s) {
        const int e = eov[3 * (size_t)vv + s];
        vsign[3 * (size_t)vv + s] = voe[2 * (size_t)e + 1] == vv ? T(1) : T(-1);
      }
    zz = cast(vt->zz, cL());
    zxu = cast(vt->zxu, eL());
    rdzw = cast(vt->rdzw, L); rdzu = cast(vt->rdzu, L); fzm = cast(vt->fzm, L); fzp = cast(vt->fzp, L);
    cf1 = (T)vt->cf1; cf2 = (T)vt->cf2; cf3 = (T)vt->cf3;
    zbc = cast(vt->zb_cell, wL() * ME);
    zb3c = cast(vt->zb3_cell, wL() * ME);
    dss = cast(vt->dss, cL());
    etp = cast(vt->etp, L); etm = cast(vt->etm, L); ewp = cast(vt->ewp, L + 1); ewm = cast(vt->ewm, L + 1);
    uinit = cast(vt->u_init, L); vinit = cast(vt->v_init, L);
    rb = cast(vt->rho_base, cL()); rtb = cast(vt->rtheta_base, cL()); exb = cast(vt->exner_base, cL());
    schedule(cfg.config_dt, cfg.config_time_integration_order, cfg.config_number_of_sub_steps,
             cfg.config_dynamics_split_steps, dyn);
    schedule(cfg.config_dt, cfg.config_time_integration_order, cfg.config_number_of_sub_steps, 1, sca);
    auto z = [](std::vector<T> &a, size_t n) { a.assign(n, T(0)); };
    for (auto *a : {&rho, &rt, &theta_m, &exner, &rho_p, &rtheta_p, &pp, &qtot, &cqw, &o_rho, &c_rho, &c_rt,
                    &s_rho, &s_rt, &c_theta, &c_rho_p, &c_rtheta_p, &s_theta, &s_rho_p, &s_rtheta_p, &s_exner,
                    &cofwr, &cofwz, &cofwt, &atri, &btri, &ctri, &alpha, &gamma, &dpdz, &tend_rho, &tend_rt,
                    &hdiv, &rho_pp, &rtheta_pp, &rtheta_pp_old, &ke, &divergence, &pv_cell, &kdiff,
                    &euler_theta})
      z(*a, cL());
    for (auto *a : {&ru, &u, &cqu, &c_ru, &s_ru, &euler_ru, &tend_ru, &ru_p, &ru_avg, &v, &pv_edge, &flux_u, &split_u})
      z(*a, eL());
    for (auto *a : {&rw, &w, &c_rw, &s_rw, &coftz, &euler_rw, &tend_rw, &rw_p, &ww_avg, &flux_w, &split_w})
      z(*a, wL());
    for (auto *a : {&vorticity, &ke_vertex, &pv_vertex}) z(*a, (size_t)L * V);
    z(q, (size_t)NT * cL()); z(o_q, (size_t)NT * cL()); z(q_stage, (size_t)NT * cL());
    return 0;
  }

  // ---------------------------------------------------------------- helpers
  T sgn(int cell, int slot) const { return csign[(size_t)cell * ME + slot]; }
  size_t ci(int k, int c) const { return (size_t)k * C + c; }
  size_t ei(int k, int e) const { return (size_t)k * E + e; }

  // solve diagnostics (diagnostics.py:95-306; v8.4.1 stored-inverse forms)
  void solve_diagnostics(const std::vector<T> &uu, int rk_step) {
    const T quarter = T(0.25);
#pragma omp parallel for
    for (int vv = 0; vv < V; ++vv)
      for (int k = 0; k < L; ++k) {
        T vort = T(0), kev = T(0);
        for (int s = 0; s < 3; ++s) {
          const int e = eov[3 * (size_t)vv + s];
          const T ue = uu[ei(k, e)];
          vort = vort + (vsign[3 * (size_t)vv + s] * dcv[e]) * ue;
          kev = kev + dcdv[e] * (ue * ue);
        }
        vort = vort * invtri[vv];
        kev = kev * (quarter * invtri[vv]);
        vorticity[(size_t)k * V + vv] = vort;
        ke_vertex[(size_t)k * V + vv] = kev;
        pv_vertex[(size_t)k * V + vv] = fvert[vv] + vort;
      }
    const T ke_fact = T(0.625);
    const T ke_rest = T(1) - ke_fact;
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc)
      for (int k = 0; k < L; ++k) {
        T d = T(0), kc = T(0), pvc = T(0);
        const T ia = invarea[cc];
        for (int s = 0; s < neoc[cc]; ++s) {
          const int e = eoc[(size_t)cc * ME + s];
          const T ue = uu[ei(k, e)];
          d = d + (sgn(cc, s) * dv[e]) * ue;
          kc = kc + quarter * (dcdv[e] * (ue * ue));
        }
        d = d * ia;
        kc = kc * ia;
        kc = kc * ke_fact;
        for (int s = 0; s < neoc[cc]; ++s) {
          const int vv = voc[(size_t)cc * ME + s];
          const int kslot = kfc[(size_t)cc * ME + s];
          T corr = ke_rest * kite[3 * (size_t)vv + kslot];
          corr = corr * ke_vertex[(size_t)k * V + vv];
          corr = corr * ia;
          kc = kc + corr;
          pvc = pvc + (kite[3 * (size_t)vv + kslot] * pv_vertex[(size_t)k * V + vv]) * ia;
        }
        divergence[ci(k, cc)] = d;
        ke[ci(k, cc)] = kc;
        pv_cell[ci(k, cc)] = pvc;
      }
    const bool recon = rk_step == 3;
    const T apvm = T(cfg.config_apvm_upwinding);
    const T scale = apvm * T(cfg.config_dt);
#pragma omp parallel for
    for (int e = 0; e < E; ++e)
      for (int k = 0; k < L; ++k) {
        if (recon) {
          T val = T(0);
          for (int s = 0; s < neoe[e]; ++s)
            val = val + woe[(size_t)e * ME2 + s] * uu[ei(k, eoe[(size_t)e * ME2 + s])];
          v[ei(k, e)] = val;
        }
        const int v0 = voe[2 * (size_t)e], v1 = voe[2 * (size_t)e + 1];
        const T pv0 = pv_vertex[(size_t)k * V + v0], pv1 = pv_vertex[(size_t)k * V + v1];
        T pve = T(0.5) * (pv0 + pv1);
        if (cfg.config_apvm_upwinding > 0.0f) {
          const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
          const T gt = (pv1 - pv0) * invdv[e];
          const T gn = (pv_cell[ci(k, c1)] - pv_cell[ci(k, c0)]) * invdc[e];
          pve = pve - scale * (v[ei(k, e)] * gt + uu[ei(k, e)] * gn);
        }
        pv_edge[ei(k, e)] = pve;
      }
  }

  // ------------------------------------------------------ moist coefficients
  void moist_coefficients() {
    const T half = T(0.5), one = T(1);
    if (NM <= 0) {
      std::fill(qtot.begin(), qtot.end(), T(0));
      std::fill(cqw.begin(), cqw.end(), one);
      std::fill(cqu.begin(), cqu.end(), one);
      return;
    }
    for (size_t i = 0; i < cL(); ++i) {
      T t = T(0);
      for (int s = 0; s < NM; ++s) t = t + q[(size_t)s * cL() + i];
      qtot[i] = t;
    }
    for (int cc = 0; cc < C; ++cc) {
      cqw[ci(0, cc)] = one;
      for (int k = 1; k < L; ++k) cqw[ci(k, cc)] = one / (one + half * (qtot[ci(k, cc)] + qtot[ci(k - 1, cc)]));
    }
    for (int e = 0; e < E; ++e) {
      const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
      for (int k = 0; k < L; ++k) {
        T t = T(0);
        for (int s = 0; s < NM; ++s)
          t = t + half * (q[(size_t)s * cL() + ci(k, c0)] + q[(size_t)s * cL() + ci(k, c1)]);
        cqu[ei(k, e)] = one / (one + t);
      }
    }
  }

  // ---------------------------------------- vertical implicit coefficients
  void vertical_coefficients(double dts, const std::vector<T> &exn, const std::vector<T> &theta,
                             const std::vector<T> &rtp) {
    const T half = T(0.5), one = T(1), g = T(kGravity);
    const T rcv = T(kRgas) / (T(kCp) - T(kRgas));
    const T c2 = T(kCp) * rcv;
    const T dts_r = T(dts);
    const T dts2 = dts_r * dts_r;
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc) {
      coftz[ci(0, cc)] = T(0);
      coftz[(size_t)L * C + cc] = T(0);
      cofwr[ci(0, cc)] = T(0); cofwz[ci(0, cc)] = T(0);
      for (int k = 1; k < L; ++k) {
        const T iz = fzm[k] * zz[ci(k, cc)] + fzp[k] * zz[ci(k - 1, cc)];
        cofwr[ci(k, cc)] = (half * g) * iz;
        cofwz[ci(k, cc)] = (((c2 * iz) * rdzu[k]) * cqw[ci(k, cc)]) *
                           (fzm[k] * exn[ci(k, cc)] + fzp[k] * exn[ci(k - 1, cc)]);
        coftz[ci(k, cc)] = fzm[k] * theta[ci(k, cc)] + fzp[k] * theta[ci(k - 1, cc)];
      }
      for (int k = 0; k < L; ++k) {
        const size_t i = ci(k, cc);
        cofwt[i] = (((((((half * rcv) * zz[i]) * g) * rb[i]) / (one + qtot[i])) * exn[i]) /
                    ((rtb[i] + rtp[i]) * exb[i]));
      }
      atri[ci(0, cc)] = T(0); btri[ci(0, cc)] = one; ctri[ci(0, cc)] = T(0);
      alpha[ci(0, cc)] = T(0); gamma[ci(0, cc)] = T(0);
      for (int k = 1; k < L; ++k) {
        const size_t i = ci(k, cc), im = ci(k - 1, cc);
        const T cofrz_k = rdzw[k], cofrz_m = rdzw[k - 1];
        T a = (((-cofwz[i]) * coftz[im]) * rdzw[k - 1]) * zz[im] + cofwr[i] * cofrz_m -
              (cofwt[im] * coftz[im]) * rdzw[k - 1];
        atri[i] = (a * etp[k - 1]) * ewp[k - 1];
        T b = (cofwz[i] * coftz[i]) * ((etp[k] * rdzw[k]) * zz[i] + (etp[k - 1] * rdzw[k - 1]) * zz[im]) -
              coftz[i] * ((etp[k] * cofwt[i]) * rdzw[k] - (etp[k - 1] * cofwt[im]) * rdzw[k - 1]) +
              cofwr[i] * (etp[k] * cofrz_k - etp[k - 1] * cofrz_m);
        btri[i] = b * ewp[k];
        const T coftz_p = coftz[ci(k + 1, cc)];
        T cc_ = (((-cofwz[i]) * coftz_p) * rdzw[k]) * zz[i] - cofwr[i] * cofrz_k + (cofwt[i] * coftz_p) * rdzw[k];
        ctri[i] = (cc_ * etp[k]) * ewp[k + 1];
      }
      ctri[ci(L - 1, cc)] = T(0);
      for (int k = 1; k < L; ++k) {
        const size_t i = ci(k, cc);
        const T den = one + dts2 * (btri[i] - atri[i] * gamma[ci(k - 1, cc)]);
        if (den == T(0)) failed = 1;
        alpha[i] = one / den;
        gamma[i] = (dts2 * ctri[i]) * alpha[i];
      }
    }
  }

  // ------------------------------------------------------------ step start
  void begin_step() {
    in_step = 1;
    o_rho = rho; o_q = q;
    c_rho = rho; c_rt = rt; c_ru = ru; c_rw = rw;
    c_theta = theta_m; c_rho_p = rho_p; c_rtheta_p = rtheta_p;
    moist_coefficients();
    solve_diagnostics(u, 3);
  }

  // -------------------------------------------------------- subcycle start
  void begin_subcycle() {
    s_rho = c_rho; s_rt = c_rt; s_ru = c_ru; s_rw = c_rw;
    s_theta = theta_m; s_rho_p = rho_p; s_rtheta_p = rtheta_p; s_exner = exner;
    c_theta = theta_m; c_rho_p = rho_p; c_rtheta_p = rtheta_p;
    vertical_coefficients(dyn[0].acoustic, s_exner, s_theta, s_rtheta_p);
    const T g = T(kGravity);
    const T one = T(1), half = T(0.5);
    if (NM <= 0) {
      for (size_t i = 0; i < cL(); ++i) dpdz[i] = (-g) * s_rho_p[i];
    } else {
      for (size_t i = 0; i < cL(); ++i) dpdz[i] = (-g) * (rb[i] * qtot[i] + s_rho_p[i] * (one + qtot[i]));
    }
#pragma omp parallel for
    for (int e = 0; e < E; ++e) {
      const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
      for (int k = 0; k < L; ++k) {
        const T a = ((pp[ci(k, c1)] - pp[ci(k, c0)]) * invdc[e]) / (half * (zz[ci(k, c1)] + zz[ci(k, c0)]));
        const T b = (half * zxu[ei(k, e)]) * (dpdz[ci(k, c0)] + dpdz[ci(k, c1)]);
        euler_ru[ei(k, e)] = (T(0) - cqu[ei(k, e)]) * (a - b);
      }
    }
    mass_flux_divergence(s_ru, tend_rho);
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc)
      for (int k = 0; k < L; ++k) {
        // (0 - div) - ..., density_tendency_f32 (cuda_horizontal.py:412-433): a zero divergence gives +0
        tend_rho[ci(k, cc)] = (T(0) - tend_rho[ci(k, cc)]) - rdzw[k] * (s_rw[(size_t)(k + 1) * C + cc] - s_rw[ci(k, cc)]);
        if (physics) tend_rho[ci(k, cc)] = tend_rho[ci(k, cc)] + phys_rho[ci(k, cc)];
      }
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc) {
      euler_rw[cc] = T(0);
      euler_rw[(size_t)L * C + cc] = T(0);
      for (int k = 1; k < L; ++k) {
        const T inner = rdzu[k] * (pp[ci(k, cc)] - pp[ci(k - 1, cc)]) -
                        (fzm[k] * dpdz[ci(k, cc)] + fzp[k] * dpdz[ci(k - 1, cc)]);
        // euler_w_f32 / euler_w_moist_v841_f32 (cuda_driver.py:955-975, 1604-1630)
        euler_rw[ci(k, cc)] = NM <= 0 ? T(0) - inner : T(0) - cqw[ci(k, cc)] * inner;
      }
    }
    if (mixing && !pools_only) horizontal_mixing();
  }
  int pools_only = 0;  // test hook: begin_subcycle stops before the mixing

  // ------------------------------------------------------ horizontal mixing
  // v8.4.1 2-D Smagorinsky mixing from the subcycle-start state, in the host
  // order of compute_dry_mixing_tendencies_v841 (cuda_horizontal_v841.py:1270-1588)
  // with the per-element arithmetic of its kernels (:205-583): viscosity,
  // del2 u into the u pool, del4 u, del2/del4 theta (the theta pool), del2/del4 w
  // added to the w pool.
  void horizontal_mixing() {
    const T one = T(1), half = T(0.5), quarter = T(0.25);
    const T len = T(cfg.config_len_disp), csm = T(cfg.config_smagorinsky_coef);
    const T strain_scale = (csm * len) * (csm * len);
    const T inv_dt = one / T(cfg.config_dt);
    const T ceiling = (T(0.01) * (len * len)) * inv_dt;
    const T h4 = T(cfg.config_visc4_2dsmag) * ((len * len) * len);
    const bool mix4 = h4 > T(0);
    std::vector<T> del2(E, one), del4(E, one);
    if (cfg.config_h_ScaleWithMesh)
      for (int e = 0; e < E; ++e) {
        const T mean = half * (density[coe[2 * (size_t)e]] + density[coe[2 * (size_t)e + 1]]);
        // binary64 pow rounded once (mixing.py:313-318, rkind_libm.py:68-74)
        del2[e] = one / T(std::pow((double)mean, (double)0.25f));
        del4[e] = one / T(std::pow((double)mean, (double)0.75f));
      }
    auto vi = [&](int k, int vx) { return (size_t)k * V + vx; };
    auto rho_e = [&](int k, int e) {
      return half * (c_rho[ci(k, coe[2 * (size_t)e])] + c_rho[ci(k, coe[2 * (size_t)e + 1])]);
    };
    // viscosity (smagorinsky_v841_f32)
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc)
      for (int k = 0; k < L; ++k) {
        T dudx = T(0), dudy = T(0), dvdx = T(0), dvdy = T(0);
        for (int s = 0; s < neoc[cc]; ++s) {
          const size_t o = (size_t)cc * ME + s;
          const int e = eoc[o];
          const T uu = u[ei(k, e)], vv = v[ei(k, e)];
          dudx = (dudx + defc2[o] * uu) - defcs[o] * vv;
          dudy = (dudy + defcs[o] * uu) - defs2[o] * vv;
          dvdx = (dvdx + defcs[o] * uu) + defc2[o] * vv;
          dvdy = (dvdy + defs2[o] * uu) + defcs[o] * vv;
        }
        const T d11 = T(2) * dudx, d22 = T(2) * dvdy, d12 = dudy + dvdx;
        const T diff = d11 - d22;
        const T strain = std::sqrt(quarter * (diff * diff) + d12 * d12);
        kdiff[ci(k, cc)] = std::fmin(strain_scale * strain, ceiling);
        // CAM floor over the top levels (native/upper_damping.rs:5-16)
        const int ncam = cfg.config_mpas_cam_coef > 0.0f ? cfg.config_number_cam_damping_levels : 0;
        if (k >= L - ncam) {
          const T base = ((T(4) * T(2.0833f)) * len) * T(cfg.config_mpas_cam_coef);
          const T floor = base * (one - T(L - 1 - k) / T(ncam));
          kdiff[ci(k, cc)] = std::fmax(kdiff[ci(k, cc)], floor);
        }
      }
    // del2 u (momentum_filter_seed_v841_f32)
    std::vector<T> delsq_u(eL());
#pragma omp parallel for
    for (int e = 0; e < E; ++e) {
      const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
      const int v0 = voe[2 * (size_t)e], v1 = voe[2 * (size_t)e + 1];
      const T idc = one / dcv[e];
      const T idv = std::fmin(one / dv[e], T(4) * idc);
      for (int k = 0; k < L; ++k) {
        const T lap = (divergence[ci(k, c1)] - divergence[ci(k, c0)]) * idc -
                      (vorticity[vi(k, v1)] - vorticity[vi(k, v0)]) * idv;
        const T kedge = half * (kdiff[ci(k, c0)] + kdiff[ci(k, c1)]);
        delsq_u[ei(k, e)] = lap;
        euler_ru[ei(k, e)] = euler_ru[ei(k, e)] + ((rho_e(k, e) * kedge) * lap) * del2[e];
      }
    }
    if (mix4) {
      // laplacian_divergence / laplacian_vorticity / momentum_filter_lap4
      std::vector<T> ddiv(cL()), dvort((size_t)L * V);
#pragma omp parallel for
      for (int cc = 0; cc < C; ++cc) {
        const T ia = one / area[cc];
        for (int k = 0; k < L; ++k) {
          T value = T(0);
          for (int s = 0; s < neoc[cc]; ++s) {
            const int e = eoc[(size_t)cc * ME + s];
            value = value + ((ia * dv[e]) * sgn(cc, s)) * delsq_u[ei(k, e)];
          }
          ddiv[ci(k, cc)] = value;
        }
      }
#pragma omp parallel for
      for (int vx = 0; vx < V; ++vx) {
        const T it = one / tri[vx];
        for (int k = 0; k < L; ++k) {
          T value = T(0);
          for (int s = 0; s < 3; ++s) {
            const int e = eov[3 * (size_t)vx + s];
            const T vs = voe[2 * (size_t)e] == vx ? T(-1) : T(1);
            value = value + ((it * dcv[e]) * vs) * delsq_u[ei(k, e)];
          }
          dvort[vi(k, vx)] = value;
        }
      }
#pragma omp parallel for
      for (int e = 0; e < E; ++e) {
        const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
        const int v0 = voe[2 * (size_t)e], v1 = voe[2 * (size_t)e + 1];
        const T idc = one / dcv[e];
        const T idv = std::fmin(one / dv[e], T(4) * idc);
        const T scale = del4[e] * h4;
        const T r_dc = (scale * T(cfg.config_del4u_div_factor)) * idc;
        const T r_dv = scale * idv;
        for (int k = 0; k < L; ++k) {
          const T diffusion = rho_e(k, e) * ((ddiv[ci(k, c1)] - ddiv[ci(k, c0)]) * r_dc -
                                             (dvort[vi(k, v1)] - dvort[vi(k, v0)]) * r_dv);
          euler_ru[ei(k, e)] = euler_ru[ei(k, e)] - diffusion;
        }
      }
    }
    // theta (theta_filter_lap2/lap4_v841_f32) and w (w_filter_lap2/lap4_v841_f32)
    std::vector<T> dtheta(cL()), dw(cL(), T(0)), tw(wL(), T(0));
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc) {
      const T ia = one / area[cc];
      for (int k = 0; k < L; ++k) {
        T lap = T(0), tend = T(0), lapw = T(0), tendw = T(0);
        for (int s = 0; s < neoc[cc]; ++s) {
          const int e = eoc[(size_t)cc * ME + s];
          const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
          const T idc = one / dcv[e];
          const T factor = ((ia * sgn(cc, s)) * dv[e]) * idc;
          const T flux = (factor * (s_theta[ci(k, c1)] - s_theta[ci(k, c0)])) * rho_e(k, e);
          lap = lap + flux;
          tend = tend + ((flux * half) * (kdiff[ci(k, c0)] + kdiff[ci(k, c1)])) * del2[e];
          if (k >= 1) {
            const T fw = (((half * ia) * sgn(cc, s)) * dv[e]) * idc;
            const T fluxw = (fw * (rho_e(k, e) + rho_e(k - 1, e))) * (w[ci(k, c1)] - w[ci(k, c0)]);
            lapw = lapw + fluxw;
            const T sum_k = ((kdiff[ci(k, c0)] + kdiff[ci(k, c1)]) + kdiff[ci(k - 1, c0)]) + kdiff[ci(k - 1, c1)];
            tendw = tendw + ((fluxw * del2[e]) * quarter) * sum_k;
          }
        }
        dtheta[ci(k, cc)] = lap;
        euler_theta[ci(k, cc)] = tend;
        dw[ci(k, cc)] = lapw;
        tw[ci(k, cc)] = tendw;
      }
    }
    if (mix4) {
#pragma omp parallel for
      for (int cc = 0; cc < C; ++cc) {
        const T ia = one / area[cc];
        for (int k = 0; k < L; ++k) {
          T tend = euler_theta[ci(k, cc)], tendw = tw[ci(k, cc)];
          for (int s = 0; s < neoc[cc]; ++s) {
            const int e = eoc[(size_t)cc * ME + s];
            const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
            const T factor = (((del4[e] * (h4 * ia)) * dv[e]) * sgn(cc, s)) * (one / dcv[e]);
            tend = tend - factor * (dtheta[ci(k, c1)] - dtheta[ci(k, c0)]);
            if (k >= 1) tendw = tendw - factor * (dw[ci(k, c1)] - dw[ci(k, c0)]);
          }
          euler_theta[ci(k, cc)] = tend;
          tw[ci(k, cc)] = tendw;
        }
      }
    }
    for (size_t i = 0; i < wL(); ++i) euler_rw[i] = euler_rw[i] + tw[i];
  }

  void mass_flux_divergence(const std::vector<T> &r_u, std::vector<T> &out) {
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc)
      for (int k = 0; k < L; ++k) {
        T acc = T(0);
        for (int s = 0; s < neoc[cc]; ++s) {
          const int e = eoc[(size_t)cc * ME + s];
          acc = acc + (sgn(cc, s) * dv[e]) * r_u[ei(k, e)];
        }
        out[ci(k, cc)] = acc * invarea[cc];
      }
  }

  static T flux4(T qm2, T qm1, T q0, T qp1, T vel) {
    return (vel * (T(7) * (q0 + qm1) - (qp1 + qm2))) / T(12);
  }
  static T flux3(T qm2, T qm1, T q0, T qp1, T vel, T coef) {
    return flux4(qm2, qm1, q0, qp1, vel) + ((coef * std::fabs(vel)) * ((qp1 - qm2) - T(3) * (q0 - qm1))) / T(12);
  }

  // ------------------------------------------------------ stage tendencies
  // vertical transport of normal momentum of edge e, all levels: out[k] =
  // -rdzw(k) * (wduz(k+1) - wduz(k)) (dynamics.py:227-270)
  void vertical_transport_u(int e, T *out) const {
    const T half = T(0.5), one = T(1);
    const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
    T wduz[HEXDYN_ROW + 1] = {};
    auto rwe = [&](int k) { return half * (c_rw[(size_t)k * C + c0] + c_rw[(size_t)k * C + c1]); };
    wduz[1] = rwe(1) * (fzm[1] * u[ei(1, e)] + fzp[1] * u[ei(0, e)]);
    for (int k = 2; k < L - 1; ++k)
      wduz[k] = flux3(u[ei(k - 2, e)], u[ei(k - 1, e)], u[ei(k, e)], u[ei(k + 1, e)], rwe(k), one);
    wduz[L - 1] = rwe(L - 1) * (fzm[L - 1] * u[ei(L - 1, e)] + fzp[L - 1] * u[ei(L - 2, e)]);
    for (int k = 0; k < L; ++k) out[k] = (-rdzw[k]) * (wduz[k + 1] - wduz[k]);
  }

  // Nonlinear PV/KE momentum terms of edge e at level k with the v8.4.1
  // perturbation-Coriolis subtraction (dynamics_v841.py:128-185: the second
  // neighbour loop takes the neighbour edge's angle and the target edge's f):
  //   (initial + rho_edge*(q - dKE/dc)) - (u*0.5)*(hdiv0 + hdiv1).
  // Without `initial` it is dynamics.py:273-305's
  //   rho_edge*(q - dKE/dc) - (u*0.5)*(hdiv0 + hdiv1).
  T momentum_terms(int e, int k, const T *initial) const {
    const T half = T(0.5);
    const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
    T qq = T(0);
    for (int s = 0; s < neoe[e]; ++s) {
      const int n = eoe[(size_t)e * ME2 + s];
      const T work_pv = half * (pv_edge[ei(k, e)] + pv_edge[ei(k, n)]);
      qq = qq + (woe[(size_t)e * ME2 + s] * u[ei(k, n)]) * work_pv;
    }
    for (int s = 0; s < neoe[e]; ++s) {
      const int n = eoe[(size_t)e * ME2 + s];
      const int na = plant == PLANT_CORIOLIS_TARGET_ANGLE ? e : n;
      const T ref_u = uinit[k] * cosang[na] - vinit[k] * sinang[na];
      const T f = plant == PLANT_CORIOLIS_NEIGHBOUR_F ? fedge[n] : fedge[e];
      qq = qq - (woe[(size_t)e * ME2 + s] * ref_u) * f;
    }
    const T rho_edge = half * (c_rho[ci(k, c0)] + c_rho[ci(k, c1)]);
    const T horiz = rho_edge * (qq - (ke[ci(k, c1)] - ke[ci(k, c0)]) * invdc[e]);
    const T lhs = initial ? *initial + horiz : horiz;
    return lhs - (u[ei(k, e)] * half) * (hdiv[ci(k, c0)] + hdiv[ci(k, c1)]);
  }

  // Upper-level Rayleigh damping of u at level k (native/upper_damping.rs:18-31):
  // t - (rho_edge*u)*((k-(L-n)+1)*inverse) on the top n levels.
  T rayleigh_u(int e, int k, T t) const {
    if (!cfg.config_rayleigh_damp_u) return t;
    const int n = cfg.config_number_rayleigh_damp_u_levels;
    if (k < L - n) return t;
    const T inverse = T(1) / (T(n) * (T(cfg.config_rayleigh_damp_u_timescale_days) * T(86400)));
    const T rho_edge = T(0.5) * (c_rho[ci(k, coe[2 * (size_t)e])] + c_rho[ci(k, coe[2 * (size_t)e + 1])]);
    return t - (rho_edge * u[ei(k, e)]) * (T(k - (L - n) + 1) * inverse);
  }

  // w -> omega tendency of cell cc at interface k, 1 <= k < L
  // (acoustic.py:246-281): r - sum_slots (zb + sign(tend_u)*zb3)*flux, then
  // times the interface zz.
  T omega(int cc, int k, T r, const std::vector<T> &tru) const {
    for (int s = 0; s < neoc[cc]; ++s) {
      const int e = eoc[(size_t)cc * ME + s];
      const T tk = tru[ei(k, e)];
      const T flux = sgn(cc, s) * (fzm[k] * tk + fzp[k] * tru[ei(k - 1, e)]);
      const size_t zi = ((size_t)k * C + cc) * ME + s;
      const T up = plant == PLANT_OMEGA_SIGN ? -copysign1(tk) : copysign1(tk);
      r = r - (zbc[zi] + up * zb3c[zi]) * flux;
    }
    return r * (fzm[k] * zz[ci(k, cc)] + fzp[k] * zz[ci(k - 1, cc)]);
  }

  void dyn_tend(int stage) {
    const T half = T(0.5), one = T(1);
    const T coef = T(cfg.config_coef_3rd_order);
    if (stage == 2 && cfg.config_time_integration_order == 3)
      vertical_coefficients(dyn[1].acoustic, s_exner, c_theta, c_rtheta_p);
    mass_flux_divergence(c_ru, hdiv);
    // momentum (vertical transport, then v8.4.1 horizontal terms, Rayleigh
    // damping, then the Euler pool)
#pragma omp parallel for
    for (int e = 0; e < E; ++e) {
      T vtu[HEXDYN_ROW];
      vertical_transport_u(e, vtu);
      for (int k = 0; k < L; ++k) {
        const T t = rayleigh_u(e, k, momentum_terms(e, k, &vtu[k]));
        tend_ru[ei(k, e)] = t + euler_ru[ei(k, e)];
        if (physics) tend_ru[ei(k, e)] = tend_ru[ei(k, e)] + phys_ru[ei(k, e)];
      }
    }
    // theta: edge fluxes (stored-inverse branch, driver.py:1003-1034)
    fth_e.assign(eL(), T(0));
    corr_e.assign(eL(), T(0));
    std::vector<T> &fth = fth_e, &corr = corr_e;
    const T sv_sign = plant == PLANT_BLEND_SIGN ? T(-1) : T(1);
#pragma omp parallel for
    for (int e = 0; e < E; ++e) {
      const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
      for (int k = 0; k < L; ++k) {
        const T rue = c_ru[ei(k, e)];
        const T sv = sv_sign * copysign1(rue);
        T et = T(0);
        for (int s = 0; s < nadv[e]; ++s) {
          const size_t j = (size_t)e * MA + s;
          const size_t jc = plant == PLANT_ADV_SLOT ? (size_t)e * MA + (s + 1) % nadv[e] : j;
          const T wgt = adv[j] + (coef * sv) * adv3[j];
          et = et + wgt * c_theta[ci(k, advc[jc])];
        }
        fth[ei(k, e)] = rue * et;
        const T dru = plant == PLANT_CORRECTION_SIGN ? rue - s_ru[ei(k, e)] : s_ru[ei(k, e)] - rue;
        corr[ei(k, e)] = ((dv[e] * dru) * half) * (s_theta[ci(k, c1)] + s_theta[ci(k, c0)]);
      }
    }
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc) {
      T vf[HEXDYN_ROW + 1] = {};
      auto th = [&](int k) { return c_theta[ci(k, cc)]; };
      auto ths = [&](int k) { return s_theta[ci(k, cc)]; };
      auto rwc = [&](int k) { return c_rw[(size_t)k * C + cc]; };
      auto rws = [&](int k) { return s_rw[(size_t)k * C + cc]; };
      // (rws - rwc): the RK correction of the vertical flux (planted mistake: reversed)
      auto drw = [&](int k) { return plant == PLANT_VERTICAL_CORRECTION ? rwc(k) - rws(k) : rws(k) - rwc(k); };
      vf[1] = rwc(1) * (fzm[1] * th(1) + fzp[1] * th(0));
      vf[1] = vf[1] + drw(1) * (fzm[1] * ths(1) + fzp[1] * ths(0));
      for (int k = 2; k < L - 1; ++k) {
        vf[k] = flux3(th(k - 2), th(k - 1), th(k), th(k + 1), rwc(k), plant == PLANT_BLEND_SIGN ? -coef : coef);
        vf[k] = vf[k] + drw(k) * (fzm[k] * ths(k) + fzp[k] * ths(k - 1));
      }
      vf[L - 1] = rws(L - 1) * (fzm[L - 1] * th(L - 1) + fzp[L - 1] * th(L - 2));
      for (int k = 0; k < L; ++k) {
        T r = T(0);
        for (int s = 0; s < neoc[cc]; ++s) {
          const int e = eoc[(size_t)cc * ME + s];
          r = r - sgn(cc, s) * fth[ei(k, e)];
        }
        if (stage > 1)
          for (int s = 0; s < neoc[cc]; ++s) {
            const int e = eoc[(size_t)cc * ME + s];
            r = r - sgn(cc, s) * corr[ei(k, e)];
          }
        r = r * invarea[cc];
        tend_rt[ci(k, cc)] = r - rdzw[k] * (vf[k + 1] - vf[k]);
        if (mixing) tend_rt[ci(k, cc)] = tend_rt[ci(k, cc)] + euler_theta[ci(k, cc)];
        if (physics) tend_rt[ci(k, cc)] = tend_rt[ci(k, cc)] + phys_rt[ci(k, cc)];
      }
    }
    // w transport (stored-inverse branch, driver.py:1140-1207), Euler pool, omega conversion
    fw_e.assign((size_t)(L + 1) * E, T(0));
    std::vector<T> &fw = fw_e;
#pragma omp parallel for
    for (int e = 0; e < E; ++e) {
      for (int k = 1; k < L; ++k) {
        const T rui = plant == PLANT_W_INTERP ? fzp[k] * c_ru[ei(k, e)] + fzm[k] * c_ru[ei(k - 1, e)]
                                              : fzm[k] * c_ru[ei(k, e)] + fzp[k] * c_ru[ei(k - 1, e)];
        const T sv = sv_sign * copysign1(rui);
        T ew = T(0);
        for (int s = 0; s < nadv[e]; ++s) {
          const size_t j = (size_t)e * MA + s;
          const size_t jc = plant == PLANT_ADV_SLOT ? (size_t)e * MA + (s + 1) % nadv[e] : j;
          const T wgt = adv[j] + (coef * sv) * adv3[j];
          ew = ew + wgt * w[(size_t)k * C + advc[jc]];
        }
        fw[(size_t)k * E + e] = rui * ew;
      }
    }
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc) {
      auto rwc = [&](int k) { return c_rw[(size_t)k * C + cc]; };
      auto ww = [&](int k) { return w[(size_t)k * C + cc]; };
      T vf[HEXDYN_ROW + 1] = {};
      vf[1] = (T(0.25) * (rwc(1) + rwc(0))) * (ww(1) + ww(0));
      for (int k = 2; k < L - 1; ++k)
        vf[k] = flux3(ww(k - 2), ww(k - 1), ww(k), ww(k + 1), half * (rwc(k) + rwc(k - 1)),
                      plant == PLANT_BLEND_SIGN ? -one : one);
      vf[L - 1] = (T(0.25) * (rwc(L - 1) + rwc(L - 2))) * (ww(L - 1) + ww(L - 2));
      tend_rw[cc] = T(0);
      tend_rw[(size_t)L * C + cc] = T(0);
      for (int k = 1; k < L; ++k) {
        T r = T(0);
        for (int s = 0; s < neoc[cc]; ++s) {
          const int e = eoc[(size_t)cc * ME + s];
          r = r - sgn(cc, s) * fw[(size_t)k * E + e];
        }
        r = r * invarea[cc];
        r = r - rdzu[k] * (vf[k + 1] - vf[k]);
        r = r + euler_rw[(size_t)k * C + cc];
        tend_rw[(size_t)k * C + cc] = omega(cc, k, r, tend_ru);
      }
    }
  }

  // ------------------------------------------------------- acoustic substep
  void acoustic_substep(int stage, int small_step) {
    const double dts = dyn[stage - 1].acoustic;
    const T dts_r = T(dts), half = T(0.5), one = T(1), g = T(kGravity);
    const T rcv = T(kRgas) / (T(kCp) - T(kRgas));
    const T c2 = T(kCp) * rcv;
    if (small_step != 1) {
#pragma omp parallel for
      for (int e = 0; e < E; ++e) {
        const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
        for (int k = 0; k < L; ++k) {
          T pg = ((rtheta_pp[ci(k, c1)] - rtheta_pp[ci(k, c0)]) * invdc[e]) / (half * (zz[ci(k, c1)] + zz[ci(k, c0)]));
          pg = (((cqu[ei(k, e)] * half) * c2) * (s_exner[ci(k, c0)] + s_exner[ci(k, c1)])) * pg;
          pg = pg + ((half * zxu[ei(k, e)]) * g) * (rho_pp[ci(k, c0)] + rho_pp[ci(k, c1)]);
          const size_t i = ei(k, e);
          ru_p[i] = ru_p[i] + dts_r * (tend_ru[i] - pg);
          ru_avg[i] = ru_avg[i] + ru_p[i];
        }
      }
    } else {
      for (size_t i = 0; i < eL(); ++i) {
        ru_p[i] = dts_r * tend_ru[i];
        ru_avg[i] = ru_p[i];
      }
    }
    if (small_step == 1) {
      std::fill(rtheta_pp_old.begin(), rtheta_pp_old.end(), T(0));
      std::fill(ww_avg.begin(), ww_avg.end(), T(0));
      std::fill(rho_pp.begin(), rho_pp.end(), T(0));
      std::fill(rtheta_pp.begin(), rtheta_pp.end(), T(0));
      std::fill(rw_p.begin(), rw_p.end(), T(0));
    } else {
      rtheta_pp_old = rtheta_pp;
    }
    const T dts2 = dts_r * dts_r;
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc) {
      T rs[HEXDYN_ROW] = {}, ts[HEXDYN_ROW] = {}, col[HEXDYN_ROW + 1] = {};
      const T ia = invarea[cc];
      for (int k = 0; k < L; ++k) {
        T r = T(0), t = T(0);
        for (int s = 0; s < neoc[cc]; ++s) {
          const int e = eoc[(size_t)cc * ME + s];
          const int e0 = coe[2 * (size_t)e], e1 = coe[2 * (size_t)e + 1];
          const T flux = (((sgn(cc, s) * dts_r) * dv[e]) * ru_p[ei(k, e)]) * ia;
          r = r - flux;
          t = t - (flux * half) * (s_theta[ci(k, e1)] + s_theta[ci(k, e0)]);
        }
        rs[k] = r; ts[k] = t;
      }
      auto rwp = [&](int k) { return rw_p[(size_t)k * C + cc]; };
      for (int k = 0; k < L; ++k) {
        const size_t i = ci(k, cc);
        rs[k] = ((rho_pp[i] + dts_r * tend_rho[i]) + rs[k]) -
                (dts_r * rdzw[k]) * (ewm[k + 1] * rwp(k + 1) - ewm[k] * rwp(k));
        ts[k] = ((rtheta_pp[i] + dts_r * tend_rt[i]) + ts[k]) -
                (dts_r * rdzw[k]) * ((ewm[k + 1] * coftz[ci(k + 1, cc)]) * rwp(k + 1) -
                                     (ewm[k] * coftz[ci(k, cc)]) * rwp(k));
      }
      for (int k = 1; k < L; ++k) ww_avg[(size_t)k * C + cc] = ww_avg[(size_t)k * C + cc] + ewm[k] * rwp(k);
      for (int k = 0; k <= L; ++k) col[k] = rwp(k);
      for (int k = 1; k < L; ++k) {
        const size_t i = ci(k, cc), im = ci(k - 1, cc);
        T x = tend_rw[(size_t)k * C + cc] -
              cofwz[i] * (((etp[k] * zz[i]) * ts[k] - (etp[k - 1] * zz[im]) * ts[k - 1]) +
                          ((etm[k] * zz[i]) * rtheta_pp[i] - (etm[k - 1] * zz[im]) * rtheta_pp[im]));
        x = x - cofwr[i] * ((etp[k] * rs[k] + etp[k - 1] * rs[k - 1]) + (etm[k] * rho_pp[i] + etm[k - 1] * rho_pp[im]));
        x = x + cofwt[i] * (etp[k] * ts[k] + etm[k] * rtheta_pp[i]);
        x = x + cofwt[im] * (etp[k - 1] * ts[k - 1] + etm[k - 1] * rtheta_pp[im]);
        col[k] = col[k] + dts_r * x;
      }
      for (int k = 1; k < L; ++k)
        col[k] = (col[k] - (dts2 * atri[ci(k, cc)]) * col[k - 1]) * alpha[ci(k, cc)];
      for (int k = L - 1; k >= 0; --k) col[k] = col[k] - gamma[ci(k, cc)] * col[k + 1];
      for (int k = 1; k < L; ++k) {
        const size_t i = ci(k, cc), im = ci(k - 1, cc);
        const T delta = s_rw[(size_t)k * C + cc] - c_rw[(size_t)k * C + cc];
        const T damp = dss[i];
        const T zi = fzm[k] * zz[i] + fzp[k] * zz[im];
        const T ri = fzm[k] * c_rho[i] + fzp[k] * c_rho[im];
        col[k] = (col[k] + delta - (((dts_r * damp) * zi) * ri) * w[(size_t)k * C + cc]) / (one + dts_r * damp) - delta;
      }
      for (int k = 0; k <= L; ++k) rw_p[(size_t)k * C + cc] = col[k];
      for (int k = 1; k < L; ++k) ww_avg[(size_t)k * C + cc] = ww_avg[(size_t)k * C + cc] + ewp[k] * col[k];
      for (int k = 0; k < L; ++k) {
        const size_t i = ci(k, cc);
        rho_pp[i] = rs[k] - (dts_r * rdzw[k]) * (ewp[k + 1] * col[k + 1] - ewp[k] * col[k]);
        rtheta_pp[i] = ts[k] - (dts_r * rdzw[k]) * ((ewp[k + 1] * coftz[ci(k + 1, cc)]) * col[k + 1] -
                                                     (ewp[k] * coftz[ci(k, cc)]) * col[k]);
      }
    }
  }

  // ---------------------------------------------------- divergence damping
  void divergence_damping(int stage) {
    const double dts = dyn[stage - 1].acoustic;
    const T base = (T(2) * T(cfg.config_smdiv)) * T(cfg.config_len_disp);
    const T coefficient = base * (T(1) / T(dts));
    const T one = T(1);
#pragma omp parallel for
    for (int e = 0; e < E; ++e) {
      const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
      for (int k = 0; k < L; ++k) {
        const T d0 = rtheta_pp[ci(k, c0)] - rtheta_pp_old[ci(k, c0)];
        const T d1 = rtheta_pp[ci(k, c1)] - rtheta_pp_old[ci(k, c1)];
        const T den = s_theta[ci(k, c0)] + s_theta[ci(k, c1)];
        ru_p[ei(k, e)] = ru_p[ei(k, e)] + ((coefficient * (d0 - d1)) * (one - T(0))) / den;
      }
    }
  }

  // --------------------------------------------------------------- recovery
  void recover(int stage) {
    const T inv_steps = T(1) / T(dyn[stage - 1].steps);
    const T two = T(2);
    for (size_t i = 0; i < cL(); ++i) {
      const T dp = s_rho_p[i] + rho_pp[i];
      const T rp = s_rtheta_p[i] + rtheta_pp[i];
      c_rho_p[i] = dp;
      c_rtheta_p[i] = rp;
      c_rho[i] = dp + rb[i];
      c_rt[i] = rp + rtb[i];
      if (!(c_rho[i] > T(0)) || !(c_rt[i] > T(0))) failed = 1;
    }
    for (size_t i = 0; i < eL(); ++i) {
      c_ru[i] = s_ru[i] + ru_p[i];
      flux_u[i] = s_ru[i] + ru_avg[i] * inv_steps;
    }
    for (size_t i = 0; i < wL(); ++i) {
      c_rw[i] = s_rw[i] + rw_p[i];
      flux_w[i] = s_rw[i] + ww_avg[i] * inv_steps;
    }
    for (int cc = 0; cc < C; ++cc) {
      c_rw[cc] = T(0);
      c_rw[(size_t)L * C + cc] = T(0);
    }
    for (size_t i = 0; i < cL(); ++i) c_theta[i] = c_rt[i] / c_rho[i];
    if (stage == 3) {
      const T cfac = T(kRgas / kP0);
      const T rcv = T(kRgas / (kCp - kRgas));
      const T R = T(kRgas);
      for (size_t i = 0; i < cL(); ++i) {
        const T arg = (zz[i] * cfac) * c_rt[i];
        if (!(arg > T(0))) failed = 1;
        const T ex = power<T>(arg, rcv);
        exner[i] = ex;
        pp[i] = (zz[i] * R) * (ex * c_rtheta_p[i] + rtb[i] * (ex - exb[i]));
      }
    }
    for (int e = 0; e < E; ++e) {
      const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
      for (int k = 0; k < L; ++k) u[ei(k, e)] = (two * c_ru[ei(k, e)]) / (c_rho[ci(k, c0)] + c_rho[ci(k, c1)]);
    }
    recover_w(c_rho, c_ru, c_rw, w);
  }

  void recover_w(const std::vector<T> &r, const std::vector<T> &r_u, const std::vector<T> &r_w, std::vector<T> &ww) {
    const T one = T(1);
#pragma omp parallel for
    for (int cc = 0; cc < C; ++cc) {
      T col[HEXDYN_ROW + 1] = {};
      for (int k = 1; k < L; ++k)
        col[k] = r_w[(size_t)k * C + cc] / (fzm[k] * zz[ci(k, cc)] + fzp[k] * zz[ci(k - 1, cc)]);
      for (int s = 0; s < neoc[cc]; ++s) {
        const int e = eoc[(size_t)cc * ME + s];
        const T sg = sgn(cc, s);
        T flux = (cf1 * r_u[ei(0, e)] + cf2 * r_u[ei(1, e)]) + cf3 * r_u[ei(2, e)];
        size_t zi = ((size_t)0 * C + cc) * ME + s;
        col[0] = col[0] + (sg * (zbc[zi] + copysign1(flux) * zb3c[zi])) * flux;
        for (int k = 1; k < L; ++k) {
          flux = fzm[k] * r_u[ei(k, e)] + fzp[k] * r_u[ei(k - 1, e)];
          zi = ((size_t)k * C + cc) * ME + s;
          col[k] = col[k] + (sg * (zbc[zi] + copysign1(flux) * zb3c[zi])) * flux;
        }
      }
      const T bottom = (cf1 * r[ci(0, cc)] + cf2 * r[ci(1, cc)]) + cf3 * r[ci(2, cc)];
      col[0] = col[0] / bottom;
      for (int k = 1; k < L; ++k) col[k] = col[k] / (fzm[k] * r[ci(k, cc)] + fzp[k] * r[ci(k - 1, cc)]);
      col[L] = T(0);
      (void)one;
      for (int k = 0; k <= L; ++k) ww[(size_t)k * C + cc] = col[k];
    }
  }

  // -------------------------------------------------------- subcycle end
  void end_subcycle(int subcycle) {
    theta_m = c_theta; rho_p = c_rho_p; rtheta_p = c_rtheta_p;
    if (subcycle == 1) {
      split_u = flux_u;
      split_w = flux_w;
    } else {
      for (size_t i = 0; i < eL(); ++i) split_u[i] = flux_u[i] + split_u[i];
      for (size_t i = 0; i < wL(); ++i) split_w[i] = flux_w[i] + split_w[i];
    }
  }

  // ------------------------------------------------------- scalar transport
  void scalars_begin() {
    const T inv = T(1) / T(cfg.config_dynamics_split_steps);
    for (size_t i = 0; i < eL(); ++i) split_u[i] = split_u[i] * inv;
    for (size_t i = 0; i < wL(); ++i) split_w[i] = split_w[i] * inv;
    q_stage = o_q;
  }

  T vflux(const T *qc, const std::vector<T> &wa, int cc, int k, T coef) const {
    const T wk = wa[(size_t)k * C + cc];
    if (k == 1 || k == L - 1) return wk * (fzm[k] * qc[ci(k, cc)] + fzp[k] * qc[ci(k - 1, cc)]);
    const T q0 = qc[ci(k - 2, cc)], q1 = qc[ci(k - 1, cc)], q2 = qc[ci(k, cc)], q3 = qc[ci(k + 1, cc)];
    const T base = (wk * (T(7) * (q2 + q1) - (q3 + q0))) / T(12);
    const T corr = ((coef * std::fabs(wk)) * ((q3 - q0) - T(3) * (q2 - q1))) / T(12);
    return base + corr;
  }

  void scalars_stage(int stage) {
    const T dt_s = T(sca[stage - 1].large);
    const T coef = T(cfg.config_coef_3rd_order);
    const T one = T(1);
    const std::vector<T> &uh = split_u;
    const std::vector<T> &wa = split_w;
    std::vector<T> out((size_t)NT * cL());
    if (stage < 3 || !cfg.config_monotonic) {
      T wn = one;
      if (cfg.config_split_dynamics_transport) {
        if (stage == 1) wn = T(1.0 / (cfg.config_time_integration_order == 3 ? 3.0 : 2.0));
        else if (stage == 2) wn = T(0.5);
      }
      std::vector<T> target(cL());
      for (size_t i = 0; i < cL(); ++i) target[i] = (one - wn) * o_rho[i] + wn * c_rho[i];
      for (int sp = 0; sp < NT; ++sp) {
        const T *qs = &q_stage[(size_t)sp * cL()];
        const T *qo = &o_q[(size_t)sp * cL()];
        std::vector<T> ev(eL());
#pragma omp parallel for
        for (int e = 0; e < E; ++e)
          for (int k = 0; k < L; ++k) {
            const T sv = copysign1(uh[ei(k, e)]);
            T r = T(0);
            for (int s = 0; s < nadv[e]; ++s) {
              const size_t j = (size_t)e * MA + s;
              r = r + (adv[j] + (coef * sv) * adv3[j]) * qs[ci(k, advc[j])];
            }
            ev[ei(k, e)] = r;
          }
#pragma omp parallel for
        for (int cc = 0; cc < C; ++cc) {
          T vf[HEXDYN_ROW + 1] = {};
          for (int k = 1; k < L; ++k) vf[k] = vflux(qs, wa, cc, k, coef);
          for (int k = 0; k < L; ++k) {
            T tc = T(0);
            for (int s = 0; s < neoc[cc]; ++s) {
              const int e = eoc[(size_t)cc * ME + s];
              tc = tc + ((-sgn(cc, s)) * uh[ei(k, e)]) * ev[ei(k, e)];
            }
            tc = tc * invarea[cc] + (physics_q ? phys_q[(size_t)sp * cL() + ci(k, cc)] : T(0));
            const size_t i = ci(k, cc);
            const T num = qo[i] * o_rho[i] + dt_s * (tc - rdzw[k] * (vf[k + 1] - vf[k]));
            out[(size_t)sp * cL() + i] = num * (one / target[i]);
          }
        }
      }
      q_stage = out;
      return;
    }
    // monotonic FCT (transport.py:1829-2175, v8.4.1 stored-inverse lane)
    std::vector<T> target(cL());
    if (cfg.config_split_dynamics_transport) {
      for (int cc = 0; cc < C; ++cc)
        for (int k = 0; k < L; ++k) {
          T hd = T(0);
          for (int s = 0; s < neoc[cc]; ++s) {
            const int e = eoc[(size_t)cc * ME + s];
            hd = hd + (((-sgn(cc, s)) * uh[ei(k, e)]) * dv[e]) * invarea[cc];
          }
          const size_t i = ci(k, cc);
          target[i] = o_rho[i] + dt_s * (hd - rdzw[k] * (wa[(size_t)(k + 1) * C + cc] - wa[(size_t)k * C + cc]));
          if (!(target[i] > T(0))) failed = 1;
        }
    } else {
      target = c_rho;
    }
    const T eps = T(1.0e-20);
    for (int sp = 0; sp < NT; ++sp) {
      const T *qs = &q_stage[(size_t)sp * cL()];
      std::vector<T> qo(cL()), mn(cL()), mx(cL()), mass(cL()), sin_(cL()), sout(cL()), vres(wL()), hf(eL()),
          upw(eL()), hres(eL());
      // old + dt*source/rho_old (transport.py:1906); a zero source adds +0
      for (size_t i = 0; i < cL(); ++i)
        qo[i] = o_q[(size_t)sp * cL() + i] + (physics_q ? (dt_s * phys_q[(size_t)sp * cL() + i]) / o_rho[i] : T(0));
#pragma omp parallel for
      for (int e = 0; e < E; ++e)
        for (int k = 0; k < L; ++k) {
          const T uu = uh[ei(k, e)];
          const T sv = copysign1(uu);
          T r;
          if (nadv[e] == 10) {
            const size_t j0 = (size_t)e * MA;
            T val = (adv[j0] + (coef * sv) * adv3[j0]) * qs[ci(k, advc[j0])];
            for (int s = 1; s < 10; ++s) {
              const size_t j = (size_t)e * MA + s;
              val = val + (adv[j] + (coef * sv) * adv3[j]) * qs[ci(k, advc[j])];
            }
            r = uu * val;
          } else {
            r = T(0);
            for (int s = 0; s < nadv[e]; ++s) {
              const size_t j = (size_t)e * MA + s;
              r = r + (uu * (adv[j] + (coef * sv) * adv3[j])) * qs[ci(k, advc[j])];
            }
          }
          hf[ei(k, e)] = r;
        }
#pragma omp parallel for
      for (int cc = 0; cc < C; ++cc) {
        for (int k = 0; k < L; ++k) {
          const int lo = std::max(0, k - 1), hi = std::min(L, k + 2);
          T a = qo[ci(lo, cc)], b = qo[ci(lo, cc)];
          for (int kk = lo + 1; kk < hi; ++kk) {
            a = pmin(a, qo[ci(kk, cc)]);
            b = pmax(b, qo[ci(kk, cc)]);
          }
          for (int s = 0; s < neoc[cc]; ++s) {
            const int nb = coc[(size_t)cc * ME + s];
            a = pmin(a, qo[ci(k, nb)]);
            b = pmax(b, qo[ci(k, nb)]);
          }
          mn[ci(k, cc)] = a;
          mx[ci(k, cc)] = b;
          mass[ci(k, cc)] = qo[ci(k, cc)] * o_rho[ci(k, cc)];
        }
        vres[cc] = dt_s * T(0);
        vres[(size_t)L * C + cc] = dt_s * T(0);
        for (int k = 1; k < L; ++k) vres[(size_t)k * C + cc] = dt_s * vflux(qs, wa, cc, k, coef);
        T fcol[HEXDYN_ROW + 1] = {};
        for (int k = 1; k < L; ++k) {
          const T vel = wa[(size_t)k * C + cc];
          fcol[k] = dt_s * (pmax0(vel) * qo[ci(k - 1, cc)] + pmin0(vel) * qo[ci(k, cc)]);
        }
        for (int k = 0; k < L - 1; ++k) mass[ci(k, cc)] = mass[ci(k, cc)] - fcol[k + 1] * rdzw[k];
        for (int k = 1; k < L; ++k) {
          mass[ci(k, cc)] = mass[ci(k, cc)] + fcol[k] * rdzw[k];
          vres[(size_t)k * C + cc] = vres[(size_t)k * C + cc] - fcol[k];
        }
        for (int k = 0; k < L; ++k) {
          const T up = vres[(size_t)(k + 1) * C + cc], dn = vres[(size_t)k * C + cc];
          sin_[ci(k, cc)] = (-rdzw[k]) * (pmin0(up) - pmax0(dn));
          sout[ci(k, cc)] = (-rdzw[k]) * (pmax0(up) - pmin0(dn));
        }
      }
#pragma omp parallel for
      for (int e = 0; e < E; ++e) {
        const int c1 = coe[2 * (size_t)e], c2 = coe[2 * (size_t)e + 1];
        for (int k = 0; k < L; ++k) {
          const T vel = uh[ei(k, e)];
          const T up = (dv[e] * dt_s) * (pmax0(vel) * qo[ci(k, c1)] + pmin0(vel) * qo[ci(k, c2)]);
          upw[ei(k, e)] = up;
          hres[ei(k, e)] = dt_s * hf[ei(k, e)] - up;
        }
      }
#pragma omp parallel for
      for (int cc = 0; cc < C; ++cc) {
        const T ia = invarea[cc];
        for (int s = 0; s < neoc[cc]; ++s) {
          const int e = eoc[(size_t)cc * ME + s];
          const T so = -sgn(cc, s);   // operator sign: first cell -1
          for (int k = 0; k < L; ++k) {
            const size_t i = ci(k, cc);
            mass[i] = mass[i] + (so * upw[ei(k, e)]) * ia;
            const T sr = (-so) * hres[ei(k, e)];
            sout[i] = sout[i] - pmax0(sr) * ia;
            sin_[i] = sin_[i] - pmin0(sr) * ia;
          }
        }
        for (int k = 0; k < L; ++k) {
          const size_t i = ci(k, cc);
          T f = (mx[i] * target[i] - mass[i]) / (sin_[i] + eps);
          sin_[i] = pmin(one, pmax0(f));
          f = (mn[i] * target[i] - mass[i]) / (sout[i] - eps);
          sout[i] = pmin(one, pmax0(f));
        }
      }
#pragma omp parallel for
      for (int e = 0; e < E; ++e) {
        const int c1 = coe[2 * (size_t)e], c2 = coe[2 * (size_t)e + 1];
        for (int k = 0; k < L; ++k) {
          const T f = hres[ei(k, e)];
          hres[ei(k, e)] = pmax0(f) * pmin(sout[ci(k, c1)], sin_[ci(k, c2)]) +
                           pmin0(f) * pmin(sin_[ci(k, c1)], sout[ci(k, c2)]);
        }
      }
#pragma omp parallel for
      for (int cc = 0; cc < C; ++cc) {
        for (int k = 1; k < L; ++k) {
          const T f = vres[(size_t)k * C + cc];
          vres[(size_t)k * C + cc] = pmax0(f) * pmin(sout[ci(k - 1, cc)], sin_[ci(k, cc)]) +
                                     pmin0(f) * pmin(sout[ci(k, cc)], sin_[ci(k - 1, cc)]);
        }
        const T ia = invarea[cc];
        for (int s = 0; s < neoc[cc]; ++s) {
          const int e = eoc[(size_t)cc * ME + s];
          const T so = -sgn(cc, s);
          for (int k = 0; k < L; ++k) mass[ci(k, cc)] = mass[ci(k, cc)] + (so * hres[ei(k, e)]) * ia;
        }
        for (int k = 0; k < L; ++k) {
          const size_t i = ci(k, cc);
          mass[i] = mass[i] - rdzw[k] * (vres[(size_t)(k + 1) * C + cc] - vres[(size_t)k * C + cc]);
          out[(size_t)sp * cL() + i] = pmax0(mass[i] / target[i]);
        }
      }
    }
    q_stage = out;
  }

  int post_clamp = 1;  // 0: transported scalars handed over unclamped (cuda_driver.py:5299-5311)

  void end_step() {
    if (cfg.config_scalar_advection && cfg.config_split_dynamics_transport && NT > 0) q = q_stage;
    // the q < 0 -> 0 where-assignment on every species (cuda_driver.py:5306-5311)
    if (post_clamp)
      for (auto &x : q)
        if (x < T(0)) x = T(0);
    rho = c_rho; rt = c_rt; ru = c_ru; rw = c_rw;
    time_s += cfg.config_dt;
    in_step = 0;
  }

  int step() {
    failed = 0;
    begin_step();
    for (int sub = 1; sub <= cfg.config_dynamics_split_steps; ++sub) {
      begin_subcycle();
      for (int st = 1; st <= 3; ++st) {
        dyn_tend(st);
        for (int ss = 1; ss <= dyn[st - 1].steps; ++ss) {
          acoustic_substep(st, ss);
          if (cfg.config_divergence_damping) divergence_damping(st);
        }
        recover(st);
        solve_diagnostics(u, st);
      }
      end_subcycle(sub);
    }
    if (cfg.config_scalar_advection && cfg.config_split_dynamics_transport && NT > 0) {
      scalars_begin();
      for (int st = 1; st <= 3; ++st) scalars_stage(st);
    }
    end_step();
    return failed ? HEXDYN_E_NONFINITE : 0;
  }

  void rebuild_saved() {
    const T cfac = T(kRgas / kP0);
    const T rcv = T(kRgas / (kCp - kRgas));
    const T R = T(kRgas);
    for (size_t i = 0; i < cL(); ++i) {
      theta_m[i] = rt[i] / rho[i];
      const T ex = power<T>((zz[i] * cfac) * rt[i], rcv);
      exner[i] = ex;
      rho_p[i] = rho[i] - rb[i];
      rtheta_p[i] = rt[i] - rtb[i];
      pp[i] = (zz[i] * R) * (ex * rtheta_p[i] + rtb[i] * (ex - exb[i]));
    }
    for (int e = 0; e < E; ++e) {
      const int c0 = coe[2 * (size_t)e], c1 = coe[2 * (size_t)e + 1];
      for (int k = 0; k < L; ++k) u[ei(k, e)] = (T(2) * ru[ei(k, e)]) / (rho[ci(k, c0)] + rho[ci(k, c1)]);
    }
    recover_w(rho, ru, rw, w);
  }

  std::vector<T> *field(int id) {
    switch (id) {
      case HEXDYN_F_RHO_ZZ: return in_step ? &c_rho : &rho;
      case HEXDYN_F_RTHETA_M: return in_step ? &c_rt : &rt;
      case HEXDYN_F_RU: return in_step ? &c_ru : &ru;
      case HEXDYN_F_RW: return in_step ? &c_rw : &rw;
      case HEXDYN_F_SCALARS: return &q;
      case HEXDYN_F_THETA_M: return in_step ? &c_theta : &theta_m;
      case HEXDYN_F_EXNER: return &exner;
      case HEXDYN_F_RHO_P: return in_step ? &c_rho_p : &rho_p;
      case HEXDYN_F_RTHETA_P: return in_step ? &c_rtheta_p : &rtheta_p;
      case HEXDYN_F_PRESSURE_P: return &pp;
      case HEXDYN_F_U: return &u;
      case HEXDYN_F_W: return &w;
      case HEXDYN_F_V: return &v;
      case HEXDYN_F_PV_EDGE: return &pv_edge;
      case HEXDYN_F_KE: return &ke;
      case HEXDYN_F_DIVERGENCE: return &divergence;
      case HEXDYN_F_VORTICITY: return &vorticity;
      case HEXDYN_F_TEND_RU: return &tend_ru;
      case HEXDYN_F_TEND_RT: return &tend_rt;
      case HEXDYN_F_TEND_RW: return &tend_rw;
      case HEXDYN_F_TEND_RHO: return &tend_rho;
      case HEXDYN_F_RU_P: return &ru_p;
      case HEXDYN_F_RW_P: return &rw_p;
      case HEXDYN_F_RHO_PP: return &rho_pp;
      case HEXDYN_F_RTHETA_PP: return &rtheta_pp;
      case HEXDYN_F_RU_AVG: return &ru_avg;
      case HEXDYN_F_WW_AVG: return &ww_avg;
      case HEXDYN_F_A_TRI: return &atri;
      case HEXDYN_F_ALPHA_TRI: return &alpha;
      case HEXDYN_F_GAMMA_TRI: return &gamma;
      case HEXDYN_F_COFWZ: return &cofwz;
      case HEXDYN_F_COFWT: return &cofwt;
      case HEXDYN_F_COFTZ: return &coftz;
      case HEXDYN_F_EULER_RU: return &euler_ru;
      case HEXDYN_F_EULER_RW: return &euler_rw;
      case HEXDYN_F_FLUX_U: return &flux_u;
      case HEXDYN_F_FLUX_W: return &flux_w;
      case HEXDYN_F_SCALARS_STAGE: return &q_stage;
      case HEXDYN_F_QTOT: return &qtot;
      case HEXDYN_F_CQW: return &cqw;
      case HEXDYN_F_CQU: return &cqu;
      case HEXDYN_F_KDIFF: return &kdiff;
      case HEXDYN_F_EULER_THETA: return &euler_theta;
      default: return nullptr;
    }
  }
};

// Working arrays by member name, for tests that feed the same operands to the
// reference's own operators (the HIP field ids expose only the published ones).
template <class O> static auto *internal(O &o, const std::string &n) {
  using V = typename std::remove_reference<decltype(o.rho)>::type;
  const std::pair<const char *, V *> table[] = {
      {"s_theta", &o.s_theta}, {"s_exner", &o.s_exner}, {"s_rho_p", &o.s_rho_p}, {"s_rtheta_p", &o.s_rtheta_p},
      {"s_rho", &o.s_rho}, {"s_rt", &o.s_rt}, {"s_ru", &o.s_ru}, {"s_rw", &o.s_rw},
      {"c_rho", &o.c_rho}, {"c_rt", &o.c_rt}, {"c_ru", &o.c_ru}, {"c_rw", &o.c_rw},
      {"c_theta", &o.c_theta}, {"c_rho_p", &o.c_rho_p}, {"c_rtheta_p", &o.c_rtheta_p},
      {"w", &o.w}, {"u", &o.u}, {"exner", &o.exner}, {"pp", &o.pp},
      {"cofwr", &o.cofwr}, {"rtheta_pp_old", &o.rtheta_pp_old}, {"tend_rho", &o.tend_rho},
      {"split_u", &o.split_u}, {"split_w", &o.split_w}, {"o_rho", &o.o_rho}, {"o_q", &o.o_q},
      {"dpdz", &o.dpdz}, {"fth", &o.fth_e}, {"corr", &o.corr_e}, {"fw", &o.fw_e}, {"hdiv", &o.hdiv},
      {"euler_ru", &o.euler_ru}, {"euler_rw", &o.euler_rw}, {"euler_theta", &o.euler_theta},
      {"kdiff", &o.kdiff}, {"v", &o.v}, {"divergence", &o.divergence}, {"vorticity", &o.vorticity},
      {"tend_ru", &o.tend_ru}, {"tend_rt", &o.tend_rt}, {"tend_rw", &o.tend_rw}};
  for (const auto &t : table)
    if (n == t.first) return t.second;
  return (V *)nullptr;
}

template <class O> static void set_physics(O &o, const float *rho, const float *ru, const float *rt,
                                           const float *q) {
  using T = typename std::remove_reference<decltype(o.rho[0])>::type;
  auto in = [](std::vector<T> &dst, const float *src, size_t n) {
    dst.assign(n, T(0));
    if (src)
      for (size_t i = 0; i < n; ++i) dst[i] = (T)src[i];
  };
  o.physics = (rho || ru || rt) ? 1 : 0;
  in(o.phys_rho, rho, o.cL());
  in(o.phys_ru, ru, o.eL());
  in(o.phys_rt, rt, o.cL());
  o.physics_q = q ? 1 : 0;
  in(o.phys_q, q, (size_t)o.NT * o.cL());
}

struct Handle {
  int precision;  // 32 or 64
  std::unique_ptr<Oracle<float>> f;
  std::unique_ptr<Oracle<double>> d;
};

template <class O> static void set_state(O &o, const hexdyn_state_view *s, int rebuild) {
  using T = typename std::remove_reference<decltype(o.rho[0])>::type;
  auto in = [](std::vector<T> &dst, const float *src) {
    if (src)
      for (size_t i = 0; i < dst.size(); ++i) dst[i] = (T)src[i];
  };
  in(o.rho, s->rho_zz); in(o.rt, s->rtheta_m); in(o.ru, s->ru); in(o.rw, s->rw); in(o.q, s->scalars);
  if (rebuild) {
    o.rebuild_saved();
  } else {
    in(o.theta_m, s->theta_m); in(o.exner, s->exner); in(o.rho_p, s->rho_p); in(o.rtheta_p, s->rtheta_p);
    in(o.pp, s->pressure_p); in(o.u, s->u); in(o.w, s->w);
  }
}

template <class O> static void get_state(O &o, const hexdyn_state_view *s) {
  auto out = [](const auto &src, float *dst) {
    if (dst)
      for (size_t i = 0; i < src.size(); ++i) dst[i] = (float)src[i];
  };
  out(o.rho, s->rho_zz); out(o.rt, s->rtheta_m); out(o.ru, s->ru); out(o.rw, s->rw); out(o.q, s->scalars);
  out(o.theta_m, s->theta_m); out(o.exner, s->exner); out(o.rho_p, s->rho_p); out(o.rtheta_p, s->rtheta_p);
  out(o.pp, s->pressure_p); out(o.u, s->u); out(o.w, s->w);
}

}  // namespace

#define DISPATCH(h, expr)                 \
  do {                                    \
    Handle *hh = (Handle *)(h);           \
    if (hh->precision == 32) {            \
      auto &o = *hh->f;                   \
      expr;                               \
    } else {                              \
      auto &o = *hh->d;                   \
      expr;                               \
    }                                     \
  } while (0)

extern "C" {

void *oracle_create(const hexdyn_mesh *m, const hexdyn_vertical *v, const hexdyn_config *c, int precision) {
  Handle *h = new Handle();
  h->precision = precision == 64 ? 64 : 32;
  int rc;
  if (h->precision == 32) {
    h->f.reset(new Oracle<float>());
    rc = h->f->init(m, v, c);
  } else {
    h->d.reset(new Oracle<double>());
    rc = h->d->init(m, v, c);
  }
  if (rc) {
    delete h;
    return nullptr;
  }
  return h;
}

void oracle_destroy(void *h) { delete (Handle *)h; }


int oracle_set_post_transport_clamp(void *h, int enable) {
  DISPATCH(h, o.post_clamp = enable ? 1 : 0);
  return 0;
}

int oracle_set_physics(void *h, const float *rho, const float *rho_u, const float *rho_theta, const float *q) {
  DISPATCH(h, set_physics(o, rho, rho_u, rho_theta, q));
  return 0;
}

int oracle_set_state(void *h, const hexdyn_state_view *s, int rebuild) {
  DISPATCH(h, set_state(o, s, rebuild));
  return 0;
}

int oracle_get_state(void *h, const hexdyn_state_view *s) {
  DISPATCH(h, get_state(o, s));
  return 0;
}

int oracle_export_field(void *h, int id, float *dst, int64_t n) {
  int rc = 0;
  DISPATCH(h, {
    auto *f = o.field(id);
    if (!f || (int64_t)f->size() != n) rc = 1;
    else for (int64_t i = 0; i < n; ++i) dst[i] = (float)(*f)[i];
  });
  return rc;
}

int oracle_import_field(void *h, int id, const float *src, int64_t n) {
  int rc = 0;
  DISPATCH(h, {
    auto *f = o.field(id);
    if (!f || (int64_t)f->size() != n) rc = 1;
    else for (int64_t i = 0; i < n; ++i) (*f)[i] = src[i];
  });
  return rc;
}

int64_t oracle_field_size(void *h, int id) {
  int64_t n = -1;
  DISPATCH(h, {
    auto *f = o.field(id);
    n = f ? (int64_t)f->size() : -1;
  });
  return n;
}

int oracle_step(void *h, int n) {
  int rc = 0;
  DISPATCH(h, for (int i = 0; i < n && rc == 0; ++i) rc = o.step());
  return rc;
}

int oracle_begin_step(void *h) { DISPATCH(h, { o.failed = 0; o.begin_step(); }); return 0; }
int oracle_begin_subcycle(void *h, int sub) { (void)sub; DISPATCH(h, o.begin_subcycle()); return 0; }
int oracle_dyn_tend(void *h, int stage) { DISPATCH(h, o.dyn_tend(stage)); return 0; }
int oracle_acoustic_substep(void *h, int stage, int ss) { DISPATCH(h, o.acoustic_substep(stage, ss)); return 0; }
int oracle_divergence_damping(void *h, int stage) { DISPATCH(h, o.divergence_damping(stage)); return 0; }
int oracle_recover(void *h, int stage) {
  int rc = 0;
  DISPATCH(h, { o.recover(stage); rc = o.failed ? HEXDYN_E_NONFINITE : 0; });
  return rc;
}
int oracle_solve_diagnostics(void *h, int rk) { DISPATCH(h, o.solve_diagnostics(o.u, rk)); return 0; }
int oracle_end_subcycle(void *h, int sub) { DISPATCH(h, o.end_subcycle(sub)); return 0; }
int oracle_scalars(void *h, int stage) {
  DISPATCH(h, {
    if (stage == 1) o.scalars_begin();
    o.scalars_stage(stage);
  });
  return 0;
}
int oracle_end_step(void *h) {
  int rc = 0;
  DISPATCH(h, { o.end_step(); rc = o.failed ? HEXDYN_E_NONFINITE : 0; });
  return rc;
}
float oracle_powf(float x, float y) { return power<float>(x, y); }

int oracle_set_plant(void *h, int code) {
  DISPATCH(h, o.plant = code);
  return 0;
}

// begin_subcycle without the horizontal mixing, then the mixing alone: the
// two halves of oracle_begin_subcycle (the tests compare the pools the mixing
// starts from and the increments it adds)
int oracle_begin_subcycle_pools(void *h) {
  DISPATCH(h, {
    o.pools_only = 1;
    o.begin_subcycle();
    o.pools_only = 0;
  });
  return 0;
}
int oracle_horizontal_mixing(void *h) {
  DISPATCH(h, if (o.mixing) o.horizontal_mixing());
  return 0;
}

// The momentum terms of every edge and level without the vertical transport
// and pools (dynamics.py:273-305's association) from the current u, rho_zz,
// pv_edge, kinetic energy and the mass-flux divergence of the current ru.
}  // extern "C"
template <class D> static int momentum_terms_out(void *h, D *out, int64_t n) {
  int rc = 0;
  DISPATCH(h, {
    if (n != (int64_t)o.eL()) {
      rc = 1;
    } else {
      o.mass_flux_divergence(o.c_ru, o.hdiv);
      for (int e = 0; e < o.E; ++e)
        for (int k = 0; k < o.L; ++k) out[o.ei(k, e)] = (D)o.momentum_terms(e, k, nullptr);
    }
  });
  return rc;
}
extern "C" {
int oracle_momentum_terms(void *h, float *out, int64_t n) { return momentum_terms_out(h, out, n); }
int oracle_momentum_terms_f64(void *h, double *out, int64_t n) { return momentum_terms_out(h, out, n); }

// The omega conversion applied to a given w tendency [L+1][C] with a given
// momentum tendency [L][E] (acoustic.py:246-281); out [L+1][C], rows 0 and L
// copied.
int oracle_omega_conversion(void *h, const float *w_tend, const float *u_tend, float *out) {
  DISPATCH(h, {
    using V = typename std::remove_reference<decltype(o.tend_ru)>::type;
    V tru(o.eL());
    for (size_t i = 0; i < o.eL(); ++i) tru[i] = u_tend[i];
    for (int cc = 0; cc < o.C; ++cc) {
      out[cc] = w_tend[cc];
      out[(size_t)o.L * o.C + cc] = w_tend[(size_t)o.L * o.C + cc];
      for (int k = 1; k < o.L; ++k) {
        using TT = typename V::value_type;
        out[(size_t)k * o.C + cc] = (float)o.omega(cc, k, (TT)w_tend[(size_t)k * o.C + cc], tru);
      }
    }
  });
  return 0;
}

// Test hooks: a working array by member name (size, export, import), and the
// implicit coefficients of the imported step-start exner, theta_m and
// rtheta_p for acoustic step dts.
int64_t oracle_internal_size(void *h, const char *name) {
  int64_t n = -1;
  DISPATCH(h, { auto *f = internal(o, name); n = f ? (int64_t)f->size() : -1; });
  return n;
}
int oracle_export_internal(void *h, const char *name, float *dst, int64_t n) {
  int rc = 0;
  DISPATCH(h, {
    auto *f = internal(o, name);
    if (!f || (int64_t)f->size() != n) rc = 1;
    else for (int64_t i = 0; i < n; ++i) dst[i] = (float)(*f)[i];
  });
  return rc;
}
int oracle_import_internal(void *h, const char *name, const float *src, int64_t n) {
  int rc = 0;
  DISPATCH(h, {
    auto *f = internal(o, name);
    if (!f || (int64_t)f->size() != n) rc = 1;
    else for (int64_t i = 0; i < n; ++i) (*f)[i] = src[i];
  });
  return rc;
}
// Test hook: the vertical transport of normal momentum (the first term of the
// momentum tendency) of the current u and rw, [L][E].
int oracle_vertical_transport_u(void *h, float *out, int64_t n) {
  int rc = 0;
  DISPATCH(h, {
    if (n != (int64_t)o.eL()) {
      rc = 1;
    } else {
      std::vector<std::remove_reference<decltype(o.u[0])>::type> col(o.L);
      for (int e = 0; e < o.E; ++e) {
        o.vertical_transport_u(e, col.data());
        for (int k = 0; k < o.L; ++k) out[(size_t)k * o.E + e] = (float)col[k];
      }
    }
  });
  return rc;
}

// The host C library's powf and the product's exner power method over many
// arguments (tests and golden-vector scripts compare them with the reference's
// GNU powf).
void oracle_libm_powf_many(const float *x, float y, float *out, int64_t n) {
#pragma omp parallel for
  for (int64_t i = 0; i < n; ++i) out[i] = ::powf(x[i], y);
}
void oracle_hexdyn_powf_many(const float *x, float y, float *out, int64_t n) {
#pragma omp parallel for
  for (int64_t i = 0; i < n; ++i) out[i] = hexdyn_powf_eval(x[i], y);
}

// RK schedule of integration.py:110-160 rounded to binary32 (large[3], acoustic[3], steps[3]).
void oracle_schedule(double dt, int order, int substeps, int splits, float *large, float *acoustic, int *steps) {
  Stage s[3];
  schedule(dt, order, substeps, splits, s);
  for (int i = 0; i < 3; ++i) {
    large[i] = (float)s[i].large;
    acoustic[i] = (float)s[i].acoustic;
    steps[i] = s[i].steps;
  }
}

// Tridiagonal coefficients b_tri and c_tri of the last vertical-coefficient evaluation.
int oracle_export_btri_ctri(void *h, float *b, float *c, int64_t n) {
  int rc = 0;
  DISPATCH(h, {
    if ((int64_t)o.btri.size() != n) rc = 1;
    else for (int64_t i = 0; i < n; ++i) { b[i] = (float)o.btri[i]; c[i] = (float)o.ctri[i]; }
  });
  return rc;
}
int oracle_export_btri_ctri_f64(void *h, double *b, double *c, int64_t n) {
  int rc = 0;
  DISPATCH(h, {
    if ((int64_t)o.btri.size() != n) rc = 1;
    else for (int64_t i = 0; i < n; ++i) { b[i] = (double)o.btri[i]; c[i] = (double)o.ctri[i]; }
  });
  return rc;
}
int oracle_export_internal_f64(void *h, const char *name, double *dst, int64_t n) {
  int rc = 0;
  DISPATCH(h, {
    auto *f = internal(o, name);
    if (!f || (int64_t)f->size() != n) rc = 1;
    else for (int64_t i = 0; i < n; ++i) dst[i] = (double)(*f)[i];
  });
  return rc;
}

int oracle_export_field_f64(void *h, int id, double *dst, int64_t n) {
  int rc = 0;
  DISPATCH(h, {
    auto *f = o.field(id);
    if (!f || (int64_t)f->size() != n) rc = 1;
    else for (int64_t i = 0; i < n; ++i) dst[i] = (double)(*f)[i];
  });
  return rc;
}
double oracle_model_time(void *h) {
  double t = 0;
  DISPATCH(h, t = o.time_s);
  return t;
}

}  // extern "C"
