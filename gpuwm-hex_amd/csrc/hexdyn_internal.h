/*
 * Internal device-side description of one dycore instance.
 *
 * Layout: every field is entity-major with the level index fastest, one
 * 64-float row (256 bytes) per entity, entities renumbered along a 3-D
 * Hilbert curve.  Interface fields (rw, w, ...) use rows 0..L of the same
 * 64-float row.  Scalars are [T][C][64].  Slot-indexed terrain metrics are
 * [C][maxEdges][64].  A wavefront owns one entity row and lane k owns level k.
 */
#ifndef HEXDYN_INTERNAL_H
#define HEXDYN_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <string>

// lanes of a level row (one wavefront); rows are stored Sizes::RS <= HX_ROW
// floats apart and allocated HX_ROW floats each
#define HX_ROW 64

namespace hexdyn {

struct Sizes {
  int C, E, V, L, ME, ME2, MA, NT, NM;
  int RS;  // floats from one level row to the next: L + 1 rounded up to 8 (at most HX_ROW)
  int TU;  // distinct advection cells per edge tile (0: the tiled flux gathers are off)
  int OCP;  // 1: in every cell, cellsOnCell slot s is the other cell of edgesOnCell slot s
  // wavefronts (rows) per block of the launch in flight: each launch passes
  // its own copy, so a kernel finds its tile with its other scalar arguments
  int WPB = 4;
};

// Edge tiles of the staged flux gathers: HX_TILE_E edges per block, at most
// HX_TILE_U distinct advection cells per tile (larger tiles turn the staged
// kernels off for the mesh).
constexpr int HX_TILE_E = 8;
constexpr int HX_TILE_U = 56;

struct Mesh {
  // connectivity (renumbered)
  const int *coe;    // [E][2]
  const int *voe;    // [E][2]
  const int *eoc;    // [C][ME]
  const int *coc;    // [C][ME]
  const int *voc;    // [C][ME]
  const int *neoc;   // [C]
  const int *kfc;    // [C][ME]
  const int *eov;    // [V][3]
  const int *eoe;    // [E][ME2]
  const int *neoe;   // [E]
  const int *nadv;   // [E]
  const int *advc;   // [E][MA]
  // tiles of HX_TILE_E consecutive edges: the distinct cells of their
  // advection stencils (padded with the tile's first cell) and each edge
  // slot's index in that list
  const int *tcell;  // [E / HX_TILE_E rounded up][HX_TILE_U]
  const int *ntcell; // [E / HX_TILE_E rounded up] distinct cells of the tile
  const int *advl;   // [E][MA]
  // geometry
  const float *csign;   // [C][ME] +1 first cell of the edge, -1 otherwise
  const float *vsign;   // [V][3]
  const float *woe;     // [E][ME2]
  const float *dc, *dv, *dcdv, *invdc, *invdv;  // [E]
  const float *invarea; // [C]
  const float *invtri;  // [V]
  const float *kite;    // [V][3]
  const float *fedge, *fvert, *cosang, *sinang;
  const float *adv, *adv3;  // [E][MA]
  const uint32_t *powf_ties;  // GNU powf correction bitmap of the exner exponent (hexdyn_powf.h)
  // per-slot copies of edge geometry, so a kernel reads them with the
  // connectivity instead of through it (csign and vsign are +-1, so the
  // signed products are exact)
  const float *csdv;    // [C][ME] csign * dv[edge]
  const float *cdcdv;   // [C][ME] dcdv[edge]
  const float *ckite;   // [C][ME] kite area of the cell's slot vertex
  const float *vdc;     // [V][3]  vsign * dc[edge]
  const float *vdcdv;   // [V][3]  dcdv[edge]
  const float *ecos, *esin;  // [E][ME2] cos/sin angle of edgesOnEdge
  const int *ocell;     // [C][ME] the other cell of the slot's edge
  // horizontal mixing (binary32 factors in the reference's operation order)
  const float *defc2, *defs2, *defcs;  // [C][ME] deformation weights
  const float *mx_fdiv;   // [C][ME] ((1/area)*dv)*sign                        divergence of delsq_u
  const float *mx_fth2;   // [C][ME] (((1/area)*sign)*dv)*(1/dc)                theta del2
  const float *mx_fw2;    // [C][ME] (((0.5*(1/area))*sign)*dv)*(1/dc)          w del2
  const float *mx_f4;     // [C][ME] ((((del4*(h4*(1/area)))*dv)*sign)*(1/dc)   theta and w del4
  const float *mx_fvort;  // [V][3]  ((1/areaTriangle)*dc)*vertex sign          vorticity of delsq_u
  const float *mx_idc, *mx_idv, *mx_rdc, *mx_rdv, *mx_del2;  // [E]
};

struct Vert {
  const float *rdzw, *rdzu, *fzm, *fzp, *etp, *etm, *ewp, *ewm, *uinit, *vinit;  // [64] each, padded
  float cf1, cf2, cf3;
  const float *zz, *dss, *rb, *rtb, *exb;  // cell rows
  const float *zxu;                        // edge rows
  const float *zb, *zb3;                   // [C][ME][64]
  const float *zmid;                       // cell rows: layer-mid heights (null without zgrid)
};

// Device buffers of the working set (all row layout).
struct Work {
  // published (outer) state
  float *rho, *rt, *ru, *rw, *q;
  // carried diagnostics (time level one), working copies updated in place during a step
  float *theta, *exner, *rho_p, *rtheta_p, *pp, *u, *w;
  // published carried diagnostics
  float *p_theta, *p_exner, *p_rho_p, *p_rtheta_p, *p_pp, *p_u, *p_w;
  // current state
  float *c_rho, *c_rt, *c_ru, *c_rw;
  // subcycle start copies
  float *s_rho, *s_rt, *s_ru, *s_rw, *s_theta, *s_rho_p, *s_rtheta_p, *s_exner;
  // moist coefficients
  float *qtot, *cqw, *cqu;
  // implicit coefficients
  float *cofwz, *cofwt, *coftz, *atri, *alpha, *gamma;
  // subcycle pools
  float *dpdz, *euler_ru, *euler_rw, *tend_rho;
  // stage tendencies and edge fluxes
  float *tend_ru, *tend_rt, *tend_rw, *hdiv, *fth, *fcorr, *fw;
  // acoustic
  float *ru_p, *ru_avg, *rw_p, *ww_avg, *rho_pp, *rtheta_pp, *rtheta_pp_old;
  // diagnostics
  float *v, *pv_edge, *ke, *divergence, *vorticity, *ke_vertex, *pv_vertex, *pv_cell;
  // fluxes
  float *flux_u, *flux_w, *split_u, *split_w;
  // transport
  float *q_stage, *q_next, *target, *ev, *fct_mn, *fct_mx, *fct_mass, *fct_sin, *fct_sout, *fct_vres,
      *fct_hres;
  // horizontal mixing: viscosity, Laplacians and the theta Euler pool
  float *kdiff, *delsq_u, *delsq_div, *delsq_vort, *delsq_theta, *delsq_w, *mix_tw, *euler_theta;
  // held physics tendencies (rows; allocated when first installed)
  float *phys_rho, *phys_ru, *phys_rt;
  float *phys_q, *q_src;   // held scalar tendency [T][C][64]; old scalars + dt*source/rho_old
  int *flag;  // sticky validation flag
};

// Every member is 4 bytes except the leading double, so the struct has no
// padding and its bytes identify its value (the step-graph cache keys on them).
struct Consts {
  double powf_near_tie;                // GNU powf correction: bitmap read only this close to a tie (ulps)
  float gravity, rcv, c2, cfac, rgas;  // binary32 constants in reference form
  float coef3, apvm_scale, divdamp_base;
  float dts[3];            // acoustic timestep per stage
  float dts_inv_coef[3];   // divergence damping coefficient per stage
  float inv_steps[3];      // 1/steps per stage
  int steps[3];
  float sca_dt[3];         // scalar transport stage timesteps
  float wn[3];             // split-transport density weights, stages 1-2
  float split_inv;         // 1/splits
  int order, splits, monotonic, divdamp, split_transport, scalar_adv;
  int apvm_on;
  int mixing, mix4;              // Smagorinsky mixing on; fourth-order part on (h4 > 0)
  int physics;                   // held physics tendencies are added
  int physics_q;                 // a held scalar tendency enters the transport
  int store_flux;                // set by every phase entry point, clear on the whole-step path:
                                 // kernels also store what only exports read between phases (the
                                 // stage's flux_u/flux_w, which stage 3 folds into the split sums;
                                 // rho_p/rtheta_m at stages 1-2, rtheta_p at stage 2; divergence and
                                 // vorticity of rk 1-2; the damped ru_p of stages 1-2; atri)
  int fused;                     // whole-step launch sequence: the next stage's theta/w edge fluxes
                                 // ride on the edge diagnostics, the first acoustic edge update on
                                 // the momentum tendency, and ru_avg is kept only where read back
  float strain_scale, ceiling;   // (c_s*len)^2, (0.01*len^2)/dt
  int ray_levels;                // Rayleigh damping of u over the top ray_levels levels (0: off)
  float ray_inverse;             // 1/(levels*(days*86400))
  int cam_levels;                // CAM kdiff floor over the top cam_levels levels (0: off)
  float cam_base;                // ((4*2.0833)*len)*cam_coef
};
static_assert(sizeof(Consts) == sizeof(double) + 46 * 4, "Consts must have no padding");

// Optional per-kernel timing (events around each launch on its stream).
void prof_mark(const char *name, hipStream_t s, int start);
extern int g_prof_on;

// Launch settings of the gather kernels (host side; kernels never read it):
// wavefronts (entities) per block, 1, 2, 4, 8 or 16, by kernel name.
struct LaunchTune {
  static constexpr int kMax = 64;
  int n = 0;
  int def = 4;  // kernels not listed
  char names[kMax][32];
  int wpb[kMax];
  int find(const char *name) const {
    for (int i = 0; i < n; ++i)
      if (std::strncmp(names[i], name, 31) == 0) return i;
    return -1;
  }
  int waves_per_block(const char *name) const {
    const int i = find(name);
    return i < 0 ? def : wpb[i];
  }
};

struct Instance {
  Sizes n;
  Mesh m;
  Vert z;
  Work b;
  Consts k;
  const LaunchTune *tune;  // host pointer to the handle's launch table
};

// A second stream (default priority) for the implicit coefficients, which
// depend on none of the kernels launched beside them: forked from the
// caller's stream and joined back to it with the two events, so in a
// captured step they are a parallel branch of the graph.  Null: one stream.
struct Side {
  hipStream_t s;
  hipEvent_t fork, join;
};

// Phase entry points, compiled once with strict binary32 numerics
// (no contraction) and once with fused multiply-adds allowed.
#define HEXDYN_PHASES(NS)                                                 \
  namespace NS {                                                          \
  int begin_step(const Instance &I, hipStream_t s);                       \
  int begin_subcycle(const Instance &I, int sub, hipStream_t s, const Side *side = nullptr); \
  int dyn_tend(const Instance &I, int stage, hipStream_t s, const Side *side = nullptr);     \
  int acoustic_substep(const Instance &I, int stage, int ss, int damp_prev, hipStream_t s); \
  int divergence_damping(const Instance &I, int stage, hipStream_t s);    \
  int recover(const Instance &I, int stage, int sub, int damp, hipStream_t s); \
  int solve_diagnostics(const Instance &I, int rk, hipStream_t s);        \
  int scalars_stage(const Instance &I, int stage, hipStream_t s);         \
  int rebuild_saved(const Instance &I, hipStream_t s);                    \
  int rebuild_cells(const Instance &I, hipStream_t s);                    \
  int clamp_scalars(const Instance &I, const float *src, float *dst, hipStream_t s); \
  }

HEXDYN_PHASES(strict)
HEXDYN_PHASES(fast)

// FP64 global integrals (hexdyn_integrals.hip)
int integrals(const Instance &I, const float *rho, const float *rt, const float *q, const float *theta,
              const float *exner, const float *u, const float *zmid, double *partial, double *out,
              hipStream_t s);

// Learned large-step corrector (hexdyn_net.hip).  Every entry point works on
// the published state rows of an Instance; the mesh connectivity must be the
// one the network was created with.
namespace net {
struct Net;
Net *create(const Instance &I, int layers, std::string *err);
void destroy(Net *n);
size_t param_count(const Net *n);
int set_params(Net &n, const float *host, hipStream_t s);
int get_vec(const Net &n, int which, float *host, hipStream_t s);  // 0 params, 1 grads
int set_scales(Net &n, const float *in_mean, const float *in_scale, const float *out_scale, hipStream_t s);
int forward(Net &n, const Instance &I, bool keep, hipStream_t s);
int target(Net &n, const Instance &Ic, const Instance &If, hipStream_t s);
int loss(Net &n, const Instance &I, double *out5, hipStream_t s);
int backward(Net &n, const Instance &I, hipStream_t s);
int adam(Net &n, float lr, float b1, float b2, float eps, int t, hipStream_t s);
int apply(Net &n, const Instance &I, float gain, hipStream_t s);
int checks(Net &n, const Instance &I, float dt, float out[3], double *mass, hipStream_t s);
int raw_features(Net &n, const Instance &I, const int *old_c, float *dst, hipStream_t s);
int export_buf(Net &n, int what, const int *old_c, const int *old_e, float *dst, hipStream_t s);
}  // namespace net

}  // namespace hexdyn

#endif
