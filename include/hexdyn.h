/*
 * hexdyn — MPAS-A v8.4.2 split-explicit nonhydrostatic dycore large step on
 * AMD MI355X (gfx950).  C ABI of libhexdyn.so.
 *
 * The library replaces the reference's dycore step object
 * (tree/src/hexcore/cuda_driver.py CudaDryDycoreDriver):
 *   hexdyn_create            <- CudaDryDycoreDriver.from_host   (cuda_driver.py:2826-3025)
 *   hexdyn_set_state         <- DeviceAtmosphere.from_host      (cuda_backend/containers.py:1108)
 *   hexdyn_get_state         <- step() download half            (cuda_driver.py:6112-6149)
 *   hexdyn_step              <- _step_device_v841 / step_device (cuda_driver.py:4854-5537, :5751)
 *   hexdyn_begin_step        <- moist coefficients + first solve_diagnostics (cuda_driver.py:5049-5070, :4232-4245)
 *   hexdyn_begin_subcycle    <- subcycle-start pools and 2-D Smagorinsky mixing (cuda_driver.py:4232-4383,
 *                               cuda_horizontal_v841.py:1270-1588)
 *   hexdyn_dyn_tend          <- stage tendencies                (cuda_driver.py:4388-4509)
 *   hexdyn_acoustic_substep  <- advance_acoustic_step_cuda_v841 (cuda_acoustic_v841.py:801-1040)
 *   hexdyn_divergence_damping<- divergence_damping_3d           (cuda_horizontal.py:1584-1660)
 *   hexdyn_recover           <- _recover_candidate + validation (cuda_driver.py:3808-3960, :4692-4718)
 *   hexdyn_solve_diagnostics <- CudaHorizontalV841.solve_diagnostics (cuda_horizontal_v841.py:1041-1236)
 *   hexdyn_end_subcycle      <- accumulate_split_flux_cuda_v841 (cuda_dynamics_v841.py:412-442)
 *   hexdyn_scalars           <- advance_scalar_transport_cuda_v841 (cuda_transport_v841.py:1988)
 *   hexdyn_end_step          <- finish_split_flux / clamp / validation read / publish (cuda_driver.py:5147-5403)
 *
 * Conventions.  Every array crossing this ABI is in MPAS logical order,
 * level first and entity fastest ([L][entity], [L+1][C] for interface fields,
 * [T][L][C] for scalars), FP32, 0-based int32 connectivity with -1 padding.
 * Static mesh/vertical/reference data are host pointers copied at create
 * time.  State views are device pointers.  Internally the library renumbers
 * entities along a space-filling curve and stores every field entity-major
 * with the level index fastest, entity rows L+1 rounded up to 8 floats
 * apart (56 at 55 levels, at most 64).
 *
 * Return codes: 0 ok, >0 refusal or numerical flag, <0 HIP error.  A set
 * validation flag makes hexdyn_step/hexdyn_end_step return
 * HEXDYN_E_NONFINITE and leaves the published state untouched.
 * Calls on one handle must not overlap; separate handles are independent.
 */
#ifndef HEXDYN_H
#define HEXDYN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HEXDYN_ABI_VERSION 4
#define HEXDYN_MAX_EDGES 7      /* nEdgesOnCell limit (stencil widths derive from it) */
#define HEXDYN_ROW 64           /* padded level row; nVertLevels + 1 <= 64 */

enum {
  HEXDYN_OK = 0,
  HEXDYN_E_INVALID = 1,       /* malformed argument */
  HEXDYN_E_NONFINITE = 2,     /* validation flag: non-finite or non-positive state */
  HEXDYN_E_REFUSED = 3,       /* configuration refusal */
  HEXDYN_E_HIP = -1
};

enum { HEXDYN_NUMERICS_STRICT = 0, HEXDYN_NUMERICS_FAST = 1 };
enum { HEXDYN_MIXING_OFF = 0, HEXDYN_MIXING_2D_SMAGORINSKY = 1 };

typedef struct hexdyn_mesh {
  int32_t n_cells, n_edges, n_vertices;
  int32_t max_edges;       /* row width of edgesOnCell/cellsOnCell/verticesOnCell */
  int32_t vertex_degree;   /* must be 3 */
  int32_t max_edges2;      /* row width of edgesOnEdge/weightsOnEdge */
  int32_t max_adv;         /* row width of advCellsForEdge/adv_coefs */
  const int32_t *cells_on_edge;        /* [E][2]  normal points cell 0 -> cell 1 */
  const int32_t *vertices_on_edge;     /* [E][2]  vertex 1 on the left of the normal */
  const int32_t *edges_on_cell;        /* [C][max_edges] counter-clockwise */
  const int32_t *cells_on_cell;        /* [C][max_edges] */
  const int32_t *vertices_on_cell;     /* [C][max_edges] read only through kite_for_cell */
  const int32_t *n_edges_on_cell;      /* [C] */
  const int32_t *kite_for_cell;        /* [C][max_edges] slot of the cell in cellsOnVertex */
  const int32_t *edges_on_vertex;      /* [V][3] */
  const int32_t *cells_on_vertex;      /* [V][3] */
  const int32_t *edges_on_edge;        /* [E][max_edges2] */
  const int32_t *n_edges_on_edge;      /* [E] */
  const float *weights_on_edge;        /* [E][max_edges2] */
  const float *dc_edge, *dv_edge;      /* [E] */
  const float *inv_dc_edge, *inv_dv_edge;          /* [E] stored inverses */
  const float *area_cell, *inv_area_cell;          /* [C] */
  const float *area_triangle, *inv_area_triangle;  /* [V] */
  const float *kite_areas_on_vertex;   /* [V][3] */
  const float *f_edge, *f_vertex;      /* [E], [V] */
  const float *cos_angle_edge, *sin_angle_edge;    /* [E] */
  const int32_t *n_adv_cells_for_edge; /* [E] */
  const int32_t *adv_cells_for_edge;   /* [E][max_adv] */
  const float *adv_coefs, *adv_coefs_3rd;          /* [E][max_adv] */
  const double *x_cell, *y_cell, *z_cell;          /* [C] positions for locality ordering (may be NULL) */
  /* Horizontal mixing inputs (read only when config_horiz_mixing is on):
   * meshDensity (NULL reads as 1) and the v8.4.1 deformation weights
   * deformation_coef_c2/s2/cs in edgesOnCell slot order, signed by the
   * cell's side of each edge (mixing_v841.py:372-455). */
  const float *mesh_density;           /* [C] */
  const float *deformation_coef_c2;    /* [C][max_edges] */
  const float *deformation_coef_s2;    /* [C][max_edges] */
  const float *deformation_coef_cs;    /* [C][max_edges] */
} hexdyn_mesh;

typedef struct hexdyn_vertical {
  int32_t n_levels;                    /* L, with L + 1 <= HEXDYN_ROW */
  const float *zz;                     /* [L][C] */
  const float *zxu;                    /* [L][E] */
  const float *rdzw, *rdzu, *fzm, *fzp;/* [L] */
  float cf1, cf2, cf3;
  const float *zb_cell, *zb3_cell;     /* [L+1][C][max_edges] */
  const float *dss;                    /* [L][C] Rayleigh w damping */
  const float *etp, *etm;              /* [L] off-centring */
  const float *ewp, *ewm;              /* [L+1] */
  const float *u_init, *v_init;        /* [L] reference winds */
  const float *rho_base, *rtheta_base, *exner_base; /* [L][C] */
  const float *zgrid;                  /* [L+1][C] interface heights (optional: potential energy) */
} hexdyn_vertical;

typedef struct hexdyn_config {
  double config_dt;
  int32_t config_time_integration_order;   /* 2 or 3 */
  int32_t config_number_of_sub_steps;      /* >= 1 */
  int32_t config_dynamics_split_steps;     /* 1 or 3 (3 needs split transport) */
  int32_t config_split_dynamics_transport; /* bool */
  int32_t config_scalar_advection;         /* bool */
  int32_t config_monotonic;                /* bool */
  int32_t config_divergence_damping;       /* bool */
  float config_coef_3rd_order;
  float config_apvm_upwinding;             /* >= 0 */
  float config_smdiv;                      /* >= 0 */
  float config_len_disp;                   /* > 0 (resolved nominal dc) */
  int32_t n_scalars;                       /* T */
  int32_t n_moist;                         /* leading mass-loading species, 0 = dry */
  int32_t numerics;                        /* HEXDYN_NUMERICS_* */
  /* 2-D Smagorinsky horizontal mixing (mixing_v841.py:71-85 V841MixingConfig);
   * config_len_disp above is its filter length too. */
  int32_t config_horiz_mixing;             /* HEXDYN_MIXING_* */
  float config_visc4_2dsmag;               /* >= 0 */
  float config_smagorinsky_coef;           /* >= 0 */
  float config_del4u_div_factor;           /* > 0 */
  int32_t config_h_ScaleWithMesh;          /* bool */
  /* CAM viscosity floor on kdiff over the top levels (mixing_v841.py:746-750,
   * cuda_upper_damping_v841.py:5-11): kdiff = max(kdiff, base*(1-(L-1-k)/n)),
   * base = ((4*2.0833)*config_len_disp)*config_mpas_cam_coef. */
  float config_mpas_cam_coef;              /* >= 0, 0 = off */
  int32_t config_number_cam_damping_levels;        /* n in 1..L when the floor is on (MPAS default 4) */
  /* Upper-level Rayleigh damping of the normal wind on every stage's
   * momentum tendency, before the Euler pool is added (cuda_driver.py:4427-4433,
   * native/upper_damping.rs:18-31): tend_ru -= (rho_edge*u)*((k-(L-n)+1)*inv),
   * inv = 1/(n*(days*86400)), for the top n levels. */
  int32_t config_rayleigh_damp_u;                  /* bool */
  int32_t config_number_rayleigh_damp_u_levels;    /* n in 1..L when on (MPAS default 6) */
  float config_rayleigh_damp_u_timescale_days;     /* > 0 when on (MPAS default 5) */
} hexdyn_config;

/* Prognostic state and the carried time-level-one diagnostics
 * (state.py:14-27 PrognosticState; driver.py:661 DrySavedDiagnostics). */
typedef struct hexdyn_state_view {
  float *rho_zz;     /* [L][C] */
  float *rtheta_m;   /* [L][C] */
  float *ru;         /* [L][E] */
  float *rw;         /* [L+1][C] */
  float *scalars;    /* [T][L][C] */
  float *theta_m;    /* [L][C] */
  float *exner;      /* [L][C] */
  float *rho_p;      /* [L][C] density perturbation */
  float *rtheta_p;   /* [L][C] */
  float *pressure_p; /* [L][C] */
  float *u;          /* [L][E] normal velocity */
  float *w;          /* [L+1][C] vertical velocity */
} hexdyn_state_view;

/* Internal fields reachable through hexdyn_export_field/hexdyn_import_field
 * (logical layout, device pointers).  Used by parity tests to compare every
 * phase against the CPU restatement.  The working fields (after
 * HEXDYN_F_W) hold what the phase entry points leave; hexdyn_step stores
 * only what a later kernel of the step reads, so after it FLUX_U, FLUX_W,
 * A_TRI and the diagnostics (V, PV_EDGE, KE, DIVERGENCE, VORTICITY, which
 * the next step's start recomputes from the published state) are not
 * current. */
enum {
  HEXDYN_F_RHO_ZZ = 0, HEXDYN_F_RTHETA_M, HEXDYN_F_RU, HEXDYN_F_RW, HEXDYN_F_SCALARS,
  HEXDYN_F_THETA_M, HEXDYN_F_EXNER, HEXDYN_F_RHO_P, HEXDYN_F_RTHETA_P, HEXDYN_F_PRESSURE_P,
  HEXDYN_F_U, HEXDYN_F_W,
  HEXDYN_F_V, HEXDYN_F_PV_EDGE, HEXDYN_F_KE, HEXDYN_F_DIVERGENCE, HEXDYN_F_VORTICITY,
  HEXDYN_F_TEND_RU, HEXDYN_F_TEND_RT, HEXDYN_F_TEND_RW, HEXDYN_F_TEND_RHO,
  HEXDYN_F_RU_P, HEXDYN_F_RW_P, HEXDYN_F_RHO_PP, HEXDYN_F_RTHETA_PP, HEXDYN_F_RU_AVG, HEXDYN_F_WW_AVG,
  HEXDYN_F_A_TRI, HEXDYN_F_ALPHA_TRI, HEXDYN_F_GAMMA_TRI, HEXDYN_F_COFWZ, HEXDYN_F_COFWT, HEXDYN_F_COFTZ,
  HEXDYN_F_EULER_RU, HEXDYN_F_EULER_RW,
  HEXDYN_F_FLUX_U, HEXDYN_F_FLUX_W, HEXDYN_F_SCALARS_STAGE,
  HEXDYN_F_QTOT, HEXDYN_F_CQW, HEXDYN_F_CQU,
  HEXDYN_F_KDIFF, HEXDYN_F_EULER_THETA,
  HEXDYN_F_COUNT
};

/* Physics tendencies held through the dynamics of the following steps until
 * replaced or cleared: the sources the reference's physics_stage_callback
 * assembles at a step's first stage and holds through every RK and dynamics
 * substep (cuda_driver.py:4388-4400).  They enter as tend_rho + rho at the
 * subcycle start (density_tendency_f32), tend_ru = (tend_ru + euler_ru) +
 * rho_u (add_pair_inplace_f32, :4443-4446) and tend_rt = (tend_rt [+ mixing
 * theta pool]) + rho_theta (:4484-4494).  Device pointers, logical layout;
 * NULL members are zero. */
typedef struct hexdyn_physics_view {
  const float *rho;        /* [L][C] */
  const float *rho_u;      /* [L][E] */
  const float *rho_theta;  /* [L][C] */
  const float *scalars;    /* [T][L][C], the transport's scalar_tendency (cuda_driver.py:5162-5163):
                              added to the flux divergence at transport stages 1-2 and to the old
                              mixing ratio as dt*source/rho_old before the limiter */
} hexdyn_physics_view;

typedef struct hexdyn_handle hexdyn_handle;

int hexdyn_abi_version(void);
int hexdyn_create(const hexdyn_mesh *mesh, const hexdyn_vertical *vertical,
                  const hexdyn_config *config, hexdyn_handle **out);
void hexdyn_destroy(hexdyn_handle *h);
int hexdyn_last_error(const hexdyn_handle *h, char *buf, size_t len);
/* Why the calling thread's last hexdyn_create failed (the handle of a
 * refused create no longer exists); empty after a successful create. */
int hexdyn_last_create_error(char *buf, size_t len);

/* State transfer (device pointers, logical layout).  With rebuild != 0 the
 * saved diagnostics are recomputed from the prognostic fields
 * (driver.py:2655 _rebuild_saved_diagnostics) and the view's diagnostic
 * pointers are ignored. */
int hexdyn_set_state(hexdyn_handle *h, const hexdyn_state_view *state, int rebuild, void *stream);
int hexdyn_get_state(hexdyn_handle *h, const hexdyn_state_view *state, void *stream);
int hexdyn_export_field(hexdyn_handle *h, int field, float *dst, void *stream);
int hexdyn_import_field(hexdyn_handle *h, int field, const float *src, void *stream);
int hexdyn_field_shape(const hexdyn_handle *h, int field, int64_t *rows, int64_t *cols);

/* Whole large steps.  Work is enqueued on `stream` (a hipStream_t, NULL for
 * the default stream); the call synchronises once per step to read the
 * 4-byte validation flag. */
int hexdyn_step(hexdyn_handle *h, int n_steps, void *stream);
double hexdyn_model_time(const hexdyn_handle *h);

/* hexdyn_step replays each large step as a captured hipGraph (default on;
 * one graph per buffer layout, captured on first use; not used while
 * per-kernel profiling is enabled).  enable = 0 launches kernel by kernel.
 * (The reference captures CUDA graphs per subcycle, cuda_subcycle_graphs.py.) */
int hexdyn_set_graphs(hexdyn_handle *h, int enable);
/* Number of captured step graphs the handle holds (one per buffer layout). */
int hexdyn_graph_count(const hexdyn_handle *h);

/* Launch shape of the gather kernels (one wavefront per cell, edge or
 * vertex row): wavefronts per block, 1, 2, 4 (the default), 8 or 16, by kernel name
 * ("k_diag_edge_flux", ...; "*" resets every kernel to the given value).
 * The column kernels (k_vcoef, k_ac_cell: 16 and 8 columns per block through LDS)
 * and the edge-tiled flux pass (k_diag_edge_flux_t: 8 edges per block
 * through LDS) are refused; setting k_diag_edge_flux selects that pass's
 * gather form (one block of whole-edge wavefronts, no LDS) instead of the
 * tiled one.  Results do not depend on it.  Drops the captured step graphs.
 * The autotuner (hexdyn/autotune.py) picks it per kernel by timing.  (The
 * reference tunes its level chunking instead, cuda_edge_hoist.py:63-101.) */
int hexdyn_set_launch(hexdyn_handle *h, const char *kernel, int waves_per_block);
/* the waves per block a kernel is launched with (-1 for a null argument) */
int hexdyn_get_launch(const hexdyn_handle *h, const char *kernel);

/* Install (view != NULL) or clear (view == NULL) held physics tendencies. */
int hexdyn_set_physics_tendencies(hexdyn_handle *h, const hexdyn_physics_view *view, void *stream);

/* enable = 0 hands the transported scalars over unclamped: the end of a step
 * skips the q < 0 -> 0 clamp, as the reference does when a
 * physics_post_transport_callback takes the scalars (cuda_driver.py:5299-5311).
 * A caller's microphysics then runs between hexdyn_scalars(3) and
 * hexdyn_end_step on the candidate scalars (hexdyn_export_field /
 * hexdyn_import_field of HEXDYN_F_SCALARS), before validation and publish.
 * Default 1 (the clamp, atm_srk3's DO_PHYSICS branch). */
int hexdyn_set_post_transport_clamp(hexdyn_handle *h, int enable);

/* Phases of one large step, in the order hexdyn_step runs them. */
int hexdyn_begin_step(hexdyn_handle *h, void *stream);
int hexdyn_begin_subcycle(hexdyn_handle *h, int subcycle, void *stream);
int hexdyn_dyn_tend(hexdyn_handle *h, int stage, void *stream);
int hexdyn_acoustic_substep(hexdyn_handle *h, int stage, int small_step, void *stream);
int hexdyn_divergence_damping(hexdyn_handle *h, int stage, void *stream);
int hexdyn_recover(hexdyn_handle *h, int stage, void *stream);
int hexdyn_solve_diagnostics(hexdyn_handle *h, int rk_step, void *stream);
int hexdyn_end_subcycle(hexdyn_handle *h, int subcycle, void *stream);
int hexdyn_scalars(hexdyn_handle *h, int stage, void *stream);
int hexdyn_end_step(hexdyn_handle *h, void *stream);

/* Per-kernel timing with HIP events recorded on the launch stream around
 * every kernel while enabled; read returns the average duration of one
 * kernel name and its launch count; kernels lists the names seen. */
int hexdyn_profile_enable(hexdyn_handle *h, int enable);
int hexdyn_profile_read(hexdyn_handle *h, const char *kernel, double *avg_ms, int64_t *launches);
int hexdyn_profile_kernels(hexdyn_handle *h, char *buf, size_t len);

/* FP64 global integrals of the published state with fixed-order
 * compensated reduction (repeatable bits): out[0] dry mass
 * sum areaCell*dzw*rho_zz, out[1] sum areaCell*dzw*rtheta_m, out[2] total
 * energy sum areaCell*dzw*rho_zz*(c_v T + g z + K), out[3+s] tracer mass of
 * scalar s.  n_out >= 3 + n_scalars.  Synchronises `stream`.
 * (No reference entry point: the reference sums rho*(area*dzw) in float64
 * with CuPy, cuda_fullstep_lts_active.py:1337-1360.) */
int hexdyn_integrals(hexdyn_handle *h, double *out, int n_out, void *stream);

/* Copy the published state (prognostic fields, carried diagnostics, model
 * time) of one handle into another built from the same mesh, vertical grid
 * and scalar count (the same internal numbering), device to device on
 * `stream`.  Handles with different timesteps may exchange states this way:
 * the learned corrector's coarse and fine handles do.  No reference
 * counterpart (the reference re-uploads a host state, cuda_driver.py:2826). */
int hexdyn_copy_state(hexdyn_handle *dst, hexdyn_handle *src, void *stream);

/* ---- Learned large-step corrector.  No reference counterpart: the
 * reference dycore has no learned component.  A graph network on the mesh
 * (message passing over cellsOnCell, edge outputs over cellsOnEdge, cell
 * inputs gathered over edgesOnCell) maps the published state after one
 * coarse large step (doubled dt, fewer acoustic substeps) to the correction
 * that brings it onto the fine-dt trajectory: corrections of u (edges),
 * w, theta_m and rho_zz (cells).  bf16 MFMA GEMMs with fp32 accumulation,
 * hand-written forward, backward and Adam kernels, fp32 master weights.
 *
 * Parameters, flat fp32 in this order (row-major [in][out] matrices):
 *   W_in [384][128], b_in [128], then per layer W_l [256][128], b_l [128],
 *   then W_out [128][320], b_out [320].
 * Inputs per cell: 6 channels x 64 rows (east and north cell wind from a
 * least-squares fit of the edge normal winds, w, theta_m, rho_zz, qv; the
 * Coriolis parameter in row n_levels of the east-wind channel), normalised
 * as (raw - in_mean) * in_scale.  Outputs per cell: 5 channels x 64 rows
 * (east wind, north wind, theta_m, rho_zz, w corrections) in units of
 * out_scale; the east and north channels share one scale per row.
 *
 * Calls on one network must not overlap; every entry point enqueues on
 * `stream`, and those that return host values synchronise it. */
typedef struct hexdyn_net hexdyn_net;
#define HEXDYN_NET_HIDDEN 128
#define HEXDYN_NET_INPUTS 384
#define HEXDYN_NET_OUTPUTS 320
enum { HEXDYN_NET_X = 0, HEXDYN_NET_Y = 1, HEXDYN_NET_TARGET_CELLS = 2, HEXDYN_NET_TARGET_EDGES = 3 };

/* A network of `layers` (1..8) message-passing layers bound to the mesh of
 * handle h (weights zero until set).  h must outlive the network; any
 * handle of the same mesh (same sizes and numbering) may be passed to the
 * calls below. */
int hexdyn_net_create(hexdyn_handle *h, int layers, hexdyn_net **out);
void hexdyn_net_destroy(hexdyn_net *net);
int64_t hexdyn_net_param_count(const hexdyn_net *net);
/* fp32 master weights from the host; resets the Adam moments */
int hexdyn_net_set_params(hexdyn_net *net, const float *host, int64_t n, void *stream);
int hexdyn_net_get_params(hexdyn_net *net, float *host, int64_t n, void *stream);
int hexdyn_net_get_grads(hexdyn_net *net, float *host, int64_t n, void *stream);
/* in_mean, in_scale [6][64]; out_scale [5][64] (host) */
int hexdyn_net_set_scales(hexdyn_net *net, const float *in_mean, const float *in_scale, const float *out_scale,
                          void *stream);
/* features of h's published state and the forward pass; keep_activations
 * = 1 keeps what hexdyn_net_backward needs */
int hexdyn_net_forward(hexdyn_net *net, hexdyn_handle *h, int keep_activations, void *stream);
/* training target: (fine - coarse) of the two handles' published u, w,
 * theta_m, rho_zz, in units of out_scale */
int hexdyn_net_set_target(hexdyn_net *net, hexdyn_handle *coarse, hexdyn_handle *fine, void *stream);
/* loss of the last forward against the target: out5 = {total, u, theta_m,
 * rho_zz, w} mean squared normalised errors; also forms the output gradient */
int hexdyn_net_loss(hexdyn_net *net, double *out5, void *stream);
int hexdyn_net_backward(hexdyn_net *net, void *stream);
/* one Adam step (step >= 1 counts from 1) on the gradients of the last backward */
int hexdyn_net_adam(hexdyn_net *net, float lr, float beta1, float beta2, float eps, int step, void *stream);
/* add gain x the last forward's correction to h's published state: the
 * density correction is shifted so that the dry mass sum area*dzw*drho is
 * zero (binary64 sums), rho*theta is formed from the corrected rho and
 * theta_m, ru from the corrected u and edge density, rw from the corrected
 * w; then the carried diagnostics are recomputed from the prognostics. */
int hexdyn_net_apply(hexdyn_net *net, hexdyn_handle *h, float gain, void *stream);
/* per-step checks of h's published state: out3 = {max |u| dt/dcEdge,
 * max |w| dt rdzw/zz, max |divergence| dt} (NaN if any value is NaN) and the
 * dry mass sum areaCell*dzw*rho_zz (binary64, fixed-order reduction), with
 * one synchronisation of `stream` */
int hexdyn_net_checks(hexdyn_net *net, hexdyn_handle *h, float *out3, double *dry_mass, void *stream);
/* raw (unnormalised) features of h's published state, [C][384] in the
 * caller's cell order, device memory */
int hexdyn_net_features(hexdyn_net *net, hexdyn_handle *h, float *dst_device, void *stream);
/* export X [C][384], Y [C][320], target cells [C][3][64] (theta_m, rho_zz,
 * w) or target edges [E][64] in the caller's order, device memory, fp32 */
int hexdyn_net_export(hexdyn_net *net, int what, float *dst_device, void *stream);
int hexdyn_net_last_error(const hexdyn_net *net, char *buf, size_t len);

/* hexdyn_integrals without allocation, synchronisation or host copy: the
 * same kernels and reduction order (bitwise the same values), the per-block
 * partial sums held in a file-scope device array of
 * HEXDYN_NOALLOC_MAX_PARTIALS doubles (8 MiB per GPU, kept until the process
 * ends) instead of the handle's workspace, and the n_out >= 3 + n_scalars
 * results written to the caller's DEVICE memory `out_device` on `stream`.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Refuses meshes with ceil(n_cells/4)*(3+n_scalars) above the
 * limit (HEXDYN_E_INVALID).  Every call on one GPU shares that GPU's array
 * (it keeps the last call's partial sums; each call writes every partial it
 * reads, so nothing is cleared): calls that overlap on one GPU (any handles,
 * any streams) race. */
#define HEXDYN_NOALLOC_MAX_PARTIALS (1 << 20)
int hexdyn_integrals_noalloc(hexdyn_handle *h, double *out_device, int n_out, void *stream);

/* The exner power function of recovery (cuda_driver.py:1466-1472: the GNU
 * powf), host copy, exposed for its tests: for y = R/cv and x in [2^-12, 4)
 * the GNU C library's bits, elsewhere the correctly rounded binary64 method.
 * The first call builds the process's correction bitmap (about 0.7 s). */
float hexdyn_powf(float x, float y);

#ifdef __cplusplus
}
#endif

#endif /* HEXDYN_H */
