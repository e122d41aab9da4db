/*
 * amun_dev.h -- C-ABI of the MI355X-native AMUN per-block finite-volume update.
 *
 * The library (amun_amd/libamun_hip.so) keeps the block arrays resident in
 * HBM in structure-of-arrays layout and runs the whole time step -- CFL
 * reduction, reconstruction, Riemann fluxes, flux divergence, SSPRK stage
 * combine, primitive recovery and the ghost-zone exchange between blocks --
 * as hand-written gfx950 kernels.  Every entry point takes plain pointers and
 * sizes; a Fortran host binds them with iso_c_binding (amun_amd/host/
 * amun_dev_mod.F90), Python with ctypes (amun_amd/solver.py).
 *
 * Return codes: 0 = success.  Configuration errors reuse the reference's
 * integer codes (100/110 equation system / equation of state,
 * equations.F90:399-830; 200-202 reconstruction, interpolations.F90:382),
 * 300 = Riemann solver / equation system mismatch, 400 = integrator,
 * 500 = bad sizes or arguments, 600 = unsupported combination on the device
 * path (for instance too few ghost zones for the stencil); HIP runtime errors
 * are reported as 1000 + hipError_t.  amun_dev_last_error() returns a message.
 *
 * Host array layout.  The reference stores every block as u0/u1/q(nv,im,jm,km)
 * in Fortran order with im = ncells + 2*nghosts (blocks.F90:1120-1191); the
 * *_blocks transfer calls take exactly that layout, blocks one after another.
 * The *_global calls take a whole uniform grid as U[nv][nz][ny][nx] (C order,
 * x fastest), which is what tests and benchmarks use.
 */
#ifndef AMUN_DEV_H
#define AMUN_DEV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* equation systems: equations.F90:319-830 (parameter "equation_system") */
#define AMUN_EQSYS_HD      1
#define AMUN_EQSYS_MHD     2
#define AMUN_EQSYS_SRHD    3   /* special-relativistic HD, equations.F90:537-600 */
#define AMUN_EQSYS_SRMHD   4   /* special-relativistic MHD, equations.F90:673-760 */
/* equations of state: equations.F90:345-395 (parameter "equation_of_state") */
#define AMUN_EOS_ADI       1
#define AMUN_EOS_ISO       2   /* isothermal (HD and MHD), equations.F90:399-830 */

/* reconstructions: interpolations.F90:237-376 (parameter "reconstruction") */
#define AMUN_BND_PERIODIC  0
#define AMUN_BND_OPEN      1   /* zero gradient                                */
#define AMUN_BND_OUTFLOW   2   /* zero gradient, normal velocity only outwards */
#define AMUN_BND_REFLECTIVE 3  /* mirror, normal velocity and field negated    */

#define AMUN_C2P_1DW       0   /* nr_iterate_srmhd_adi_1dw, equations.F90:5665  */
#define AMUN_C2P_2DWV      1   /* nr_iterate_srmhd_adi_2dwv, equations.F90:5819 */
#define AMUN_C2P_2DWU      2   /* nr_iterate_srmhd_adi_2dwu, equations.F90:6020 */

#define AMUN_REC_TVD       1
#define AMUN_REC_WENO3     2   /* reconstruct_weno3, interpolations.F90:1469   */
#define AMUN_REC_LIMO3     3   /* reconstruct_limo3, interpolations.F90:1598   */
#define AMUN_REC_PPM       4   /* reconstruct_ppm, interpolations.F90:1746     */
#define AMUN_REC_WENO5Z    5
#define AMUN_REC_WENO5YC   6   /* reconstruct_weno5yc, interpolations.F90:2107 */
#define AMUN_REC_WENO5NS   7   /* reconstruct_weno5ns, interpolations.F90:2278 */
#define AMUN_REC_CRWENO5Z  8   /* reconstruct_crweno5z, interpolations.F90:2477 */
#define AMUN_REC_CRWENO5YC 9   /* reconstruct_crweno5yc, interpolations.F90:2870 */
#define AMUN_REC_CRWENO5NS 10  /* reconstruct_crweno5ns, interpolations.F90:3266 */
#define AMUN_REC_MP5       11  /* reconstruct_mp5, interpolations.F90:3675     */
#define AMUN_REC_MP7       12  /* reconstruct_mp7, interpolations.F90:3799     */
#define AMUN_REC_CRMP5     13  /* reconstruct_crmp5, interpolations.F90:3930   */
#define AMUN_REC_CRMP5LD   14  /* reconstruct_crmp5ld, interpolations.F90:4107 */
#define AMUN_REC_CRMP7     15  /* reconstruct_crmp7, interpolations.F90:4278   */
#define AMUN_REC_MGP       21  /* interfaces_mgp, interpolations.F90:978       */
#define AMUN_REC_GP        20  /* Gaussian process, reconstruct_gp interpolations.F90:4561 */

/* TVD limiters: interpolations.F90:379-398 (parameter "limiter") */
#define AMUN_LIM_ZERO      0
#define AMUN_LIM_MINMOD    1
#define AMUN_LIM_MC        2
#define AMUN_LIM_SUPERBEE  3
#define AMUN_LIM_VANLEER   4
#define AMUN_LIM_VANALBADA 5

/* Riemann solvers: schemes.F90:114-522 (parameter "riemann_solver") */
#define AMUN_RS_HLL        1
#define AMUN_RS_HLLC       2
#define AMUN_RS_HLLD       3
#define AMUN_RS_ROE        4   /* Roe, riemann_hd_adi_roe schemes.F90:1541 */
#define AMUN_RS_HLLDM      5   /* HLLD-M of the isothermal MHD system,
                                  riemann_mhd_iso_hlldm schemes.F90:2388 */

/* integrators: evolution.F90:156-200 (parameter "time_advance") */
#define AMUN_TA_EULER      0
#define AMUN_TA_RK2        1   /* SSPRK(2,2), evolve_rk2 evolution.F90:628 */
#define AMUN_TA_RK3        2   /* SSPRK(3,3), evolve_rk3 evolution.F90:1005 */
#define AMUN_TA_SSPRK_M2   3   /* SSPRK(m,2), evolve_ssprk2 evolution.F90:791 */
#define AMUN_TA_SSPRK43    4   /* SSPRK(4,3), evolve_ssprk34 evolution.F90:1228 */
#define AMUN_TA_SSPRK53    5   /* SSPRK(5,3), evolve_ssprk35 evolution.F90:1492 */

/* arithmetic of reconstruction + fluxes; the conserved state is always FP64 */
#define AMUN_PREC_FP64     0
#define AMUN_PREC_MIXED    1   /* FP32 reconstruction/fluxes, FP64 Kahan state */

/* Run configuration.  Field names follow the reference parameter names. */
typedef struct amun_cfg {
  int32_t ndims;            /* 2 or 3 (compile-time NDIMS in the reference) */
  int32_t eqsys;            /* AMUN_EQSYS_*                                  */
  int32_t eos;              /* AMUN_EOS_*                                    */
  int32_t nv;               /* filled in by amun_dev_init                    */
  int32_t ncells;           /* "ncells": cells per block edge                */
  int32_t nghosts;          /* "nghosts": ghost layers of host block arrays  */
  int32_t xblocks, yblocks, zblocks;  /* "xblocks" ...: base-level blocks    */
  int32_t reconstruction;   /* AMUN_REC_*                                    */
  int32_t limiter;          /* AMUN_LIM_*  ("limiter")                       */
  int32_t riemann_solver;   /* AMUN_RS_*                                     */
  int32_t time_advance;     /* AMUN_TA_*                                     */
  int32_t precision;        /* AMUN_PREC_*                                   */
  int32_t fix_positivity;   /* "fix_positivity" (0 = off, default)           */
  int32_t clip_extrema;     /* "clip_extrema"   (0 = off, default)           */
  double  gamma;            /* "gamma"                                       */
  double  cfl;              /* "cfl"                                         */
  double  eps;              /* "eps" of the WENO weights (default epsilon)   */
  double  alpha;            /* "alpha": GLM psi decay, psi *= exp(-alpha*cfl)*/
  double  xmin, xmax, ymin, ymax, zmin, zmax;   /* domain bounds             */
  /* primitive-variable solver of the relativistic systems (equations.F90:
   * 166-201, 303-318): Newton tolerance, iteration limit, extra iterations
   * after convergence and the pressure floor of the brackets */
  double  tolerance;        /* "tolerance" (1e-10)                            */
  double  pmin;             /* "pmin"      (1e-16)                            */
  int32_t nr_maxit;         /* "nr_maxit"  (100)                              */
  int32_t nr_extra;         /* "nr_extra"  (2)                                */
  /* Gaussian-process reconstruction and its monotonicity-preserving limiter
   * (interpolations.F90:87-106, 206-221) */
  double  sgp;              /* "sgp": kernel width in cells (9)               */
  double  kappa;            /* "kappa" (1), used as min(kappa, (1-cfl)/cfl)   */
  double  kbeta;            /* "kbeta" (1)                                    */
  int32_t ngp;              /* "ngp": stencil points (5)                      */
  int32_t stages;           /* "stages" of SSPRK(m,2), clamped to [2, 9] (2)  */
  /* limiter of the coarse-to-fine prolongation in the ghost fill of
   * multi-level meshes ("prolongation_limiter", interpolations.F90:165, 203,
   * 412-428): AMUN_LIM_MINMOD, _MC, _SUPERBEE, _VANLEER or _ZERO (minmod) */
  int32_t prol_limiter;
  /* "primitive_solver" of the relativistic system (equations.F90:309,
   * 747-789): AMUN_C2P_1DW (default), AMUN_C2P_2DWV, AMUN_C2P_2DWU */
  int32_t primitive_solver;
  /* "limo3_rad" (0.5): the radius of the LimO3 smoothness indicator in cell
   * sizes, and "ppm_const" (1.25): the curvature ratio of the PPM extremum
   * limiter (interpolations.F90:83, 106, 209, 212) */
  double  limo3_rad;
  double  ppm_const;
  /* "xlbndry", "xubndry", "ylbndry", "yubndry", "zlbndry", "zubndry"
   * (coordinates.F90, boundaries.F90:1089-1251, 5024-5637): AMUN_BND_*;
   * an axis is periodic on both sides or on neither */
  int32_t bnd_type[6];
  /* "mlp_limiting" (1 = on): the multi-dimensional limiting process after
   * every reconstruction (interpolations.F90:204, 844, 955, 1140,
   * 1171-1321); "glm_source_terms" eglm/heglm = 1/2 (sources.F90:128-141,
   * 479-539) for the magnetic systems; the "viscosity" and "resistivity"
   * coefficients (>= 0; sources.F90:145-149, 340-598) of the viscous and
   * resistive source terms and their dt term (evolution.F90:478-483).
   * Outside the device path, refused with 600 when on:
   * "fix_unphysical_cells" of update_primitive_variables (equations.F90:905,
   * 1168).  Gradients (amun_dev_tape_*) refuse mlp_limiting and the source
   * terms. */
  int32_t mlp_limiting;
  int32_t fix_unphysical_cells;
  int32_t glm_source_terms;
  int32_t reserved_cfg;
  double  viscosity;
  double  resistivity;
  /* "csnd": the sound speed of the isothermal systems (equations.F90:176,
   * 847-851; 1 by default) */
  double  csnd;
  /* "state_variables" of the relativistic systems (schemes.F90:133, 383-400):
   * 0 = primitive (default), 1 = "4vec" (reconstruction of rho W and the
   * four-velocity; refused with 600 on the device path) */
  int32_t states_4vec;
  int32_t reserved_cfg2;
  /* uniform gravity (0, "gacc", 0) of the Rayleigh-Taylor problem
   * (gravity.F90:141-150, 283-330; "gacc" -1 by default): on for
   * problem = rt, rayleightaylor or rayleigh-taylor */
  int32_t gravity;
  int32_t reserved_cfg3;
  double  gacc;
} amun_cfg;

/* Refinement criterion of the leaf blocks (initialize_refinement,
 * refinement.F90:99-190): the second-derivative error of the primitive
 * variables in varmask (bit p: primitive variable p; the reference selects
 * them by name from "refinement_variables"), the vorticity (vort != 0) and
 * the current density (jabs != 0, magnetic systems) against their
 * thresholds. */
typedef struct amun_refine_cfg {
  double  crefmin, crefmax;   /* "crefmin" (0.2), "crefmax" (0.8)             */
  double  epsref;             /* "epsref" (0.01)                              */
  double  vortmin, vortmax;   /* "vortmin" (1e-3), "vortmax" (0.1)            */
  double  jabsmin, jabsmax;   /* "jabsmin" (1e-3), "jabsmax" (0.1)            */
  int32_t varmask;            /* default: density and pressure                */
  int32_t vort, jabs;
  int32_t minlev, maxlev;     /* level limits of update_mesh (0: none)        */
  int32_t reserved;
} amun_refine_cfg;

typedef struct amun_ctx amun_ctx;

/* One independent uniform grid of a batch: base-level block counts and
 * domain bounds (the "xblocks".. and "xmin".. parameters of one run). */
typedef struct amun_grid {
  int32_t xblocks, yblocks, zblocks, reserved;
  double  xmin, xmax, ymin, ymax, zmin, zmax;
} amun_grid;

/* Fill *cfg with the reference defaults (equations.F90:166-200,
 * interpolations.F90:80-118, evolution.F90:65-80, coordinates.F90:42-80). */
void amun_dev_default_cfg(amun_cfg* cfg);

/* Create a device context for one run: validates the configuration like the
 * initialize_* routines (equations.F90:257, interpolations.F90:144,
 * schemes.F90:114, evolution.F90:115), builds the base-level periodic block
 * topology of setup_domain_default (domains.F90:136-500) and allocates the
 * device block pool.  Replaces the module state set up by those routines. */
int  amun_dev_init(const amun_cfg* cfg, amun_ctx** ctx);

/* Batched form of amun_dev_init: nruns independent periodic grids that share
 * the numerical configuration of cfg (its block-count and bound fields are
 * ignored) but each have their own block counts, cell sizes, time step, time
 * and step counter.  All their blocks live in one pool, run after run in the
 * order of grids, and every sweep of every stage is one launch over the
 * whole pool.  No reference counterpart: the reference runs one grid per
 * process (driver.F90:157-657). */
int  amun_dev_init_batch(const amun_cfg* cfg, int32_t nruns, const amun_grid* grids,
                         amun_ctx** ctx);

/* A multi-level (AMR) mesh of nblk leaf blocks on the periodic base grid of
 * cfg: meta[4 b] = level (1 = base), meta[4 b + 1..3] = first interior cell
 * (i0, j0, k0) of block b in cells of its level.  Replaces the mesh state
 * the reference's block lists and neighbour pointers hold after
 * generate_mesh / update_mesh (mesh.F90:431-740, blocks.F90) with the mesh
 * frozen, and makes every step do boundary_variables with restriction and
 * prolongation (boundaries.F90:372-479, 5637-6614), the flux correction of
 * boundary_fluxes (494-1060), per-level cell sizes in update_increment
 * (evolution.F90:1898-1977) and the finest-level dx in new_time_step
 * (evolution.F90:379-493).  The mesh must be 2:1 balanced (700/701
 * otherwise); nghosts must be even and at most ncells / 2; HD and MHD only.
 * State moves with amun_dev_upload_blocks / amun_dev_download_blocks in
 * meta order.  No gradients, no batching on such a context. */
int  amun_dev_init_amr(const amun_cfg* cfg, int32_t nblk, const int32_t* meta,
                       amun_ctx** ctx);

/* info[0..5]: real blocks, virtual blocks (same-level stand-ins of the
 * ghost regions at other levels), levels, restricted and prolonged ghost
 * cells per stage, coarse faces with flux correction. */
int  amun_dev_amr_info(amun_ctx* ctx, int32_t* info);

/* Number of runs of the context, and the first pool block and block count
 * of run r (its blocks in amun_dev_block_coords / *_blocks order). */
int  amun_dev_nruns(amun_ctx* ctx, int32_t* nruns);
int  amun_dev_run_blocks(amun_ctx* ctx, int32_t r, int32_t* first, int32_t* count);

/* Upload / download the global array U[nv][nz][ny][nx] of run r alone. */
int  amun_dev_upload_run(amun_ctx* ctx, int32_t r, const double* u);
int  amun_dev_download_run(amun_ctx* ctx, int32_t r, double* u, double* q);

/* Time, last dt and step counter of every run (arrays of nruns; any may be
 * NULL). */
int  amun_dev_run_state(amun_ctx* ctx, double* time, double* dt, int64_t* step);

/* amun_dev_integrals for run r alone. */
int  amun_dev_integrals_run(amun_ctx* ctx, int32_t r, double out[8]);
/* amun_dev_statistics for run r alone. */
int  amun_dev_statistics_run(amun_ctx* ctx, int32_t r, double out[21]);

/* With enable != 0, steps are captured once into HIP graphs (one per cmax
 * slot, primitive buffer and dtnext) and replayed with hipGraphLaunch.  The
 * captured kernels are the ones launched otherwise; no memset is captured. */
int  amun_dev_set_graph(amun_ctx* ctx, int32_t enable);

/* Launch-shape autotuner: runs nsteps timed steps of the live state for every
 * candidate shape (workgroup of 64, 128 or 256 pencils; plain or XCD-grouped
 * workgroup order), keeps the fastest per sweep direction and restores the
 * state, so the run continues as if the call had not happened.  choice (may
 * be NULL) receives ndims pairs (workgroup size, XCD grouping).
 * amun_dev_set_launch sets such pairs directly. */
int  amun_dev_autotune(amun_ctx* ctx, int32_t nsteps, int32_t* choice);

/* One launch per stage (no reference counterpart; the reference's evolve_*
 * calls update_fluxes once per stage, evolution.F90:628-1208): with
 * enable != 0 every stage runs its x, y (and z) sweeps in one kernel whose
 * workgroups hold whole blocks, with a workgroup barrier between the
 * directions.  For the TVD and WENO5-Z reconstructions of the HD and MHD
 * systems when ncells^(ndims-1) divides 256 (600 otherwise); the results
 * are bitwise those of the per-direction launches. */
int  amun_dev_set_stage_fusion(amun_ctx* ctx, int32_t enable);
int  amun_dev_set_launch(amun_ctx* ctx, const int32_t* choice);

/* Release every device allocation of the context (finalize_* routines). */
void amun_dev_finalize(amun_ctx* ctx);

/* Upload conserved variables.  _blocks: nblk host blocks u0(nv,im,jm,km)
 * (blocks.F90:1120-1191 layout; ghosts are ignored), all runs of a batch.
 * _global: U[nv][nz][ny][nx] of a single-run context.
 * Both recompute the primitive variables (update_primitive_variables,
 * equations.F90:1088) on the device. */
int  amun_dev_upload_blocks(amun_ctx* ctx, const double* u);
int  amun_dev_upload_global(amun_ctx* ctx, const double* u);

/* Download conserved (and optionally primitive, q may be NULL) variables.
 * _blocks fills the ghost layers too, as after boundary_variables
 * (boundaries.F90:372-479), all runs of a batch; _global needs a single run. */
int  amun_dev_download_blocks(amun_ctx* ctx, double* u, double* q);
int  amun_dev_download_global(amun_ctx* ctx, double* u, double* q);

/* new_time_step (evolution.F90:379-493): cmax reduction over all interior
 * cells and dt = min(cfl*dx_min/cmax, dtnext) when dtnext > 0.  Every run of
 * a batch gets its own; dt and cmax returned here are those of run 0. */
int  amun_dev_new_time_step(amun_ctx* ctx, double dtnext, double* dt,
                            double* cmax);

/* advance (evolution.F90:292-363) for a single-level mesh: new_time_step,
 * then every stage of the selected integrator (update_fluxes, update_increment,
 * stage combine, psi decay, update_variables), all on the device. */
int  amun_dev_step(amun_ctx* ctx, double dtnext, double* dt, double* cmax);

/* Run nsteps steps back to back on the device stream (the host reads dt only
 * at the end); dts (may be NULL) receives the dt of every step of every run,
 * dts[step * nruns + run]. */
int  amun_dev_run(amun_ctx* ctx, int32_t nsteps, double dtnext, double* dts);

/* Timing helper: queue nsteps steps on the context stream with HIP events
 * around them (and, if kernel_ms has room for *nkernel entries, after every
 * kernel of every step); elapsed_ms = device time of all steps, kernel_ms[k] =
 * average duration of the k-th kernel of a step.  No reference counterpart
 * (the reference times steps with system_clock, driver.F90:730-793). */
int  amun_dev_timed_run(amun_ctx* ctx, int32_t nsteps, double dtnext,
                        double* elapsed_ms, double* kernel_ms,
                        int32_t nkernel_max, int32_t* nkernel);

/* Volume integrals of store_integrals (integrals.F90:254-485), FP64 on the
 * device: out[0]=mass, [1..3]=momentum, [4]=total energy, [5]=kinetic,
 * [6]=magnetic energy, [7]=internal energy (summed over the runs of a
 * batch). */
int  amun_dev_integrals(amun_ctx* ctx, double out[8]);
/* The statistics half of store_integrals (integrals.F90:331-431, the
 * "statistics" file): out = mean, min, max of density, pressure, |v|, |B|,
 * psi, the sonic Mach number sqrt(rho)|v|/sqrt(gamma p) and the Alfvenic
 * sqrt(rho)|v|/max(eps, |B|), from the primitive state; the means weigh
 * cells by their volume (levels included) over the domain volume.  |B|,
 * psi and M_A are 0 without a magnetic field.  In the mixed mode the
 * primitives are the stored FP32 ones. */
int  amun_dev_statistics(amun_ctx* ctx, double out[21]);

/* Simulation time and step counter kept by the context (evolution.F90:75);
 * run 0 of a batch. */
int  amun_dev_get_time(amun_ctx* ctx, double* time, int64_t* step);

/* Number of blocks, and the base-level block coordinates (nblk x 3 int32,
 * 0-based) in the order used by the *_blocks transfers. */
int  amun_dev_nblocks(amun_ctx* ctx, int32_t* nblk);
int  amun_dev_block_coords(amun_ctx* ctx, int32_t* coords);

/* Wait until all work queued on the context's stream has finished. */
int  amun_dev_sync(amun_ctx* ctx);

/* Number of non-positive density or pressure values met by the TVD
 * reconstruction so far (the reference prints a warning for each one and
 * zeroes the slopes, interpolations.F90:814-822). */
int  amun_dev_positivity_events(amun_ctx* ctx, uint64_t* n);

/* Events of the relativistic primitive solver so far, each of which the
 * reference reports on stderr and survives (cons2prim_srmhd_adi,
 * equations.F90:4740-4777; nr_iterate_srmhd_adi_1dw 5766-5770):
 * n[0] = cells whose bracketing failed (pressure set to 0, the other
 * primitives keep their previous values), n[1] = cells recovered with a
 * non-positive pressure (set to 0), n[2] = cells whose Newton iteration
 * ended without reaching the tolerance.  All zero for HD/MHD. */
int  amun_dev_c2p_events(amun_ctx* ctx, uint64_t n[3]);

/* The refinement criterion of every leaf block of the current state
 * (check_refinement_criterion, refinement.F90:257-640, as update_mesh calls
 * it, mesh.F90:1545-1550): crit[b] = -1 (may derefine), 0 (keep) or 1
 * (refine) in the block order of the context; with cfg->minlev / maxlev
 * above 0 the level limits of update_mesh are applied.  ind (may be NULL)
 * receives three floats per block: the largest second-derivative error of
 * the selected variables, the vorticity and the current-density magnitude.
 * The ghost cells it reads are refilled from the current state first. */
int  amun_dev_refinement(amun_ctx* ctx, const amun_refine_cfg* cfg, int32_t* crit, float* ind);

/* Moves a multi-level context onto the leaf set meta (nblk blocks, layout as
 * amun_dev_init_amr) without a state transfer through the host: the device
 * side of update_mesh's derefine_selected_blocks / refine_selected_blocks
 * (mesh.F90:756-833, 2069-2257).  Every new block is an old leaf (its data
 * kept), a child of an old leaf (prolong_block, mesh.F90:1031-1257: the
 * parent's conserved cells with limited slopes, the prolongation limiter,
 * density and energy slopes halved until positive) or the parent of 2^d old
 * leaves (restrict_block, 1258-1454: the mean of the covering cells); 702
 * otherwise, 700/701 when meta is not 2:1 balanced.  The handle, its stream,
 * time, step, dt, event counters and launch settings stay; q and the maximum
 * speed are recomputed as update_variables does after update_mesh
 * (evolution.F90:321-340).  A forward tape is dropped.  Replaces the
 * reference's mesh.F90 update_mesh data path (prolong_block /
 * restrict_block called from refine/derefine_selected_blocks); which blocks
 * change is the caller's choice (amun_dev_refinement gives the criterion). */
int  amun_dev_amr_remesh(amun_ctx* ctx, int32_t nblk, const int32_t* meta);

/* Pencils of the compact reconstructions (CRWENO5, CRMP) whose Thomas solve
 * met a zero pivot so far.  The reference's tridiag prints "Could not find
 * solution!" there and leaves the solution undefined (algebra.F90:1145-1152);
 * the device counts the pencil and gives it its cell values as face states
 * (first order). */
int  amun_dev_tridiag_events(amun_ctx* ctx, uint64_t* n);

/* The context's HIP stream (a hipStream_t), so that a caller can order its
 * own device work after the solver's. */
void* amun_dev_stream(amun_ctx* ctx);

/* Gradients of a scalar objective of the final state (no reference
 * counterpart: the reference computes no derivatives).  Single-run FP64
 * contexts with TVD reconstruction and the HD or MHD system.
 * _tape_forward: nsteps steps from the current state, as amun_dev_run, with a
 *   checkpoint of the conserved state every interval steps and the dt and
 *   GLM cmax of every step recorded (_tape_steps returns them).
 * _tape_backward: given lam = dJ/dU of the final state (global array
 *   U[nv][nz][ny][nx]), lam0 = dJ/dU of the taped initial state and
 *   *dgamma = dJ/dgamma, by the adjoint of every stage (reverse order,
 *   steps recomputed from the checkpoints).  dt and cmax are held at their
 *   taped values: the derivative is that of the map with the step sizes
 *   frozen, as are the limiter branches and positivity halvings.
 * _tape_tangent: the matching tangent-linear run: du = dU(final) for an
 *   initial perturbation du0 and dgamma.
 * _run_fixed: nsteps ordinary steps with dt and cmax imposed per step (the
 *   nonlinear map that the tangent and adjoint linearise; finite differences).
 * The live state is left as it was. */
int  amun_dev_tape_forward(amun_ctx* ctx, int32_t nsteps, int32_t interval, double dtnext);
int  amun_dev_tape_steps(amun_ctx* ctx, double* dt, double* cmax);
int  amun_dev_tape_backward(amun_ctx* ctx, const double* lam, double* lam0, double* dgamma);
int  amun_dev_tape_tangent(amun_ctx* ctx, const double* du0, double dgamma, double* du);
int  amun_dev_run_fixed(amun_ctx* ctx, int32_t nsteps, const double* dt, const double* cmax);

/* Unit level, for parity tests of the device routines on vectors of states
 * (the seams riemann_iface, schemes.F90:66-71, and prim2cons_iface /
 * cons2prim_iface / fluxspeed_iface, equations.F90:70-108), in the
 * arithmetic type of cfg->precision.  States are (nv, n) in Fortran order,
 * i.e. n consecutive vectors of nv doubles.
 * _riemann: f = riemann(ql, qr) with the GLM cmax given.
 * _eqs: u = prim2cons(q), f/cm/cp = fluxspeed(q, u), q2 = cons2prim(u) in
 * FP64; q2 enters holding the previous primitives (kept by a failed
 * relativistic recovery). */
int  amun_dev_unit_riemann(const amun_cfg* cfg, int32_t n, double cmax, const double* ql,
                           const double* qr, double* f);
int  amun_dev_unit_eqs(const amun_cfg* cfg, int32_t n, double cmax, const double* q,
                       double* u, double* q2, double* f, double* cm, double* cp);
/* _reconstruct: the 1D scheme of cfg->reconstruction (reconstruct_states of
 * interfaces_dir, interpolations.F90:867-959, e.g. reconstruct_ppm at
 * 1746-1913) on npencil pencils fp of n cells (8 <= n <= 4096) with cell
 * size h: fl[i] / fr[i-1] are the states at the upper / lower face of cell
 * i, written for the cells whose window (the sweeps' stencil) lies in the
 * pencil; other entries keep their input values.  Not for TVD, whose slopes
 * couple the directions (error 600). */
int  amun_dev_unit_reconstruct(const amun_cfg* cfg, int32_t n, int32_t npencil, double h,
                               const double* fp, double* fl, double* fr);
/* _cons2prim: q = cons2prim(u) in FP64 on n conserved vectors
 * (cons2prim_iface, equations.F90:70-108; for the relativistic system the
 * primitive solver of cfg->primitive_solver); q enters holding the previous
 * primitives (kept but p = 0 by a failed recovery); events (may be NULL)
 * gets the counts of amun_dev_c2p_events. */
int  amun_dev_unit_cons2prim(const amun_cfg* cfg, int32_t n, const double* u, double* q,
                             int64_t* events);

/* No counterpart in the reference: amun_dev_unit_cons2prim for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  d_u (n conserved vectors) and d_q (n
 * primitive vectors, entering with the previous primitives) are DEVICE
 * arrays in the layout of amun_dev_unit_cons2prim; d_events (may be NULL)
 * is a device array of 3 int64 that receives the counts of
 * amun_dev_c2p_events for this call.  The event counters live in a static
 * device array of the library (256 bytes per GPU, held until the process
 * ends) instead of a hipMalloc'd buffer; the call allocates nothing, clears
 * the counters with a kernel, and only enqueues kernels on stream (a
 * hipStream_t; NULL = the null stream) without synchronising.  Calls that
 * overlap on one GPU share those counters and race.  n is limited only by
 * int32 (the counters do not grow with n). */
int  amun_dev_cons2prim_noalloc(const amun_cfg* cfg, int32_t n, const double* d_u, double* d_q,
                                int64_t* d_events, void* stream);

/* Message of the last failing call of this thread. */
const char* amun_dev_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* AMUN_DEV_H */
