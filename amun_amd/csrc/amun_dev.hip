// amun_dev.hip -- host side of the C-ABI of include/amun_dev.h for the MI355X
// (gfx950) AMUN per-block update: contexts, transfers, the step and gradient
// drivers.  The device code is in kernels.h (see there for the data layout and
// the sweep design); the kernel tables of the reconstructions are built in
// inst_tvd.hip, inst_weno5z.hip and inst_gp.hip.
#include "kernels.h"
#include "remesh.h"
#include "sources.h"

#include <array>
#include <map>

// ===========================================================================
// context
// ===========================================================================
struct amun_ctx {
  amun_cfg cfg;
  RunConst rc;
  int nv;
  size_t nstate;            // doubles per state array (all runs)
  double *U0 = nullptr, *U1 = nullptr;
  void *Q[2] = {nullptr, nullptr}, *D = nullptr;   // element size qsize
  float* C = nullptr;       // Kahan compensation of U0 (mixed mode)
  size_t qsize = 8;
  int* nbr = nullptr;
  DevState* st = nullptr;
  RunState* rs = nullptr;   // per-run step state
  int* brun = nullptr;      // run of every block
  double* trace = nullptr;  // dt ring, [tsteps][nruns]
  int tsteps = 0;
  double* scratch = nullptr;   // reductions / small transfers
  SpeedPart* spart = nullptr;  // per-workgroup maximum speeds of the last sweep
  int nspart = 0;              // entries of spart (workgroups of >= 64 lanes)
  double* ugp = nullptr;       // MGP weights (RunConst::ugp points here)
  void* src = nullptr;         // GLM source products per stage (element size qsize)
  void* vsrc = nullptr;        // viscous and resistive source products per stage
  void* q4 = nullptr;          // four-velocity states of the stage input (relativistic)
  int vis = 0;                 // 1 viscosity, 2 resistivity (SweepArgs::vis)
  int fuse = 0;                // one k_stage launch per stage (amun_dev_set_stage_fusion)
  void* mlps = nullptr;        // multi-dimensional limiting: slopes of the interior cells
  void* mlpg = nullptr;        // and of the cells beyond the pencil ends (k_mlp)
  int qcur = 0;                // Q[qcur] holds the primitive state of U0
  long long nsteps = 0;        // steps launched (slot = nsteps & 1)
  hipStream_t stream = nullptr;
  std::vector<amun_grid> grids;     // one per run
  std::vector<int32_t> run_b0;      // first block of every run (+ total)
  std::vector<int32_t> coords;      // block coordinates inside its run
  Dispatch disp;
  // launch shapes of the sweeps (workgroup size, XCD-aware order), per
  // direction; set by amun_dev_autotune
  int wg[3] = {256, 256, 256};
  int swz[3] = {0, 0, 0};
  // captured steps: one graph per (cmax slot, primitive buffer, dtnext)
  bool use_graph = false;
  struct Graph { int slot, qcur; double dtnext; hipGraphExec_t exec; };
  std::vector<Graph> graphs;
  // forward tape of amun_dev_tape_forward: checkpoints of U every interval
  // steps, dt and cmax of every step
  struct Tape {
    int nsteps = 0, interval = 0;
    double dtnext = 0.0;
    std::vector<double> dt, cmax;
    double* ck = nullptr;       // ceil(nsteps / interval) states
    float* ckc = nullptr;       // their Kahan compensations (mixed mode)
    double* dtc = nullptr;      // device copy of (dt, cmax) per step
  } tape;
  // multi-level mesh (amun_dev_init_amr): real blocks [0, nreal) of the
  // pools, virtual blocks [nreal, nreal + nvirt) of the primitive pools
  struct Amr {
    bool on = false;
    int nreal = 0, nvirt = 0, nlev = 1, nfix = 0;
    std::vector<int32_t> meta;          // level, i0, j0, k0 per real block
    int* blev = nullptr;                // level - 1 per real block
    int4* rtask = nullptr;              // restrictions
    int4* ptask = nullptr;              // prolongations
    int4* ftask[3] = {nullptr, nullptr, nullptr};   // fine fluxes per direction
    int nr = 0, np = 0, nf[3] = {0, 0, 0};
    int* ffix = nullptr;                // correction slot per block face
    double* FB = nullptr;               // averaged fine fluxes per slot
    DevState* dummy = nullptr;          // events of the fine-flux recomputation
  } amr;
  // ghost blocks of the non-periodic domain sides
  struct Bnd {
    bool on = false;
    int n = 0;                          // virtual blocks
    int4* task = nullptr;               // {virtual, source, sides}
  } bnd;
};

static int amr_fill(amun_ctx* c, int qcur, int slot);

// calls f(P{}) with the physics of the configuration's equation system
template <class F>
static void by_system(const amun_cfg& cfg, F&& f) {
  const bool iso = cfg.eos == AMUN_EOS_ISO;
  if (cfg.eqsys == AMUN_EQSYS_HD) {
    if (iso) f(amun::HDI{});
    else f(amun::HD{});
  } else if (cfg.eqsys == AMUN_EQSYS_MHD) {
    if (iso) f(amun::MHDI{});
    else f(amun::MHD{});
  } else if (cfg.eqsys == AMUN_EQSYS_SRHD) {
    f(amun::SRHD{});
  } else {
    f(amun::SRMHD{});
  }
}

// the systems with a magnetic field (and GLM psi)
static bool has_field(const amun_cfg& c) {
  return c.eqsys == AMUN_EQSYS_MHD || c.eqsys == AMUN_EQSYS_SRMHD;
}

// the special-relativistic systems
static bool relativistic(const amun_cfg& c) {
  return c.eqsys == AMUN_EQSYS_SRHD || c.eqsys == AMUN_EQSYS_SRMHD;
}

// the pressure (energy) variable, -1 for the isothermal systems
static int pres_index(const amun_ctx* c) {
  return c->cfg.eos == AMUN_EOS_ISO ? -1 : c->nv - 1;
}

// variables of a configuration validate() accepts
static int cfg_nv(const amun_cfg& c) {
  const int nv = has_field(c) ? 9 : 5;
  return c.eos == AMUN_EOS_ISO ? nv - 1 : nv;
}

// the ghost blocks of the non-periodic sides from the primitive state Q[qcur]
static int bnd_fill(amun_ctx* c, int qcur) {
  if (!c->bnd.on) return 0;
  const long long n = (long long)c->bnd.n * c->rc.ncell;
  const dim3 grid((unsigned)((n + 255) / 256)), bl(256);
  const int ibx = has_field(c->cfg) ? 4 : -1;
  by_system(c->cfg, [&](auto tag) {
    using P = decltype(tag);
    if (c->qsize == sizeof(double))
      hipLaunchKernelGGL((k_bnd_fill<P::nv, double>), grid, bl, 0, c->stream, c->bnd.task,
                         c->bnd.n, static_cast<double*>(c->Q[qcur]), c->rc, 1, ibx);
    else
      hipLaunchKernelGGL((k_bnd_fill<P::nv, float>), grid, bl, 0, c->stream, c->bnd.task,
                         c->bnd.n, static_cast<float*>(c->Q[qcur]), c->rc, 1, ibx);
  });
  HIP_TRY(hipGetLastError());
  return 0;
}

namespace {

bool pick(Dispatch& d, const amun_cfg& c) {
  // one kernel-table unit per reconstruction, equation system and precision
  const int prec = c.precision == AMUN_PREC_MIXED ? AMUN_PREC_MIXED : AMUN_PREC_FP64;
  const int eq = c.eqsys + (c.eos == AMUN_EOS_ISO ? EQ_ISO : 0);
#define AMUN_PICK_CASE(name, rec, eqk, pr) \
  if (c.reconstruction == (rec) && eq == (eqk) && prec == (pr)) return name(d, c);
  AMUN_INST_LIST(AMUN_PICK_CASE)
#undef AMUN_PICK_CASE
  return false;
}

int validate(const amun_cfg& c, int* nv) {
  if (c.ndims != 2 && c.ndims != 3) return fail(500, "ndims must be 2 or 3");
  if (c.eqsys == AMUN_EQSYS_HD || c.eqsys == AMUN_EQSYS_SRHD) *nv = 5;
  else if (c.eqsys == AMUN_EQSYS_MHD || c.eqsys == AMUN_EQSYS_SRMHD) *nv = 9;
  else return fail(100, "unsupported equation_system");
  if (c.eos != AMUN_EOS_ADI && c.eos != AMUN_EOS_ISO)
    return fail(110, "unsupported equation_of_state");
  if (c.eos == AMUN_EOS_ISO) {
    // equations.F90:399-830: isothermal HD and MHD, no energy equation
    if (relativistic(c))
      return fail(110, "the relativistic systems are adiabatic only");
    if (!(c.csnd > 0.0)) return fail(500, "csnd must be positive");
    *nv -= 1;
  }
  switch (c.reconstruction) {
  case AMUN_REC_TVD: case AMUN_REC_WENO5Z: case AMUN_REC_WENO5YC: case AMUN_REC_WENO5NS:
  case AMUN_REC_MP5: case AMUN_REC_MP7: case AMUN_REC_GP: case AMUN_REC_WENO3:
  case AMUN_REC_LIMO3: case AMUN_REC_PPM: case AMUN_REC_CRWENO5Z: case AMUN_REC_CRWENO5YC:
  case AMUN_REC_CRWENO5NS: case AMUN_REC_CRMP5: case AMUN_REC_CRMP5LD: case AMUN_REC_CRMP7:
  case AMUN_REC_MGP:
    break;
  default: return fail(200, "unsupported reconstruction");
  }
  if (c.reconstruction == AMUN_REC_GP && c.ngp != 5)
    return fail(600, "the Gaussian-process reconstruction runs with ngp = 5 on the device path");
  if (c.reconstruction == AMUN_REC_MGP && c.ngp != 3 && c.ngp != 5)
    return fail(600, "the multi-dimensional GP runs with ngp = 3 or 5 on the device path");
  if ((c.reconstruction == AMUN_REC_GP || c.reconstruction == AMUN_REC_MGP) && !(c.sgp > 0.0))
    return fail(202, "sgp must be positive");
  if (c.limiter < 0 || c.limiter > 5) return fail(200, "unsupported limiter");
  if (relativistic(c)) {
    if (!(c.tolerance > 0.0) || c.nr_maxit < 1 || c.nr_extra < 0 || !(c.pmin >= 0.0))
      return fail(500, "tolerance, nr_maxit, nr_extra or pmin out of range");
    if (c.primitive_solver < AMUN_C2P_1DW || c.primitive_solver > AMUN_C2P_2DWU)
      return fail(110, "unsupported primitive_solver");
  } else {
    const bool iso = c.eos == AMUN_EOS_ISO;
    bool rs_ok = c.riemann_solver == AMUN_RS_HLL || c.riemann_solver == AMUN_RS_HLLC ||
                 (c.riemann_solver == AMUN_RS_HLLD && c.eqsys == AMUN_EQSYS_MHD) ||
                 (c.riemann_solver == AMUN_RS_HLLDM && c.eqsys == AMUN_EQSYS_MHD && iso) ||
                 c.riemann_solver == AMUN_RS_ROE;
    if (!rs_ok) return fail(300, "riemann_solver does not match the equation system");
  }
  if (c.time_advance < AMUN_TA_EULER || c.time_advance > AMUN_TA_SSPRK53)
    return fail(400, "unsupported time_advance");
  if (c.precision != AMUN_PREC_FP64 && c.precision != AMUN_PREC_MIXED)
    return fail(500, "unsupported precision");
  if (c.fix_positivity || c.clip_extrema)
    return fail(600, "fix_positivity / clip_extrema are not supported on the device path");
  if (c.fix_unphysical_cells)
    return fail(600, "fix_unphysical_cells is not supported on the device path");
  if (c.glm_source_terms < 0 || c.glm_source_terms > 2)
    return fail(500, "glm_source_terms must be 0 (none), 1 (eglm) or 2 (heglm)");
  if (!(c.viscosity >= 0.0) || !(c.resistivity >= 0.0))
    return fail(500, "viscosity and resistivity must be non-negative");
  if (c.ncells < 4 || c.ncells > 64 || c.ncells % 4)
    return fail(500, "ncells must be a multiple of 4 in [4, 64] (4x4 brick layout)");
  // interior formulas at the faces of cells -1 and ncells (the reference's
  // lower-order end stencils then stay in the outer ghost layers)
  int need = 2;
  switch (c.reconstruction) {
  case AMUN_REC_WENO5Z: case AMUN_REC_WENO5YC: case AMUN_REC_WENO5NS: case AMUN_REC_MP7:
  case AMUN_REC_PPM: need = 4; break;
  case AMUN_REC_GP: case AMUN_REC_MP5: need = 3; break;
  case AMUN_REC_MGP: need = (c.ngp - 1) / 2 + 1; break;
  }
  if (c.ncells + 2 * c.nghosts > 96)
    return fail(600, "ncells + 2 nghosts above 96");
  // with two ghost layers the reference's multi-dimensional limiting reads
  // the face state of each block's outermost cell, which is q itself only
  // after interfaces_tvd and interfaces_mgp
  if (c.mlp_limiting && c.nghosts == 2 && c.reconstruction != AMUN_REC_TVD &&
      c.reconstruction != AMUN_REC_MGP)
    return fail(600, "mlp_limiting with nghosts = 2 is built for the TVD and MGP "
                     "reconstructions (use nghosts >= 3)");
  if (c.nghosts > c.ncells)
    return fail(600, "nghosts above ncells (ghost reads reach one neighbour block only)");
  if (c.nghosts < need)
    return fail(600, "nghosts too small for the reconstruction stencil (WENO5, MP7 and PPM "
                     "need 4, GP and MP5 3, TVD, WENO3 and LimO3 2)");
  if (c.ncells < 3) return fail(500, "ncells too small for the stencil");
  for (int n = 0; n < 3; n++) {
    const int a = c.bnd_type[2 * n], b = c.bnd_type[2 * n + 1];
    if (a < AMUN_BND_PERIODIC || a > AMUN_BND_REFLECTIVE || b < AMUN_BND_PERIODIC ||
        b > AMUN_BND_REFLECTIVE)
      return fail(600, "boundary type not supported on the device path (periodic, open, "
                       "outflow, reflective)");
    if ((a == AMUN_BND_PERIODIC) != (b == AMUN_BND_PERIODIC))
      return fail(600, "an axis is periodic on both sides or on neither");
  }
  if ((c.bnd_type[0] || c.bnd_type[2] || (c.ndims == 3 && c.bnd_type[4])) && relativistic(c))
    return fail(600, "non-periodic sides are built for the HD and MHD systems");
  if (!(c.gamma > 1.0) || !(c.cfl > 0.0)) return fail(500, "gamma must exceed 1 and cfl be positive");
  return 0;
}

Dispatch& dispatch_of(amun_ctx* ctx) { return ctx->disp; }

// prepare_gp (interpolations.F90:4438-4547) for five points: the weights
// x^T A^-1 of a Gaussian process with the integrated squared-exponential
// kernel of width sgp, A the covariance of the cell averages, x that of the
// face value; symmetrised inverse, as the reference.  Extended precision
// (the reference works in its widest real kind and rounds the result).
void gp_weights(double sgp, double* cgp) {
  typedef long double L;
  const int n = 5, m = 2;
  const L sig = (L)(std::sqrt(2.0) * sgp);
  const L fc = (L)(0.5 * std::sqrt(3.14159265358979323846264338327950)) * sig;
  L g[5][10], x[5];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++) {
      if (j < n) {
        const L zl = ((L)(i - j) - 0.5L) / sig, zr = ((L)(i - j) + 0.5L) / sig;
        g[i][j] = fc * (std::erf(zr) - std::erf(zl));
      } else {
        g[i][j] = (i + n == j) ? 1.0L : 0.0L;
      }
    }
  // Gauss-Jordan (algebra.F90:1198-1305; the matrix is positive definite)
  for (int k = 0; k < n - 1; k++)
    for (int j = k + 1; j < n; j++) {
      const L t = g[j][k] / g[k][k];
      for (int i = k; i < 2 * n; i++) g[j][i] -= t * g[k][i];
    }
  for (int i = 0; i < n; i++) {
    const L t = g[i][i];
    for (int j = i; j < 2 * n; j++) g[i][j] /= t;
  }
  for (int k = n - 2; k >= 0; k--)
    for (int i = 0; i <= k; i++) {
      const L t = g[i][k + 1];
      for (int j = k; j < 2 * n; j++) g[i][j] -= g[k + 1][j] * t;
    }
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++)
      if (i != j) {
        const L a = 0.5L * (g[i][j + n] + g[j][i + n]);
        g[i][j + n] = a;
        g[j][i + n] = a;
      }
  for (int j = 0; j < n; j++) {
    const L z = (L)(0.5 - (double)(j - m)) / sig;
    x[j] = std::exp(-(z * z));
  }
  for (int j = 0; j < n; j++) {
    L acc = 0.0L;
    for (int i = 0; i < n; i++) acc += x[i] * g[i][j + n];
    cgp[j] = (double)acc;
  }
}

// prepare_mgp in quad precision (gp_weights.cpp, built by the host compiler)
extern "C" bool amun_mgp_weights_quad(int nd, int ngp, double sgp, double* out);
bool mgp_weights(const amun_cfg& cfg, std::vector<double>& out) {
  const int n = cfg.ndims == 3 ? cfg.ngp * cfg.ngp * cfg.ngp : cfg.ngp * cfg.ngp;
  out.assign((size_t)2 * cfg.ndims * n, 0.0);
  return amun_mgp_weights_quad(cfg.ndims, cfg.ngp, cfg.sgp, out.data());
}

int ssprk_stages(const amun_cfg& cfg) { return std::max(2, std::min(9, (int)cfg.stages)); }

// run constants of a configuration (block counts are filled in by the caller)
RunConst make_rc(const amun_cfg& cfg) {
  RunConst rc;
  std::memset(&rc, 0, sizeof(rc));
  const int nd = cfg.ndims, nc = cfg.ncells;
  rc.nd = nd;
  rc.nc = nc;
  rc.nruns = 1;
  rc.ncell = (nd == 3) ? nc * nc * nc : nc * nc;
  rc.limiter = cfg.limiter;
  rc.prol_limiter = cfg.prol_limiter;
  rc.gamma = cfg.gamma;
  rc.gammam1 = cfg.gamma - 1.0;
  rc.gammam1i = 1.0 / rc.gammam1;
  rc.csnd = cfg.csnd;
  rc.csnd2 = cfg.csnd * cfg.csnd;                 // equations.F90:849
  // the viscous time-step term takes both coefficients whatever the
  // system (evolution.F90:478-483)
  rc.nu = cfg.viscosity;
  rc.eta = cfg.resistivity;
  rc.grav = cfg.gravity ? 1 : 0;
  rc.gacc = cfg.gacc;
  rc.v4 = cfg.states_4vec && relativistic(cfg) ? 1 : 0;
  // initialize_evolution (evolution.F90:166-200): SSPRK(m,2) with m clamped
  // to [2, 9], SSPRK(4,3) and SSPRK(5,3) step at a multiple of the CFL
  // number, and the psi decay follows the scaled number
  double cfl = cfg.cfl;
  if (cfg.time_advance == AMUN_TA_SSPRK_M2) cfl = (ssprk_stages(cfg) - 1) * cfg.cfl;
  if (cfg.time_advance == AMUN_TA_SSPRK43) cfl = 2.0 * cfg.cfl;
  if (cfg.time_advance == AMUN_TA_SSPRK53) cfl = 2.65062919143939 * cfg.cfl;
  rc.cfl = cfl;
  // the LimO3 scheme raises the module's eps to at least 1e-12
  // (interpolations.F90:248); every user of eps then sees the raised value
  rc.eps = cfg.reconstruction == AMUN_REC_LIMO3 ? std::fmax(1.0e-12, cfg.eps) : cfg.eps;
  for (int n = 0; n < 6; n++) rc.bnd[n] = cfg.bnd_type[n];
  rc.ng = cfg.nghosts;
  rc.rad = cfg.limo3_rad;
  rc.ppm_const = cfg.ppm_const;
  rc.decay = std::exp(-cfg.alpha * cfl);
  rc.sc.gamma = cfg.gamma;
  rc.sc.gammaxi = rc.gammam1 / cfg.gamma;       // equations.F90:843
  rc.sc.rgammaxi = 1.0 / rc.sc.gammaxi;
  rc.sc.tol = cfg.tolerance;
  rc.sc.pmin = cfg.pmin;
  rc.sc.nrmax = cfg.nr_maxit;
  rc.sc.nrext = cfg.nr_extra;
  rc.sc.solver = cfg.primitive_solver;
  rc.kappa = std::fmin(cfg.kappa, (1.0 - cfg.cfl) / cfg.cfl);   // interpolations.F90:217
  rc.kbeta = cfg.kbeta;
  if (cfg.reconstruction == AMUN_REC_GP) gp_weights(cfg.sgp, rc.cgp);
  return rc;
}

int validate_grid(const amun_cfg& c, const amun_grid& g) {
  if (g.xblocks < 1 || g.yblocks < 1 || (c.ndims == 3 && g.zblocks < 1))
    return fail(500, "block counts must be positive");
  if (!(g.xmax > g.xmin) || !(g.ymax > g.ymin) || (c.ndims == 3 && !(g.zmax > g.zmin)))
    return fail(500, "domain bounds must be increasing");
  return 0;
}

// LDS bytes of the sweep along dir (chunked pencil I/O of the x and y
// sweeps: primitive slots, increment slots, kLdsLanes lanes each)
size_t sweep_lds(const amun_ctx* c, int dir, int /*wg*/) {
  const bool chunk = dir < 2, last = dir == c->rc.nd - 1;
  const bool chunk_d = c->qsize == 8;   // sweep_chunk_d<T>()
  const bool din = chunk && dir > 0 && chunk_d, dout = chunk && !last && chunk_d;
  const int nbuf = (chunk ? 1 : 0) + ((din || dout) ? 1 : 0);
  return (size_t)nbuf * 4 * c->nv * kLdsLanes * c->qsize;
}

// the slopes of the multi-dimensional limiting of the stage input Q[qcur]
// (k_mlp)
template <class P, typename T>
static void mlp_launch_t(amun_ctx* c, int qcur) {
  const RunConst& rc = c->rc;
  const int nd = rc.nd, npen = nd == 3 ? rc.nc * rc.nc : rc.nc;
  const long long n = (long long)rc.nblk * (rc.ncell + nd * 3 * npen);
  const dim3 grid((unsigned)((n + 255) / 256)), bl(256);
  const T* Q = static_cast<const T*>(c->q4 ? c->q4 : c->Q[qcur]);
  T* S = static_cast<T*>(c->mlps);
  T* G = static_cast<T*>(c->mlpg);
  if (nd == 3)
    hipLaunchKernelGGL((k_mlp<P, 3, T>), grid, bl, 0, c->stream, Q, c->nbr, rc, S, G);
  else
    hipLaunchKernelGGL((k_mlp<P, 2, T>), grid, bl, 0, c->stream, Q, c->nbr, rc, S, G);
}

static int mlp_launch(amun_ctx* c, int qcur) {
  const bool f64 = c->qsize == sizeof(double);
  by_system(c->cfg, [&](auto tag) {
    using P = decltype(tag);
    f64 ? mlp_launch_t<P, double>(c, qcur) : mlp_launch_t<P, float>(c, qcur);
  });
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

const char* amun_dev_last_error(void) { return g_last_error.c_str(); }

void amun_dev_default_cfg(amun_cfg* c) {
  std::memset(c, 0, sizeof(*c));
  c->ndims = 3;
  c->eqsys = AMUN_EQSYS_HD;
  c->eos = AMUN_EOS_ADI;
  c->ncells = 8;
  c->nghosts = 2;
  c->xblocks = c->yblocks = c->zblocks = 1;
  c->reconstruction = AMUN_REC_TVD;
  c->limiter = AMUN_LIM_MINMOD;
  c->riemann_solver = AMUN_RS_HLL;
  c->time_advance = AMUN_TA_RK2;
  c->precision = AMUN_PREC_FP64;
  c->gamma = 5.0 / 3.0;
  c->csnd = 1.0;
  c->cfl = 0.5;
  c->eps = 2.220446049250313e-16;
  c->alpha = 2.0;
  c->xmin = c->ymin = c->zmin = 0.0;
  c->xmax = c->ymax = c->zmax = 1.0;
  c->tolerance = 1.0e-10;
  c->pmin = 1.0e-16;
  c->nr_maxit = 100;
  c->nr_extra = 2;
  c->sgp = 9.0;
  c->kappa = 1.0;
  c->kbeta = 1.0;
  c->ngp = 5;
  c->stages = 2;
  c->prol_limiter = AMUN_LIM_MINMOD;
  c->limo3_rad = 0.5;
  c->ppm_const = 1.25;
}

static int init_device(amun_ctx* c, const std::vector<int32_t>& nbr,
                       const std::vector<int32_t>& brun, const std::vector<RunState>& hrs,
                       int nqblk, amun_ctx** out);

int amun_dev_init_batch(const amun_cfg* cfg, int32_t nruns, const amun_grid* grids,
                        amun_ctx** out) {
  if (!cfg || !out || !grids || nruns < 1) return fail(500, "bad argument");
  *out = nullptr;
  int nv = 0;
  int r = validate(*cfg, &nv);
  if (r) return r;
  for (int g = 0; g < nruns; g++)
    if ((r = validate_grid(*cfg, grids[g]))) return r;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail(1000 + (int)hipErrorNoDevice, "no HIP device");

  amun_ctx* c = new amun_ctx();
  c->cfg = *cfg;
  c->cfg.nv = nv;
  c->nv = nv;
  c->grids.assign(grids, grids + nruns);
  const int nd = cfg->ndims, nc = cfg->ncells;
  RunConst& rc = c->rc;
  rc = make_rc(*cfg);
  rc.nruns = nruns;

  // runs one after another in the pool; every run is a periodic base-level
  // block grid (setup_domain_default wiring) with its own cell sizes
  // (coordinates.F90:330-348) and its own step state
  std::vector<RunState> hrs(nruns);
  c->run_b0.assign(nruns + 1, 0);
  for (int g = 0; g < nruns; g++) {
    const amun_grid& G = grids[g];
    const int kr = (nd == 3) ? G.zblocks : 1;
    c->run_b0[g + 1] = c->run_b0[g] + G.xblocks * G.yblocks * kr;
    RunState& R = hrs[g];
    std::memset(&R, 0, sizeof(R));
    std::memcpy(&R.cmax_bits[0], &kTiny, 8);
    std::memcpy(&R.cmax_bits[1], &kTiny, 8);
    R.dt = 1.0;
    const double adx = (G.xmax - G.xmin) / (G.xblocks * nc);
    const double ady = (G.ymax - G.ymin) / (G.yblocks * nc);
    const double adz = (nd == 3) ? (G.zmax - G.zmin) / (kr * nc) : 1.0;
    R.dxi[0] = 1.0 / adx;
    R.dxi[1] = 1.0 / ady;
    R.dxi[2] = (nd == 3) ? 1.0 / adz : 1.0;
    R.h[0] = adx;
    R.h[1] = ady;
    R.h[2] = adz;
    R.dxmin = std::fmin(adx, ady);
    if (nd == 3) R.dxmin = std::fmin(R.dxmin, adz);
  }
  rc.nblk = c->run_b0[nruns];
  const bool bspec = cfg->bnd_type[0] || cfg->bnd_type[2] || (nd == 3 && cfg->bnd_type[4]);
  if (bspec && nruns != 1) {
    delete c;
    return fail(600, "non-periodic sides need a single-run context");
  }
  // virtual ghost blocks across non-periodic sides: key (source, sides)
  std::map<long long, int> vmap;
  std::vector<int4> btask;
  std::vector<int32_t> nbr((size_t)rc.nblk * 27, -1), brun(rc.nblk);
  c->coords.resize((size_t)rc.nblk * 3);
  for (int g = 0; g < nruns; g++) {
    const int ir = grids[g].xblocks, jr = grids[g].yblocks;
    const int kr = (nd == 3) ? grids[g].zblocks : 1, b0 = c->run_b0[g];
    for (int bz = 0; bz < kr; bz++)
      for (int by = 0; by < jr; by++)
        for (int bx = 0; bx < ir; bx++) {
          const int b = b0 + (bz * jr + by) * ir + bx;
          brun[b] = g;
          c->coords[b * 3 + 0] = bx;
          c->coords[b * 3 + 1] = by;
          c->coords[b * 3 + 2] = bz;
          for (int dk = -1; dk <= 1; dk++)
            for (int dj = -1; dj <= 1; dj++)
              for (int di = -1; di <= 1; di++) {
                if (nd == 2 && dk != 0) continue;
                const int pos[3] = {bx + di, by + dj, bz + dk}, cnt[3] = {ir, jr, kr};
                const int off[3] = {di, dj, dk};
                int sides = 0, src[3];
                for (int a = 0; a < 3; a++) {
                  src[a] = pos[a];
                  if (bspec && cfg->bnd_type[2 * a] && (pos[a] < 0 || pos[a] >= cnt[a])) {
                    sides |= (off[a] < 0 ? 1 : 2) << (2 * a);
                    src[a] = pos[a] - off[a];
                  }
                  src[a] = (src[a] + cnt[a]) % cnt[a];
                }
                int id = b0 + (src[2] * jr + src[1]) * ir + src[0];
                if (sides) {
                  const long long key = (long long)id * 64 + sides;
                  auto it = vmap.find(key);
                  if (it == vmap.end()) {
                    const int v = rc.nblk + (int)btask.size();
                    vmap[key] = v;
                    btask.push_back(make_int4(v, id, sides, 0));
                    id = v;
                  } else {
                    id = it->second;
                  }
                }
                nbr[(size_t)b * 27 + (dk + 1) * 9 + (dj + 1) * 3 + (di + 1)] = id;
              }
        }
  }
  const int nqblk = rc.nblk + (int)btask.size();
  int r2 = init_device(c, nbr, brun, hrs, nqblk, out);
  if (r2 || btask.empty()) return r2;
  amun_ctx* cc = *out;
  cc->bnd.on = true;
  cc->bnd.n = (int)btask.size();
  HIP_TRY(hipMalloc((void**)&cc->bnd.task, btask.size() * sizeof(int4)));
  HIP_TRY(hipMemcpy(cc->bnd.task, btask.data(), btask.size() * sizeof(int4),
                    hipMemcpyHostToDevice));
  return 0;
}

// Device buffers, tables and kernel selection of a context whose host-side
// layout (runs, blocks, neighbour table) is set up; the primitive pools hold
// nqblk blocks (the real blocks first, then the virtual blocks of a
// multi-level mesh).
static int init_device(amun_ctx* c, const std::vector<int32_t>& nbr,
                       const std::vector<int32_t>& brun, const std::vector<RunState>& hrs,
                       int nqblk, amun_ctx** out) {
  const amun_cfg* cfg = &c->cfg;
  const RunConst& rc = c->rc;
  const int nruns = rc.nruns, nv = c->nv, nd = cfg->ndims;
  c->tsteps = (nruns == 1) ? kTrace : std::max(64, kTrace / nruns);

  c->nstate = (size_t)rc.nblk * nv * rc.ncell;
  const size_t bytes = c->nstate * sizeof(double);
  c->qsize = (cfg->precision == AMUN_PREC_MIXED) ? sizeof(float) : sizeof(double);
  const size_t qbytes = (size_t)nqblk * nv * rc.ncell * c->qsize;
  hipError_t e;
#define ALLOC(ptr, sz)                                                       \
  e = hipMalloc((void**)&(ptr), (sz));                                       \
  if (e != hipSuccess) {                                                     \
    amun_dev_finalize(c);                                                    \
    return fail(1000 + (int)e, std::string("hipMalloc: ") + hipGetErrorString(e)); \
  }
  ALLOC(c->U0, bytes);
  ALLOC(c->U1, bytes);
  ALLOC(c->Q[0], qbytes);
  ALLOC(c->Q[1], qbytes);
  ALLOC(c->D, qbytes);
  if (cfg->precision == AMUN_PREC_MIXED) {
    ALLOC(c->C, c->nstate * sizeof(float));
  }
  ALLOC(c->nbr, nbr.size() * sizeof(int32_t));
  ALLOC(c->brun, brun.size() * sizeof(int32_t));
  ALLOC(c->st, sizeof(DevState));
  ALLOC(c->rs, (size_t)nruns * sizeof(RunState));
  ALLOC(c->trace, (size_t)nruns * c->tsteps * sizeof(double));
  ALLOC(c->scratch, (size_t)std::max(4096 * 8, 2 * nruns) * sizeof(double));
  {
    const long long npen = nd == 3 ? (long long)rc.nc * rc.nc : rc.nc;
    c->nspart = (int)((rc.nblk * npen + 63) / 64);
    ALLOC(c->spart, (size_t)c->nspart * sizeof(SpeedPart));
  }
  if (cfg->mlp_limiting) {
    const int gs = 3 * (nd == 3 ? rc.nc * rc.nc : rc.nc);
    ALLOC(c->mlps, (size_t)rc.nblk * nv * nd * rc.ncell * c->qsize);
    ALLOC(c->mlpg, (size_t)rc.nblk * nv * nd * gs * c->qsize);
  }
  c->vis = (cfg->viscosity > 0.0 ? 1 : 0) | (cfg->resistivity > 0.0 && has_field(*cfg) ? 2 : 0);
  if (rc.v4) ALLOC(c->q4, (size_t)rc.nblk * nv * rc.ncell * c->qsize);
  if (c->vis) {
    // the nine viscous and resistive source products of every cell (k_visc_src)
    ALLOC(c->vsrc, (size_t)rc.nblk * 9 * rc.ncell * c->qsize);
  }
  if (cfg->glm_source_terms > 0 && has_field(*cfg)) {
    // the eight GLM source products of every cell (k_glm_src)
    ALLOC(c->src, (size_t)rc.nblk * 8 * rc.ncell * c->qsize);
  }
  if (is_compact(cfg->reconstruction)) {
    // face states of one direction: [block][var][cell -1 .. nc][pencil]
    const size_t npen = rc.nd == 3 ? (size_t)rc.nc * rc.nc : (size_t)rc.nc;
    const size_t fsz = (size_t)rc.nblk * nv * (rc.nc + 2) * npen * c->qsize;
    ALLOC(c->rc.fsl, fsz);
    ALLOC(c->rc.fsr, fsz);
  }
  if (cfg->reconstruction == AMUN_REC_MGP) {
    std::vector<double> w;
    if (!mgp_weights(*cfg, w)) {
      amun_dev_finalize(c);
      return fail(201, "could not invert the MGP covariance matrix");
    }
    ALLOC(c->ugp, w.size() * sizeof(double));
    e = hipMemcpy(c->ugp, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      amun_dev_finalize(c);
      return fail(1000 + (int)e, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    c->rc.ugp = c->ugp;
    c->rc.mgp = (cfg->ngp - 1) / 2;
    c->rc.dgp = (int)(w.size() / (2 * rc.nd));
  }
#undef ALLOC
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    amun_dev_finalize(c);
    return fail(1000 + (int)e, "hipStreamCreate");
  }
  e = hipMemcpy(c->nbr, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(c->brun, brun.data(), brun.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(c->rs, hrs.data(), hrs.size() * sizeof(RunState), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(c->st, 0, sizeof(DevState));
  if (e == hipSuccess) e = hipMemset(c->U0, 0, bytes);
  // primitive arrays start at zero: a relativistic cell whose recovery fails
  // keeps the previous primitive values, as the reference's q array does
  if (e == hipSuccess) e = hipMemset(c->Q[0], 0, qbytes);
  if (e == hipSuccess) e = hipMemset(c->Q[1], 0, qbytes);
  if (e != hipSuccess) {
    amun_dev_finalize(c);
    return fail(1000 + (int)e, "initial copies");
  }
  if (!pick(dispatch_of(c), c->cfg)) {
    amun_dev_finalize(c);
    return fail(600, "no kernel instantiation for this configuration");
  }
  // the x and y sweeps stage pencil chunks in LDS: allow the largest
  // workgroup's share (above the 64 KiB default of dynamic LDS)
  for (int dir = 0; dir < nd; dir++) {
    const size_t lds = sweep_lds(c, dir, 256);
    if (!lds) continue;
    e = hipFuncSetAttribute((const void*)c->disp.sweep[dir],
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      amun_dev_finalize(c);
      return fail(1000 + (int)e, "hipFuncSetAttribute (LDS)");
    }
  }
  *out = c;
  return 0;
}

int amun_dev_init(const amun_cfg* cfg, amun_ctx** out) {
  if (!cfg || !out) return fail(500, "null argument");
  amun_grid g;
  std::memset(&g, 0, sizeof(g));
  g.xblocks = cfg->xblocks;
  g.yblocks = cfg->yblocks;
  g.zblocks = (cfg->ndims == 3) ? cfg->zblocks : 1;
  g.xmin = cfg->xmin; g.xmax = cfg->xmax;
  g.ymin = cfg->ymin; g.ymax = cfg->ymax;
  g.zmin = cfg->zmin; g.zmax = cfg->zmax;
  if (cfg->ndims != 3) { g.zmin = 0.0; g.zmax = 1.0; }
  return amun_dev_init_batch(cfg, 1, &g, out);
}

// A multi-level mesh of leaf blocks on the periodic base grid of cfg.
// meta[4 b .. 4 b + 3] = level (1 = base) and the first interior cell
// (i0, j0, k0) of block b in cells of its level.  The mesh must be 2:1
// balanced across faces, edges and corners (as AMUN's refinement keeps it).
int amun_dev_init_amr(const amun_cfg* cfg, int32_t nblk, const int32_t* meta, amun_ctx** out) {
  if (!cfg || !meta || !out || nblk < 1) return fail(500, "bad argument");
  *out = nullptr;
  int nv = 0;
  int r = validate(*cfg, &nv);
  if (r) return r;
  if (relativistic(*cfg))
    return fail(600, "multi-level meshes are not supported for the relativistic systems");
  if (cfg->bnd_type[0] || cfg->bnd_type[2] || (cfg->ndims == 3 && cfg->bnd_type[4]))
    return fail(600, "multi-level meshes are built for periodic domains");
  if (is_compact(cfg->reconstruction))
    return fail(600, "the compact schemes are built for single-level meshes");
  const int nd = cfg->ndims, nc = cfg->ncells, ng = cfg->nghosts;
  if (ng % 2 || 2 * ng > nc)
    return fail(600, "multi-level meshes need an even nghosts of at most ncells / 2");
  amun_grid G;
  std::memset(&G, 0, sizeof(G));
  G.xblocks = cfg->xblocks; G.yblocks = cfg->yblocks; G.zblocks = nd == 3 ? cfg->zblocks : 1;
  G.xmin = cfg->xmin; G.xmax = cfg->xmax; G.ymin = cfg->ymin; G.ymax = cfg->ymax;
  G.zmin = nd == 3 ? cfg->zmin : 0.0; G.zmax = nd == 3 ? cfg->zmax : 1.0;
  if ((r = validate_grid(*cfg, G))) return r;
  const int base[3] = {G.xblocks, G.yblocks, G.zblocks};

  // leaf table per level
  int nlev = 1;
  for (int b = 0; b < nblk; b++) {
    const int L = meta[4 * b];
    if (L < 1 || L > 16) return fail(500, "block level out of range");
    nlev = std::max(nlev, L);
    for (int d = 0; d < 3; d++)
      if (meta[4 * b + 1 + d] % nc || meta[4 * b + 1 + d] < 0)
        return fail(500, "block offsets must be non-negative multiples of ncells");
  }
  auto nb_at = [&](int L, int d) -> long { return (d < nd) ? (long)base[d] << (L - 1) : 1; };
  std::vector<std::vector<int>> lut(nlev + 2);
  for (int L = 1; L <= nlev; L++)
    lut[L].assign((size_t)nb_at(L, 0) * nb_at(L, 1) * nb_at(L, 2), -1);
  auto bid = [&](int L, long x, long y, long z) -> long {   // block position id
    return (z * nb_at(L, 1) + y) * nb_at(L, 0) + x;
  };
  for (int b = 0; b < nblk; b++) {
    const int L = meta[4 * b];
    const long x = meta[4 * b + 1] / nc, y = meta[4 * b + 2] / nc, z = meta[4 * b + 3] / nc;
    if (x >= nb_at(L, 0) || y >= nb_at(L, 1) || z >= nb_at(L, 2))
      return fail(500, "block outside the domain");
    int& e = lut[L][bid(L, x, y, z)];
    if (e >= 0) return fail(500, "two blocks at one position");
    e = b;
  }
  auto wrapl = [](long v, long n) { v %= n; return v < 0 ? v + n : v; };
  // leaf covering cell (x, y, z) of level L (wrapped), -1 if none at L
  auto leaf = [&](int L, long x, long y, long z) -> int {
    if (L < 1 || L > nlev) return -1;
    return lut[L][bid(L, x / nc, y / nc, z / nc)];
  };

  // neighbour table with virtual blocks for positions without a same-level leaf
  std::vector<int32_t> nbr((size_t)nblk * 27, -1);
  std::map<std::array<long, 4>, int> vmap;
  std::vector<std::array<long, 4>> vpos;          // level, block x, y, z
  std::vector<std::vector<unsigned char>> vmask;  // cells to fill
  const int ncell = (nd == 3) ? nc * nc * nc : nc * nc;
  for (int b = 0; b < nblk; b++) {
    const int L = meta[4 * b];
    const long p0[3] = {meta[4 * b + 1] / nc, meta[4 * b + 2] / nc, meta[4 * b + 3] / nc};
    for (int dk = -1; dk <= 1; dk++)
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          if (nd == 2 && dk != 0) continue;
          const long x = wrapl(p0[0] + di, nb_at(L, 0)), y = wrapl(p0[1] + dj, nb_at(L, 1));
          const long z = wrapl(p0[2] + dk, nb_at(L, 2));
          int& slot = nbr[(size_t)b * 27 + (dk + 1) * 9 + (dj + 1) * 3 + (di + 1)];
          const int same = lut[L][bid(L, x, y, z)];
          if (same >= 0) { slot = same; continue; }
          const std::array<long, 4> key = {L, x, y, z};
          auto it = vmap.find(key);
          int v;
          if (it == vmap.end()) {
            v = (int)vpos.size();
            vmap[key] = v;
            vpos.push_back(key);
            vmask.emplace_back(ncell, 0);
          } else {
            v = it->second;
          }
          slot = nblk + v;
          // the cells within nghosts of block b
          const int d3[3] = {di, dj, dk};
          int lo[3], hi[3];
          for (int d = 0; d < 3; d++) {
            if (d >= nd) { lo[d] = 0; hi[d] = 1; continue; }
            lo[d] = d3[d] > 0 ? 0 : (d3[d] < 0 ? nc - ng : 0);
            hi[d] = d3[d] > 0 ? ng : nc;
          }
          for (int k = lo[2]; k < hi[2]; k++)
            for (int j = lo[1]; j < hi[1]; j++)
              for (int i = lo[0]; i < hi[0]; i++)
                vmask[v][((size_t)k * nc + j) * nc + i] = 1;
        }
  }
  const int nvirt = (int)vpos.size();
  auto pack = [](int i, int j, int k) { return i | (j << 8) | (k << 16); };

  // ghost tasks of the virtual blocks
  std::vector<int4> rtask, ptask;
  for (int v = 0; v < nvirt; v++) {
    const int L = (int)vpos[v][0];
    for (int k = 0; k < (nd == 3 ? nc : 1); k++)
      for (int j = 0; j < nc; j++)
        for (int i = 0; i < nc; i++) {
          if (!vmask[v][((size_t)k * nc + j) * nc + i]) continue;
          const long x = vpos[v][1] * nc + i, y = vpos[v][2] * nc + j;
          const long z = nd == 3 ? vpos[v][3] * nc + k : 0;
          const int F = leaf(L + 1, 2 * x, 2 * y, 2 * z);
          if (F >= 0) {
            const int fi = (int)(2 * x - meta[4 * F + 1]), fj = (int)(2 * y - meta[4 * F + 2]);
            const int fk = nd == 3 ? (int)(2 * z - meta[4 * F + 3]) : 0;
            rtask.push_back(make_int4(nblk + v, pack(i, j, k), F, pack(fi, fj, fk)));
            continue;
          }
          const int C = leaf(L - 1, x / 2, y / 2, z / 2);
          if (C < 0) return fail(700, "mesh is not 2:1 balanced (or has a hole)");
          const int ci = (int)(x / 2 - meta[4 * C + 1]), cj = (int)(y / 2 - meta[4 * C + 2]);
          const int ck = nd == 3 ? (int)(z / 2 - meta[4 * C + 3]) : 0;
          // the slope stencil must only meet real blocks or restricted cells
          for (int d = 0; d < nd; d++)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
              int q[3] = {ci, cj, ck};
              q[d] += sgn;
              if (q[d] >= 0 && q[d] < nc) continue;
              const int di = q[0] < 0 ? -1 : (q[0] >= nc ? 1 : 0);
              const int dj = q[1] < 0 ? -1 : (q[1] >= nc ? 1 : 0);
              const int dk = q[2] < 0 ? -1 : (q[2] >= nc ? 1 : 0);
              const int nb = nbr[(size_t)C * 27 + (dk + 1) * 9 + (dj + 1) * 3 + (di + 1)];
              if (nb < nblk) continue;
              const int w = nb - nblk;
              const long cx = vpos[w][1] * nc + (q[0] - di * nc);
              const long cy = vpos[w][2] * nc + (q[1] - dj * nc);
              const long cz = nd == 3 ? vpos[w][3] * nc + (q[2] - dk * nc) : 0;
              const bool restricted = leaf(L, 2 * cx, 2 * cy, 2 * cz) >= 0;
              const size_t lc = ((size_t)(q[2] - dk * nc) * nc + (q[1] - dj * nc)) * nc + (q[0] - di * nc);
              if (!restricted || !vmask[w][lc])
                return fail(701, "prolongation stencil outside the restricted ghost cells");
            }
          ptask.push_back(make_int4(nblk + v, pack(i, j, k), C,
                                    pack(ci, cj, ck) | (int)((x & 1) << 24) |
                                        (int)((y & 1) << 25) | (int)((z & 1) << 26)));
        }
  }

  // flux-correction slots: coarse faces whose neighbours are finer
  std::vector<int32_t> ffix((size_t)nblk * 6, -1);
  std::vector<int4> ftask[3];
  int nfix = 0;
  for (int b = 0; b < nblk; b++) {
    const int L = meta[4 * b];
    const long o[3] = {meta[4 * b + 1], meta[4 * b + 2], meta[4 * b + 3]};
    for (int n = 0; n < nd; n++)
      for (int sd = 0; sd < 2; sd++) {
        long cc[3] = {o[0], o[1], o[2]};
        cc[n] = sd == 0 ? o[n] - 1 : o[n] + nc;
        for (int d = 0; d < nd; d++) cc[d] = wrapl(cc[d], nb_at(L, d) * nc);
        if (leaf(L, cc[0], cc[1], cc[2]) >= 0) continue;
        if (leaf(L + 1, 2 * cc[0], 2 * cc[1], 2 * cc[2]) < 0) continue;   // coarser
        const int fs = nfix++;
        ffix[(size_t)b * 6 + 2 * n + sd] = fs;
        const int t1 = n == 0 ? 1 : 0, t2 = n == 2 ? 1 : 2;
        for (int c2 = 0; c2 < (nd == 3 ? nc : 1); c2++)
          for (int c1 = 0; c1 < nc; c1++) {
            long e[3] = {0, 0, 0};
            e[n] = sd == 0 ? 2 * cc[n] + 1 : 2 * cc[n];
            e[t1] = 2 * (o[t1] + c1);
            if (nd == 3) e[t2] = 2 * (o[t2] + c2);
            const int F = leaf(L + 1, e[0], e[1], e[2]);
            if (F < 0) return fail(700, "mesh is not 2:1 balanced");
            const int f1 = (int)(e[t1] - meta[4 * F + 1 + t1]);
            const int f2 = nd == 3 ? (int)(e[t2] - meta[4 * F + 1 + t2]) : 0;
            ftask[n].push_back(make_int4(fs, c1 + nc * c2, F,
                                         f1 | (f2 << 8) | ((sd == 0 ? nc : 0) << 16)));
          }
      }
  }

  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail(1000 + (int)hipErrorNoDevice, "no HIP device");
  amun_ctx* c = new amun_ctx();
  c->cfg = *cfg;
  c->cfg.nv = nv;
  c->nv = nv;
  c->grids.assign(1, G);
  RunConst& rc = c->rc;
  rc = make_rc(*cfg);
  rc.nruns = 1;
  rc.nblk = nblk;
  c->run_b0 = {0, nblk};
  c->coords.assign((size_t)nblk * 3, 0);
  for (int b = 0; b < nblk; b++)
    for (int d = 0; d < 3; d++) c->coords[b * 3 + d] = meta[4 * b + 1 + d] / nc;
  // cell sizes: level 1 in dxi, 2^(level-1) per block in the sweeps; dt from
  // the finest level present (coordinates.F90:330-356, evolution.F90:467-476)
  std::vector<RunState> hrs(1);
  RunState& R = hrs[0];
  std::memset(&R, 0, sizeof(R));
  std::memcpy(&R.cmax_bits[0], &kTiny, 8);
  std::memcpy(&R.cmax_bits[1], &kTiny, 8);
  R.dt = 1.0;
  const int ff = 1 << (nlev - 1);
  const double adx1 = (G.xmax - G.xmin) / (G.xblocks * nc);
  const double ady1 = (G.ymax - G.ymin) / (G.yblocks * nc);
  const double adz1 = nd == 3 ? (G.zmax - G.zmin) / (G.zblocks * nc) : 1.0;
  R.dxi[0] = 1.0 / adx1;
  R.dxi[1] = 1.0 / ady1;
  R.dxi[2] = nd == 3 ? 1.0 / adz1 : 1.0;
  R.h[0] = adx1;
  R.h[1] = ady1;
  R.h[2] = adz1;
  R.dxmin = std::fmin((G.xmax - G.xmin) / (G.xblocks * (nc * ff)),
                      (G.ymax - G.ymin) / (G.yblocks * (nc * ff)));
  if (nd == 3) R.dxmin = std::fmin(R.dxmin, (G.zmax - G.zmin) / (G.zblocks * (nc * ff)));
  std::vector<int32_t> brun(nblk, 0), blev(nblk);
  for (int b = 0; b < nblk; b++) blev[b] = meta[4 * b] - 1;
  auto& A = c->amr;
  A.on = true;
  A.nreal = nblk;
  A.nvirt = nvirt;
  A.nlev = nlev;
  A.nfix = nfix;
  A.meta.assign(meta, meta + 4 * (size_t)nblk);
  A.nr = (int)rtask.size();
  A.np = (int)ptask.size();
  for (int d = 0; d < 3; d++) A.nf[d] = (int)ftask[d].size();
  if ((r = init_device(c, nbr, brun, hrs, nblk + nvirt, out))) return r;
  // init_device has freed c on failure; from here c is live
  hipError_t e = hipSuccess;
  auto up = [&](void** dst, const void* src, size_t bytes) {
    if (e != hipSuccess || !bytes) return;
    e = hipMalloc(dst, bytes);
    if (e == hipSuccess) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  };
  up((void**)&A.blev, blev.data(), blev.size() * sizeof(int32_t));
  up((void**)&A.rtask, rtask.data(), rtask.size() * sizeof(int4));
  up((void**)&A.ptask, ptask.data(), ptask.size() * sizeof(int4));
  for (int d = 0; d < 3; d++)
    up((void**)&A.ftask[d], ftask[d].data(), ftask[d].size() * sizeof(int4));
  up((void**)&A.ffix, ffix.data(), ffix.size() * sizeof(int32_t));
  if (e == hipSuccess)
    e = hipMalloc((void**)&A.FB, (size_t)std::max(nfix, 1) * nv * (nd == 3 ? nc * nc : nc) *
                                     sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&A.dummy, sizeof(DevState));
  if (e == hipSuccess) e = hipMemset(A.dummy, 0, sizeof(DevState));
  if (e != hipSuccess) {
    amun_dev_finalize(c);
    *out = nullptr;
    return fail(1000 + (int)e, "multi-level mesh tables");
  }
  return 0;
}

// blocks and virtual blocks of a multi-level mesh, its correction slots
int amun_dev_amr_info(amun_ctx* c, int32_t* info) {
  if (!c || !info) return fail(500, "null argument");
  info[0] = c->amr.on ? c->amr.nreal : c->rc.nblk;
  info[1] = c->amr.nvirt;
  info[2] = c->amr.nlev;
  info[3] = c->amr.nr;
  info[4] = c->amr.np;
  info[5] = c->amr.nfix;
  return 0;
}

static void drop_graphs(amun_ctx* c) {
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
}

void amun_dev_finalize(amun_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graphs(c);
  hipFree(c->tape.ck);
  hipFree(c->tape.ckc);
  hipFree(c->tape.dtc);
  hipFree(c->U0);
  hipFree(c->U1);
  hipFree(c->Q[0]);
  hipFree(c->Q[1]);
  hipFree(c->D);
  hipFree(c->C);
  hipFree(c->nbr);
  hipFree(c->brun);
  hipFree(c->st);
  hipFree(c->rs);
  hipFree(c->trace);
  hipFree(c->scratch);
  hipFree(c->spart);
  hipFree(c->rc.fsl);
  hipFree(c->ugp);
  hipFree(c->src);
  hipFree(c->vsrc);
  hipFree(c->q4);
  hipFree(c->mlps);
  hipFree(c->mlpg);
  hipFree(c->rc.fsr);
  hipFree(c->amr.blev);
  hipFree(c->bnd.task);
  hipFree(c->amr.rtask);
  hipFree(c->amr.ptask);
  for (int d = 0; d < 3; d++) hipFree(c->amr.ftask[d]);
  hipFree(c->amr.ffix);
  hipFree(c->amr.FB);
  hipFree(c->amr.dummy);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

// primitive variables and maximum speed of runs [r0, r0 + nr) after their
// conserved state changed on the host side (zero_c: the FP32 compensation
// of U0 restarts from zero)
static int after_upload(amun_ctx* c, int r0, int nr, bool zero_c = true) {
  Dispatch& d = dispatch_of(c);
  const int slot = (int)(c->nsteps & 1);
  const int b0 = c->run_b0[r0], nb = c->run_b0[r0 + nr] - b0;
  hipLaunchKernelGGL(k_reset_slot, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0,
                     c->stream, c->rs, slot, r0, nr);
  const long long n = (long long)nb * c->rc.ncell;
  // in place over the current primitive array (update_primitive_variables
  // overwrites pdata%q)
  hipLaunchKernelGGL(d.prim, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     c->U0, c->Q[c->qcur], c->rc, c->st, c->rs, c->brun, slot, b0, nb);
  if (c->C && zero_c)
    hipLaunchKernelGGL(k_zero_f32, dim3(2048), dim3(256), 0, c->stream,
                       c->C + (size_t)b0 * c->nv * c->rc.ncell,
                       (long long)nb * c->nv * c->rc.ncell);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

// global array U[nv][nz][ny][nx] of run r <-> its blocks (h: the run's
// segment of the pool, FP64 layout of lidx<8>)
static void scatter_run(const amun_ctx* c, int r, const double* u, double* h) {
  const RunConst& rc = c->rc;
  const int nc = rc.nc, nv = c->nv, b0 = c->run_b0[r], b1 = c->run_b0[r + 1];
  const amun_grid& G = c->grids[r];
  const size_t nx = (size_t)G.xblocks * nc, ny = (size_t)G.yblocks * nc;
  const size_t nz = (rc.nd == 3) ? (size_t)G.zblocks * nc : 1;
  const int kn = (rc.nd == 3) ? nc : 1;
  for (int b = b0; b < b1; b++) {
    const size_t bx = c->coords[b * 3], by = c->coords[b * 3 + 1], bz = c->coords[b * 3 + 2];
    for (int v = 0; v < nv; v++)
      for (int k = 0; k < kn; k++)
        for (int j = 0; j < nc; j++) {
          const double* src = u + ((v * nz + bz * kn + k) * ny + by * nc + j) * nx + bx * nc;
          double* dst = &h[((size_t)(b - b0) * nv + v) * rc.ncell];
          for (int i = 0; i < nc; i++) dst[lidx<8>(nc, rc.nd, i, j, k)] = src[i];
        }
  }
}

static void gather_run(const amun_ctx* c, int r, const double* h, double* u, int es) {
  const RunConst& rc = c->rc;
  const int nc = rc.nc, nv = c->nv, b0 = c->run_b0[r], b1 = c->run_b0[r + 1];
  const amun_grid& G = c->grids[r];
  const size_t nx = (size_t)G.xblocks * nc, ny = (size_t)G.yblocks * nc;
  const size_t nz = (rc.nd == 3) ? (size_t)G.zblocks * nc : 1;
  const int kn = (rc.nd == 3) ? nc : 1;
  for (int b = b0; b < b1; b++) {
    const size_t bx = c->coords[b * 3], by = c->coords[b * 3 + 1], bz = c->coords[b * 3 + 2];
    for (int v = 0; v < nv; v++)
      for (int k = 0; k < kn; k++)
        for (int j = 0; j < nc; j++) {
          double* dst = u + ((v * nz + bz * kn + k) * ny + by * nc + j) * nx + bx * nc;
          const double* src = &h[((size_t)(b - b0) * nv + v) * rc.ncell];
          if (es == 8)
            for (int i = 0; i < nc; i++) dst[i] = src[lidx<8>(nc, rc.nd, i, j, k)];
          else
            for (int i = 0; i < nc; i++) dst[i] = src[lidx<4>(nc, rc.nd, i, j, k)];
        }
  }
}

// Multi-level meshes: interiors of u from U0, q with its ghost layers
// (virtual blocks filled from the current state), and the ghost layers of u
// as prim2cons of q's (update_ghost_cells, boundaries.F90:6633-6736).
static int download_blocks_amr(amun_ctx* c, double* u, double* q) {
  const RunConst& rc = c->rc;
  const int ng = c->cfg.nghosts, nc = rc.nc, nv = c->nv, nd = rc.nd;
  const int im = nc + 2 * ng, km = (nd == 3) ? im : 1, kg = (nd == 3) ? ng : 0;
  const size_t per = (size_t)im * im * km;
  int r = c->amr.on ? amr_fill(c, c->qcur, (int)(c->nsteps & 1)) : bnd_fill(c, c->qcur);
  if (r) return r;
  const size_t n = (size_t)rc.nblk * per * nv;
  std::vector<double> qh(n);
  double* buf = nullptr;
  HIP_TRY(hipMalloc((void**)&buf, n * sizeof(double)));
  hipLaunchKernelGGL(c->disp.pack_q, dim3(2048), dim3(256), 0, c->stream,
                     (const void*)c->Q[c->qcur], buf, c->nbr, rc, ng);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(qh.data(), buf, n * sizeof(double), hipMemcpyDeviceToHost);
  hipFree(buf);
  if (e != hipSuccess) return fail(1000 + (int)e, "download_blocks");
  std::vector<double> h(c->nstate);
  HIP_TRY(hipMemcpy(h.data(), c->U0, c->nstate * sizeof(double), hipMemcpyDeviceToHost));
  const bool mhd = has_field(c->cfg), iso = c->cfg.eos == AMUN_EOS_ISO;
  const double g1 = c->cfg.gamma - 1.0;
  for (int b = 0; b < rc.nblk; b++)
    for (int k = 0; k < km; k++)
      for (int j = 0; j < im; j++)
        for (int i = 0; i < im; i++) {
          const size_t cell = (size_t)b * per + ((size_t)k * im + j) * im + i;
          double* uo = u + cell * nv;
          const double* qq = qh.data() + cell * nv;
          const int li = i - ng, lj = j - ng, lk = k - kg;
          if (li >= 0 && li < nc && lj >= 0 && lj < nc && lk >= 0 && lk < (nd == 3 ? nc : 1)) {
            for (int v = 0; v < nv; v++)
              uo[v] = h[((size_t)b * nv + v) * rc.ncell + lidx<8>(nc, nd, li, lj, lk)];
          } else {
            // prim2cons_hd_adi / prim2cons_mhd_adi (equations.F90:1762, 2815)
            uo[0] = qq[0];
            uo[1] = qq[0] * qq[1];
            uo[2] = qq[0] * qq[2];
            uo[3] = qq[0] * qq[3];
            const double ek = 0.5 * ((uo[1] * qq[1] + uo[2] * qq[2]) + uo[3] * qq[3]);
            if (iso) {
              // prim2cons_hd_iso / _mhd_iso (equations.F90:1360, 2220)
              for (int v = 4; v < nv; v++) uo[v] = qq[v];
            } else if (mhd) {
              for (int v = 4; v < 8; v++) uo[v] = qq[v];
              const double em = 0.5 * ((qq[4] * qq[4] + qq[5] * qq[5]) + qq[6] * qq[6]);
              uo[8] = (qq[8] / g1 + ek) + em;
            } else {
              uo[4] = qq[4] / g1 + ek;
            }
          }
        }
  if (q) std::memcpy(q, qh.data(), n * sizeof(double));
  return 0;
}

int amun_dev_upload_run(amun_ctx* c, int32_t r, const double* u) {
  if (!c || !u) return fail(500, "null argument");
  if (c->amr.on)
    return fail(600, "multi-level meshes transfer block arrays (amun_dev_upload_blocks)");
  if (r < 0 || r >= c->rc.nruns) return fail(500, "run index out of range");
  const size_t per = (size_t)c->nv * c->rc.ncell;
  const size_t n = (size_t)(c->run_b0[r + 1] - c->run_b0[r]) * per;
  std::vector<double> h(n);
  scatter_run(c, r, u, h.data());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(c->U0 + (size_t)c->run_b0[r] * per, h.data(), n * sizeof(double),
                    hipMemcpyHostToDevice));
  return after_upload(c, r, 1);
}

int amun_dev_upload_global(amun_ctx* c, const double* u) {
  if (!c || !u) return fail(500, "null argument");
  if (c->rc.nruns != 1) return fail(500, "a batch of runs is uploaded with amun_dev_upload_run");
  return amun_dev_upload_run(c, 0, u);
}

int amun_dev_upload_blocks(amun_ctx* c, const double* u) {
  if (!c || !u) return fail(500, "null argument");
  const RunConst& rc = c->rc;
  const int nc = rc.nc, nv = c->nv, ng = c->cfg.nghosts;
  const int im = nc + 2 * ng, km = (rc.nd == 3) ? im : 1, kn = (rc.nd == 3) ? nc : 1;
  const int kg = (rc.nd == 3) ? ng : 0;
  const size_t per = (size_t)nv * im * im * km;
  std::vector<double> h(c->nstate);
  for (int b = 0; b < rc.nblk; b++)
    for (int k = 0; k < kn; k++)
      for (int j = 0; j < nc; j++)
        for (int i = 0; i < nc; i++)
          for (int v = 0; v < nv; v++)
            h[(((size_t)b * nv + v) * rc.ncell) + lidx<8>(nc, rc.nd, i, j, k)] =
                u[b * per + ((((size_t)(k + kg) * im + (j + ng)) * im + (i + ng)) * nv) + v];
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(c->U0, h.data(), c->nstate * sizeof(double), hipMemcpyHostToDevice));
  return after_upload(c, 0, rc.nruns);
}

int amun_dev_download_run(amun_ctx* c, int32_t r, double* u, double* q) {
  if (!c || !u) return fail(500, "null argument");
  if (c->amr.on)
    return fail(600, "multi-level meshes transfer block arrays (amun_dev_download_blocks)");
  if (r < 0 || r >= c->rc.nruns) return fail(500, "run index out of range");
  HIP_TRY(hipStreamSynchronize(c->stream));
  const size_t per = (size_t)c->nv * c->rc.ncell;
  const size_t off = (size_t)c->run_b0[r] * per;
  const size_t n = (size_t)(c->run_b0[r + 1] - c->run_b0[r]) * per;
  std::vector<double> h(n);
  HIP_TRY(hipMemcpy(h.data(), c->U0 + off, n * sizeof(double), hipMemcpyDeviceToHost));
  gather_run(c, r, h.data(), u, 8);
  if (q) {
    if (c->qsize == sizeof(double)) {
      HIP_TRY(hipMemcpy(h.data(), (const double*)c->Q[c->qcur] + off, n * sizeof(double),
                        hipMemcpyDeviceToHost));
      gather_run(c, r, h.data(), q, 8);
    } else {
      std::vector<float> hf(n);
      HIP_TRY(hipMemcpy(hf.data(), (const float*)c->Q[c->qcur] + off, n * sizeof(float),
                        hipMemcpyDeviceToHost));
      for (size_t m = 0; m < n; m++) h[m] = hf[m];
      gather_run(c, r, h.data(), q, 4);
    }
  }
  return 0;
}

int amun_dev_download_global(amun_ctx* c, double* u, double* q) {
  if (!c || !u) return fail(500, "null argument");
  if (c->rc.nruns != 1) return fail(500, "a batch of runs is downloaded with amun_dev_download_run");
  return amun_dev_download_run(c, 0, u, q);
}

int amun_dev_download_blocks(amun_ctx* c, double* u, double* q) {
  if (!c || !u) return fail(500, "null argument");
  // ghost layers from virtual blocks: u ghosts as prim2cons of q's
  if (c->amr.on || c->bnd.on) return download_blocks_amr(c, u, q);
  Dispatch& d = dispatch_of(c);
  const int ng = c->cfg.nghosts, nc = c->rc.nc;
  const int im = nc + 2 * ng, km = (c->rc.nd == 3) ? im : 1;
  const size_t n = (size_t)c->rc.nblk * im * im * km * c->nv;
  double* buf = nullptr;
  HIP_TRY(hipMalloc((void**)&buf, n * sizeof(double)));
  for (int pass = 0; pass < 2; pass++) {
    double* dst = pass == 0 ? u : q;
    if (!dst) continue;
    if (pass == 0)
      hipLaunchKernelGGL(d.pack_u, dim3(2048), dim3(256), 0, c->stream, (const void*)c->U0,
                         buf, c->nbr, c->rc, ng);
    else
      hipLaunchKernelGGL(d.pack_q, dim3(2048), dim3(256), 0, c->stream,
                         (const void*)c->Q[c->qcur], buf, c->nbr, c->rc, ng);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(dst, buf, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      hipFree(buf);
      return fail(1000 + (int)e, "download_blocks");
    }
  }
  HIP_TRY(hipFree(buf));
  return 0;
}

static int stages_per_step(const amun_ctx* c) {
  switch (c->cfg.time_advance) {
  case AMUN_TA_EULER: return 1;
  case AMUN_TA_RK2: return 2;
  case AMUN_TA_SSPRK_M2: return ssprk_stages(c->cfg);
  case AMUN_TA_SSPRK43: return 4;
  case AMUN_TA_SSPRK53: return 5;
  default: return 3;
  }
}

// The stages of one step: combine kind (see combine() in kernels.h), ds =
// fac * dt, the state weights ca, cb of kinds 5-7 and whether the stage
// writes U0 (the step state) or U1.
struct StageDef { int kind; double fac, ca, cb; bool to_u0; };

static std::vector<StageDef> stage_table(const amun_ctx* c) {
  switch (c->cfg.time_advance) {
  case AMUN_TA_EULER:                               // evolve_euler
    return {{0, 1.0, 0, 0, true}};
  case AMUN_TA_RK2:                                 // evolve_rk2
    return {{0, 1.0, 0, 0, false}, {3, 1.0, 0, 0, true}};
  case AMUN_TA_SSPRK_M2: {                          // evolve_ssprk2
    // U1 = U0 + ds L(U0) is the reference's copy U1 = U0 then U1 + ds L(U1)
    const int m = ssprk_stages(c->cfg);
    const double ft = 1.0 / (m - 1), fl = 1.0 / m, fr = 1.0 - fl;
    std::vector<StageDef> t = {{0, ft, 0, 0, false}};
    for (int n = 2; n < m; n++) t.push_back({4, ft, 0, 0, false});
    t.push_back({5, ft, fl, fr, true});
    return t;
  }
  case AMUN_TA_SSPRK43: {                           // evolve_ssprk34
    const double b1 = 1.0 / 2.0, b3 = 1.0 / 6.0, a31 = 2.0 / 3.0, a33 = 1.0 / 3.0;
    return {{0, b1, 0, 0, false}, {4, b1, 0, 0, false}, {6, b3, a31, a33, false},
            {4, b1, 0, 0, true}};
  }
  case AMUN_TA_SSPRK53: {                           // evolve_ssprk35
    const double b1 = 3.77268915331368e-01, b3 = 2.42995220537396e-01;
    const double b4 = 2.38458932846290e-01, b5 = 2.87632146308408e-01;
    const double a31 = 3.55909775063327e-01, a33 = 6.44090224936673e-01;
    const double a41 = 3.67933791638137e-01, a44 = 6.32066208361863e-01;
    const double a53 = 2.37593836598569e-01, a55 = 7.62406163401431e-01;
    return {{0, b1, 0, 0, false}, {4, b1, 0, 0, false}, {6, b3, a31, a33, false},
            {6, b4, a41, a44, true}, {7, b5, a53, a55, true}};
  }
  default:                                          // evolve_rk3
    return {{0, 1.0, 0, 0, false}, {1, 1.0 / 4.0, 0, 0, false}, {2, 2.0 / 3.0, 0, 0, true}};
  }
}

static int kernels_per_step(const amun_ctx* c) {
  return stages_per_step(c) * (c->fuse ? 1 : c->rc.nd) + 1;
}

// Enqueue one whole step for cmax slot `slot` with the primitive state in
// Q[qcur]; host-side state is left alone.  With ev != nullptr an event is
// recorded after every kernel.
// Multi-level meshes, before the sweeps of a stage: the virtual blocks of the
// stage input Q[qcur] (restrictions, then prolongations), then the averaged
// fine fluxes of the coarse faces with finer neighbours.
static int amr_fill(amun_ctx* c, int qcur, int slot) {
  auto& A = c->amr;
  const dim3 bl(256);
  if (c->qsize == sizeof(double)) {
    double* Q = static_cast<double*>(c->Q[qcur]);
    if (A.nr)
      hipLaunchKernelGGL(k_amr_restrict<double>, dim3((A.nr + 255) / 256), bl, 0, c->stream,
                         A.rtask, A.nr, Q, c->rc, c->nv);
    if (A.np)
      hipLaunchKernelGGL(k_amr_prolong<double>, dim3((A.np + 255) / 256), bl, 0, c->stream,
                         A.ptask, A.np, Q, c->nbr, c->rc, c->nv, pres_index(c), c->st);
  } else {
    float* Q = static_cast<float*>(c->Q[qcur]);
    if (A.nr)
      hipLaunchKernelGGL(k_amr_restrict<float>, dim3((A.nr + 255) / 256), bl, 0, c->stream,
                         A.rtask, A.nr, Q, c->rc, c->nv);
    if (A.np)
      hipLaunchKernelGGL(k_amr_prolong<float>, dim3((A.np + 255) / 256), bl, 0, c->stream,
                         A.ptask, A.np, Q, c->nbr, c->rc, c->nv, pres_index(c), c->st);
  }
  // the slopes of the multi-dimensional limiting read the filled virtual
  // blocks and feed the fine-face fluxes
  if (c->mlps) {
    int r = mlp_launch(c, qcur);
    if (r) return r;
  }
  for (int d = 0; d < c->rc.nd; d++)
    if (A.nf[d])
      hipLaunchKernelGGL(c->disp.fineflux[d], dim3((A.nf[d] + 255) / 256), bl, 0, c->stream,
                         A.ftask[d], A.nf[d], (const void*)c->Q[qcur], c->nbr, c->rc, c->rs,
                         slot, A.FB, A.dummy, (const int*)A.blev, (const void*)c->mlps,
                         (const void*)c->mlpg);
  HIP_TRY(hipGetLastError());
  return 0;
}

// the GLM source products of the stage input Q[qcur] (k_glm_src)
static int glm_sources(amun_ctx* c, int qcur) {
  const RunConst& rc = c->rc;
  const long long n = (long long)rc.nblk * rc.ncell;
  const dim3 grid((unsigned)((n + 255) / 256)), bl(256);
  const int* blev = c->amr.on ? c->amr.blev : nullptr;
  by_system(c->cfg, [&](auto tag) {
    using P = decltype(tag);
    if constexpr (P::mhd) {
      if (c->qsize == sizeof(double))
        hipLaunchKernelGGL((k_glm_src<P, double>), grid, bl, 0, c->stream,
                           static_cast<const double*>(c->Q[qcur]), c->nbr, rc, c->rs, c->brun,
                           blev, static_cast<double*>(c->src));
      else
        hipLaunchKernelGGL((k_glm_src<P, float>), grid, bl, 0, c->stream,
                           static_cast<const float*>(c->Q[qcur]), c->nbr, rc, c->rs, c->brun,
                           blev, static_cast<float*>(c->src));
    }
  });
  HIP_TRY(hipGetLastError());
  return 0;
}

// the four-velocity states of the stage input Q[qcur] (k_to_4vec)
static int to_4vec(amun_ctx* c, int qcur) {
  const long long n = (long long)c->rc.nblk * c->rc.ncell;
  const dim3 grid((unsigned)((n + 255) / 256)), bl(256);
  by_system(c->cfg, [&](auto tag) {
    using P = decltype(tag);
    if constexpr (P::sr) {
      if (c->qsize == sizeof(double))
        hipLaunchKernelGGL((k_to_4vec<P, double>), grid, bl, 0, c->stream,
                           static_cast<const double*>(c->Q[qcur]),
                           static_cast<double*>(c->q4), n, c->rc.ncell);
      else
        hipLaunchKernelGGL((k_to_4vec<P, float>), grid, bl, 0, c->stream,
                           static_cast<const float*>(c->Q[qcur]), static_cast<float*>(c->q4),
                           n, c->rc.ncell);
    }
  });
  HIP_TRY(hipGetLastError());
  return 0;
}

// the viscous and resistive source products of the stage input Q[qcur]
// (k_visc_src)
static int visc_sources(amun_ctx* c, int qcur) {
  const RunConst& rc = c->rc;
  const long long n = (long long)rc.nblk * rc.ncell;
  const dim3 grid((unsigned)((n + 255) / 256)), bl(256);
  const int* blev = c->amr.on ? c->amr.blev : nullptr;
  by_system(c->cfg, [&](auto tag) {
    using P = decltype(tag);
    auto go = [&](auto ndtag, auto ttag) {
      constexpr int ND = decltype(ndtag)::value;
      using T = decltype(ttag);
      hipLaunchKernelGGL((k_visc_src<P, ND, T>), grid, bl, 0, c->stream,
                         static_cast<const T*>(c->Q[qcur]), c->nbr, rc, c->rs, c->brun, blev,
                         c->vis, static_cast<T*>(c->vsrc));
    };
    const bool f64 = c->qsize == sizeof(double);
    if (rc.nd == 3) {
      if (f64) go(std::integral_constant<int, 3>{}, double{});
      else go(std::integral_constant<int, 3>{}, float{});
    } else {
      if (f64) go(std::integral_constant<int, 2>{}, double{});
      else go(std::integral_constant<int, 2>{}, float{});
    }
  });
  HIP_TRY(hipGetLastError());
  return 0;
}

static int enqueue_step(amun_ctx* c, double dtnext, int slot, int qcur, hipEvent_t* ev,
                        double* rec = nullptr) {
  Dispatch& d = dispatch_of(c);
  const RunConst& rc = c->rc;
  const int npen = (rc.nd == 3) ? rc.nc * rc.nc : rc.nc;
  const long long nthreads = (long long)rc.nblk * npen;

  const std::vector<StageDef> stages = stage_table(c);
  if (rec)   // stage inputs: U^n, then every intermediate stage output
    HIP_TRY(hipMemcpyAsync(rec, c->U0, c->nstate * sizeof(double), hipMemcpyDeviceToDevice,
                           c->stream));
  for (size_t s = 0; s < stages.size(); s++) {
    if (c->amr.on) {
      int r = amr_fill(c, qcur, slot);
      if (r) return r;
    } else {
      if (c->bnd.on) {
        int r = bnd_fill(c, qcur);
        if (r) return r;
      }
      if (c->q4) {
        int r = to_4vec(c, qcur);
        if (r) return r;
      }
      if (c->mlps) {
        int r = mlp_launch(c, qcur);
        if (r) return r;
      }
    }
    if (c->src) {
      int r = glm_sources(c, qcur);
      if (r) return r;
    }
    if (c->vsrc) {
      int r = visc_sources(c, qcur);
      if (r) return r;
    }
    SweepArgs a;
    a.vsrc = c->vsrc;
    a.vis = c->vis;
    a.Qold = c->Q[qcur];
    a.mlps = c->mlps;
    a.mlpg = c->mlpg;
    a.src = c->src;
    a.glm = c->cfg.glm_source_terms;
    a.blev = c->amr.on ? c->amr.blev : nullptr;
    a.ffix = c->amr.on ? c->amr.ffix : nullptr;
    a.FB = c->amr.on ? c->amr.FB : nullptr;
    a.Q = c->q4 ? c->q4 : c->Q[qcur];
    a.U0 = c->U0;
    a.U1 = c->U1;
    a.Uout = stages[s].to_u0 ? c->U0 : c->U1;
    a.Qout = c->Q[qcur ^ 1];
    a.D = c->D;
    a.C = c->C;
    a.nbr = c->nbr;
    a.st = c->st;
    a.rs = c->rs;
    a.brun = c->brun;
    a.spart = c->spart;
    a.slot = slot;
    a.kind = stages[s].kind;
    a.last = (s + 1 == stages.size());
    a.comp = stages[s].to_u0;
    a.fac = stages[s].fac;
    a.ca = stages[s].ca;
    a.cb = stages[s].cb;
    a.dtnext = dtnext;
    if (c->fuse) {
      // every workgroup of 256 lanes holds whole blocks in each direction
      a.first = (s == 0);
      a.swz = c->swz[0];
      const dim3 grid((unsigned)((nthreads + 255) / 256)), block(256);
      const size_t lds = std::max(sweep_lds(c, 0, 256), sweep_lds(c, 1, 256));
      hipLaunchKernelGGL(d.stage, grid, block, lds, c->stream, a, rc);
      if (a.last)
        hipLaunchKernelGGL(k_reduce_speed, dim3(1), dim3(1024), 0, c->stream, c->spart,
                           (int)grid.x, c->rs, slot ^ 1, rc.nruns);
      if (ev) HIP_TRY(hipEventRecord(*ev++, c->stream));
    }
    for (int dir = 0; dir < rc.nd && !c->fuse; dir++) {
      const int wg = c->wg[dir];
      a.first = (s == 0 && dir == 0);
      if (d.compact[dir]) {
        const long long nt = (long long)rc.nblk * (rc.nd == 3 ? rc.nc * rc.nc : rc.nc) * c->nv;
        hipLaunchKernelGGL(d.compact[dir], dim3((unsigned)((nt + 63) / 64)), dim3(64), 0,
                           c->stream, (const void*)(c->q4 ? c->q4 : c->Q[qcur]), c->nbr, rc,
                           c->st);
      }
      a.swz = c->swz[dir];
      const dim3 grid((unsigned)((nthreads + wg - 1) / wg)), block(wg);
      hipLaunchKernelGGL(d.sweep[dir], grid, block, sweep_lds(c, dir, wg), c->stream, a, rc);
      if (a.last && dir == rc.nd - 1)
        hipLaunchKernelGGL(k_reduce_speed, dim3(1), dim3(1024), 0, c->stream, c->spart,
                           (int)grid.x, c->rs, slot ^ 1, rc.nruns);
      if (ev) HIP_TRY(hipEventRecord(*ev++, c->stream));
    }
    qcur ^= 1;
    if (rec && s + 1 < stages.size())
      HIP_TRY(hipMemcpyAsync(rec + (s + 1) * c->nstate, stages[s].to_u0 ? c->U0 : c->U1,
                             c->nstate * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  }
  hipLaunchKernelGGL(k_end_step, dim3((unsigned)((rc.nruns + 255) / 256)), dim3(256), 0,
                     c->stream, c->rs, c->trace, c->tsteps, rc, slot, dtnext);
  if (ev) HIP_TRY(hipEventRecord(*ev++, c->stream));
  HIP_TRY(hipGetLastError());
  return 0;
}

// One step: replayed from a captured graph when graphs are on (and no
// per-kernel events are wanted), launched kernel by kernel otherwise.
static int launch_step(amun_ctx* c, double dtnext, hipEvent_t* ev = nullptr) {
  const int slot = (int)(c->nsteps & 1), qcur = c->qcur;
  if (c->use_graph && !ev) {
    hipGraphExec_t exec = nullptr;
    for (auto& g : c->graphs)
      if (g.slot == slot && g.qcur == qcur && g.dtnext == dtnext) exec = g.exec;
    if (!exec) {
      hipGraph_t graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int r = enqueue_step(c, dtnext, slot, qcur, nullptr);
      hipError_t e = hipStreamEndCapture(c->stream, &graph);
      if (r) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
      }
      if (e != hipSuccess) return fail(1000 + (int)e, "hipStreamEndCapture");
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(1000 + (int)e, "hipGraphInstantiate");
      if (c->graphs.size() >= 8) drop_graphs(c);
      c->graphs.push_back({slot, qcur, dtnext, exec});
    }
    HIP_TRY(hipGraphLaunch(exec, c->stream));
  } else {
    int r = enqueue_step(c, dtnext, slot, qcur, ev);
    if (r) return r;
  }
  if (stages_per_step(c) & 1) c->qcur ^= 1;
  c->nsteps++;
  return 0;
}

static int peek(amun_ctx* c, double dtnext) {
  const int slot = (int)(c->nsteps & 1);
  hipLaunchKernelGGL(k_peek_dt, dim3((unsigned)((c->rc.nruns + 255) / 256)), dim3(256), 0,
                     c->stream, c->rs, c->rc, slot, dtnext, c->scratch);
  HIP_TRY(hipGetLastError());
  return 0;
}

int amun_dev_new_time_step(amun_ctx* c, double dtnext, double* dt, double* cmax) {
  if (!c) return fail(500, "null argument");
  int r = peek(c, dtnext);
  if (r) return r;
  double h[2];
  HIP_TRY(hipMemcpyAsync(h, c->scratch, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (dt) *dt = h[0];
  if (cmax) *cmax = h[1];
  return 0;
}

int amun_dev_step(amun_ctx* c, double dtnext, double* dt, double* cmax) {
  if (!c) return fail(500, "null argument");
  int r = peek(c, dtnext);
  if (!r) r = launch_step(c, dtnext);
  if (r) return r;
  double h[2];
  HIP_TRY(hipMemcpyAsync(h, c->scratch, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (dt) *dt = h[0];
  if (cmax) *cmax = h[1];
  return 0;
}

int amun_dev_run(amun_ctx* c, int32_t nsteps, double dtnext, double* dts) {
  if (!c || nsteps < 0) return fail(500, "bad argument");
  if (nsteps > c->tsteps) return fail(500, "too many steps for the dt trace in one call");
  std::vector<long long> s0(c->rc.nruns);
  {
    std::vector<RunState> h(c->rc.nruns);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(h.data(), c->rs, h.size() * sizeof(RunState), hipMemcpyDeviceToHost));
    for (int g = 0; g < c->rc.nruns; g++) s0[g] = h[g].step;
  }
  for (int s = 0; s < nsteps; s++) {
    int r = launch_step(c, dtnext);
    if (r) return r;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (dts && nsteps > 0) {
    const int nr = c->rc.nruns;
    std::vector<double> tr((size_t)nr * c->tsteps);
    HIP_TRY(hipMemcpy(tr.data(), c->trace, tr.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int s = 0; s < nsteps; s++)
      for (int g = 0; g < nr; g++)
        dts[(size_t)s * nr + g] = tr[(size_t)((s0[g] + s) % c->tsteps) * nr + g];
  }
  return 0;
}

static int timed_steps(amun_ctx* c, int32_t nsteps, double dtnext, double* elapsed_ms,
                       double* kernel_ms, int nk_want) {
  const int nk = kernels_per_step(c);
  const bool per_kernel = kernel_ms && nk_want >= nk;
  std::vector<hipEvent_t> ev((size_t)(per_kernel ? nk * nsteps : 0) + 2);
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
  int r = 0;
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipEventRecord(ev[0], c->stream));
  for (int s = 0; s < nsteps && !r; s++)
    r = launch_step(c, dtnext, per_kernel ? &ev[1 + (size_t)s * nk] : nullptr);
  hipEvent_t last = ev[ev.size() - 1];
  if (!r) {
    HIP_TRY(hipEventRecord(last, c->stream));
    HIP_TRY(hipEventSynchronize(last));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ev[0], last));
    *elapsed_ms = ms;
    if (per_kernel) {
      for (int k = 0; k < nk; k++) kernel_ms[k] = 0.0;
      for (int s = 0; s < nsteps; s++)
        for (int k = 0; k < nk; k++) {
          hipEvent_t a = (s == 0 && k == 0) ? ev[0] : ev[(size_t)s * nk + k];
          hipEvent_t b = ev[1 + (size_t)s * nk + k];
          float t = 0.f;
          HIP_TRY(hipEventElapsedTime(&t, a, b));
          kernel_ms[k] += t / nsteps;
        }
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return r;
}

int amun_dev_timed_run(amun_ctx* c, int32_t nsteps, double dtnext, double* elapsed_ms,
                       double* kernel_ms, int32_t nkernel_max, int32_t* nkernel) {
  if (!c || nsteps < 1 || !elapsed_ms) return fail(500, "bad argument");
  if (nkernel) *nkernel = kernels_per_step(c);
  return timed_steps(c, nsteps, dtnext, elapsed_ms, kernel_ms, kernel_ms ? nkernel_max : 0);
}

int amun_dev_set_graph(amun_ctx* c, int32_t enable) {
  if (!c) return fail(500, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (!enable) drop_graphs(c);
  c->use_graph = enable != 0;
  return 0;
}

int amun_dev_set_stage_fusion(amun_ctx* c, int32_t on) {
  if (!c) return fail(500, "null argument");
  if (on) {
    const int npen = c->rc.nd == 3 ? c->rc.nc * c->rc.nc : c->rc.nc;
    if (!c->disp.stage)
      return fail(600, "the fused stage is built for the TVD and WENO5-Z reconstructions of "
                       "the non-relativistic systems");
    if (256 % npen)
      return fail(600, "the fused stage needs whole blocks per workgroup of 256 lanes "
                       "(ncells^(ndims-1) dividing 256)");
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->fuse = on ? 1 : 0;
  drop_graphs(c);
  return 0;
}

int amun_dev_autotune(amun_ctx* c, int32_t nsteps, int32_t* choice) {
  if (!c || nsteps < 1) return fail(500, "bad argument");
  if (c->fuse)
    return fail(600, "autotune chooses the per-direction launches: call it with the fused "
                     "stage off");
  const int nd = c->rc.nd, nr = c->rc.nruns;
  HIP_TRY(hipStreamSynchronize(c->stream));
  // the candidates run real steps on the live state, which is restored
  // afterwards: conserved state, primitive state, compensation, run and
  // event state, and the host-side step counters
  const size_t bytes = c->nstate * sizeof(double), qbytes = c->nstate * c->qsize;
  double *u0 = nullptr, *qs = nullptr;
  float* cs = nullptr;
  RunState* rs = nullptr;
  DevState* st = nullptr;
  auto release = [&]() {
    hipFree(u0); hipFree(qs); hipFree(cs); hipFree(rs); hipFree(st);
  };
  hipError_t e = hipMalloc((void**)&u0, bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&qs, qbytes);
  if (e == hipSuccess && c->C) e = hipMalloc((void**)&cs, c->nstate * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&rs, nr * sizeof(RunState));
  if (e == hipSuccess) e = hipMalloc((void**)&st, sizeof(DevState));
  if (e != hipSuccess) {
    release();
    return fail(1000 + (int)e, "autotune: hipMalloc");
  }
  const long long nsteps0 = c->nsteps;
  const int qcur0 = c->qcur;
  auto save = [&](bool back) -> hipError_t {
    hipError_t x = hipSuccess;
    auto cp = [&](void* a, void* b, size_t n) {
      if (x == hipSuccess) x = hipMemcpyAsync(back ? a : b, back ? b : a, n,
                                              hipMemcpyDeviceToDevice, c->stream);
    };
    cp(c->U0, u0, bytes);
    cp(c->Q[qcur0], qs, qbytes);
    if (c->C) cp(c->C, cs, c->nstate * sizeof(float));
    cp(c->rs, rs, nr * sizeof(RunState));
    cp(c->st, st, sizeof(DevState));
    if (x == hipSuccess) x = hipStreamSynchronize(c->stream);
    if (back) {
      c->nsteps = nsteps0;
      c->qcur = qcur0;
    }
    return x;
  };
  if ((e = save(false)) != hipSuccess) {
    release();
    return fail(1000 + (int)e, "autotune: save");
  }
  const bool graphs = c->use_graph;
  c->use_graph = false;
  const int wgs[3] = {64, 128, 256};
  double best[3] = {1e300, 1e300, 1e300};
  int bwg[3] = {256, 256, 256}, bsw[3] = {0, 0, 0};
  const int nk = kernels_per_step(c);
  std::vector<double> kms(nk);
  int r = 0;
  // the chunked x and y sweeps stage kLdsLanes lanes of LDS per workgroup
  // whatever its size, so only full workgroups fill a CU there
  for (int w = 0; w < 3 && !r; w++)
    for (int z = 0; z < 2 && !r; z++) {
      for (int dir = 0; dir < nd; dir++) {
        c->wg[dir] = dir < 2 ? kLdsLanes : wgs[w];
        c->swz[dir] = z;
      }
      double ms = 0.0;
      r = timed_steps(c, 1, 2.0e30, &ms, nullptr, 0);
      if (!r) r = timed_steps(c, nsteps, 2.0e30, &ms, kms.data(), nk);
      if (!r && (e = save(true)) != hipSuccess) r = fail(1000 + (int)e, "autotune: restore");
      if (r) break;
      // per-direction time summed over the stages of a step
      for (int dir = 0; dir < nd; dir++) {
        double t = 0.0;
        for (int s = 0; s < stages_per_step(c); s++) t += kms[(size_t)s * nd + dir];
        if (t < best[dir]) { best[dir] = t; bwg[dir] = c->wg[dir]; bsw[dir] = z; }
      }
    }
  for (int dir = 0; dir < nd; dir++) { c->wg[dir] = bwg[dir]; c->swz[dir] = bsw[dir]; }
  drop_graphs(c);
  c->use_graph = graphs;
  release();
  if (r) return r;
  if (choice)
    for (int dir = 0; dir < nd; dir++) {
      choice[2 * dir] = bwg[dir];
      choice[2 * dir + 1] = bsw[dir];
    }
  return 0;
}

int amun_dev_set_launch(amun_ctx* c, const int32_t* choice) {
  if (!c || !choice) return fail(500, "null argument");
  for (int dir = 0; dir < c->rc.nd; dir++) {
    const int wg = choice[2 * dir];
    if (wg != 64 && wg != 128 && wg != 256) return fail(500, "workgroup size must be 64, 128 or 256");
    if (dir < 2 && wg != kLdsLanes)
      return fail(500, "the x and y sweeps run workgroups of 256 lanes");
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int dir = 0; dir < c->rc.nd; dir++) {
    c->wg[dir] = choice[2 * dir];
    c->swz[dir] = choice[2 * dir + 1] != 0;
  }
  drop_graphs(c);
  return 0;
}

int amun_dev_sync(amun_ctx* c) {
  if (!c) return fail(500, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

static int integrals_range(amun_ctx* c, int r0, int nr, double out[8]) {
  Dispatch& d = dispatch_of(c);
  const int nwg = 1024;
  std::vector<RunState> hrs(c->rc.nruns);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(hrs.data(), c->rs, hrs.size() * sizeof(RunState), hipMemcpyDeviceToHost));
  for (int m = 0; m < 8; m++) out[m] = 0.0;
  for (int g = r0; g < r0 + nr; g++) {
    const int b0 = c->run_b0[g], nb = c->run_b0[g + 1] - b0;
    hipLaunchKernelGGL(d.integ, dim3(nwg), dim3(256), 0, c->stream, c->U0, c->rc, c->scratch,
                       b0, nb, c->amr.on ? (const int*)c->amr.blev : nullptr);
    HIP_TRY(hipGetLastError());
    std::vector<double> part((size_t)nwg * 8);
    HIP_TRY(hipMemcpyAsync(part.data(), c->scratch, part.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int w = 0; w < nwg; w++)
      for (int m = 0; m < 8; m++) acc[m] += part[(size_t)w * 8 + m];
    const RunState& R = hrs[g];
    const double dvol = 1.0 / (R.dxi[0] * R.dxi[1] * (c->rc.nd == 3 ? R.dxi[2] : 1.0));
    const double f[8] = {1, 1, 1, 1, 1, 0.5, 0.5, 1};
    for (int m = 0; m < 8; m++) out[m] += acc[m] * f[m] * dvol;
  }
  return 0;
}

int amun_dev_integrals(amun_ctx* c, double out[8]) {
  if (!c || !out) return fail(500, "null argument");
  return integrals_range(c, 0, c->rc.nruns, out);
}

int amun_dev_integrals_run(amun_ctx* c, int32_t r, double out[8]) {
  if (!c || !out) return fail(500, "null argument");
  if (r < 0 || r >= c->rc.nruns) return fail(500, "run index out of range");
  return integrals_range(c, r, 1, out);
}

// statistics of runs r0 .. r0+nr-1 of the state U0 (its primitives Q[qcur]):
// mean, minimum and maximum of the seven quantities of store_integrals; the
// means weigh every cell with its volume and divide by the runs' volume
static int statistics_range(amun_ctx* c, int r0, int nr, double out[21]) {
  Dispatch& d = dispatch_of(c);
  const int nwg = 1024;
  const double big = 1.7976931348623157e308;
  HIP_TRY(hipStreamSynchronize(c->stream));
  double sum[7] = {0, 0, 0, 0, 0, 0, 0}, lo[7], hi[7], wsum = 0.0;
  for (int m = 0; m < 7; m++) { lo[m] = big; hi[m] = -big; }
  std::vector<int32_t> blev;
  if (c->amr.on) {
    blev.resize(c->rc.nblk);
    HIP_TRY(hipMemcpy(blev.data(), c->amr.blev, blev.size() * sizeof(int32_t),
                      hipMemcpyDeviceToHost));
  }
  for (int g = r0; g < r0 + nr; g++) {
    const int b0 = c->run_b0[g], nb = c->run_b0[g + 1] - b0;
    hipLaunchKernelGGL(d.stats, dim3(nwg), dim3(256), 0, c->stream, (const void*)c->Q[c->qcur],
                       c->rc, c->scratch, b0, nb,
                       c->amr.on ? (const int*)c->amr.blev : nullptr);
    HIP_TRY(hipGetLastError());
    std::vector<double> part((size_t)nwg * 21);
    HIP_TRY(hipMemcpyAsync(part.data(), c->scratch, part.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int w = 0; w < nwg; w++)
      for (int m = 0; m < 7; m++) {
        const double* p = &part[(size_t)w * 21 + 3 * m];
        sum[m] += p[0];
        lo[m] = p[1] < lo[m] ? p[1] : lo[m];
        hi[m] = p[2] > hi[m] ? p[2] : hi[m];
      }
    // the leaves tile the domain: their weights add up to the base cells
    for (int b = b0; b < b0 + nb; b++)
      wsum += (c->amr.on ? std::ldexp(1.0, -c->rc.nd * blev[b]) : 1.0) * c->rc.ncell;
  }
  // without a magnetic field the reference reports zeros for |B|, psi, M_A
  const bool mhd = has_field(c->cfg), iso = c->cfg.eos == AMUN_EOS_ISO;
  for (int m = 0; m < 7; m++) {
    const bool zero = (!mhd && (m == 3 || m == 4 || m == 6)) || (iso && m == 1);
    out[3 * m] = zero ? 0.0 : sum[m] / wsum;
    out[3 * m + 1] = zero ? 0.0 : lo[m];
    out[3 * m + 2] = zero ? 0.0 : hi[m];
  }
  return 0;
}

int amun_dev_statistics(amun_ctx* c, double out[21]) {
  if (!c || !out) return fail(500, "null argument");
  return statistics_range(c, 0, c->rc.nruns, out);
}

int amun_dev_statistics_run(amun_ctx* c, int32_t r, double out[21]) {
  if (!c || !out) return fail(500, "null argument");
  if (r < 0 || r >= c->rc.nruns) return fail(500, "run index out of range");
  return statistics_range(c, r, 1, out);
}

int amun_dev_run_state(amun_ctx* c, double* time, double* dt, int64_t* step) {
  if (!c) return fail(500, "null argument");
  std::vector<RunState> h(c->rc.nruns);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(h.data(), c->rs, h.size() * sizeof(RunState), hipMemcpyDeviceToHost));
  for (int g = 0; g < c->rc.nruns; g++) {
    if (time) time[g] = h[g].time;
    if (dt) dt[g] = h[g].dt;
    if (step) step[g] = h[g].step;
  }
  return 0;
}

int amun_dev_get_time(amun_ctx* c, double* time, int64_t* step) {
  if (!c) return fail(500, "null argument");
  RunState h;
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(&h, c->rs, sizeof(RunState), hipMemcpyDeviceToHost));
  if (time) *time = h.time;
  if (step) *step = h.step;
  return 0;
}

int amun_dev_nblocks(amun_ctx* c, int32_t* nblk) {
  if (!c || !nblk) return fail(500, "null argument");
  *nblk = c->rc.nblk;
  return 0;
}

int amun_dev_nruns(amun_ctx* c, int32_t* nruns) {
  if (!c || !nruns) return fail(500, "null argument");
  *nruns = c->rc.nruns;
  return 0;
}

int amun_dev_run_blocks(amun_ctx* c, int32_t r, int32_t* first, int32_t* count) {
  if (!c) return fail(500, "null argument");
  if (r < 0 || r >= c->rc.nruns) return fail(500, "run index out of range");
  if (first) *first = c->run_b0[r];
  if (count) *count = c->run_b0[r + 1] - c->run_b0[r];
  return 0;
}

int amun_dev_block_coords(amun_ctx* c, int32_t* coords) {
  if (!c || !coords) return fail(500, "null argument");
  std::memcpy(coords, c->coords.data(), c->coords.size() * sizeof(int32_t));
  return 0;
}

int amun_dev_positivity_events(amun_ctx* c, uint64_t* n) {
  if (!c || !n) return fail(500, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  DevState h;
  HIP_TRY(hipMemcpy(&h, c->st, sizeof(DevState), hipMemcpyDeviceToHost));
  *n = h.npos;
  return 0;
}

// Refill the stand-ins of the ghost cells (virtual blocks of level changes
// and of non-periodic sides) from the current primitive state Q[qcur], as
// boundary_variables leaves them after a stage; the positivity events of
// this fill go to a scratch counter (the steps count their own fills).
static int refill_ghosts(amun_ctx* c) {
  const RunConst& rc = c->rc;
  if (c->amr.on) {
    auto& A = c->amr;
    const dim3 bl(256);
    if (c->qsize == sizeof(double)) {
      double* Q = static_cast<double*>(c->Q[c->qcur]);
      if (A.nr)
        hipLaunchKernelGGL(k_amr_restrict<double>, dim3((A.nr + 255) / 256), bl, 0, c->stream,
                           A.rtask, A.nr, Q, rc, c->nv);
      if (A.np)
        hipLaunchKernelGGL(k_amr_prolong<double>, dim3((A.np + 255) / 256), bl, 0, c->stream,
                           A.ptask, A.np, Q, c->nbr, rc, c->nv, c->nv - 1, A.dummy);
    } else {
      float* Q = static_cast<float*>(c->Q[c->qcur]);
      if (A.nr)
        hipLaunchKernelGGL(k_amr_restrict<float>, dim3((A.nr + 255) / 256), bl, 0, c->stream,
                           A.rtask, A.nr, Q, rc, c->nv);
      if (A.np)
        hipLaunchKernelGGL(k_amr_prolong<float>, dim3((A.np + 255) / 256), bl, 0, c->stream,
                           A.ptask, A.np, Q, c->nbr, rc, c->nv, c->nv - 1, A.dummy);
    }
    HIP_TRY(hipGetLastError());
  } else if (c->bnd.on) {
    int r = bnd_fill(c, c->qcur);
    if (r) return r;
  }
  return 0;
}

int amun_dev_refinement(amun_ctx* c, const amun_refine_cfg* cfg, int32_t* crit, float* ind) {
  if (!c || !cfg || !crit) return fail(500, "null argument");
  const RunConst& rc = c->rc;
  const int nb = c->amr.on ? c->amr.nreal : rc.nblk;
  // the ghost cells the criterion reads (one layer beyond the interior, two
  // for the central differences) must hold the current state
  if (int r = refill_ghosts(c)) return r;
  RefineArgs ra;
  ra.crefmin = cfg->crefmin;
  ra.crefmax = cfg->crefmax;
  ra.epsref = cfg->epsref;
  ra.vortmin = cfg->vortmin;
  ra.vortmax = cfg->vortmax;
  ra.jabsmin = cfg->jabsmin;
  ra.jabsmax = cfg->jabsmax;
  ra.varmask = cfg->varmask;
  ra.vort = cfg->vort;
  ra.jabs = cfg->jabs;
  int* dcrit = nullptr;
  float* dind = nullptr;
  HIP_TRY(hipMalloc((void**)&dcrit, (size_t)nb * sizeof(int)));
  hipError_t e = hipMalloc((void**)&dind, (size_t)nb * 3 * sizeof(float));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(c->disp.refine, dim3((unsigned)nb), dim3(256), 0, c->stream,
                       (const void*)c->Q[c->qcur], c->nbr, rc, ra, dcrit, dind);
    e = hipGetLastError();
  }
  std::vector<int> hc(nb);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(hc.data(), dcrit, (size_t)nb * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess && ind)
    e = hipMemcpy(ind, dind, (size_t)nb * 3 * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(dcrit);
  (void)hipFree(dind);
  if (e != hipSuccess) return fail(1000 + (int)e, std::string("refinement: ") + hipGetErrorString(e));
  // the level limits of update_mesh (mesh.F90:1547-1550)
  for (int b = 0; b < nb; b++) {
    int r = hc[b];
    if (cfg->minlev > 0 || cfg->maxlev > 0) {
      const int lev = c->amr.on ? c->amr.meta[4 * b] : 1;
      if (cfg->minlev > 0 && lev < cfg->minlev) r = 1;
      if (cfg->minlev > 0 && lev == cfg->minlev) r = std::max(0, r);
      if (cfg->maxlev > 0 && lev == cfg->maxlev) r = std::min(0, r);
      if (cfg->maxlev > 0 && lev > cfg->maxlev) r = -1;
    }
    crit[b] = r;
  }
  return 0;
}

// The data movement of update_mesh (mesh.F90:756-833) onto the leaf set
// meta: a block kept from the old mesh is copied, a child of an old leaf is
// prolonged from it (prolong_block, mesh.F90:1031-1257), the parent of 2^d
// old leaves is restricted from them (restrict_block, 1258-1454), all in one
// launch from the old pools to a new context's; the new context then takes
// the old one's place (stream, time, step, dt, event counters, launch
// shapes and graph setting kept) and recomputes the primitive state and the
// maximum speed as update_variables does after update_mesh
// (evolution.F90:321-340).
int amun_dev_amr_remesh(amun_ctx* c, int32_t nblk, const int32_t* meta) {
  if (!c || !meta || nblk < 1) return fail(500, "bad argument");
  if (!c->amr.on) return fail(600, "mesh updates need a multi-level context (amun_dev_init_amr)");
  const int nc = c->rc.nc, nd = c->rc.nd;
  // the old leaves by level and position (in blocks of their level)
  std::map<std::array<long, 4>, int> old;
  const auto& M = c->amr.meta;
  for (int b = 0; b < c->amr.nreal; b++)
    old[{M[4 * b], M[4 * b + 1] / nc, M[4 * b + 2] / nc, M[4 * b + 3] / nc}] = b;
  auto find = [&](long L, long x, long y, long z) {
    auto it = old.find({L, x, y, z});
    return it == old.end() ? -1 : it->second;
  };
  std::vector<int4> task(nblk);
  std::vector<int32_t> kids;
  for (int b = 0; b < nblk; b++) {
    const long L = meta[4 * b];
    for (int d = 0; d < 3; d++)
      if (meta[4 * b + 1 + d] < 0 || meta[4 * b + 1 + d] % nc)
        return fail(500, "block offsets must be non-negative multiples of ncells");
    const long x = meta[4 * b + 1] / nc, y = meta[4 * b + 2] / nc;
    const long z = nd == 3 ? meta[4 * b + 3] / nc : 0;
    int src = find(L, x, y, z);
    if (src >= 0) {
      task[b] = make_int4(REMESH_COPY, src, 0, 0);
      continue;
    }
    src = find(L - 1, x / 2, y / 2, z / 2);
    if (src >= 0) {
      task[b] = make_int4(REMESH_PROLONG, src, (int)((x & 1) | (y & 1) << 1 | (z & 1) << 2), 0);
      continue;
    }
    task[b] = make_int4(REMESH_RESTRICT, (int)kids.size(), 0, 0);
    for (int q = 0; q < 8; q++) {
      if (nd == 2 && q >= 4) {
        const int32_t same_k = kids[kids.size() - 4];   // never read in 2D
        kids.push_back(same_k);
        continue;
      }
      const int k = find(L + 1, 2 * x + (q & 1), 2 * y + (q >> 1 & 1), 2 * z + (q >> 2 & 1));
      if (k < 0)
        return fail(702, "a block of the new mesh is neither an old leaf, a child of one, "
                         "nor the parent of 2^d old leaves");
      kids.push_back(k);
    }
  }
  // the prolongations read one cell beyond the old parents' interiors
  if (int r = refill_ghosts(c)) return r;
  HIP_TRY(hipStreamSynchronize(c->stream));
  amun_ctx* n = nullptr;
  if (int r = amun_dev_init_amr(&c->cfg, nblk, meta, &n)) return r;
  int4* dtask = nullptr;
  int32_t* dkids = nullptr;
  RunState R;
  hipError_t e = hipMalloc((void**)&dtask, task.size() * sizeof(int4));
  if (e == hipSuccess) e = hipMalloc((void**)&dkids, std::max<size_t>(kids.size(), 1) * sizeof(int32_t));
  if (e == hipSuccess)
    e = hipMemcpy(dtask, task.data(), task.size() * sizeof(int4), hipMemcpyHostToDevice);
  if (e == hipSuccess && !kids.empty())
    e = hipMemcpy(dkids, kids.data(), kids.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  // events so far carry over; the prolongations add theirs
  if (e == hipSuccess)
    e = hipMemcpyAsync(n->st, c->st, sizeof(DevState), hipMemcpyDeviceToDevice, n->stream);
  if (e == hipSuccess) {
    const long long cells = (long long)nblk * n->rc.ncell;
    const dim3 grid((unsigned)((cells + 255) / 256)), bl(256);
    const int ien = pres_index(c);
    if (c->qsize == sizeof(double))
      hipLaunchKernelGGL(k_remesh<double>, grid, bl, 0, n->stream, dtask, dkids, nblk, c->U0,
                         c->C, static_cast<const double*>(c->Q[c->qcur]), c->nbr, c->rc, n->U0,
                         n->C, c->nv, ien, n->st);
    else
      hipLaunchKernelGGL(k_remesh<float>, grid, bl, 0, n->stream, dtask, dkids, nblk, c->U0,
                         c->C, static_cast<const float*>(c->Q[c->qcur]), c->nbr, c->rc, n->U0,
                         n->C, c->nv, ien, n->st);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  // time, step and dt of the run; cell sizes and dxmin are the new mesh's
  if (e == hipSuccess) e = hipMemcpy(&R, c->rs, sizeof(RunState), hipMemcpyDeviceToHost);
  if (e == hipSuccess) {
    RunState Rn;
    e = hipMemcpy(&Rn, n->rs, sizeof(RunState), hipMemcpyDeviceToHost);
    Rn.dt = R.dt;
    Rn.time = R.time;
    Rn.step = R.step;
    Rn.fix_dt = R.fix_dt;
    Rn.fix_cmax = R.fix_cmax;
    if (e == hipSuccess) e = hipMemcpy(n->rs, &Rn, sizeof(RunState), hipMemcpyHostToDevice);
  }
  (void)hipFree(dtask);
  (void)hipFree(dkids);
  if (e != hipSuccess) {
    amun_dev_finalize(n);
    return fail(1000 + (int)e, std::string("mesh update: ") + hipGetErrorString(e));
  }
  for (int d = 0; d < 3; d++) {
    n->wg[d] = c->wg[d];
    n->swz[d] = c->swz[d];
  }
  n->use_graph = c->use_graph;
  // the new context takes the caller's handle and stream
  std::swap(*c, *n);
  std::swap(c->stream, n->stream);
  amun_dev_finalize(n);
  return after_upload(c, 0, 1, false);
}

int amun_dev_tridiag_events(amun_ctx* c, uint64_t* n) {
  if (!c || !n) return fail(500, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  DevState h;
  HIP_TRY(hipMemcpy(&h, c->st, sizeof(DevState), hipMemcpyDeviceToHost));
  *n = h.tridiag;
  return 0;
}

int amun_dev_c2p_events(amun_ctx* c, uint64_t n[3]) {
  if (!c || !n) return fail(500, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  DevState h;
  HIP_TRY(hipMemcpy(&h, c->st, sizeof(DevState), hipMemcpyDeviceToHost));
  n[0] = h.c2p[0];
  n[1] = h.c2p[1];
  n[2] = h.c2p[2];
  return 0;
}

void* amun_dev_stream(amun_ctx* c) { return c ? (void*)c->stream : nullptr; }

// ---------------------------------------------------------------------------
// gradients: forward tape, tangent-linear and adjoint runs
// ---------------------------------------------------------------------------
static int tape_check(amun_ctx* c) {
  if (c->rc.nruns != 1) return fail(600, "gradients need a single-run context");
  if (c->amr.on) return fail(600, "gradients are not supported on multi-level meshes");
  if (c->bnd.on) return fail(600, "gradients are built for periodic domains");
  if (!c->disp.has_adj)
    return fail(600, "gradients need TVD reconstruction and the HD or MHD system");
  if (c->src) return fail(600, "gradients do not include the GLM source terms");
  if (c->vsrc) return fail(600, "gradients do not include the viscous and resistive sources");
  if (c->cfg.mlp_limiting)
    return fail(600, "gradients do not include the multi-dimensional limiting");
  return 0;
}

int amun_dev_tape_forward(amun_ctx* c, int32_t nsteps, int32_t interval, double dtnext) {
  if (!c || nsteps < 1 || interval < 1) return fail(500, "bad argument");
  int r = tape_check(c);
  if (r) return r;
  auto& T = c->tape;
  hipFree(T.ck);
  hipFree(T.ckc);
  hipFree(T.dtc);
  T.ck = nullptr;
  T.ckc = nullptr;
  T.dtc = nullptr;
  const int nck = (nsteps + interval - 1) / interval;
  HIP_TRY(hipMalloc((void**)&T.ck, (size_t)nck * c->nstate * sizeof(double)));
  if (c->C) HIP_TRY(hipMalloc((void**)&T.ckc, (size_t)nck * c->nstate * sizeof(float)));
  HIP_TRY(hipMalloc((void**)&T.dtc, (size_t)2 * nsteps * sizeof(double)));
  T.nsteps = nsteps;
  T.interval = interval;
  T.dtnext = dtnext;
  for (int s = 0; s < nsteps; s++) {
    if (s % interval == 0) {
      HIP_TRY(hipMemcpyAsync(T.ck + (size_t)(s / interval) * c->nstate, c->U0,
                             c->nstate * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      if (c->C)
        HIP_TRY(hipMemcpyAsync(T.ckc + (size_t)(s / interval) * c->nstate, c->C,
                               c->nstate * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    }
    hipLaunchKernelGGL(k_tape, dim3(1), dim3(1), 0, c->stream, c->rs, c->rc,
                       (int)(c->nsteps & 1), dtnext, T.dtc + 2 * s);
    r = launch_step(c, dtnext);
    if (r) return r;
  }
  std::vector<double> h((size_t)2 * nsteps);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(h.data(), T.dtc, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  T.dt.resize(nsteps);
  T.cmax.resize(nsteps);
  for (int s = 0; s < nsteps; s++) {
    T.dt[s] = h[2 * s];
    T.cmax[s] = h[2 * s + 1];
  }
  return 0;
}

int amun_dev_tape_steps(amun_ctx* c, double* dt, double* cmax) {
  if (!c) return fail(500, "null argument");
  for (int s = 0; s < c->tape.nsteps; s++) {
    if (dt) dt[s] = c->tape.dt[s];
    if (cmax) cmax[s] = c->tape.cmax[s];
  }
  return 0;
}

int amun_dev_run_fixed(amun_ctx* c, int32_t nsteps, const double* dt, const double* cmax) {
  if (!c || nsteps < 0 || !dt || !cmax) return fail(500, "bad argument");
  for (int s = 0; s < nsteps; s++) {
    if (!(dt[s] > 0.0) || !(cmax[s] > 0.0)) return fail(500, "dt and cmax must be positive");
    hipLaunchKernelGGL(k_set_fixed, dim3(1), dim3(1), 0, c->stream, c->rs, dt[s], cmax[s]);
    int r = launch_step(c, 2.0e30);
    if (r) return r;
  }
  hipLaunchKernelGGL(k_set_fixed, dim3(1), dim3(1), 0, c->stream, c->rs, 0.0, 0.0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

namespace {

struct GradBufs {
  std::vector<double*> all;
  double *seg = nullptr, *Q = nullptr, *MU = nullptr, *LT = nullptr, *GC = nullptr;
  double *LY = nullptr, *LU0 = nullptr, *LX = nullptr, *GP = nullptr;
  double *LS = nullptr;          // tangent / adjoint of a later stage's X_src
  double *AL[3] = {nullptr, nullptr, nullptr}, *AR[3] = {nullptr, nullptr, nullptr};
  double *GF[3] = {nullptr, nullptr, nullptr};
  ~GradBufs() { for (double* p : all) (void)hipFree(p); }
  hipError_t get(double** p, size_t n) {
    hipError_t e = hipMalloc((void**)p, n * sizeof(double) + 64);
    if (e == hipSuccess) all.push_back(*p);
    return e;
  }
};

// Linear form of stage k: Y = a X_src + b X_k + c dt L(X_k), with X_0 = U^n
// the step state and X_{k+1} = Y_k; src is 0 except for the last stage of
// SSPRK(5,3), whose combine takes the output of its third stage (X_3).
struct StageCoef { double a, b, c; int src; };

// the stage whose input a later stage combines with (0 if none)
static int stage_src(const std::vector<StageCoef>& co) {
  int src = 0;
  for (const auto& x : co) src = std::max(src, x.src);
  return src;
}

std::vector<StageCoef> stage_coefs(const amun_ctx* c) {
  const auto t = stage_table(c);
  std::vector<StageCoef> co;
  for (size_t k = 0; k < t.size(); k++) {
    const StageDef& s = t[k];
    switch (s.kind) {
    case 0: case 4: co.push_back({0.0, 1.0, s.fac, 0}); break;
    case 1: co.push_back({0.75, 0.25, s.fac, 0}); break;
    case 2: co.push_back({1.0 / 3.0, 2.0 / 3.0, s.fac, 0}); break;
    case 3: co.push_back({0.5, 0.5, 0.5 * s.fac, 0}); break;
    case 5: co.push_back({s.ca, s.cb, s.cb * s.fac, 0}); break;
    case 6: co.push_back({s.ca, s.cb, s.fac, 0}); break;
    default: co.push_back({s.ca, s.cb, s.fac, (int)k - 1}); break;
    }
  }
  return co;
}

// recompute steps [s0, s1) from the checkpoint of s0, storing the stage
// inputs of every step into seg ([step - s0][stage][state])
int recompute_segment(amun_ctx* c, int s0, int s1, double* seg) {
  auto& T = c->tape;
  const int ns = stages_per_step(c);
  HIP_TRY(hipMemcpyAsync(c->U0, T.ck + (size_t)(s0 / T.interval) * c->nstate,
                         c->nstate * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  int r = after_upload(c, 0, 1);
  if (r) return r;
  if (c->C)   // the step state of the mixed mode is U0 - C
    HIP_TRY(hipMemcpyAsync(c->C, T.ckc + (size_t)(s0 / T.interval) * c->nstate,
                           c->nstate * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  for (int s = s0; s < s1; s++) {
    hipLaunchKernelGGL(k_set_fixed, dim3(1), dim3(1), 0, c->stream, c->rs, T.dt[s], T.cmax[s]);
    r = enqueue_step(c, 2.0e30, (int)(c->nsteps & 1), c->qcur, nullptr,
                     seg + (size_t)(s - s0) * ns * c->nstate);
    if (r) return r;
    if (ns & 1) c->qcur ^= 1;
    c->nsteps++;
  }
  hipLaunchKernelGGL(k_set_fixed, dim3(1), dim3(1), 0, c->stream, c->rs, 0.0, 0.0);
  HIP_TRY(hipGetLastError());
  return 0;
}

// global array of the single run <-> pool layout (device buffer dst/src)
int global_to_pool(amun_ctx* c, const double* g, double* dpool) {
  std::vector<double> h(c->nstate);
  scatter_run(c, 0, g, h.data());
  HIP_TRY(hipMemcpy(dpool, h.data(), c->nstate * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int pool_to_global(amun_ctx* c, const double* dpool, double* g) {
  std::vector<double> h(c->nstate);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(h.data(), dpool, c->nstate * sizeof(double), hipMemcpyDeviceToHost));
  gather_run(c, 0, h.data(), g, 8);
  return 0;
}

}  // namespace

// Save and restore the live state around a gradient computation, which
// recomputes from the checkpoints inside the context: conserved state, run
// state (time, step, dt) and the host step counters.
struct GradKeep {
  double* u = nullptr;
  float* cp = nullptr;        // Kahan compensation (mixed mode)
  RunState rs;
  long long nsteps = 0;
  int qcur = 0;
};

static int grad_state(amun_ctx* c, GradKeep& k, bool back) {
  if (!back) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMalloc((void**)&k.u, c->nstate * sizeof(double)));
    HIP_TRY(hipMemcpy(k.u, c->U0, c->nstate * sizeof(double), hipMemcpyDeviceToDevice));
    if (c->C) {
      HIP_TRY(hipMalloc((void**)&k.cp, c->nstate * sizeof(float)));
      HIP_TRY(hipMemcpy(k.cp, c->C, c->nstate * sizeof(float), hipMemcpyDeviceToDevice));
    }
    HIP_TRY(hipMemcpy(&k.rs, c->rs, sizeof(RunState), hipMemcpyDeviceToHost));
    k.nsteps = c->nsteps;
    k.qcur = c->qcur;
    return 0;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(c->U0, k.u, c->nstate * sizeof(double), hipMemcpyDeviceToDevice));
  HIP_TRY(hipMemcpy(c->rs, &k.rs, sizeof(RunState), hipMemcpyHostToDevice));
  HIP_TRY(hipFree(k.u));
  k.u = nullptr;
  c->nsteps = k.nsteps;
  c->qcur = k.qcur;
  int r = after_upload(c, 0, 1);
  if (!r && k.cp) {
    HIP_TRY(hipMemcpy(c->C, k.cp, c->nstate * sizeof(float), hipMemcpyDeviceToDevice));
    HIP_TRY(hipFree(k.cp));
    k.cp = nullptr;
  }
  return r;
}

int amun_dev_tape_tangent(amun_ctx* c, const double* du0, double dgamma, double* du) {
  if (!c || !du0 || !du) return fail(500, "null argument");
  int r = tape_check(c);
  if (r) return r;
  auto& T = c->tape;
  if (T.nsteps < 1) return fail(500, "no forward tape: call amun_dev_tape_forward first");
  Dispatch& d = c->disp;
  const RunConst& rc = c->rc;
  const int ns = stages_per_step(c), nd = rc.nd;
  const auto co = stage_coefs(c);
  const size_t n = c->nstate;
  GradBufs B;
  HIP_TRY(B.get(&B.seg, (size_t)T.interval * ns * n));
  HIP_TRY(B.get(&B.Q, n));
  HIP_TRY(B.get(&B.LX, n));      // tangent of the stage input
  HIP_TRY(B.get(&B.LU0, n));     // tangent of U^n
  HIP_TRY(B.get(&B.LT, n));      // tangent of L
  HIP_TRY(B.get(&B.LS, n));
  for (int k = 0; k < nd; k++) HIP_TRY(B.get(&B.AL[k], n));   // face flux tangents
  const int ksrc = stage_src(co);
  GradKeep keep;
  if ((r = grad_state(c, keep, false))) return r;
  const long long nn = (long long)n;
  const dim3 gc((unsigned)(((long long)rc.nblk * rc.ncell + 255) / 256)), bl(256);
  if ((r = global_to_pool(c, du0, B.LU0))) return r;
  RunState h;
  HIP_TRY(hipMemcpy(&h, c->rs, sizeof(RunState), hipMemcpyDeviceToHost));
  for (int s0 = 0; s0 < T.nsteps && !r; s0 += T.interval) {
    const int s1 = std::min(T.nsteps, s0 + T.interval);
    if ((r = recompute_segment(c, s0, s1, B.seg))) break;
    for (int s = s0; s < s1; s++) {
      const double* X0 = B.seg + (size_t)(s - s0) * ns * n;
      // tangent of the stage input: dU^n for the first stage
      hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, 1.0, B.LU0, 0.0, B.LX);
      for (int k = 0; k < ns; k++) {
        const double* X = X0 + (size_t)k * n;
        if (ksrc > 0 && k == ksrc)
          hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, 1.0, B.LX, 0.0, B.LS);
        hipLaunchKernelGGL(d.adj_prim, gc, bl, 0, c->stream, X, B.Q, rc);
        for (int dd = 0; dd < nd; dd++)
          hipLaunchKernelGGL(d.tlm_face[dd], gc, bl, 0, c->stream, X, B.Q, B.LX, dgamma,
                             B.AL[dd], c->nbr, rc, T.cmax[s]);
        hipLaunchKernelGGL(d.tlm_cell, gc, bl, 0, c->stream, B.AL[0], B.AL[1],
                           nd == 3 ? B.AL[2] : B.AL[1], B.LT, c->nbr, rc, h.dxi[0], h.dxi[1],
                           h.dxi[2]);
        // dY = a dX_src + b dX + c dt dL
        hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, co[k].b, B.LX, 0.0, B.LX);
        if (co[k].a != 0.0)
          hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, co[k].a,
                             co[k].src ? B.LS : B.LU0, 1.0, B.LX);
        hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, co[k].c * T.dt[s], B.LT,
                           1.0, B.LX);
      }
      if (c->cfg.eqsys == AMUN_EQSYS_MHD)
        hipLaunchKernelGGL(k_scale_var, dim3(2048), bl, 0, c->stream, B.LX, c->nv, 7, rc.decay, rc);
      hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, 1.0, B.LX, 0.0, B.LU0);
    }
    HIP_TRY(hipGetLastError());
  }
  if (!r) r = pool_to_global(c, B.LU0, du);
  int r2 = grad_state(c, keep, true);
  return r ? r : r2;
}

int amun_dev_tape_backward(amun_ctx* c, const double* lam, double* lam0, double* dgamma) {
  if (!c || !lam || !lam0) return fail(500, "null argument");
  int r = tape_check(c);
  if (r) return r;
  auto& T = c->tape;
  if (T.nsteps < 1) return fail(500, "no forward tape: call amun_dev_tape_forward first");
  Dispatch& d = c->disp;
  const RunConst& rc = c->rc;
  const int ns = stages_per_step(c), nd = rc.nd;
  const auto co = stage_coefs(c);
  const size_t n = c->nstate, ncell = (size_t)rc.nblk * rc.ncell;
  const int npart = 1024;
  GradBufs B;
  HIP_TRY(B.get(&B.seg, (size_t)T.interval * ns * n));
  HIP_TRY(B.get(&B.Q, n));
  HIP_TRY(B.get(&B.MU, n));
  HIP_TRY(B.get(&B.LT, n));
  HIP_TRY(B.get(&B.LY, n));
  HIP_TRY(B.get(&B.LU0, n));
  HIP_TRY(B.get(&B.LS, n));
  HIP_TRY(B.get(&B.GC, ncell));
  const int ksrc = stage_src(co);
  HIP_TRY(B.get(&B.GP, npart));
  for (int k = 0; k < nd; k++) {
    HIP_TRY(B.get(&B.AL[k], n));
    HIP_TRY(B.get(&B.AR[k], n));
    HIP_TRY(B.get(&B.GF[k], ncell));
  }
  HIP_TRY(hipMemset(B.GP, 0, npart * sizeof(double)));
  GradKeep keep;
  if ((r = grad_state(c, keep, false))) return r;
  const long long nn = (long long)n;
  const dim3 gc((unsigned)(((long long)ncell + 255) / 256)), bl(256);
  if ((r = global_to_pool(c, lam, B.LU0))) return r;     // lambda of U^N
  RunState h;
  HIP_TRY(hipMemcpy(&h, c->rs, sizeof(RunState), hipMemcpyDeviceToHost));
  const int nseg = (T.nsteps + T.interval - 1) / T.interval;
  for (int g = nseg - 1; g >= 0 && !r; g--) {
    const int s0 = g * T.interval, s1 = std::min(T.nsteps, s0 + T.interval);
    if ((r = recompute_segment(c, s0, s1, B.seg))) break;
    for (int s = s1 - 1; s >= s0; s--) {
      const double* X0 = B.seg + (size_t)(s - s0) * ns * n;
      // lambda of the last stage output: psi decay of the step
      hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, 1.0, B.LU0, 0.0, B.LY);
      if (c->cfg.eqsys == AMUN_EQSYS_MHD)
        hipLaunchKernelGGL(k_scale_var, dim3(2048), bl, 0, c->stream, B.LY, c->nv, 7, rc.decay, rc);
      hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, 0.0, B.LY, 0.0, B.LU0);
      if (ksrc > 0)
        hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, 0.0, B.LY, 0.0, B.LS);
      for (int k = ns - 1; k >= 0; k--) {
        const double* X = X0 + (size_t)k * n;
        // MU = c dt lambda_Y, LT = (dL/dX)^T MU
        hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, co[k].c * T.dt[s], B.LY,
                           0.0, B.MU);
        hipLaunchKernelGGL(d.adj_prim, gc, bl, 0, c->stream, X, B.Q, rc);
        for (int dd = 0; dd < nd; dd++)
          hipLaunchKernelGGL(d.adj_face[dd], gc, bl, 0, c->stream, B.Q, B.MU, B.AL[dd], B.AR[dd],
                             B.GF[dd], c->nbr, rc, h.dxi[dd], T.cmax[s]);
        const int z = nd == 3 ? 2 : 1;
        hipLaunchKernelGGL(d.adj_cell, gc, bl, 0, c->stream, X, B.Q, B.AL[0], B.AR[0], B.AL[1],
                           B.AR[1], B.AL[z], B.AR[z], B.GF[0], B.GF[1], B.GF[z], B.LT, B.GC,
                           c->nbr, rc);
        hipLaunchKernelGGL(k_sum_acc, dim3(npart), bl, 0, c->stream, B.GC, (long long)ncell, B.GP);
        // lambda of X_src gets a lambda_Y; lambda of X = b lambda_Y + LT
        // (+ what a later stage passed to it as its X_src)
        if (co[k].a != 0.0)
          hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, co[k].a, B.LY, 1.0,
                             co[k].src ? B.LS : B.LU0);
        hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, co[k].b, B.LY, 0.0, B.MU);
        hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, 1.0, B.LT, 1.0, B.MU);
        if (ksrc > 0 && k == ksrc)
          hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, 1.0, B.LS, 1.0, B.MU);
        if (k == 0)
          hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, 1.0, B.MU, 1.0, B.LU0);
        else
          hipLaunchKernelGGL(k_axpby, dim3(2048), bl, 0, c->stream, nn, 1.0, B.MU, 0.0, B.LY);
      }
    }
    HIP_TRY(hipGetLastError());
  }
  if (!r) r = pool_to_global(c, B.LU0, lam0);
  if (!r && dgamma) {
    std::vector<double> gp(npart);
    HIP_TRY(hipMemcpy(gp.data(), B.GP, npart * sizeof(double), hipMemcpyDeviceToHost));
    double acc = 0.0;
    for (int w = 0; w < npart; w++) acc += gp[w];
    *dgamma = acc;
  }
  int r2 = grad_state(c, keep, true);
  return r ? r : r2;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the allocation-free cons2prim: its event counters live in a static device
// array instead of a hipMalloc'd DevState
// ---------------------------------------------------------------------------
namespace {
constexpr int kScratchWords = 32;   // 256 bytes: one DevState and room to spare
static_assert(sizeof(DevState) <= kScratchWords * sizeof(unsigned long long),
              "DevState must fit the static scratch array");
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

// device address of scratch on every device the library has run on (the
// array's host-side name is not a device address)
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(500, "device index beyond the scratch table");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(1000, "hipGetSymbolAddress returned a null address for scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

__global__ void k_scratch_clear(unsigned long long* p, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = 0ull;
}

__global__ void k_scratch_events(const DevState* st, int64_t* out) {
  const int i = (int)threadIdx.x;
  if (i < 3) out[i] = (int64_t)st->c2p[i];
}
}  // namespace

extern "C" int amun_dev_cons2prim_noalloc(const amun_cfg* cfg, int32_t n, const double* d_u,
                                          double* d_q, int64_t* d_events, void* stream) {
  if (!cfg) return fail(500, "null argument");
  int nv = 0;
  int r = validate(*cfg, &nv);
  if (r) return r;
  if (n < 1 || !d_u || !d_q) return fail(500, "bad argument: n >= 1 and device arrays d_u, d_q");
  Dispatch d;
  if (!pick(d, *cfg)) return fail(600, "no kernel instantiation for this configuration");
  const RunConst rc = make_rc(*cfg);
  void* sp = nullptr;
  r = scratch_of_current_device(&sp);
  if (r) return r;
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned long long* words = static_cast<unsigned long long*>(sp);
  hipLaunchKernelGGL(k_scratch_clear, dim3(1), dim3(kScratchWords), 0, s, words, kScratchWords);
  hipLaunchKernelGGL(d.unit_c2p, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, n, d_u, d_q,
                     rc, static_cast<DevState*>(sp));
  if (d_events)
    hipLaunchKernelGGL(k_scratch_events, dim3(1), dim3(64), 0, s,
                       static_cast<const DevState*>(sp), d_events);
  HIP_TRY(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// unit level (tests): device routines on host vectors
// ---------------------------------------------------------------------------
namespace {
struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() { for (void* x : p) (void)hipFree(x); }
  template <typename X>
  hipError_t alloc(X** out, size_t n) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, n * sizeof(X) + 8);
    if (e == hipSuccess) { p.push_back(v); *out = static_cast<X*>(v); }
    return e;
  }
};

int unit_setup(const amun_cfg* cfg, Dispatch& d, RunConst& rc) {
  if (!cfg) return fail(500, "null argument");
  int nv = 0;
  int r = validate(*cfg, &nv);
  if (r) return r;
  if (!pick(d, *cfg)) return fail(600, "no kernel instantiation for this configuration");
  rc = make_rc(*cfg);
  return 0;
}
}  // namespace

extern "C" {

int amun_dev_unit_riemann(const amun_cfg* cfg, int32_t n, double cmax, const double* ql,
                          const double* qr, double* f) {
  Dispatch d;
  RunConst rc;
  int r = unit_setup(cfg, d, rc);
  if (r) return r;
  if (n < 1 || !ql || !qr || !f) return fail(500, "bad argument");
  const int nv = cfg_nv(*cfg);
  const size_t m = (size_t)n * nv;
  DevBufs B;
  double *dl, *dr, *df;
  HIP_TRY(B.alloc(&dl, m));
  HIP_TRY(B.alloc(&dr, m));
  HIP_TRY(B.alloc(&df, m));
  HIP_TRY(hipMemcpy(dl, ql, m * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dr, qr, m * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(d.unit_riemann, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, 0, n, cmax,
                     dl, dr, df, rc);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(f, df, m * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int amun_dev_unit_cons2prim(const amun_cfg* cfg, int32_t n, const double* u, double* q,
                            int64_t* events) {
  Dispatch d;
  RunConst rc;
  int r = unit_setup(cfg, d, rc);
  if (r) return r;
  if (n < 1 || !u || !q) return fail(500, "bad argument");
  const int nv = cfg_nv(*cfg);
  const size_t m = (size_t)n * nv;
  DevBufs B;
  double *du, *dq;
  DevState* st;
  HIP_TRY(B.alloc(&du, m));
  HIP_TRY(B.alloc(&dq, m));
  HIP_TRY(B.alloc(&st, 1));
  HIP_TRY(hipMemset(st, 0, sizeof(DevState)));
  HIP_TRY(hipMemcpy(du, u, m * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dq, q, m * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(d.unit_c2p, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, 0, n, du, dq,
                     rc, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(q, dq, m * sizeof(double), hipMemcpyDeviceToHost));
  if (events) {
    DevState h;
    HIP_TRY(hipMemcpy(&h, st, sizeof(DevState), hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; k++) events[k] = (int64_t)h.c2p[k];
  }
  return 0;
}

int amun_dev_unit_reconstruct(const amun_cfg* cfg, int32_t n, int32_t npencil, double h,
                              const double* fp, double* fl, double* fr) {
  Dispatch d;
  RunConst rc;
  int r = unit_setup(cfg, d, rc);
  if (r) return r;
  if (!d.unit_recon)
    return fail(600, "TVD (slopes coupled across directions) and the compact schemes (whole "
                     "block pencils) have no pencil form here");
  if (n < 8 || n > 4096 || npencil < 1 || !fp || !fl || !fr || !(h > 0.0))
    return fail(500, "bad argument");
  const size_t m = (size_t)n * npencil;
  DevBufs B;
  double *df, *dl, *dr;
  HIP_TRY(B.alloc(&df, m));
  HIP_TRY(B.alloc(&dl, m));
  HIP_TRY(B.alloc(&dr, m));
  HIP_TRY(hipMemcpy(df, fp, m * sizeof(double), hipMemcpyHostToDevice));
  // entries outside the window range keep what the caller passed in
  HIP_TRY(hipMemcpy(dl, fl, m * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dr, fr, m * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(d.unit_recon, dim3((unsigned)((npencil + 63) / 64)), dim3(64), 0, 0, n,
                     npencil, h, df, dl, dr, rc);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(fl, dl, m * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(fr, dr, m * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int amun_dev_unit_eqs(const amun_cfg* cfg, int32_t n, double cmax, const double* q, double* u,
                      double* q2, double* f, double* cm, double* cp) {
  Dispatch d;
  RunConst rc;
  int r = unit_setup(cfg, d, rc);
  if (r) return r;
  if (n < 1 || !q || !u || !q2 || !f || !cm || !cp) return fail(500, "bad argument");
  const int nv = cfg_nv(*cfg);
  const size_t m = (size_t)n * nv;
  DevBufs B;
  double *dq, *du, *dq2, *df, *dm, *dp;
  DevState* st;
  HIP_TRY(B.alloc(&dq, m));
  HIP_TRY(B.alloc(&du, m));
  HIP_TRY(B.alloc(&dq2, m));
  HIP_TRY(B.alloc(&df, m));
  HIP_TRY(B.alloc(&dm, (size_t)n));
  HIP_TRY(B.alloc(&dp, (size_t)n));
  HIP_TRY(B.alloc(&st, 1));
  HIP_TRY(hipMemset(st, 0, sizeof(DevState)));
  HIP_TRY(hipMemcpy(dq, q, m * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dq2, q2, m * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(d.unit_eqs, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, 0, n, cmax,
                     dq, du, dq2, df, dm, dp, rc, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(u, du, m * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(q2, dq2, m * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(f, df, m * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(cm, dm, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(cp, dp, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
