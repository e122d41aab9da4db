/*
 * amun_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded CPU restatement of the AMUN per-block
 * finite-volume update, written for this project to serve as the checker of
 * the HIP kernels.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py load it (liboracle.so, through ctypes); the product path
 * never does.
 *
 * It follows the reference structure and operation order closely, so that on
 * the x86-64 baseline ISA (no FMA, -ffp-contract=off) it reproduces the
 * reference binaries built by oracle/build_ref.sh to round-off:
 *   - block arrays u0, u1, q (nv,im,jm,km) and f (NDIMS,nv,im,jm,km) in
 *     Fortran order with ghost layers (blocks.F90:1120-1191);
 *   - update_flux_hd_adi / update_flux_mhd_adi (schemes.F90:1080-1223,
 *     2970-3138) over full pencils with the component rotation;
 *   - interfaces_tvd (interpolations.F90:729-848), interfaces_dir (867-959)
 *     with reconstruct_weno5z (1937-2083), limiters (4684-4877);
 *   - riemann_hd_adi_hll (schemes.F90:1248), _hllc (1366-1513),
 *     riemann_mhd_adi_hll (3163), _hlld (3596-4203);
 *   - prim2cons / cons2prim / fluxspeed / maxspeed (equations.F90:1762-2044,
 *     2815-3131), update_primitive_variables (1088-1148);
 *   - the special-relativistic MHD system: prim2cons_srmhd_adi,
 *     cons2prim_srmhd_adi with the 1D(W) Newton solver and its brackets,
 *     fluxspeed_srmhd_adi, maxspeed_srmhd_adi (equations.F90:4590-5163,
 *     5249-5788), riemann_srmhd_adi_hll (schemes.F90:5244-5355) and the root
 *     finders quadratic, quadratic_normalized, cubic, cubic_normalized,
 *     quartic, sort (algebra.F90:86-1085);
 *   - update_increment (evolution.F90:1898-1977), evolve_euler/rk2/rk3
 *     (517, 628-770, 1005-1208), new_time_step (379-493);
 *   - same-level ghost exchange of q (boundaries.F90:372-479 with the face,
 *     edge and corner copies 1273, 2447, 3867) on the periodic base-level
 *     block grid of setup_domain_default (domains.F90:136-500), followed by
 *     update_ghost_cells (boundaries.F90:6633);
 *   - on multi-level meshes: restriction (block_face/edge/corner_restrict,
 *     boundaries.F90:5637, 5919, 6256) and limited prolongation
 *     (block_face/edge/corner_prolong, 5762, 6071, 6351) of ghost cells, the
 *     coarse-face flux correction (boundary_fluxes 494-1060 with
 *     block_update_flux 6539-6614), per-level cell sizes
 *     (coordinates.F90:330-356) and the finest-level CFL step.
 *
 * Pins (see tests/): golden vectors written by the reference binaries
 * (tests/golden/), round trips, consistency F(u,u)=F(u), exact discrete
 * conservation, uniform-state preservation, polynomial exactness of WENO5-Z,
 * polynomial residuals and numpy.roots for the root finders, the
 * prim -> cons -> prim round trip of the relativistic primitive solver.
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/amun_dev.h"

/* ------------------------------------------------------------------------ */
/* configuration derived from amun_cfg, mirroring the module variables       */
/* ------------------------------------------------------------------------ */
typedef struct {
  amun_cfg c;
  int nd, nv, nc, ng;
  int im, jm, km;          /* block sizes incl. ghosts                       */
  int ib, ie, jb, je, kb, ke, ibl, ieu, jbl, jeu, kbl, keu;  /* 0-based      */
  int idn, ivx, ivy, ivz, imx, imy, imz, ibx, iby, ibz, ibp, ipr, ien;
  int    mhd, sr;          /* magnetic field present / relativistic system   */
  int    iso;              /* isothermal equation of state (no energy)       */
  int    srhd;             /* special-relativistic HD (o->sr is set too)      */
  double csnd, csnd2;      /* isothermal sound speed and its square          */
  double gammam1, gammam1i, gammaxi;
  double cflx;             /* cfl as adjusted by initialize_evolution        */
  int    stages;           /* m of SSPRK(m,2)                                */
  double decay;            /* exp(-alpha*cfl), evolution.F90:211-213         */
  double adx, ady, adz, adxi, adyi, adzi;
  double h[3];             /* cell sizes of the block being reconstructed   */
  double cmax, cmax2;      /* equations module state                         */
  long   npos_events;      /* "Positive variable is not positive" prints     */
  int    amr_err;          /* multi-level mesh the ghost fill cannot serve   */
  long   c2p_fail, c2p_pneg, c2p_noconv;   /* SR primitive solver messages */
  double kappa;            /* min(kappa, (1 - cfl)/cfl), interpolations.F90:217 */
  int    mgp;              /* (ngp - 1) / 2                                  */
  double cgp[16];          /* Gaussian-process weights                      */
  int    dgp;              /* ngp^nd, the MGP stencil                        */
  double ugp[750];         /* MGP weights [dir][side][stencil], ngp <= 5     */
  /* cells where interfaces_mgp fell back to the TVD slopes, and a hash of
   * their positions in visiting order (which fallbacks happened, not only how
   * many): a record for tests, no effect on the update */
  unsigned long long mgp_fallbacks, mgp_hash;
} ocfg;

static int prepare_gp(ocfg* o);
static int prepare_mgp(ocfg* o);

/* MIN and MAX as the reference's compiler evaluates them: a < b ? a : b and
 * a > b ? a : b, so a NaN first operand yields the second and a NaN second
 * operand propagates (wave speeds of a negative reconstructed pressure) */
static double fmin_r(double a, double b) { return a < b ? a : b; }
static double fmax_r(double a, double b) { return a > b ? a : b; }


static int oc_setup(ocfg* o, const amun_cfg* c)
{
  memset(o, 0, sizeof(*o));
  o->c  = *c;
  o->nd = c->ndims;
  o->nc = c->ncells;
  o->ng = c->nghosts;
  if (o->nd != 2 && o->nd != 3) return 500;
  if (c->eos != AMUN_EOS_ADI && c->eos != AMUN_EOS_ISO) return 110;
  if (c->eos == AMUN_EOS_ISO) {
    /* equations.F90:399-830: the isothermal systems carry no energy; the
     * sound speed csnd is a parameter (equations.F90:176, 847-851) */
    if (c->eqsys != AMUN_EQSYS_HD && c->eqsys != AMUN_EQSYS_MHD) return 110;
    o->iso = 1;
    o->csnd = c->csnd;
    o->csnd2 = c->csnd * c->csnd;
    o->mhd = (c->eqsys == AMUN_EQSYS_MHD);
    o->nv = o->mhd ? 8 : 4;
    o->idn = 0; o->ivx = 1; o->ivy = 2; o->ivz = 3;
    o->imx = 1; o->imy = 2; o->imz = 3;
    if (o->mhd) { o->ibx = 4; o->iby = 5; o->ibz = 6; o->ibp = 7; }
    else o->ibx = o->iby = o->ibz = o->ibp = -1;
    o->ipr = o->ien = -1;
  } else if (c->eqsys == AMUN_EQSYS_HD) {
    o->nv = 5;
    o->idn = 0; o->ivx = 1; o->ivy = 2; o->ivz = 3;
    o->imx = 1; o->imy = 2; o->imz = 3;
    o->ibx = o->iby = o->ibz = o->ibp = -1;
    o->ipr = o->ien = 4;
  } else if (c->eqsys == AMUN_EQSYS_SRHD) {
    /* equations.F90:537-600: the special-relativistic HD system */
    o->sr = o->srhd = 1;
    o->nv = 5;
    o->idn = 0; o->ivx = 1; o->ivy = 2; o->ivz = 3;
    o->imx = 1; o->imy = 2; o->imz = 3;
    o->ibx = o->iby = o->ibz = o->ibp = -1;
    o->ipr = o->ien = 4;
  } else if (c->eqsys == AMUN_EQSYS_MHD || c->eqsys == AMUN_EQSYS_SRMHD) {
    o->mhd = 1;
    o->sr  = (c->eqsys == AMUN_EQSYS_SRMHD);
    o->nv = 9;
    o->idn = 0; o->ivx = 1; o->ivy = 2; o->ivz = 3;
    o->imx = 1; o->imy = 2; o->imz = 3;
    o->ibx = 4; o->iby = 5; o->ibz = 6; o->ibp = 7;
    o->ipr = o->ien = 8;
  } else {
    return 100;
  }
  o->im = o->nc + 2 * o->ng;
  o->jm = o->nc + 2 * o->ng;
  o->km = (o->nd == 3) ? o->nc + 2 * o->ng : 1;
  o->ib = o->ng; o->ie = o->ng + o->nc - 1;
  o->jb = o->ng; o->je = o->ng + o->nc - 1;
  o->ibl = o->ib - 1; o->ieu = o->ie + 1;
  o->jbl = o->jb - 1; o->jeu = o->je + 1;
  if (o->nd == 3) {
    o->kb = o->ng; o->ke = o->ng + o->nc - 1;
    o->kbl = o->kb - 1; o->keu = o->ke + 1;
  } else {
    o->kb = o->ke = o->kbl = o->keu = 0;
  }
  o->gammam1  = c->gamma - 1.0;
  o->gammam1i = 1.0 / o->gammam1;
  o->gammaxi  = o->gammam1 / c->gamma;       /* equations.F90:843 */
  o->kappa    = fmin_r(c->kappa, (1.0 - c->cfl) / c->cfl);   /* interpolations.F90:217 */
  if (c->reconstruction == AMUN_REC_GP) {
    int r = prepare_gp(o);
    if (r) return r;
  }
  if (c->reconstruction == AMUN_REC_MGP) {
    int r = prepare_mgp(o);
    if (r) return r;
    if (o->ng < o->mgp + 1) return 600;   /* the stencil of the cells next to the block */
  }
  /* initialize_evolution (evolution.F90:156-200): the m-stage and
   * four/five-stage SSPRK schemes run at a multiple of the CFL number */
  o->cflx  = c->cfl;
  o->stages = c->stages < 2 ? 2 : (c->stages > 9 ? 9 : c->stages);
  if (c->time_advance == AMUN_TA_SSPRK_M2) o->cflx = (o->stages - 1) * c->cfl;
  if (c->time_advance == AMUN_TA_SSPRK43) o->cflx = 2.0 * c->cfl;
  if (c->time_advance == AMUN_TA_SSPRK53) o->cflx = 2.65062919143939 * c->cfl;
  o->decay = exp(-c->alpha * o->cflx);
  {
    int ir = c->xblocks, jr = c->yblocks, kr = (o->nd == 3) ? c->zblocks : 1;
    double xlen = c->xmax - c->xmin, ylen = c->ymax - c->ymin;
    double zlen = c->zmax - c->zmin;
    o->adx = xlen / (ir * o->nc);
    o->ady = ylen / (jr * o->nc);
    o->adz = (o->nd == 3) ? zlen / (kr * o->nc) : 1.0;
    o->adxi = 1.0 / o->adx;
    o->adyi = 1.0 / o->ady;
    o->adzi = (o->nd == 3) ? 1.0 / o->adz : 1.0;
    o->h[0] = o->adx; o->h[1] = o->ady; o->h[2] = o->adz;
  }
  /* the LimO3 branch of initialize_interpolations raises eps
   * (interpolations.F90:248) */
  if (c->reconstruction == AMUN_REC_LIMO3) o->c.eps = fmax_r(1.0e-12, c->eps);
  return 0;
}

/* bound of the positivity halving loops of interfaces_tvd and the
 * prolongation: a finite slope falls below qc within 2098 halvings, the
 * bound only ends an infinite one (on which the reference loops forever) */
#define OMAXHALV 2200

#define QI(o, p, i, j, k) ((((size_t)(k) * (o)->jm + (j)) * (o)->im + (i)) * (o)->nv + (p))

/* ------------------------------------------------------------------------ */
/* limiters, interpolations.F90:4684-4877                                     */
/* ------------------------------------------------------------------------ */
static double lim(int kind, double x, double a, double b)
{
  double c;
  switch (kind) {
  case AMUN_LIM_MINMOD:     /* 4731 */
    return 0.5 * (copysign(x, a) + copysign(x, b)) * fmin_r(fabs(a), fabs(b));
  case AMUN_LIM_MC:         /* 4765 */
    return (copysign(x, a) + copysign(x, b))
         * fmin_r(fmin_r(fabs(a), fabs(b)), 0.25 * fabs(a + b));
  case AMUN_LIM_SUPERBEE:   /* 4799 */
    return 0.5 * (copysign(x, a) + copysign(x, b))
         * fmax_r(fmin_r(2.0 * fabs(a), fabs(b)), fmin_r(fabs(a), 2.0 * fabs(b)));
  case AMUN_LIM_VANLEER:    /* 4834 */
    c = a * b;
    if (c > 0.0) return 2.0 * x * c / (a + b);
    return 0.0;
  case AMUN_LIM_VANALBADA:  /* 4873: eps = epsilon(1d0) default */
    return x * a * b * (a + b) / fmax_r(2.220446049250313e-16, a * a + b * b);
  default:                  /* limiter_zero 4700 */
    return 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* reconstruct_weno5z, interpolations.F90:1937-2083, one pencil of length n   */
/* ------------------------------------------------------------------------ */
/* reconstruct_weno5z (interpolations.F90:1937-2083) and, with yc != 0,
 * reconstruct_weno5yc (2107-2254), whose tau is the squared fourth
 * difference instead of |beta_r - beta_l| */
static void weno5z_v(const ocfg* o, int n, const double* f, double* fl,
                     double* fr, double* dfm, double* dfp, double* df2, int yc)
{
  const double c1 = 13.0 / 12.0, c2 = 0.25;
  const double dl = 0.1, dc = 0.6, dr = 0.3;
  const double a11 = 2.0 / 6.0, a12 = -7.0 / 6.0, a13 = 11.0 / 6.0;
  const double a21 = -1.0 / 6.0, a22 = 5.0 / 6.0, a23 = 2.0 / 6.0;
  const double a31 = 2.0 / 6.0, a32 = 5.0 / 6.0, a33 = -1.0 / 6.0;
  const double eps = o->c.eps;
  int i;
  for (i = 0; i < n - 1; i++) {
    dfp[i]     = f[i + 1] - f[i];
    dfm[i + 1] = dfp[i];
  }
  dfm[0]     = dfp[0];
  dfp[n - 1] = dfm[n - 1];
  for (i = 0; i < n; i++) {
    double d = dfp[i] - dfm[i];
    df2[i] = c1 * (d * d);
  }
  for (i = 2; i < n - 2; i++) {
    double bl, bc, br, tt, al, ac, ar, wl, wc, wr, ww, ql, qc, qr, t;
    t  = 3.0 * dfm[i] - dfm[i - 1];
    bl = df2[i - 1] + c2 * (t * t);
    t  = dfp[i] + dfm[i];
    bc = df2[i] + c2 * (t * t);
    t  = 3.0 * dfp[i] - dfp[i + 1];
    br = df2[i + 1] + c2 * (t * t);
    if (yc) {
      t  = ((6.0 * f[i] - 4.0 * (f[i - 1] + f[i + 1])) + (f[i - 2] + f[i + 2]));
      tt = t * t;
    } else {
      tt = fabs(br - bl);
    }
    al = 1.0 + tt / (bl + eps);
    ac = 1.0 + tt / (bc + eps);
    ar = 1.0 + tt / (br + eps);

    wl = dl * al; wc = dc * ac; wr = dr * ar;
    ww = (wl + wr) + wc;
    wl = wl / ww; wr = wr / ww; wc = 1.0 - (wl + wr);
    ql = a11 * f[i - 2] + a12 * f[i - 1] + a13 * f[i];
    qc = a21 * f[i - 1] + a22 * f[i] + a23 * f[i + 1];
    qr = a31 * f[i] + a32 * f[i + 1] + a33 * f[i + 2];
    fl[i] = (wl * ql + wr * qr) + wc * qc;

    wl = dl * ar; wc = dc * ac; wr = dr * al;
    ww = (wl + wr) + wc;
    wl = wl / ww; wr = wr / ww; wc = 1.0 - (wl + wr);
    ql = a11 * f[i + 2] + a12 * f[i + 1] + a13 * f[i];
    qc = a21 * f[i + 1] + a22 * f[i] + a23 * f[i - 1];
    qr = a31 * f[i] + a32 * f[i - 1] + a33 * f[i - 2];
    fr[i - 1] = (wl * ql + wr * qr) + wc * qc;
  }
  {
    double df;
    fl[0] = 0.5 * (f[0] + f[1]);
    df    = lim(o->c.limiter, 0.5, dfm[1], dfp[1]);
    fr[0] = f[1] - df;
    fl[1] = f[1] + df;
    i     = n - 2;
    df    = lim(o->c.limiter, 0.5, dfm[i], dfp[i]);
    fr[i - 1] = f[i] - df;
    fl[i]     = f[i] + df;
    fr[i]     = 0.5 * (f[i] + f[n - 1]);
    fl[n - 1] = f[n - 1];
    fr[n - 1] = f[n - 1];
  }
}

static void weno5z(const ocfg* o, int n, const double* f, double* fl,
                   double* fr, double* dfm, double* dfp, double* df2)
{
  weno5z_v(o, n, f, fl, fr, dfm, dfp, df2, 0);
}

/* reconstruct_weno5ns, interpolations.F90:2278-2441 */
static void weno5ns(const ocfg* o, int n, const double* f, double* fl,
                    double* fr, double* dfm, double* dfp, double* df2)
{
  const double dl = 0.1, dc = 0.6, dr = 0.3, xi = 4.0e-01;
  const double a11 = 2.0 / 6.0, a12 = -7.0 / 6.0, a13 = 11.0 / 6.0;
  const double a21 = -1.0 / 6.0, a22 = 5.0 / 6.0, a23 = 2.0 / 6.0;
  const double a31 = 2.0 / 6.0, a32 = 5.0 / 6.0, a33 = -1.0 / 6.0;
  const double eps = o->c.eps;
  int i;
  for (i = 0; i < n - 1; i++) {
    dfp[i]     = f[i + 1] - f[i];
    dfm[i + 1] = dfp[i];
  }
  dfm[0]     = dfp[0];
  dfp[n - 1] = dfm[n - 1];
  for (i = 0; i < n; i++) df2[i] = 0.5 * fabs(dfp[i] - dfm[i]);
  for (i = 2; i < n - 2; i++) {
    double df, lq, bl, bc, br, l3, zt, al, ac, ar, aa, wl, wc, wr, ql, qc, qr, t, s;
    df = fabs(dfp[i]);
    lq = xi * df;
    bl = df2[i - 1] + xi * fabs(2.0 * dfm[i] - dfm[i - 1]);
    bc = df2[i] + lq;
    br = df2[i + 1] + lq;
    l3 = (df * df) * df;
    t  = bl - br;
    s  = l3 / (1.0 + l3);
    zt = 0.5 * (t * t + s * s);
    t = bl + eps; al = dl * (1.0 + zt / (t * t));
    t = bc + eps; ac = dc * (1.0 + zt / (t * t));
    t = br + eps; ar = dr * (1.0 + zt / (t * t));
    aa = (al + ar) + ac;
    wl = al / aa; wr = ar / aa; wc = 1.0 - (wl + wr);
    ql = a11 * f[i - 2] + a12 * f[i - 1] + a13 * f[i];
    qc = a21 * f[i - 1] + a22 * f[i] + a23 * f[i + 1];
    qr = a31 * f[i] + a32 * f[i + 1] + a33 * f[i + 2];
    fl[i] = (wl * ql + wr * qr) + wc * qc;

    df = fabs(dfm[i]);
    lq = xi * df;
    bl = df2[i + 1] + xi * fabs(2.0 * dfp[i] - dfp[i + 1]);
    bc = df2[i] + lq;
    br = df2[i - 1] + lq;
    l3 = (df * df) * df;
    t  = bl - br;
    s  = l3 / (1.0 + l3);
    zt = 0.5 * (t * t + s * s);
    t = bl + eps; al = dl * (1.0 + zt / (t * t));
    t = bc + eps; ac = dc * (1.0 + zt / (t * t));
    t = br + eps; ar = dr * (1.0 + zt / (t * t));
    aa = (al + ar) + ac;
    wl = al / aa; wr = ar / aa; wc = 1.0 - (wl + wr);
    ql = a11 * f[i + 2] + a12 * f[i + 1] + a13 * f[i];
    qc = a21 * f[i + 1] + a22 * f[i] + a23 * f[i - 1];
    qr = a31 * f[i] + a32 * f[i - 1] + a33 * f[i - 2];
    fr[i - 1] = (wl * ql + wr * qr) + wc * qc;
  }
  {
    double df;
    fl[0] = 0.5 * (f[0] + f[1]);
    df    = lim(o->c.limiter, 0.5, dfm[1], dfp[1]);
    fr[0] = f[1] - df;
    fl[1] = f[1] + df;
    i     = n - 2;
    df    = lim(o->c.limiter, 0.5, dfm[i], dfp[i]);
    fr[i - 1] = f[i] - df;
    fl[i]     = f[i] + df;
    fr[i]     = 0.5 * (f[i] + f[n - 1]);
    fl[n - 1] = f[n - 1];
    fr[n - 1] = f[n - 1];
  }
}

/* fix_positivity, interpolations.F90:4987-5064 (active if fix_positivity) */
static void fix_positivity(int n, const double* f, double* fl, double* fr)
{
  int i;
  for (i = 0; i < n; i++) {
    if (fl[i] <= 0.0) {
      int im1 = (i - 1 > 0) ? i - 1 : 0;
      fl[i]   = f[i];
      fr[im1] = f[i];
    }
    if (fr[i] <= 0.0) {
      int ip1 = (i + 1 < n - 1) ? i + 1 : n - 1;
      fl[ip1] = f[ip1];
      fr[i]   = f[ip1];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Gaussian-process reconstruction                                           */
/* ------------------------------------------------------------------------ */
/* invert, algebra.F90:1198-1305: Gauss-Jordan elimination in quad precision
 * (max_real_kind, algebra.F90:42); returns 0 when the matrix is singular */
static int invert_q(int n, const __float128* m, __float128* r)
{
  __float128 *gm, t;
  int i, j, k, flag, ok = 0;
  gm = (__float128*)malloc((size_t)n * 2 * n * sizeof(__float128));
  if (!gm) return 0;
#define g(i, j) gm[(size_t)(i) * 2 * n + (j)]
  for (i = 0; i < n; i++)
    for (j = 0; j < 2 * n; j++)
      g(i, j) = (j < n) ? m[i * n + j] : ((i + n == j) ? 1.0Q : 0.0Q);
  for (k = 0; k < n - 1; k++) {
    if (g(k, k) == 0.0Q) {
      flag = 0;
      for (i = k + 1; i < n; i++) {
        if (g(i, k) != 0.0Q) {
          for (j = 0; j < 2 * n; j++) g(k, j) = g(k, j) + g(i, j);
          flag = 1;
          break;
        }
      }
      if (!flag) goto done;
    }
    for (j = k + 1; j < n; j++) {
      t = g(j, k) / g(k, k);
      for (i = k; i < 2 * n; i++) g(j, i) = g(j, i) - t * g(k, i);
    }
  }
  for (i = 0; i < n; i++)
    if (g(i, i) == 0.0Q) goto done;
  for (i = 0; i < n; i++) {
    t = g(i, i);
    for (j = i; j < 2 * n; j++) g(i, j) = g(i, j) / t;
  }
  for (k = n - 2; k >= 0; k--)
    for (i = 0; i <= k; i++) {
      t = g(i, k + 1);
      for (j = k; j < 2 * n; j++) g(i, j) = g(i, j) - g(k + 1, j) * t;
    }
  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++) r[i * n + j] = g(i, j + n);
  ok = 1;
done:
#undef g
  free(gm);
  return ok;
}

/* prepare_mgp, interpolations.F90:558-711: the multi-dimensional Gaussian
 * process over the ngp^nd stencil; covariance of the cell averages and the
 * face vectors (xgp) in quad precision, ugp(:, s, d) = xgp(:, s, d) A^-1
 * (no symmetrisation here, unlike prepare_gp).  Stencil index runs i fastest,
 * then j, then k. */
static int prepare_mgp(ocfg* o)
{
  const int ngp = o->c.ngp, mgp = (ngp - 1) / 2, nd = o->nd;
  const int dgp = nd == 3 ? ngp * ngp * ngp : ngp * ngp;
  __float128 *cov, *inv, *xgp, sig, fc;
  int i, j, a1, b1, c1, a2, b2, c2, s, d, m;
  if (ngp < 1 || ngp > 5 || ngp % 2 == 0) return 202;
  cov = (__float128*)malloc(((size_t)2 * dgp * dgp + 6 * dgp) * sizeof(__float128));
  if (!cov) return 500;
  inv = cov + (size_t)dgp * dgp; xgp = inv + (size_t)dgp * dgp;
  sig = (__float128)(sqrt(2.0) * o->c.sgp);
  fc  = (__float128)(0.5 * sqrt(3.14159265358979323846264338327950)) * sig;
#define ERFD(t) (erfq((((__float128)(1.0 * (t))) + 0.5Q) / sig) - erfq((((__float128)(1.0 * (t))) - 0.5Q) / sig))
  i = 0;
  for (c1 = (nd == 3 ? -mgp : 0); c1 <= (nd == 3 ? mgp : 0); c1++)
    for (b1 = -mgp; b1 <= mgp; b1++)
      for (a1 = -mgp; a1 <= mgp; a1++, i++) {
        j = 0;
        for (c2 = (nd == 3 ? -mgp : 0); c2 <= (nd == 3 ? mgp : 0); c2++)
          for (b2 = -mgp; b2 <= mgp; b2++)
            for (a2 = -mgp; a2 <= mgp; a2++, j++) {
              __float128 v = fc * ERFD(a1 - a2) * ERFD(b1 - b2);
              if (nd == 3) v = v * ERFD(c1 - c2);
              cov[(size_t)i * dgp + j] = v;
            }
      }
  i = 0;
  for (c1 = (nd == 3 ? -mgp : 0); c1 <= (nd == 3 ? mgp : 0); c1++)
    for (b1 = -mgp; b1 <= mgp; b1++)
      for (a1 = -mgp; a1 <= mgp; a1++, i++) {
        const __float128 xl = ((__float128)(1.0 * a1) - 0.5Q) / sig, xr = ((__float128)(1.0 * a1) + 0.5Q) / sig;
        const __float128 yl = ((__float128)(1.0 * b1) - 0.5Q) / sig, yr = ((__float128)(1.0 * b1) + 0.5Q) / sig;
        const __float128 zl = ((__float128)(1.0 * c1) - 0.5Q) / sig, zr = ((__float128)(1.0 * c1) + 0.5Q) / sig;
        const __float128 fx = erfq(xr) - erfq(xl), fy = erfq(yr) - erfq(yl);
        const __float128 fz = nd == 3 ? erfq(zr) - erfq(zl) : 1.0Q;
        if (nd == 3) {
          xgp[0 * dgp + i] = expq(-(xl * xl)) * fy * fz;
          xgp[1 * dgp + i] = expq(-(xr * xr)) * fy * fz;
          xgp[2 * dgp + i] = expq(-(yl * yl)) * fx * fz;
          xgp[3 * dgp + i] = expq(-(yr * yr)) * fx * fz;
          xgp[4 * dgp + i] = expq(-(zl * zl)) * fx * fy;
          xgp[5 * dgp + i] = expq(-(zr * zr)) * fx * fy;
        } else {
          xgp[0 * dgp + i] = expq(-(xl * xl)) * fy;
          xgp[1 * dgp + i] = expq(-(xr * xr)) * fy;
          xgp[2 * dgp + i] = expq(-(yl * yl)) * fx;
          xgp[3 * dgp + i] = expq(-(yr * yr)) * fx;
        }
      }
#undef ERFD
  if (!invert_q(dgp, cov, inv)) { free(cov); return 201; }
  for (d = 0; d < nd; d++)
    for (s = 0; s < 2; s++)
      for (m = 0; m < dgp; m++) {
        __float128 acc = 0.0Q;
        for (i = 0; i < dgp; i++) acc = acc + xgp[(2 * d + s) * dgp + i] * inv[(size_t)i * dgp + m];
        o->ugp[((size_t)(d * 2 + s)) * dgp + m] = (double)acc;
      }
  o->dgp = dgp;
  o->mgp = mgp;
  free(cov);
  return 0;
}

/* prepare_gp, interpolations.F90:4438-4547: the weights cgp = x^T A^-1 of
 * the Gaussian process with an integrated squared-exponential kernel of width
 * sgp over ngp cells; quad precision as in the reference, with its mixed
 * precision constants (sqrt(2) sgp and sqrt(pi)/2 are formed in double) */
static int prepare_gp(ocfg* o)
{
  const int ngp = o->c.ngp, mgp = (ngp - 1) / 2;
  __float128 cov[256], agp[256], xgp[16], sig, fc, zl, zr;
  int i, j, ip, jp;
  if (ngp < 1 || ngp > 15 || ngp % 2 == 0) return 202;
  sig = (__float128)(sqrt(2.0) * o->c.sgp);
  fc  = (__float128)(0.5 * sqrt(3.14159265358979323846264338327950)) * sig;
  for (ip = -mgp, i = 0; ip <= mgp; ip++, i++)
    for (jp = -mgp, j = 0; jp <= mgp; jp++, j++) {
      zl = ((__float128)(1.0 * (ip - jp)) - 0.5Q) / sig;
      zr = ((__float128)(1.0 * (ip - jp)) + 0.5Q) / sig;
      cov[i * ngp + j] = fc * (erfq(zr) - erfq(zl));
    }
  if (!invert_q(ngp, cov, agp)) return 202;
  for (j = 0; j < ngp; j++)
    for (i = 0; i < ngp; i++)
      if (i != j) {
        fc = 0.5Q * (agp[i * ngp + j] + agp[j * ngp + i]);
        agp[i * ngp + j] = fc;
        agp[j * ngp + i] = fc;
      }
  for (jp = -mgp, j = 0; jp <= mgp; jp++, j++) {
    zr = (__float128)(0.5 - 1.0 * jp) / sig;
    xgp[j] = expq(-(zr * zr));
  }
  for (j = 0; j < ngp; j++) {
    __float128 acc = 0.0Q;
    for (i = 0; i < ngp; i++) acc = acc + xgp[i] * agp[i * ngp + j];
    o->cgp[j] = (double)acc;
  }
  o->mgp = mgp;
  return 0;
}

/* minmod, minmod4, median (interpolations.F90:4892-4975) */
static double mmod(double a, double b)
{
  return (copysign(0.5, a) + copysign(0.5, b)) * fmin_r(fabs(a), fabs(b));
}
static double mmod4(double a, double b, double c, double d) { return mmod(mmod(a, b), mmod(c, d)); }
static double median3(double a, double b, double c) { return a + mmod(b - a, c - a); }

/* mp_limiting, interpolations.F90:5209-5284 (Suresh & Huynh 1997) on the
 * left-face values v of pencil f */
static void mp_limiting(const ocfg* o, int n, const double* f, double* v, double* dm)
{
  const double kappa = o->kappa, kbeta = o->c.kbeta, eps = o->c.eps;
  int i;
  /* dm[k] holds dm(k) of the reference, k = 0 .. n + 2 */
  dm[0] = 0.0;
  dm[1] = 0.0;
  for (i = 2; i <= n; i++) dm[i] = f[i - 1] - f[i - 2];
  dm[n + 1] = 0.0;
  dm[n + 2] = 0.0;
  for (i = 1; i <= n - 1; i++) {
    double df, ds, fmp, dm1, dc0, dp1, dc4, dml, dmr, fmd, ful, flc, fmx, fmn;
    const double fi = f[i - 1];
    if (dm[i] * dm[i + 1] >= 0.0) df = kappa * dm[i];
    else                          df = kbeta * dm[i];
    fmp = fi + mmod(dm[i + 1], df);
    ds  = (v[i - 1] - fi) * (v[i - 1] - fmp);
    if (ds > eps) {
      dm1 = dm[i] - dm[i - 1];
      dc0 = dm[i + 1] - dm[i];
      dp1 = dm[i + 2] - dm[i + 1];
      dc4 = 4.0 * dc0;
      dml = 0.5 * mmod4(dc4 - dm1, 4.0 * dm1 - dc0, dc0, dm1);
      dmr = 0.5 * mmod4(dc4 - dp1, 4.0 * dp1 - dc0, dc0, dp1);
      fmd = (fi + 0.5 * dm[i + 1]) - dmr;
      ful = fi + df;
      flc = (fi + 0.5 * df) + dml;
      fmx = fmax_r(fmin_r(fmin_r(fi, f[i]), fmd), fmin_r(fmin_r(fi, ful), flc));
      fmn = fmin_r(fmax_r(fmax_r(fi, f[i]), fmd), fmax_r(fmax_r(fi, ful), flc));
      v[i - 1] = median3(v[i - 1], fmn, fmx);
    }
  }
}

/* one direction of reconstruct_gp: u(1..n) of pencil f, mp-limited */
static void gp_pass(const ocfg* o, int n, const double* f, double* u, double* dm)
{
  static const double ce5[5] = {2.0 / 6.0e+01, -1.3e+01 / 6.0e+01, 4.70e+01 / 6.0e+01,
                                2.70e+01 / 6.0e+01, -3.0 / 6.0e+01};
  static const double ce3[3] = {-1.0 / 6.0, 5.0 / 6.0, 2.0 / 6.0};
  const int mgp = o->mgp, ngp = o->c.ngp;
  int i, m;
  u[0] = 5.0e-01 * f[0] + 5.0e-01 * f[1];
  u[1] = (ce3[0] * f[0] + ce3[1] * f[1]) + ce3[2] * f[2];
  for (i = 3; i <= mgp; i++)
    u[i - 1] = (((ce5[0] * f[i - 3] + ce5[1] * f[i - 2]) + ce5[2] * f[i - 1]) + ce5[3] * f[i])
               + ce5[4] * f[i + 1];
  for (i = n + 1 - mgp; i <= n - 2; i++)
    u[i - 1] = (((ce5[0] * f[i - 3] + ce5[1] * f[i - 2]) + ce5[2] * f[i - 1]) + ce5[3] * f[i])
               + ce5[4] * f[i + 1];
  u[n - 2] = (ce3[0] * f[n - 3] + ce3[1] * f[n - 2]) + ce3[2] * f[n - 1];
  u[n - 1] = f[n - 1];
  for (i = 1 + mgp; i <= n - mgp; i++) {
    const double fi = f[i - 1];
    double acc = o->cgp[0] * (f[i - 1 - mgp] - fi);
    for (m = 1; m < ngp; m++) acc = acc + o->cgp[m] * (f[i - 1 - mgp + m] - fi);
    u[i - 1] = acc + fi;
  }
  mp_limiting(o, n, f, u, dm);
}

/* one upwind pass of reconstruct_mp5 / reconstruct_mp7 (interpolations.F90
 * 3675-3770, 3799-3901): polynomial face values (fifth or seventh order in
 * the interior, lower order near the ends), then mp_limiting */
static void mp_pass(const ocfg* o, int n, const double* f, double* u, double* dm, int order)
{
  static const double ce7[7] = {-3.0 / 4.2e+02, 2.5e+01 / 4.2e+02, -1.01e+02 / 4.2e+02,
                                3.19e+02 / 4.2e+02, 2.14e+02 / 4.2e+02, -3.8e+01 / 4.2e+02,
                                4.0 / 4.2e+02};
  static const double ce5[5] = {2.0 / 6.0e+01, -1.3e+01 / 6.0e+01, 4.70e+01 / 6.0e+01,
                                2.70e+01 / 6.0e+01, -3.0 / 6.0e+01};
  static const double ce3[3] = {-1.0 / 6.0, 5.0 / 6.0, 2.0 / 6.0};
  int i, m;
  if (order == 7) {
    for (i = 4; i <= n - 3; i++) {
      double acc = ce7[0] * f[i - 4];
      for (m = 1; m < 7; m++) acc = acc + ce7[m] * f[i - 4 + m];
      u[i - 1] = acc;
    }
  } else {
    for (i = 3; i <= n - 2; i++) {
      double acc = ce5[0] * f[i - 3];
      for (m = 1; m < 5; m++) acc = acc + ce5[m] * f[i - 3 + m];
      u[i - 1] = acc;
    }
  }
  u[0] = 5.0e-01 * f[0] + 5.0e-01 * f[1];
  u[1] = (ce3[0] * f[0] + ce3[1] * f[1]) + ce3[2] * f[2];
  if (order == 7) {
    double acc = ce5[0] * f[0];
    for (m = 1; m < 5; m++) acc = acc + ce5[m] * f[m];
    u[2] = acc;
    acc = ce5[0] * f[n - 5];
    for (m = 1; m < 5; m++) acc = acc + ce5[m] * f[n - 5 + m];
    u[n - 3] = acc;
  }
  u[n - 2] = (ce3[0] * f[n - 3] + ce3[1] * f[n - 2]) + ce3[2] * f[n - 1];
  u[n - 1] = f[n - 1];
  mp_limiting(o, n, f, u, dm);
}

static void reconstruct_mp(const ocfg* o, int n, const double* f, double* fl, double* fr,
                           double* w1, double* w2, double* w3, int order)
{
  int i;
  mp_pass(o, n, f, fl, w3, order);
  for (i = 0; i < n; i++) w1[i] = f[n - 1 - i];
  mp_pass(o, n, w1, w2, w3, order);
  for (i = 0; i < n - 1; i++) fr[i] = w2[n - 2 - i];
  i = n - 2;
  fl[0] = 0.5 * (f[0] + f[1]);
  fr[i] = 0.5 * (f[i] + f[n - 1]);
  fl[n - 1] = f[n - 1];
  fr[n - 1] = f[n - 1];
}

/* reconstruct_weno3, interpolations.F90:1469-1574: third-order WENO with
 * the cell size squared in the smoothness weights */
static void reconstruct_weno3(int n, double h, const double* f, double* fl, double* fr)
{
  const double dp = 2.0 / 3.0, dm = 1.0 / 3.0, h2 = h * h;
  double dfl, dfr, ww, bp, bm, ap, am, wp, wm, fp, fm;
  int i;
  for (i = 1; i < n - 1; i++) {
    dfl = f[i] - f[i - 1];
    dfr = f[i + 1] - f[i];
    ww = (dfr - dfl) * (dfr - dfl);
    bp = dfr * dfr;
    bm = dfl * dfl;
    ap = 1.0 + ww / (bp + h2);
    am = 1.0 + ww / (bm + h2);
    wp = dp * ap;
    wm = dm * am;
    ww = 2.0 * (wp + wm);
    fp = f[i] + f[i + 1];
    fm = -f[i - 1] + 3.0 * f[i];
    fl[i] = (wp * fp + wm * fm) / ww;
    wp = dp * am;
    wm = dm * ap;
    ww = 2.0 * (wp + wm);
    fp = f[i] + f[i - 1];
    fm = -f[i + 1] + 3.0 * f[i];
    fr[i - 1] = (wp * fp + wm * fm) / ww;
  }
  i = n - 2;
  fl[0] = 0.5 * (f[0] + f[1]);
  fr[i] = 0.5 * (f[i] + f[n - 1]);
  fl[n - 1] = f[n - 1];
  fr[n - 1] = f[n - 1];
}

/* the third-order limited slope of reconstruct_limo3 (eq. 3.13 of Cada &
 * Torrilhon 2009, halved): th is the slope ratio */
static double limo3_phi(double th, double xl, double xi)
{
  double f1 = (2.0 + th) / 6.0, f2;
  if (th >= 0.0) f2 = fmax_r(0.0, fmin_r(fmin_r(f1, th), 0.8));
  else           f2 = fmax_r(0.0, fmin_r(f1, -0.25 * th));
  return xl * f1 + xi * f2;
}

/* reconstruct_limo3, interpolations.F90:1598-1719: LimO3 with the switch
 * between the limited and the plain third-order slope over a radius of
 * limo3_rad cells */
static void reconstruct_limo3(const ocfg* o, int n, double h, const double* f, double* fl,
                              double* fr)
{
  const double eps = o->c.eps, rdx = o->c.limo3_rad * h, rdx2 = rdx * rdx;
  double dfl, dfr, et, xi, xl;
  int i;
  for (i = 1; i < n - 1; i++) {
    dfl = f[i] - f[i - 1];
    dfr = f[i + 1] - f[i];
    et = (dfl * dfl + dfr * dfr) / rdx2;
    xi = fmax_r(0.0, 0.5 * fmin_r(2.0, 1.0 + (et - 1.0) / eps));
    xl = 1.0 - xi;
    if (fabs(dfr) > eps) fl[i] = f[i] + dfr * limo3_phi(dfl / dfr, xl, xi);
    else                 fl[i] = f[i];
    if (fabs(dfl) > eps) fr[i - 1] = f[i] - dfl * limo3_phi(dfr / dfl, xl, xi);
    else                 fr[i - 1] = f[i];
  }
  i = n - 2;
  fl[0] = 0.5 * (f[0] + f[1]);
  fr[i] = 0.5 * (f[i] + f[n - 1]);
  fl[n - 1] = f[n - 1];
  fr[n - 1] = f[n - 1];
}

/* reconstruct_ppm, interpolations.F90:1746-1913: fourth-order face values
 * limited as in Colella & Sekora 2008, then the extremum and monotonicity
 * limiters of the cell states.  The neighbour face difference dfm of the
 * monotonicity test is taken between the same face (im2 = max(1, i - 1) in
 * the reference), so it is zero; kept as written. */
static void reconstruct_ppm(const ocfg* o, int n, const double* f, double* fl, double* fr,
                            double* fi, double* df2)
{
  const double pc = o->c.ppm_const;
  double df2c, df2m, df2p, df2s, df2l, dfm, dfp, dcm, dcp, alm, alp, amx, sgn, del;
  int i, cm, cp, ext, im2;
  fi[0] = 0.5 * (f[0] + f[1]);
  for (i = 1; i < n - 2; i++)
    fi[i] = (7.0 * (f[i] + f[i + 1]) - (f[i - 1] + f[i + 2])) / 12.0;
  fi[n - 2] = 0.5 * (f[n - 2] + f[n - 1]);
  fi[n - 1] = f[n - 1];
  df2[0] = 0.0;
  for (i = 1; i < n - 1; i++) df2[i] = (f[i + 1] + f[i - 1]) - 2.0 * f[i];
  df2[n - 1] = 0.0;
  for (i = 1; i < n - 2; i++) {
    if ((f[i + 1] - fi[i]) * (fi[i] - f[i]) < 0.0) {
      df2c = 3.0 * ((f[i + 1] + f[i]) - 2.0 * fi[i]);
      df2m = df2[i];
      df2p = df2[i + 1];
      if (fmin_r(fmin_r(df2c, df2m), df2p) * fmax_r(fmax_r(df2c, df2m), df2p) > 0.0)
        df2l = copysign(fmin_r(pc * fmin_r(fabs(df2m), fabs(df2p)), fabs(df2c)), df2c);
      else
        df2l = 0.0;
      fi[i] = 0.5 * (f[i] + f[i + 1]) - df2l / 6.0;
    }
  }
  for (i = 1; i < n - 1; i++) {
    alm = fi[i - 1] - f[i];
    alp = fi[i] - f[i];
    cm = fabs(alm) >= 2.0 * fabs(alp);
    cp = fabs(alp) >= 2.0 * fabs(alm);
    ext = 0;
    if (alm * alp > 0.0) {
      ext = 1;
    } else if (cm || cp) {
      im2 = i - 1 > 0 ? i - 1 : 0;
      dfm = fi[i - 1] - fi[im2];
      dfp = fi[i + 1] - fi[i];
      dcm = f[i] - f[i - 1];
      dcp = f[i + 1] - f[i];
      if (fmin_r(fabs(dfm), fabs(dfp)) >= fmin_r(fabs(dcm), fabs(dcp))) ext = dfm * dfp < 0.0;
      else                                                         ext = dcm * dcp < 0.0;
    }
    if (ext) {
      df2s = 6.0 * (alm + alp);
      df2m = df2[i - 1];
      df2c = df2[i];
      df2p = df2[i + 1];
      if (fmin_r(fmin_r(fmin_r(df2s, df2c), df2m), df2p) * fmax_r(fmax_r(fmax_r(df2s, df2c), df2m), df2p) > 0.0) {
        df2l = copysign(fmin_r(pc * fmin_r(fmin_r(fabs(df2m), fabs(df2c)), fabs(df2p)), fabs(df2s)), df2s);
        if (fabs(df2l) > 0.0) {
          alm = alm * df2l / df2s;
          alp = alp * df2l / df2s;
        } else {
          alm = 0.0;
          alp = 0.0;
        }
      } else {
        alm = 0.0;
        alp = 0.0;
      }
    } else {
      if (cp) {
        sgn = copysign(1.0, alp);
        amx = -(0.25 * (alp * alp) / (alp + alm));
        dfp = f[i + 1] - f[i];
        if (sgn * amx >= sgn * dfp) {
          del = dfp * (dfp - alm);
          if (del >= 0.0) alp = -2.0 * (dfp + sgn * sqrt(del));
          else            alp = -2.0 * alm;
        }
      }
      if (cm) {
        sgn = copysign(1.0, alm);
        amx = -(0.25 * (alm * alm) / (alp + alm));
        dfm = f[i - 1] - f[i];
        if (sgn * amx >= sgn * dfm) {
          del = dfm * (dfm - alp);
          if (del >= 0.0) alm = -2.0 * (dfm + sgn * sqrt(del));
          else            alm = -2.0 * alp;
        }
      }
    }
    fr[i - 1] = f[i] + alm;
    fl[i] = f[i] + alp;
  }
  fl[0] = fi[0];
  fr[n - 2] = fi[n - 1];
  fl[n - 1] = fi[n - 1];
  fr[n - 1] = f[n - 1];
}

/* tridiag, algebra.F90:1115-1174: the Thomas algorithm for the system with
 * sub-diagonal l, diagonal d, super-diagonal u; g is workspace.  A zero
 * pivot stops it as the reference does (x left partly solved). */
static void tridiag(int n, const double* l, const double* d, const double* u, const double* r,
                    double* x, double* g)
{
  double t = d[0];
  int i;
  x[0] = r[0] / t;
  for (i = 1; i < n; i++) {
    g[i] = u[i - 1] / t;
    t = d[i] - l[i] * g[i];
    if (t == 0.0) return;
    x[i] = (r[i] - l[i] * x[i - 1]) / t;
  }
  for (i = n - 2; i >= 0; i--) x[i] = x[i] - g[i + 1] * x[i + 1];
}

/* reconstruct_crweno5z / _crweno5yc / _crweno5ns (interpolations.F90:
 * 2477-2834, 2870-3230, 3266-3646): compact WENO (Ghosh & Baeder 2012).
 * The nonlinear weights of each cell make a tridiagonal system for the face
 * values; the rows within ng cells of the pencil ends hold the explicit
 * WENO5 values (the weights of the variant) and the end rows the averages.
 * variant 0: Z, 1: YC, 2: NS (separate weights for the two states). */
static void reconstruct_crweno(const ocfg* o, int n, const double* f, double* fl, double* fr,
                               int variant)
{
  const double c1 = 13.0 / 12.0, c2 = 0.25, cl = 1.0 / 9.0, cc = 5.0 / 9.0, cr = 1.0 / 3.0;
  const double dl = 0.1, dc = 0.6, dr = 0.3, dq = 0.5, xi = 0.4, eps = o->c.eps;
  const double a11 = 2.0 / 6.0, a12 = -7.0 / 6.0, a13 = 11.0 / 6.0;
  const double a21 = -1.0 / 6.0, a22 = 5.0 / 6.0, a23 = 2.0 / 6.0;
  const double a31 = 2.0 / 6.0, a32 = 5.0 / 6.0, a33 = -1.0 / 6.0;
  const int ng = o->ng;
  double *dfm, *dfp, *df2, *al, *ac, *ar, *a, *b, *c, *r, *u, *g;
  double bl, bc, br, tt, wl, wc, wr, ww, ql, qc, qr;
  int i, s;
  dfm = (double*)calloc(19 * (size_t)n, sizeof(double));
  dfp = dfm + n; df2 = dfp + n;
  al = df2 + n; ac = al + 2 * n; ar = ac + 2 * n;        /* [s * n + i] */
  a = ar + 2 * n; b = a + 2 * n; c = b + 2 * n; r = c + 2 * n;
  u = r + 2 * n; g = u + n;
  for (i = 0; i < n - 1; i++) { dfp[i] = f[i + 1] - f[i]; dfm[i + 1] = dfp[i]; }
  dfm[0] = dfp[0];
  dfp[n - 1] = dfm[n - 1];
  for (i = 0; i < n; i++)
    df2[i] = variant == 2 ? 0.5 * fabs(dfp[i] - dfm[i])
                          : c1 * ((dfp[i] - dfm[i]) * (dfp[i] - dfm[i]));
  for (i = 1; i < n - 1; i++) {
    if (variant == 2) {
      double df, lq, l3, zt;
      df = fabs(dfp[i]); lq = xi * df;
      bl = df2[i - 1] + xi * fabs(2.0 * dfm[i] - dfm[i - 1]);
      bc = df2[i] + lq; br = df2[i + 1] + lq;
      l3 = (df * df) * df;
      zt = 0.5 * ((bl - br) * (bl - br) + (l3 / (1.0 + l3)) * (l3 / (1.0 + l3)));
      al[i] = 1.0 + zt / ((bl + eps) * (bl + eps));
      ac[i] = 1.0 + zt / ((bc + eps) * (bc + eps));
      ar[i] = 1.0 + zt / ((br + eps) * (br + eps));
      df = fabs(dfm[i]); lq = xi * df;
      bl = df2[i - 1] + lq; bc = df2[i] + lq;
      br = df2[i + 1] + xi * fabs(2.0 * dfp[i] - dfp[i + 1]);
      l3 = (df * df) * df;
      zt = 0.5 * ((bl - br) * (bl - br) + (l3 / (1.0 + l3)) * (l3 / (1.0 + l3)));
      al[n + i] = 1.0 + zt / ((bl + eps) * (bl + eps));
      ac[n + i] = 1.0 + zt / ((bc + eps) * (bc + eps));
      ar[n + i] = 1.0 + zt / ((br + eps) * (br + eps));
    } else {
      const double t1 = 3.0 * dfm[i] - dfm[i - 1], t2 = dfp[i] + dfm[i];
      const double t3 = 3.0 * dfp[i] - dfp[i + 1];
      bl = df2[i - 1] + c2 * (t1 * t1);
      bc = df2[i] + c2 * (t2 * t2);
      br = df2[i + 1] + c2 * (t3 * t3);
      if (variant == 1) {
        const int im2 = i - 2 > 0 ? i - 2 : 0, ip2 = i + 2 < n - 1 ? i + 2 : n - 1;
        const double d4 = (6.0 * f[i] + (f[im2] + f[ip2])) - 4.0 * (f[i - 1] + f[i + 1]);
        tt = d4 * d4;
      } else {
        tt = fabs(br - bl);
      }
      al[i] = 1.0 + tt / (bl + eps);
      ac[i] = 1.0 + tt / (bc + eps);
      ar[i] = 1.0 + tt / (br + eps);
      al[n + i] = al[i]; ac[n + i] = ac[i]; ar[n + i] = ar[i];
    }
  }
  /* compact rows i = ng .. n-ng+1 (1-based) */
  for (i = ng - 1; i <= n - ng; i++) {
    wl = cl * al[i]; wc = cc * ac[i]; wr = cr * ar[i];
    ww = (wl + wr) + wc; wl = wl / ww; wr = wr / ww; wc = 1.0 - (wl + wr);
    a[i] = 2.0 * wl + wc; b[i] = wl + 2.0 * (wc + wr); c[i] = wr;
    r[i] = ((wl * f[i - 1] + (5.0 * (wl + wc) + wr) * f[i]) + (wc + 5.0 * wr) * f[i + 1]) * dq;
    wl = cl * ar[n + i]; wc = cc * ac[n + i]; wr = cr * al[n + i];
    ww = (wl + wr) + wc; wl = wl / ww; wr = wr / ww; wc = 1.0 - (wl + wr);
    a[n + i] = wr; b[n + i] = wl + 2.0 * (wc + wr); c[n + i] = 2.0 * wl + wc;
    r[n + i] = ((wl * f[i + 1] + (5.0 * (wl + wc) + wr) * f[i]) + (wc + 5.0 * wr) * f[i - 1]) * dq;
  }
  /* explicit WENO5 rows of the left states: 2..ng and n-ng..n-1 (1-based) */
  for (s = 0; s < 2; s++) {
    const int lo = s == 0 ? 1 : n - ng - 1, hi = s == 0 ? ng - 1 : n - 2;
    for (i = lo; i <= hi; i++) {
      const int im2 = i - 2 > 0 ? i - 2 : 0, ip2 = i + 2 < n - 1 ? i + 2 : n - 1;
      wl = dl * al[i]; wc = dc * ac[i]; wr = dr * ar[i];
      ww = (wl + wr) + wc; wl = wl / ww; wr = wr / ww; wc = 1.0 - (wl + wr);
      ql = (a11 * f[im2] + a12 * f[i - 1]) + a13 * f[i];
      qc = (a21 * f[i - 1] + a22 * f[i]) + a23 * f[i + 1];
      qr = (a31 * f[i] + a32 * f[i + 1]) + a33 * f[ip2];
      a[i] = 0.0; b[i] = 1.0; c[i] = 0.0;
      r[i] = (wl * ql + wr * qr) + wc * qc;
    }
  }
  a[0] = 0.0; b[0] = 1.0; c[0] = 0.0; r[0] = 0.5 * (f[0] + f[1]);
  a[n - 1] = 0.0; b[n - 1] = 1.0; c[n - 1] = 0.0; r[n - 1] = f[n - 1];
  /* explicit rows of the right states: 2..ng+1 and n-ng+1..n-1 (1-based) */
  for (s = 0; s < 2; s++) {
    const int lo = s == 0 ? 1 : n - ng, hi = s == 0 ? ng : n - 2;
    for (i = lo; i <= hi; i++) {
      const int im2 = i - 2 > 0 ? i - 2 : 0, ip2 = i + 2 < n - 1 ? i + 2 : n - 1;
      wl = dl * ar[n + i]; wc = dc * ac[n + i]; wr = dr * al[n + i];
      ww = (wl + wr) + wc; wl = wl / ww; wr = wr / ww; wc = 1.0 - (wl + wr);
      ql = (a11 * f[ip2] + a12 * f[i + 1]) + a13 * f[i];
      qc = (a21 * f[i + 1] + a22 * f[i]) + a23 * f[i - 1];
      qr = (a31 * f[i] + a32 * f[i - 1]) + a33 * f[im2];
      a[n + i] = 0.0; b[n + i] = 1.0; c[n + i] = 0.0;
      r[n + i] = (wl * ql + wr * qr) + wc * qc;
    }
  }
  a[n] = 0.0; b[n] = 1.0; c[n] = 0.0; r[n] = f[0];
  a[2 * n - 1] = 0.0; b[2 * n - 1] = 1.0; c[2 * n - 1] = 0.0;
  r[2 * n - 1] = 0.5 * (f[n - 2] + f[n - 1]);
  tridiag(n, a, b, c, r, u, g);
  for (i = 0; i < n; i++) fl[i] = u[i];
  tridiag(n, a + n, b + n, c + n, r + n, u, g);
  for (i = 0; i < n - 1; i++) fr[i] = u[i + 1];
  i = n - 2;
  fl[0] = 0.5 * (f[0] + f[1]);
  fr[i] = 0.5 * (f[i] + f[n - 1]);
  fl[n - 1] = f[n - 1];
  fr[n - 1] = f[n - 1];
  free(dfm);
}

/* right-hand side of the compact MP systems on the pencil x: the compact
 * rows ng .. n-ng+1 (1-based), then the explicit rows of the ends */
static void crmp_rhs(int n, int ng, int variant, const double* x, double* r)
{
  static const double ci5[3] = {1.0 / 30.0, 19.0 / 30.0, 10.0 / 30.0};
  static const double ci5ld[4] = {3.0 / 120.0, 67.0 / 120.0, 49.0 / 120.0, 1.0 / 120.0};
  static const double ci7[5] = {-1.0 / 420.0, 19.0 / 420.0, 239.0 / 420.0, 159.0 / 420.0,
                                4.0 / 420.0};
  static const double ce7[7] = {-3.0 / 420.0, 25.0 / 420.0, -101.0 / 420.0, 319.0 / 420.0,
                                214.0 / 420.0, -38.0 / 420.0, 4.0 / 420.0};
  static const double ce5[5] = {2.0 / 60.0, -13.0 / 60.0, 47.0 / 60.0, 27.0 / 60.0, -3.0 / 60.0};
  static const double ce3[3] = {-1.0 / 6.0, 5.0 / 6.0, 2.0 / 6.0};
  /* stencil reach of the compact rows; rows whose stencil leaves the pencil
   * (ng < 3) are overwritten by the explicit rows below */
  const int lo = variant == 2 ? 2 : 1, hi = variant == 0 ? 1 : 2;
  int i, m;
  for (i = ng - 1; i <= n - ng; i++) {
    double s = 0.0;
    if (i < lo || i + hi > n - 1) continue;
    if (variant == 0)      for (m = 0; m < 3; m++) s = m ? s + ci5[m] * x[i - 1 + m] : ci5[0] * x[i - 1];
    else if (variant == 1) for (m = 0; m < 4; m++) s = m ? s + ci5ld[m] * x[i - 1 + m] : ci5ld[0] * x[i - 1];
    else                   for (m = 0; m < 5; m++) s = m ? s + ci7[m] * x[i - 2 + m] : ci7[0] * x[i - 2];
    r[i] = s;
  }
  r[0] = 0.5 * x[0] + 0.5 * x[1];
  r[1] = (ce3[0] * x[0] + ce3[1] * x[1]) + ce3[2] * x[2];
#define SUMC(c, k, i0) { double s = c[0] * x[i0]; for (m = 1; m < k; m++) s = s + c[m] * x[i0 + m]; r[i] = s; }
  if (variant < 2) {
    for (i = 2; i <= ng - 1; i++) SUMC(ce5, 5, i - 2)
    for (i = n - ng - 1; i <= n - 3; i++) SUMC(ce5, 5, i - 2)
  } else {
    i = 2; SUMC(ce5, 5, 0)
    for (i = 3; i <= ng - 1; i++) SUMC(ce7, 7, i - 3)
    for (i = n - ng - 1; i <= n - 4; i++) SUMC(ce7, 7, i - 3)
    i = n - 3; SUMC(ce5, 5, n - 5)
  }
#undef SUMC
  r[n - 2] = (ce3[0] * x[n - 3] + ce3[1] * x[n - 2]) + ce3[2] * x[n - 1];
  r[n - 1] = x[n - 1];
}

/* reconstruct_crmp5 / _crmp5ld / _crmp7 (interpolations.F90:3930-4427):
 * compact fifth/seventh-order face values from a tridiagonal system of
 * constant coefficients (identity rows within ng cells of the pencil ends),
 * then mp_limiting; the right states from the reversed pencil.  variant 0:
 * crmp5, 1: crmp5ld, 2: crmp7. */
static void reconstruct_crmp(const ocfg* o, int n, const double* f, double* fl, double* fr,
                             int variant)
{
  static const double di[3][3] = {{0.3, 0.6, 0.1}, {0.25, 0.6, 0.15},
                                  {2.0 / 7.0, 4.0 / 7.0, 1.0 / 7.0}};
  const int ng = o->ng;
  double *a, *b, *c, *r, *u, *g, *fi, *dm;
  int i;
  a = (double*)calloc(9 * (size_t)n + 8, sizeof(double));
  b = a + n; c = b + n; r = c + n; u = r + n; g = u + n; fi = g + n; dm = fi + n;
  for (i = 0; i < n; i++) {
    const int id = i >= ng && i <= n - ng - 2;
    a[i] = id ? di[variant][0] : 0.0;
    b[i] = id ? di[variant][1] : 1.0;
    c[i] = id ? di[variant][2] : 0.0;
  }
  crmp_rhs(n, ng, variant, f, r);
  tridiag(n, a, b, c, r, u, g);
  mp_limiting(o, n, f, u, dm);
  for (i = 0; i < n; i++) fl[i] = u[i];
  for (i = 0; i < n; i++) fi[i] = f[n - 1 - i];
  crmp_rhs(n, ng, variant, fi, r);
  tridiag(n, a, b, c, r, u, g);
  mp_limiting(o, n, fi, u, dm);
  for (i = 0; i < n - 1; i++) fr[i] = u[n - 2 - i];
  i = n - 2;
  fl[0] = 0.5 * (f[0] + f[1]);
  fr[i] = 0.5 * (f[i] + f[n - 1]);
  fl[n - 1] = f[n - 1];
  fr[n - 1] = f[n - 1];
  free(a);
}

/* reconstruct_gp, interpolations.F90:4561-4668 */
static void reconstruct_gp(const ocfg* o, int n, const double* f, double* fl, double* fr,
                           double* w1, double* w2, double* w3)
{
  int i;
  gp_pass(o, n, f, fl, w3);
  for (i = 0; i < n; i++) w1[i] = f[n - 1 - i];
  gp_pass(o, n, w1, w2, w3);
  for (i = 0; i < n - 1; i++) fr[i] = w2[n - 2 - i];
  i = n - 2;
  fl[0] = 0.5 * (f[0] + f[1]);
  fr[i] = 0.5 * (f[i] + f[n - 1]);
  fl[n - 1] = f[n - 1];
  fr[n - 1] = f[n - 1];
}

/* ------------------------------------------------------------------------ */
/* reconstruct_interfaces (schemes.F90:591-633) for one block                 */
/* qs layout: [(d*2 + s)][k][j][i][p]                                         */
/* ------------------------------------------------------------------------ */
/* reconstruct_states of interfaces_dir: the selected 1D scheme */
static void recon_dir(const ocfg* o, int n, double h, const double* f, double* fl, double* fr,
                      double* w1, double* w2, double* w3)
{
  switch (o->c.reconstruction) {
  case AMUN_REC_WENO3: reconstruct_weno3(n, h, f, fl, fr); break;
  case AMUN_REC_CRWENO5Z: reconstruct_crweno(o, n, f, fl, fr, 0); break;
  case AMUN_REC_CRWENO5YC: reconstruct_crweno(o, n, f, fl, fr, 1); break;
  case AMUN_REC_CRWENO5NS: reconstruct_crweno(o, n, f, fl, fr, 2); break;
  case AMUN_REC_CRMP5: reconstruct_crmp(o, n, f, fl, fr, 0); break;
  case AMUN_REC_CRMP5LD: reconstruct_crmp(o, n, f, fl, fr, 1); break;
  case AMUN_REC_CRMP7: reconstruct_crmp(o, n, f, fl, fr, 2); break;
  case AMUN_REC_LIMO3: reconstruct_limo3(o, n, h, f, fl, fr); break;
  case AMUN_REC_PPM: reconstruct_ppm(o, n, f, fl, fr, w1, w2); break;
  case AMUN_REC_GP: reconstruct_gp(o, n, f, fl, fr, w1, w2, w3); break;
  case AMUN_REC_WENO5YC: weno5z_v(o, n, f, fl, fr, w1, w2, w3, 1); break;
  case AMUN_REC_WENO5NS: weno5ns(o, n, f, fl, fr, w1, w2, w3); break;
  case AMUN_REC_MP5: reconstruct_mp(o, n, f, fl, fr, w1, w2, w3, 5); break;
  case AMUN_REC_MP7: reconstruct_mp(o, n, f, fl, fr, w1, w2, w3, 7); break;
  default: weno5z(o, n, f, fl, fr, w1, w2, w3);
  }
}

/* mlp_limiting, interpolations.F90:1171-1321: the multi-dimensional
 * limiting process over the face states of variable p of one block, every
 * cell (neighbour indices clamped to the block array): minmod slopes in
 * every direction; where their absolute sum exceeds the distance dqc of the
 * cell value to the extremes of its 3^d box, slopes from the central
 * differences scaled to dqc and redistributed between the directions (the
 * 2D and the cyclic 3D formulas) with the minmod slopes' signs; then the
 * cell's two face states of each direction are reset to q +- dq when
 * either lies further from q than |dq|. */
static void mlp_limiting(const ocfg* o, const double* q, int p, double* qs)
{
  const int im = o->im, jm = o->jm, km = o->km, nv = o->nv, nd = o->nd;
  const size_t ncell = (size_t)im * jm * km;
  int i, j, k, m, n, a, b, c;
#define QS(d, s, i, j, k) qs[((size_t)((d) * 2 + (s)) * ncell + ((size_t)(k) * jm + (j)) * im + (i)) * nv + p]
#define Q(i, j, k) q[QI(o, p, i, j, k)]
  for (k = 0; k < km; k++) {
    const int km1 = nd == 3 ? (k - 1 > 0 ? k - 1 : 0) : k;
    const int kp1 = nd == 3 ? (k + 1 < km - 1 ? k + 1 : km - 1) : k;
    for (j = 0; j < jm; j++) {
      const int jm1 = j - 1 > 0 ? j - 1 : 0, jp1 = j + 1 < jm - 1 ? j + 1 : jm - 1;
      for (i = 0; i < im; i++) {
        const int im1 = i - 1 > 0 ? i - 1 : 0, ip1 = i + 1 < im - 1 ? i + 1 : im - 1;
        const double qc = Q(i, j, k);
        double dq[3] = {0.0, 0.0, 0.0}, dqm[3] = {0.0, 0.0, 0.0}, ap[3] = {0.0, 0.0, 0.0};
        double qmn, qmx, dqc, sa, fc, fl, dql, dqr;
        dq[0] = lim(AMUN_LIM_MINMOD, 0.5, qc - Q(im1, j, k), Q(ip1, j, k) - qc);
        dq[1] = lim(AMUN_LIM_MINMOD, 0.5, qc - Q(i, jm1, k), Q(i, jp1, k) - qc);
        if (nd == 3) dq[2] = lim(AMUN_LIM_MINMOD, 0.5, qc - Q(i, j, km1), Q(i, j, kp1) - qc);
        qmn = qmx = qc;
        for (c = km1; c <= kp1; c++)
          for (b = jm1; b <= jp1; b++)
            for (a = im1; a <= ip1; a++) {
              const double v = Q(a, b, c);
              qmn = v < qmn ? v : qmn;
              qmx = v > qmx ? v : qmx;
            }
        dqc = fmin_r(qmx - qc, qc - qmn);
        sa = fabs(dq[0]) + fabs(dq[1]);
        if (nd == 3) sa = sa + fabs(dq[2]);
        if (sa > dqc) {
          double sm;
          dqm[0] = fabs(Q(ip1, j, k) - Q(im1, j, k));
          dqm[1] = fabs(Q(i, jp1, k) - Q(i, jm1, k));
          if (nd == 3) dqm[2] = fabs(Q(i, j, kp1) - Q(i, j, km1));
          sm = fabs(dqm[0]) + fabs(dqm[1]);
          if (nd == 3) sm = sm + fabs(dqm[2]);
          fc = dqc / fmax_r(1.0e-16, sm);
          for (m = 0; m < nd; m++) ap[m] = fc * fabs(dqm[m]);
          if (nd == 3) {
            double hh[3], uu[3];
            for (n = 0; n < 3; n++) {
              const int np1 = (n + 1) % 3, np2 = (n + 2) % 3;
              hh[n] = fmax_r(ap[n] - fabs(dq[n]), 0.0);
              uu[n] = ap[n] - hh[n];
              uu[np1] = ap[np1] + 0.5 * hh[n];
              uu[np2] = ap[np2] + 0.5 * hh[n];
              fc = hh[n] / (hh[n] + 1.0e-16);
              fl = fc * (fmax_r(ap[np1] - fabs(dq[np1]), 0.0) - fmax_r(ap[np2] - fabs(dq[np2]), 0.0));
              ap[n] = uu[n];
              ap[np1] = uu[np1] - fl;
              ap[np2] = uu[np2] + fl;
            }
          } else {
            fl = fmax_r(ap[0] - fabs(dq[0]), 0.0) - fmax_r(ap[1] - fabs(dq[1]), 0.0);
            ap[0] = ap[0] - fl;
            ap[1] = ap[1] + fl;
          }
          for (m = 0; m < nd; m++) dq[m] = copysign(fabs(ap[m]), dq[m]);
        }
        dql = QS(0, 0, i, j, k) - qc;
        dqr = QS(0, 1, im1, j, k) - qc;
        if (fmax_r(fabs(dql), fabs(dqr)) > fabs(dq[0])) {
          QS(0, 0, i, j, k) = qc + dq[0];
          QS(0, 1, im1, j, k) = qc - dq[0];
        }
        dql = QS(1, 0, i, j, k) - qc;
        dqr = QS(1, 1, i, jm1, k) - qc;
        if (fmax_r(fabs(dql), fabs(dqr)) > fabs(dq[1])) {
          QS(1, 0, i, j, k) = qc + dq[1];
          QS(1, 1, i, jm1, k) = qc - dq[1];
        }
        if (nd == 3) {
          dql = QS(2, 0, i, j, k) - qc;
          dqr = QS(2, 1, i, j, km1) - qc;
          if (fmax_r(fabs(dql), fabs(dqr)) > fabs(dq[2])) {
            QS(2, 0, i, j, k) = qc + dq[2];
            QS(2, 1, i, j, km1) = qc - dq[2];
          }
        }
      }
    }
  }
#undef QS
#undef Q
}

static void reconstruct_interfaces(ocfg* o, const double* q, double* qs)
{
  const int im = o->im, jm = o->jm, km = o->km, nv = o->nv, nd = o->nd;
  const size_t ncell = (size_t)im * jm * km;
  int p, d, s, i, j, k;
#define QS(d, s, i, j, k, p) qs[((size_t)((d) * 2 + (s)) * ncell + ((size_t)(k) * jm + (j)) * im + (i)) * nv + (p)]
  /* copy ghost zones: qi = q everywhere first, the interior region is
   * overwritten below (interpolations.F90:767-778) */
  for (d = 0; d < nd; d++)
    for (s = 0; s < 2; s++)
      memcpy(&qs[(size_t)(d * 2 + s) * ncell * nv], q, ncell * nv * sizeof(double));

  for (p = 0; p < nv; p++) {
    int positive = (p == o->idn || p == o->ipr);
    if (o->c.reconstruction == AMUN_REC_MGP) {
      /* interfaces_mgp, interpolations.F90:978-1144: the
       * 2 nd face values of each cell from the GP weights; where any of them
       * leaves the range of its two cells by more than eps, all of them fall
       * back to the limited TVD slopes with the positivity halving */
      const int mg = o->mgp, dgp = o->dgp;
      double u[125];
      for (k = o->kbl; k <= o->keu; k++)
        for (j = o->jbl; j <= o->jeu; j++)
          for (i = o->ibl; i <= o->ieu; i++) {
            const double qc = q[QI(o, p, i, j, k)];
            double fv[6], dq[3] = {0.0, 0.0, 0.0}, dql, dqr;
            int m = 0, a, bb, cc, fl = 0, ds;
            for (cc = (nd == 3 ? -mg : 0); cc <= (nd == 3 ? mg : 0); cc++)
              for (bb = -mg; bb <= mg; bb++)
                for (a = -mg; a <= mg; a++) u[m++] = q[QI(o, p, i + a, j + bb, k + cc)] - qc;
            for (ds = 0; ds < 2 * nd; ds++) {
              const double* w = &o->ugp[(size_t)ds * dgp];
              double acc = 0.0;
              for (m = 0; m < dgp; m++) acc = acc + w[m] * u[m];
              fv[ds] = acc + qc;
            }
            fl = ((fv[0] - q[QI(o, p, i + 1, j, k)]) * (fv[0] - qc) > o->c.eps) ||
                 ((fv[1] - q[QI(o, p, i - 1, j, k)]) * (fv[1] - qc) > o->c.eps) ||
                 ((fv[2] - q[QI(o, p, i, j + 1, k)]) * (fv[2] - qc) > o->c.eps) ||
                 ((fv[3] - q[QI(o, p, i, j - 1, k)]) * (fv[3] - qc) > o->c.eps);
            if (nd == 3)
              fl = fl || ((fv[4] - q[QI(o, p, i, j, k + 1)]) * (fv[4] - qc) > o->c.eps) ||
                   ((fv[5] - q[QI(o, p, i, j, k - 1)]) * (fv[5] - qc) > o->c.eps);
            if (fl) {
              o->mgp_fallbacks++;
              o->mgp_hash = o->mgp_hash * 1000003ull +
                            (unsigned long long)(1 + i + 64 * (j + 64 * (k + 64 * p)));
              dql = qc - q[QI(o, p, i - 1, j, k)];
              dqr = q[QI(o, p, i + 1, j, k)] - qc;
              dq[0] = lim(o->c.limiter, 0.5, dql, dqr);
              dql = qc - q[QI(o, p, i, j - 1, k)];
              dqr = q[QI(o, p, i, j + 1, k)] - qc;
              dq[1] = lim(o->c.limiter, 0.5, dql, dqr);
              if (nd == 3) {
                dql = qc - q[QI(o, p, i, j, k - 1)];
                dqr = q[QI(o, p, i, j, k + 1)] - qc;
                dq[2] = lim(o->c.limiter, 0.5, dql, dqr);
              }
              if (positive) {
                if (qc > 0.0) {
                  int it;
                  for (it = 0; it < OMAXHALV; it++) {
                    double sa = fabs(dq[0]) + fabs(dq[1]);
                    if (nd == 3) sa += fabs(dq[2]);
                    if (!(qc <= sa)) break;
                    dq[0] *= 0.5; dq[1] *= 0.5; dq[2] *= 0.5;
                  }
                } else {
                  o->npos_events++;
                  dq[0] = dq[1] = dq[2] = 0.0;
                }
              }
              for (ds = 0; ds < nd; ds++) { fv[2 * ds] = qc + dq[ds]; fv[2 * ds + 1] = qc - dq[ds]; }
            }
            QS(0, 0, i, j, k, p)     = fv[0];
            QS(0, 1, i - 1, j, k, p) = fv[1];
            QS(1, 0, i, j, k, p)     = fv[2];
            QS(1, 1, i, j - 1, k, p) = fv[3];
            if (nd == 3) {
              QS(2, 0, i, j, k, p)     = fv[4];
              QS(2, 1, i, j, k - 1, p) = fv[5];
            }
          }
    } else if (o->c.reconstruction == AMUN_REC_TVD) {
      /* interfaces_tvd, interpolations.F90:729-848 */
      for (k = o->kbl; k <= o->keu; k++)
        for (j = o->jbl; j <= o->jeu; j++)
          for (i = o->ibl; i <= o->ieu; i++) {
            double dql, dqr, dq[3] = {0.0, 0.0, 0.0};
            double qc = q[QI(o, p, i, j, k)];
            dql = qc - q[QI(o, p, i - 1, j, k)];
            dqr = q[QI(o, p, i + 1, j, k)] - qc;
            dq[0] = lim(o->c.limiter, 0.5, dql, dqr);
            dql = qc - q[QI(o, p, i, j - 1, k)];
            dqr = q[QI(o, p, i, j + 1, k)] - qc;
            dq[1] = lim(o->c.limiter, 0.5, dql, dqr);
            if (nd == 3) {
              dql = qc - q[QI(o, p, i, j, k - 1)];
              dqr = q[QI(o, p, i, j, k + 1)] - qc;
              dq[2] = lim(o->c.limiter, 0.5, dql, dqr);
            }
            if (positive) {
              if (qc > 0.0) {
                int it;
                for (it = 0; it < OMAXHALV; it++) {
                  double sa = fabs(dq[0]) + fabs(dq[1]);
                  if (nd == 3) sa += fabs(dq[2]);
                  if (!(qc <= sa)) break;
                  dq[0] *= 0.5; dq[1] *= 0.5; dq[2] *= 0.5;
                }
              } else {
                o->npos_events++;
                dq[0] = dq[1] = dq[2] = 0.0;
              }
            }
            QS(0, 0, i, j, k, p)     = qc + dq[0];
            QS(0, 1, i - 1, j, k, p) = qc - dq[0];
            QS(1, 0, i, j, k, p)     = qc + dq[1];
            QS(1, 1, i, j - 1, k, p) = qc - dq[1];
            if (nd == 3) {
              QS(2, 0, i, j, k, p)     = qc + dq[2];
              QS(2, 1, i, j, k - 1, p) = qc - dq[2];
            }
          }
    } else {
      /* interfaces_dir, interpolations.F90:867-959 */
      int nmax = im > jm ? im : jm;
      double *fp, *fl, *fr, *w1, *w2, *w3;
      if (km > nmax) nmax = km;
      fp = (double*)malloc(8 * (size_t)(nmax + 4) * sizeof(double));
      fl = fp + nmax; fr = fl + nmax; w1 = fr + nmax; w2 = w1 + nmax; w3 = w2 + nmax;
      for (k = o->kbl; k <= o->keu; k++) {
        for (j = o->jbl; j <= o->jeu; j++) {
          for (i = 0; i < im; i++) fp[i] = q[QI(o, p, i, j, k)];
          recon_dir(o, im, o->h[0], fp, fl, fr, w1, w2, w3);
          if (positive && o->c.fix_positivity) fix_positivity(im, fp, fl, fr);
          for (i = 0; i < im; i++) { QS(0, 0, i, j, k, p) = fl[i]; QS(0, 1, i, j, k, p) = fr[i]; }
        }
        for (i = o->ibl; i <= o->ieu; i++) {
          for (j = 0; j < jm; j++) fp[j] = q[QI(o, p, i, j, k)];
          recon_dir(o, jm, o->h[1], fp, fl, fr, w1, w2, w3);
          if (positive && o->c.fix_positivity) fix_positivity(jm, fp, fl, fr);
          for (j = 0; j < jm; j++) { QS(1, 0, i, j, k, p) = fl[j]; QS(1, 1, i, j, k, p) = fr[j]; }
        }
      }
      if (nd == 3) {
        for (j = o->jbl; j <= o->jeu; j++)
          for (i = o->ibl; i <= o->ieu; i++) {
            for (k = 0; k < km; k++) fp[k] = q[QI(o, p, i, j, k)];
            recon_dir(o, km, o->h[2], fp, fl, fr, w1, w2, w3);
            if (positive && o->c.fix_positivity) fix_positivity(km, fp, fl, fr);
            for (k = 0; k < km; k++) { QS(2, 0, i, j, k, p) = fl[k]; QS(2, 1, i, j, k, p) = fr[k]; }
          }
      }
      free(fp);
    }
    if (o->c.mlp_limiting) mlp_limiting(o, q, p, qs);
  }
#undef QS
}

/* ------------------------------------------------------------------------ */
/* polynomial roots, algebra.F90:86-1085.  The operation order of every     */
/* expression follows the Fortran source as compiled (left to right;        */
/* x**n = ((x*x)*x)..., a**(1/3) = pow(a, 1/3)).                             */
/* ------------------------------------------------------------------------ */
static const double alg_pi2 = 6.28318530717958647692528676655900;  /* constants.F90:40 */
static const double alg_c0  = 1.0 / 3.0;

static double fsign(double a, double b) { return copysign(fabs(a), b); }

/* sort, algebra.F90:1040-1085: insertion sort of x[0..n-1], ascending */
static void alg_sort(int n, double* x)
{
  int k, l;
  double y;
  if (n < 2) return;
  for (k = 1; k < n; k++) {
    y = x[k];
    for (l = k - 1; l >= 0; l--) {
      if (x[l] <= y) break;
      x[l + 1] = x[l];
    }
    x[l + 1] = y;
  }
}

/* quadratic, algebra.F90:86-242: a[0] + a[1] x + a[2] x^2 = 0 */
static int alg_quadratic(const double* a, double* x)
{
  double b[3], bh, dl, dr, tm, mx;
  int nr;
  mx = fmax_r(fmax_r(fabs(a[0]), fabs(a[1])), fabs(a[2]));
  b[0] = a[0] / mx; b[1] = a[1] / mx; b[2] = a[2] / mx;
  if (b[2] != 0.0) {
    if (b[0] != 0.0) {
      bh = 0.5 * b[1];
      dl = bh * bh - b[2] * b[0];
      if (dl > 0.0) {
        dr = sqrt(dl);
        if (b[1] > 0.0) {
          tm   = -bh - dr;
          x[0] = tm / b[2];
          x[1] = b[0] / tm;
        } else if (b[1] < 0.0) {
          tm   = -bh + dr;
          x[0] = b[0] / tm;
          x[1] = tm / b[2];
        } else {
          x[1] = dr / b[2];
          x[0] = -x[1];
        }
        nr = 2;
      } else if (dl == 0.0) {
        x[0] = -0.5 * b[1] / b[2];
        x[1] = x[0];
        nr = 2;
      } else {
        x[0] = 0.0; x[1] = 0.0;
        nr = 0;
      }
    } else {
      tm = -b[1] / b[2];
      if (tm >= 0.0) { x[0] = 0.0; x[1] = tm; }
      else           { x[0] = tm;  x[1] = 0.0; }
      nr = 2;
    }
  } else {
    x[1] = 0.0;
    if (b[1] != 0.0) { x[0] = -b[0] / b[1]; nr = 1; }
    else             { x[0] = 0.0;          nr = 0; }
  }
  return nr;
}

/* quadratic_normalized, algebra.F90:262-372: a[0] + a[1] x + x^2 = 0 */
static int alg_quadratic_normalized(const double* a, double* x)
{
  double bh, dl, dr, tm;
  int nr;
  if (a[0] != 0.0) {
    bh = 0.5 * a[1];
    dl = bh * bh - a[0];
    if (dl > 0.0) {
      dr = sqrt(dl);
      if (a[1] > 0.0) {
        tm   = bh + dr;
        x[0] = -a[0] / tm;
        x[1] = -tm;
      } else if (a[1] < 0.0) {
        tm   = bh - dr;
        x[0] = -tm;
        x[1] = -a[0] / tm;
      } else {
        x[0] =  dr;
        x[1] = -dr;
      }
      nr = 2;
    } else if (dl == 0.0) {
      x[0] = -bh; x[1] = -bh;
      nr = 2;
    } else {
      x[0] = 0.0; x[1] = 0.0;
      nr = 0;
    }
  } else {
    x[0] = 0.0;
    x[1] = -a[1];
    nr = 2;
  }
  return nr;
}

/* cubic, algebra.F90:398-614: a[0] + a[1] x + a[2] x^2 + a[3] x^3 = 0 */
static int alg_cubic(const double* a, double* x)
{
  double ac, bc, qc, cc, dc, ec, rc, th, cs, a2, b3;
  int nr;
  if (a[3] != 0.0) {
    if (a[0] != 0.0) {
      ac = (2.0 * (a[2] * a[2] * a[2]) - 9.0 * a[3] * a[2] * a[1])
           + 27.0 * (a[3] * a[3]) * a[0];
      bc = a[2] * a[2] - 3.0 * a[3] * a[1];
      a2 = ac * ac;
      b3 = bc * bc * bc;
      qc = a2 - 4.0 * b3;
      if (qc > 0.0) {
        cc = 0.5 * (sqrt(qc) + ac);
        dc = fsign(1.0, cc) * pow(fabs(cc), alg_c0);
        if (cc == 0.0) {
          ec = b3 / a2;
          ec = 2.0 * ec * (1.0 / sqrt(1.0 - 4.0 * ec) + 1.0 / ac);
        } else {
          ec = b3 / cc;
        }
        ec = fsign(1.0, ec) * pow(fabs(ec), alg_c0);
        x[0] = -(alg_c0 * ((a[2] + dc) + ec)) / a[3];
        x[1] = 0.0;
        x[2] = 0.0;
        nr = 1;
      } else if (qc == 0.0) {
        if (bc != 0.0) {
          x[0] = 0.5 * (9.0 * a[3] * a[0] - a[2] * a[1]) / bc;
          x[1] = x[0];
          x[2] = ((-(9.0 * a[3] * a[0]) + 4.0 * a[2] * a[1])
                  - (a[2] * a[2] * a[2]) / a[3]) / bc;
        } else {
          x[0] = x[1] = x[2] = -(alg_c0 * a[2]) / a[3];
        }
        nr = 3;
      } else {
        double fb;
        th = 3.0 * a[3];
        qc = bc / (th * th);
        rc = 0.5 * ac / (th * th * th);
        fb = fabs(bc);
        cs = fsign(0.5, ac * bc) * sqrt((ac * ac) / (fb * fb * fb));
        if (cs <= 1.0) {
          th = acos(cs);
          ac = -2.0 * sqrt(qc);
          bc = -(alg_c0 * (a[2] / a[3]));
          x[0] = ac * cos(alg_c0 * th) + bc;
          x[1] = ac * cos(alg_c0 * (th + alg_pi2)) + bc;
          x[2] = ac * cos(alg_c0 * (th - alg_pi2)) + bc;
          nr = 3;
        } else {
          ac = -(fsign(1.0, rc) * pow(fabs(rc) + sqrt(rc * rc - qc * qc * qc), alg_c0));
          if (ac != 0.0) bc = qc / ac;
          else           bc = 0.0;
          x[0] = (ac + bc) - alg_c0 * (a[2] / a[3]);
          x[1] = 0.0;
          x[2] = 0.0;
          nr = 1;
        }
      }
    } else {
      x[2] = 0.0;
      nr = alg_quadratic(a + 1, x) + 1;
    }
  } else {
    x[2] = 0.0;
    nr = alg_quadratic(a, x);
  }
  alg_sort(nr, x);
  return nr;
}

/* cubic_normalized, algebra.F90:636-796: a[0] + a[1] x + a[2] x^2 + x^3 = 0 */
static int alg_cubic_normalized(const double* a, double* x)
{
  const double c1 = 1.0 / 2.7e+01, c2 = 1.0 / 4.0;
  double ac, bc, qc, a3, b2, aa, bb, sq, cs, ph;
  int nr;
  if (a[0] != 0.0) {
    ac = alg_c0 * (a[2] * a[2] - 3.0 * a[1]);
    bc = c1 * ((2.0 * (a[2] * a[2] * a[2]) - 9.0 * a[2] * a[1]) + 2.7e+01 * a[0]);
    a3 = c1 * (ac * ac * ac);
    b2 = c2 * (bc * bc);
    qc = b2 - a3;
    if (qc > 0.0) {
      sq = sqrt(qc);
      ph = -0.5 * bc + sq;
      aa = fsign(1.0, ph) * pow(fabs(ph), alg_c0);
      ph = -0.5 * bc - sq;
      bb = fsign(1.0, ph) * pow(fabs(ph), alg_c0);
      x[0] = (aa + bb) - (alg_c0 * a[2]);
      x[1] = 0.0;
      x[2] = 0.0;
      nr = 1;
    } else if (qc < 0.0) {
      if (bc > 0.0)      cs = -sqrt(b2 / a3);
      else if (bc < 0.0) cs =  sqrt(b2 / a3);
      else               cs = 0.0;
      ph = acos(cs);
      aa = 2.0 * sqrt(alg_c0 * ac);
      x[0] = aa * cos(alg_c0 * (ph - alg_pi2));
      x[1] = aa * cos(alg_c0 * ph);
      x[2] = aa * cos(alg_c0 * (ph + alg_pi2));
      x[0] -= alg_c0 * a[2]; x[1] -= alg_c0 * a[2]; x[2] -= alg_c0 * a[2];
      nr = 3;
    } else {
      if (bc > 0.0) {
        aa = sqrt(alg_c0 * ac);
        x[0] = -2.0 * aa; x[1] = aa; x[2] = aa;
      } else if (bc < 0.0) {
        aa = sqrt(alg_c0 * ac);
        x[0] = 2.0 * aa; x[1] = -aa; x[2] = -aa;
      } else {
        x[0] = x[1] = x[2] = 0.0;
      }
      x[0] -= alg_c0 * a[2]; x[1] -= alg_c0 * a[2]; x[2] -= alg_c0 * a[2];
      nr = 3;
    }
  } else {
    x[2] = 0.0;
    nr = alg_quadratic_normalized(a + 1, x) + 1;
  }
  alg_sort(nr, x);
  return nr;
}

/* quartic, algebra.F90:820-1017 (Hacke's method through the resolvent
 * cubic; its smallest root z(1) is taken, cubic sorting ascending):
 * a[0] + a[1] x + ... + a[4] x^4 = 0 */
static int alg_quartic(const double* a, double* x)
{
  double pc, qc, rc, zm, k2, l2, kc, lc, a42, a52;
  double b[4], y[2], c[3], z[3];
  int nr, m, n;
  if (a[4] != 0.0) {
    if (a[0] != 0.0) {
      a42 = a[3] * a[3];
      a52 = a[4] * a[4];
      pc = (-(3.0 * a42) + 8.0 * a[4] * a[2]) / (8.0 * a52);
      qc = ((a[3] * a[3] * a[3] - 4.0 * a[4] * a[3] * a[2]) + 8.0 * a52 * a[1])
           / (8.0 * (a[4] * a[4] * a[4]));
      rc = (((-(3.0 * (a42 * a[3] * a[3])) + 16.0 * a[4] * a42 * a[2])
             - 64.0 * a52 * a[3] * a[1]) + 256.0 * (a[4] * a[4] * a[4]) * a[0])
           / (256.0 * (a52 * a[4] * a[4]));
      b[3] =  8.0;
      b[2] = -4.0 * pc;
      b[1] = -8.0 * rc;
      b[0] =  4.0 * pc * rc - qc * qc;
      nr = alg_cubic(b, z);
      if (nr > 0) {
        zm = z[0];
        k2 = 2.0 * zm - pc;
        l2 = zm * zm - rc;
        if (k2 > 0.0 || l2 > 0.0) {
          if (k2 > l2) {
            kc = sqrt(k2);
            lc = 0.5 * qc / kc;
          } else if (l2 > k2) {
            lc = sqrt(l2);
            kc = 0.5 * qc / lc;
          } else {
            kc = sqrt(k2);
            lc = sqrt(l2);
          }
          c[2] = 1.0; c[1] = -kc; c[0] = zm + lc;
          m = alg_quadratic(c, y);
          for (n = 0; n < m; n++) x[n] = y[n] - 0.25 * a[3] / a[4];
          nr = m;
          c[2] = 1.0; c[1] = kc; c[0] = zm - lc;
          m = alg_quadratic(c, y);
          for (n = 0; n < m; n++) x[nr + n] = y[n] - 0.25 * a[3] / a[4];
          nr = nr + m;
          for (n = nr; n < 4; n++) x[n] = 0.0;
        } else {
          x[0] = x[1] = x[2] = x[3] = 0.0;
          nr = 0;
        }
      } else {
        x[0] = x[1] = x[2] = x[3] = 0.0;
      }
    } else {
      x[3] = 0.0;
      nr = alg_cubic(a + 1, x) + 1;
    }
  } else {
    x[3] = 0.0;
    nr = alg_cubic(a, x);
  }
  alg_sort(nr, x);
  return nr;
}

/* ------------------------------------------------------------------------ */
/* special-relativistic MHD, adiabatic (equations.F90:4590-5163)              */
/* ------------------------------------------------------------------------ */
/* prim2cons_srmhd_adi, equations.F90:4590-4653 */
static void prim2cons_sr(const ocfg* o, const double* q, double* u)
{
  double vv, bb, vb, vm, vs, ww, wt;
  vv = (q[o->ivx] * q[o->ivx] + q[o->ivy] * q[o->ivy]) + q[o->ivz] * q[o->ivz];
  bb = (q[o->ibx] * q[o->ibx] + q[o->iby] * q[o->iby]) + q[o->ibz] * q[o->ibz];
  vb = (q[o->ivx] * q[o->ibx] + q[o->ivy] * q[o->iby]) + q[o->ivz] * q[o->ibz];
  vm = 1.0 - vv;
  vs = sqrt(vm);
  ww = (q[o->idn] + q[o->ipr] / o->gammaxi) / vm;
  wt = ww + bb;
  u[o->idn] = q[o->idn] / vs;
  u[o->imx] = wt * q[o->ivx] - vb * q[o->ibx];
  u[o->imy] = wt * q[o->ivy] - vb * q[o->iby];
  u[o->imz] = wt * q[o->ivz] - vb * q[o->ibz];
  u[o->ibx] = q[o->ibx];
  u[o->iby] = q[o->iby];
  u[o->ibz] = q[o->ibz];
  u[o->ibp] = q[o->ibp];
  u[o->ien] = ((wt - q[o->ipr]) - u[o->idn]) - 0.5 * (vm * bb + vb * vb);
}

/* nr_pressure_srmhd_adi_1d, equations.F90:5314-5350 */
static void nr_pressure(double mm, double bb, double mb, double en, double w,
                        double* p, double* dp)
{
  double wt = w + bb, wd = wt * wt, ss = mb * mb, fn = (mm * bb - ss) / wd;
  *p = (w - en) + 0.5 * (bb + fn);
  if (dp) *dp = 1.0 - fn / wt;
}

/* nr_velocity_srmhd_adi_1d, equations.F90:5249-5290 (without dv) */
static double nr_velocity(double mm, double bb, double mb, double w)
{
  double ss = mb * mb, ww = w * w, wt = w + bb, wt2 = wt * wt;
  return (mm * ww + (w + wt) * ss) / (ww * wt2);
}

/* nr_function_srmhd_adi_1d, equations.F90:5371-5430 */
static void nr_function(const ocfg* o, double mm, double bb, double mb, double en,
                        double dn, double w, double* f, double* df)
{
  double pr, dp, gm2, gm, dg, ww, wt, wd, ss, sw, ws, fn, dv, ds;
  ww  = w * w;
  wt  = w + bb;
  wd  = wt * wt;
  ss  = mb * mb;
  sw  = ss / ww;
  ws  = (w + wt) * sw;
  fn  = (mm * bb - ss) / wd;
  dv  = (wd - mm) - ws;
  ds  = sqrt(dv);
  gm2 = wd / dv;
  gm  = wt / ds;
  pr  = o->gammaxi * (w - gm * dn) / gm2;
  *f  = ((w - pr) - en) + 0.5 * (bb + fn);
  if (df) {
    dg  = (1.0 - wt * ((wt - sw) + ws / w) / dv) / ds;
    dp  = o->gammaxi * (1.0 - (2.0 * w / gm - dn) * dg) / gm2;
    *df = (1.0 - dp) - fn / wt;
  }
}

/* nr_initial_brackets_srmhd_adi, equations.F90:5459-5635 */
static int nr_initial_brackets(const ocfg* o, double mm, double bb, double mb,
                               double en, double dn, double* wl, double* wu,
                               double* wc, double* fl, double* fu)
{
  const double tol = o->c.tolerance, pmin = o->c.pmin;
  const int nrmax = o->c.nr_maxit;
  double dd, ss, ec, f, df, dw, err, a[5], x[4];
  int it, nr, keep;
  dd = dn * dn;
  ss = mb * mb;
  ec = en + pmin;
  *wu = en + pmin;
  /* lower bracket: largest root of the cubic and quartic positivity
   * conditions of the pressure and of 1 - v^2 */
  a[2] = 2.5 * bb - ec;
  a[1] = 2.0 * (bb - ec) * bb;
  a[0] = 0.5 * (((bb - 2.0 * ec) * bb + mm) * bb - ss);
  nr = alg_cubic_normalized(a, x);
  if (nr > 0) {
    *wl = x[nr - 1];
    a[4] = 1.0;
    a[3] = 2.0 * bb;
    a[2] = (bb * bb - dd) - mm;
    a[1] = -(2.0 * (ss + dd * bb));
    a[0] = -(ss + dd * bb) * bb;
    nr = alg_quartic(a, x);
    if (nr > 0) *wl = fmax_r(*wl, x[nr - 1]);
  } else {
    keep = 1;
    it = nrmax;
    *wl = *wu;
    while (keep) {
      nr_pressure(mm, bb, mb, ec, *wl, &f, &df);
      dw  = f / df;
      *wl = *wl - dw;
      err = fabs(dw / *wl);
      it  = it - 1;
      keep = (err > tol) && it > 0;
    }
  }
  nr_function(o, mm, bb, mb, en, dn, *wl, fl, NULL);
  if (*fl <= 0.0) {
    if (*wu <= *wl) *wu = 2.0 * *wl;
    nr_function(o, mm, bb, mb, en, dn, *wu, fu, NULL);
    it = nrmax;
    keep = *fl * *fu > 0.0;
    while (keep) {
      it  = it - 1;
      *wl = *wu;
      *fl = *fu;
      *wu = 2.0 * *wu;
      nr_function(o, mm, bb, mb, en, dn, *wu, fu, NULL);
      keep = (*fl * *fu > 0.0) && it > 0;
    }
    if (it > 0 && *fu >= 0.0) {
      *wc = *wl - *fl * (*wu - *wl) / (*fu - *fl);
      return 1;
    }
    return 0;
  }
  return 0;
}

/* nr_iterate_srmhd_adi_1dw, equations.F90:5665-5788; returns info, and
 * sets *noconv when the iteration stopped above the tolerance */
static int nr_iterate_1dw(const ocfg* o, double mm, double bb, double mb,
                          double en, double dn, double* w, double* vv,
                          int* noconv)
{
  const double tol = o->c.tolerance;
  double wl, wu, fl, fu, f, df, dw, err = 0.0;
  int it, cn, keep, info;
  *noconv = 0;
  info = nr_initial_brackets(o, mm, bb, mb, en, dn, &wl, &wu, w, &fl, &fu);
  if (info) {
    keep = 1;
    it = o->c.nr_maxit;
    cn = o->c.nr_extra;
    while (keep) {
      nr_function(o, mm, bb, mb, en, dn, *w, &f, &df);
      if (f > fl && f < 0.0) { wl = *w; fl = f; }
      if (f < fu && f > 0.0) { wu = *w; fu = f; }
      dw  = f / df;
      *w  = *w - dw;
      err = fabs(dw / *w);
      if (err < tol) {
        keep = cn > 0;
        cn = cn - 1;
      } else {
        keep = it > 0;
      }
      if (*w < wl || *w > wu) *w = 0.5 * (wl + wu);
      it = it - 1;
    }
    if (err >= tol) *noconv = 1;
    *vv = nr_velocity(mm, bb, mb, *w);
  }
  return info;
}

/* the Newton step of the two-variable solvers: the 2x2 system
 * (dfw dfx; dgw dgx) (dw, dx) = (f, g) solved by its inverse, as
 * nr_iterate_srmhd_adi_2dwv / _2dwu write it */
static void nr_solve2(double f, double g, double dfw, double dfx, double dgw, double dgx,
                      double* dw, double* dx)
{
  const double det = dfw * dgx - dfx * dgw;
  const double jfw = dgx / det, jgw = -dfx / det, jfx = -dgw / det, jgx = dfw / det;
  *dw = f * jfw + g * jgw;
  *dx = f * jfx + g * jgx;
}

/* nr_iterate_srmhd_adi_2dwv (equations.F90:5819-6000) and _2dwu (6020-6210):
 * Newton on (W, v^2) or (W, u^2 = v^2 / (1 - v^2)) from the bracketed
 * start of 1D(W).  Without a bracket the lower one is the solution (and
 * counts as a bracketing event); an iterate below it or with an unphysical
 * speed fails (info 0). */
static int nr_iterate_2d(ocfg* o, int wu_form, double mm, double bb, double mb, double en,
                         double dn, double* w, double* vv, int* noconv)
{
  const double tol = o->c.tolerance, gx = o->gammaxi;
  double wl, wu, fl, fu, f, g, dfw, dfx, dgw, dgx, dw, dx, err = 0.0, x;
  int it, cn, keep;
  *noconv = 0;
  if (!nr_initial_brackets(o, mm, bb, mb, en, dn, &wl, &wu, w, &fl, &fu)) {
    o->c2p_fail++;
    *w = wl;
    *vv = nr_velocity(mm, bb, mb, *w);
    return 1;
  }
  *vv = nr_velocity(mm, bb, mb, *w);
  x = wu_form ? *vv / (1.0 - *vv) : *vv;
  keep = 1;
  it = o->c.nr_maxit;
  cn = o->c.nr_extra;
  while (keep) {
    if (!wu_form) {
      const double mw = (mb / *w) * (mb / *w), wt = *w + bb, wt2 = wt * wt;
      const double vm = 1.0 - x, vs = sqrt(vm);
      f = ((*w - en) - gx * (*w * vm - dn * vs)) + 0.5 * ((1.0 + x) * bb - mw);
      g = (x * wt2 - (wt + *w) * mw) - mm;
      dfw = (1.0 - gx * vm) + mw / *w;
      dfx = -(gx * (0.5 * dn / vs - *w)) + 0.5 * bb;
      dgw = 2.0 * wt * (x + mw / *w);
      dgx = wt2;
    } else {
      const double up = 1.0 + x, gm = sqrt(up), ss = mb * mb, ww = *w * *w;
      const double wt = *w + bb, wd = *w - gm * dn, wp = wt / up;
      f = ((*w - en) - gx * wd / up) + 0.5 * (bb * (1.0 + x / up) - ss / ww);
      g = (wp * wt * x - (*w + wt) * ss / ww) - mm;
      dfw = (1.0 - gx / up) + ss / ww / *w;
      dfx = 0.5 * (gx * (*w + wd) + bb) / (up * up);
      dgw = 2.0 * wt * (x / up + ss / ww / *w);
      dgx = wp * wp;
    }
    nr_solve2(f, g, dfw, dfx, dgw, dgx, &dw, &dx);
    *w = *w - dw;
    x = x - dx;
    if (*w < wl) return 0;
    if (wu_form ? x < 0.0 : (x < 0.0 || x >= 1.0)) return 0;
    err = fmax_r(fabs(dw / *w), fabs(dx));
    if (err < tol) {
      keep = cn > 0;
      cn = cn - 1;
    } else {
      keep = it > 0;
    }
    it = it - 1;
  }
  *vv = wu_form ? x / (1.0 + x) : x;
  if (err >= tol) *noconv = 1;
  return 1;
}

/* cons2prim_srmhd_adi, equations.F90:4671-4790.  On failure only the
 * pressure is written (0), the other entries of q keep what they held. */
static void cons2prim_sr(ocfg* o, const double* u, double* q)
{
  double mm, mb, bb, en, dn, w = 0.0, wt, vv = 0.0, vm, vs, fc;
  int info, noconv;
  mm = (u[o->imx] * u[o->imx] + u[o->imy] * u[o->imy]) + u[o->imz] * u[o->imz];
  mb = (u[o->imx] * u[o->ibx] + u[o->imy] * u[o->iby]) + u[o->imz] * u[o->ibz];
  bb = (u[o->ibx] * u[o->ibx] + u[o->iby] * u[o->iby]) + u[o->ibz] * u[o->ibz];
  en = u[o->ien] + u[o->idn];
  dn = u[o->idn];
  if (o->c.primitive_solver == AMUN_C2P_2DWV || o->c.primitive_solver == AMUN_C2P_2DWU)
    info = nr_iterate_2d(o, o->c.primitive_solver == AMUN_C2P_2DWU, mm, bb, mb, en, dn, &w, &vv,
                         &noconv);
  else
    info = nr_iterate_1dw(o, mm, bb, mb, en, dn, &w, &vv, &noconv);
  if (noconv) o->c2p_noconv++;
  if (info) {
    vm = 1.0 - vv;
    vs = sqrt(vm);
    wt = w + bb;
    fc = mb / w;
    q[o->idn] = dn * vs;
    q[o->ivx] = (u[o->imx] + fc * u[o->ibx]) / wt;
    q[o->ivy] = (u[o->imy] + fc * u[o->iby]) / wt;
    q[o->ivz] = (u[o->imz] + fc * u[o->ibz]) / wt;
    q[o->ibx] = u[o->ibx];
    q[o->iby] = u[o->iby];
    q[o->ibz] = u[o->ibz];
    q[o->ibp] = u[o->ibp];
    q[o->ipr] = (w - en) + 0.5 * (bb + (bb * mm - mb * mb) / (wt * wt));
    if (q[o->ipr] <= 0.0) {
      o->c2p_pneg++;
      q[o->ipr] = 0.0;
    }
  } else {
    o->c2p_fail++;
    q[o->ipr] = 0.0;
  }
}


/* ---------------------------------------------------------------------- */
/* the special-relativistic HD system (equations.F90:3478-4566)            */
/* ---------------------------------------------------------------------- */
/* prim2cons_srhd_adi, equations.F90:3478-3532 */
static void prim2cons_srhd(const ocfg* o, const double* q, double* u)
{
  double vv, vm, vs, ww;
  vv = (q[o->ivx] * q[o->ivx] + q[o->ivy] * q[o->ivy]) + q[o->ivz] * q[o->ivz];
  vm = 1.0 - vv;
  vs = sqrt(vm);
  ww = (q[o->idn] + q[o->ipr] / o->gammaxi) / vm;
  u[o->idn] = q[o->idn] / vs;
  u[o->imx] = ww * q[o->ivx];
  u[o->imy] = ww * q[o->ivy];
  u[o->imz] = ww * q[o->ivz];
  u[o->ien] = (ww - q[o->ipr]) - u[o->idn];
}

/* nr_function_srhd_adi_1d, equations.F90:3853-3881 */
static void nr_function_srhd(const ocfg* o, double mm, double en, double dn, double w,
                             double* f, double* df)
{
  const double gx = o->gammaxi;
  double vv, vm, vs;
  vv = mm / (w * w);
  vm = 1.0 - vv;
  vs = sqrt(vm);
  *f = ((1.0 - gx * vm) * w + gx * dn * vs) - en;
  if (df) *df = 1.0 - gx * (1.0 + (1.0 - dn / (vs * w)) * vv);
}

/* nr_iterate_srhd_adi_1dw, equations.F90:3911-4091: brackets from
 * sqrt(|m|^2 + D^2) + gammaxi pmin and E + pmin (doubled until the signs
 * differ), a secant start and bracketed Newton.  Where the lower bracket is
 * an exact root the reference leaves v^2 undefined; here it is |m|^2 / W^2
 * of that root. */
static int nr_iterate_srhd_1dw(const ocfg* o, double mm, double en, double dn, double* w,
                               double* vv, int* noconv)
{
  const double tol = o->c.tolerance, pmin = o->c.pmin, gx = o->gammaxi;
  double wl, wu, fl, fu, f, df, dw, err = 0.0;
  int it, cn, keep;
  *noconv = 0;
  wl = sqrt(mm + dn * dn);
  if (!(en > wl)) return 0;
  wl = wl + gx * pmin;
  wu = en + pmin;
  nr_function_srhd(o, mm, en, dn, wl, &fl, NULL);
  if (fl < 0.0) {
    if (wu <= wl) wu = 2.0 * wl;
    nr_function_srhd(o, mm, en, dn, wu, &fu, NULL);
    it = o->c.nr_maxit;
    keep = fl * fu > 0.0;
    while (keep) {
      it = it - 1;
      wl = wu;
      fl = fu;
      wu = 2.0 * wu;
      nr_function_srhd(o, mm, en, dn, wu, &fu, NULL);
      keep = (fl * fu > 0.0) && it > 0;
    }
    if (!(it > 0 && fu >= 0.0)) return 0;
    *w = wl - fl * (wu - wl) / (fu - fl);
    keep = 1;
    it = o->c.nr_maxit;
    cn = o->c.nr_extra;
    while (keep) {
      nr_function_srhd(o, mm, en, dn, *w, &f, &df);
      if (f > fl && f < 0.0) { wl = *w; fl = f; }
      if (f < fu && f > 0.0) { wu = *w; fu = f; }
      dw = f / df;
      *w = *w - dw;
      err = fabs(dw / *w);
      if (err < tol) {
        keep = cn > 0;
        cn = cn - 1;
      } else {
        keep = it > 0;
      }
      if (*w < wl || *w > wu) *w = 0.5 * (wl + wu);
      it = it - 1;
    }
    if (err >= tol) *noconv = 1;
    *vv = mm / (*w * *w);
    return 1;
  }
  if (fl == 0.0) {
    *w = wl;
    *vv = mm / (wl * wl);
    return 1;
  }
  return 0;
}

/* nr_iterate_srhd_adi_2dwv (equations.F90:4128-4330) and _2dwu (4360-4566):
 * brackets as the 1D(W) solver's (the lower one fixed, the upper doubled
 * while not above it, then both moved up while the signs agree), a secant
 * start, then Newton on (W, v^2) or (W, u^2); an iterate below the lower
 * bracket or with an unphysical speed fails */
static int nr_iterate_srhd_2d(const ocfg* o, int wu_form, double mm, double en, double dn,
                              double* w, double* vv, int* noconv)
{
  const double tol = o->c.tolerance, pmin = o->c.pmin, gx = o->gammaxi;
  double wl, wu, fl, fu, x, ww, f, g, dfw, dfx, dgw, dgx, det, jfw, jgw, jfx, jgx, dw, dx;
  double err = 0.0;
  int it, cn, keep;
  *noconv = 0;
  wl = sqrt(mm + dn * dn) + gx * pmin;
  wu = en + pmin;
  keep = wl >= wu;
  it = o->c.nr_maxit;
  while (keep) {
    wu = 2.0 * wu;
    it = it - 1;
    keep = (wl >= wu) && it > 0;
  }
  if (it <= 0) return 0;
  nr_function_srhd(o, mm, en, dn, wl, &fl, NULL);
  nr_function_srhd(o, mm, en, dn, wu, &fu, NULL);
  keep = fl * fu > 0.0;
  it = o->c.nr_maxit;
  while (keep) {
    wl = wu;
    fl = fu;
    wu = 2.0 * wu;
    nr_function_srhd(o, mm, en, dn, wu, &fu, NULL);
    it = it - 1;
    keep = (fl * fu > 0.0) && it > 0;
  }
  if (it <= 0) return 0;
  *w = wl - fl * (wu - wl) / (fu - fl);
  x = wu_form ? mm / (*w * *w - mm) : mm / (*w * *w);
  keep = 1;
  it = o->c.nr_maxit;
  cn = o->c.nr_extra;
  while (keep) {
    const double gd = gx * dn;
    ww = *w * *w;
    if (!wu_form) {
      const double vm = 1.0 - x, vs = sqrt(vm), gv = 1.0 - gx * vm;
      f = (gv * *w - en) + gd * vs;
      g = x * ww - mm;
      dfw = gv;
      dfx = gx * *w - 0.5 * gd / vs;
      dgw = 2.0 * x * *w;
      dgx = ww;
    } else {
      const double up = 1.0 + x, gm = sqrt(up);
      f = ((up - gx) * *w - up * en) + gm * gd;
      g = x * ww - up * mm;
      dfw = up - gx;
      dfx = (*w - en) + 0.5 * gd / gm;
      dgw = 2.0 * x * *w;
      dgx = ww - mm;
    }
    det = dfw * dgx - dfx * dgw;
    jfw = dgx / det;
    jgw = -dfx / det;
    jfx = -dgw / det;
    jgx = dfw / det;
    dw = f * jfw + g * jgw;
    dx = f * jfx + g * jgx;
    *w = *w - dw;
    x = x - dx;
    if (*w < wl) return 0;
    if (wu_form ? x < 0.0 : (x < 0.0 || x >= 1.0)) return 0;
    err = fmax_r(fabs(dw / *w), fabs(dx));
    if (err < tol) {
      keep = cn > 0;
      cn = cn - 1;
    } else {
      keep = it > 0;
    }
    it = it - 1;
  }
  *vv = wu_form ? x / (1.0 + x) : x;
  if (err >= tol) *noconv = 1;
  return 1;
}

/* cons2prim_srhd_adi, equations.F90:3550-3639.  On failure only the
 * pressure is written (0), the other entries of q keep what they held. */
static void cons2prim_srhd(ocfg* o, const double* u, double* q)
{
  double mm, en, dn, w = 0.0, vv = 0.0, vm, vs;
  int info, noconv;
  mm = (u[o->imx] * u[o->imx] + u[o->imy] * u[o->imy]) + u[o->imz] * u[o->imz];
  en = u[o->ien] + u[o->idn];
  dn = u[o->idn];
  if (o->c.primitive_solver == AMUN_C2P_2DWV || o->c.primitive_solver == AMUN_C2P_2DWU)
    info = nr_iterate_srhd_2d(o, o->c.primitive_solver == AMUN_C2P_2DWU, mm, en, dn, &w, &vv,
                              &noconv);
  else
    info = nr_iterate_srhd_1dw(o, mm, en, dn, &w, &vv, &noconv);
  if (noconv) o->c2p_noconv++;
  if (info) {
    vm = 1.0 - vv;
    vs = sqrt(vm);
    q[o->idn] = dn * vs;
    q[o->ivx] = u[o->imx] / w;
    q[o->ivy] = u[o->imy] / w;
    q[o->ivz] = u[o->imz] / w;
    q[o->ipr] = w - en;
  } else {
    o->c2p_fail++;
    q[o->ipr] = 0.0;
  }
}

/* fluxspeed_srhd_adi, equations.F90:3665-3731 */
static void fluxspeed_srhd(const ocfg* o, const double* q, const double* u, double* f,
                           double* cm, double* cp)
{
  double ww, c2, vv, ss, fc, cc;
  f[o->idn] = u[o->idn] * q[o->ivx];
  f[o->imx] = u[o->imx] * q[o->ivx] + q[o->ipr];
  f[o->imy] = u[o->imy] * q[o->ivx];
  f[o->imz] = u[o->imz] * q[o->ivx];
  f[o->ien] = u[o->imx] - f[o->idn];
  ww = q[o->idn] + q[o->ipr] / o->gammaxi;
  c2 = o->c.gamma * q[o->ipr] / ww;
  vv = (q[o->ivx] * q[o->ivx] + q[o->ivy] * q[o->ivy]) + q[o->ivz] * q[o->ivz];
  ss = c2 * (1.0 - vv) / (1.0 - c2);
  fc = 1.0 + ss;
  cc = sqrt(ss * (fc - q[o->ivx] * q[o->ivx]));
  *cm = (q[o->ivx] - cc) / fc;
  *cp = (q[o->ivx] + cc) / fc;
}

/* maxspeed_srhd_adi, equations.F90:3746-3820, for one cell */
static double maxspeed_srhd(const ocfg* o, const double* q)
{
  double vv, v, ww, c2, ss, fc, cc;
  vv = (q[o->ivx] * q[o->ivx] + q[o->ivy] * q[o->ivy]) + q[o->ivz] * q[o->ivz];
  v = sqrt(vv);
  ww = q[o->idn] + q[o->ipr] / o->gammaxi;
  c2 = o->c.gamma * q[o->ipr] / ww;
  ss = c2 * (1.0 - vv) / (1.0 - c2);
  fc = 1.0 + ss;
  cc = sqrt(ss * (fc - vv));
  return (v + cc) / fc;
}

/* fluxspeed_srmhd_adi, equations.F90:4821-5101 (Mignone & Bodo 2006 fluxes;
 * speeds from the quartic, quadratic or bi-quadratic characteristic
 * polynomial) */
static void fluxspeed_sr(const ocfg* o, const double* q, const double* u,
                         double* f, double* cm, double* cp)
{
  const double g = o->c.gamma;
  double vv, bb, vb, vm, bx, by, bz, b2, pm, pt;
  double rh, vs, c2, v1, v2, fa, fb, fc, fd, fe, ff, fg, cc, gn, rt, ca, zm, zp;
  double a[5], x[4];
  int nr, n;
  vv = (q[o->ivx] * q[o->ivx] + q[o->ivy] * q[o->ivy]) + q[o->ivz] * q[o->ivz];
  bb = (q[o->ibx] * q[o->ibx] + q[o->iby] * q[o->iby]) + q[o->ibz] * q[o->ibz];
  vb = (q[o->ivx] * q[o->ibx] + q[o->ivy] * q[o->iby]) + q[o->ivz] * q[o->ibz];
  vm = 1.0 - vv;
  bx = q[o->ibx] * vm + vb * q[o->ivx];
  by = q[o->iby] * vm + vb * q[o->ivy];
  bz = q[o->ibz] * vm + vb * q[o->ivz];
  b2 = bb * vm + vb * vb;
  pm = 0.5 * b2;
  pt = q[o->ipr] + pm;
  f[o->idn] = u[o->idn] * q[o->ivx];
  f[o->imx] = (u[o->imx] * q[o->ivx] - q[o->ibx] * bx) + pt;
  f[o->imy] = u[o->imy] * q[o->ivx] - q[o->ibx] * by;
  f[o->imz] = u[o->imz] * q[o->ivx] - q[o->ibx] * bz;
  f[o->ibx] = q[o->ibp];
  f[o->ibp] = o->cmax2 * q[o->ibx];
  f[o->iby] = q[o->ivx] * q[o->iby] - q[o->ibx] * q[o->ivy];
  f[o->ibz] = q[o->ivx] * q[o->ibz] - q[o->ibx] * q[o->ivz];
  f[o->ien] = u[o->imx] - f[o->idn];
  if (!cm) return;
  x[0] = x[1] = x[2] = x[3] = 0.0;
  if (vv > 0.0) {
    rh = q[o->idn] + q[o->ipr] / o->gammaxi;
    vs = sqrt(vm);
    if (q[o->ibx] != 0.0) {
      c2 = g * q[o->ipr] / rh;
      v1 = fabs(q[o->ivx]);
      v2 = v1 * v1;
      fa = rh * (1.0 - c2);
      fb = c2 * (rh - vb * vb) + b2;
      fc = fsign(1.0, q[o->ivx]) * q[o->ibx] * vs;
      fd = c2 * fc * fc;
      fe = 1.0 - v2;
      ff = c2 * vb * fc;
      fg = v1 * vs;
      a[4] = fa + fb * vm;
      a[3] = 2.0 * (ff * vm + fb * fg);
      a[2] = (-(fd * vm) + 4.0 * ff * fg) - fb * fe;
      a[1] = -(2.0 * (fd * fg + fe * ff));
      a[0] = fd * fe;
      nr = alg_quartic(a, x);
      for (n = 0; n < nr; n++)
        x[n] = fsign(1.0, q[o->ivx]) * (fabs(v1) + x[n] * vs);
    } else {
      c2 = g * q[o->ipr] / rh;
      cc = (1.0 - c2) / vm;
      gn = b2 - c2 * vb * vb;
      a[2] = rh * (c2 + cc) + gn;
      a[1] = -(2.0 * rh * cc * q[o->ivx]);
      a[0] = rh * (cc * (q[o->ivx] * q[o->ivx]) - c2) - gn;
      nr = alg_quadratic(a, x);
    }
  } else {
    rh = q[o->idn] + q[o->ipr] / o->gammaxi;
    vs = sqrt(vm);
    rt = rh + b2;
    c2 = g * q[o->ipr] / rh;
    ca = q[o->ibx] * vs + vb * q[o->ivx] / vs;
    ca = ca * ca;
    a[2] = 1.0;
    a[1] = -(((rh + ca) * c2 + b2) / rt);
    a[0] = c2 * ca / rt;
    nr = alg_quadratic(a, x);
    if (nr > 0) {
      zm = fmin_r(x[0], x[1]);
      zp = fmax_r(x[0], x[1]);
      if (zm >= 0.0) {
        zm = sqrt(zm);
        zp = sqrt(zp);
        x[0] = zp; x[1] = zm; x[2] = -zm; x[3] = -zp;
        nr = 4;
      } else if (zp >= 0.0) {
        zp = sqrt(zp);
        x[0] = zp; x[1] = -zp;
        nr = 2;
      } else {
        x[0] = x[1] = x[2] = x[3] = 0.0;
        nr = 0;
      }
    }
  }
  if (nr > 1) {
    double mn = x[0], mx = x[0];
    for (n = 1; n < nr; n++) { mn = fmin_r(mn, x[n]); mx = fmax_r(mx, x[n]); }
    *cm = mn;
    *cp = mx;
  } else {
    *cm = -1.0;
    *cp =  1.0;
  }
}

/* ------------------------------------------------------------------------ */
/* equation of state, equations.F90:1762-1958 (HD) and 2815-3045 (MHD)        */
/* ------------------------------------------------------------------------ */
static void prim2cons(const ocfg* o, const double* q, double* u)
{
  double ek, ei, em;
  if (o->srhd) { prim2cons_srhd(o, q, u); return; }
  if (o->sr) { prim2cons_sr(o, q, u); return; }
  u[o->idn] = q[o->idn];
  u[o->imx] = q[o->idn] * q[o->ivx];
  u[o->imy] = q[o->idn] * q[o->ivy];
  u[o->imz] = q[o->idn] * q[o->ivz];
  if (o->iso) {      /* prim2cons_hd_iso 1360, prim2cons_mhd_iso 2220 */
    if (o->mhd) {
      u[o->ibx] = q[o->ibx];
      u[o->iby] = q[o->iby];
      u[o->ibz] = q[o->ibz];
      u[o->ibp] = q[o->ibp];
    }
  } else if (o->c.eqsys == AMUN_EQSYS_HD) {
    ek = 0.5 * (u[o->imx] * q[o->ivx] + u[o->imy] * q[o->ivy] + u[o->imz] * q[o->ivz]);
    ei = o->gammam1i * q[o->ipr];
    u[o->ien] = ei + ek;
  } else {
    u[o->ibx] = q[o->ibx];
    u[o->iby] = q[o->iby];
    u[o->ibz] = q[o->ibz];
    u[o->ibp] = q[o->ibp];
    ei = o->gammam1i * q[o->ipr];
    ek = 0.5 * (u[o->imx] * q[o->ivx] + u[o->imy] * q[o->ivy] + u[o->imz] * q[o->ivz]);
    em = 0.5 * ((q[o->ibx] * q[o->ibx] + q[o->iby] * q[o->iby]) + q[o->ibz] * q[o->ibz]);
    u[o->ien] = (ei + ek) + em;
  }
}

static void cons2prim(ocfg* o, const double* u, double* q)
{
  double ek, ei, em;
  if (o->srhd) { cons2prim_srhd(o, u, q); return; }
  if (o->sr) { cons2prim_sr(o, u, q); return; }
  q[o->idn] = u[o->idn];
  q[o->ivx] = u[o->imx] / u[o->idn];
  q[o->ivy] = u[o->imy] / u[o->idn];
  q[o->ivz] = u[o->imz] / u[o->idn];
  if (o->iso) {      /* cons2prim_hd_iso 1421, cons2prim_mhd_iso 2285 */
    if (o->mhd) {
      q[o->ibx] = u[o->ibx];
      q[o->iby] = u[o->iby];
      q[o->ibz] = u[o->ibz];
      q[o->ibp] = u[o->ibp];
    }
  } else if (o->c.eqsys == AMUN_EQSYS_HD) {
    ek = 0.5 * (u[o->imx] * q[o->ivx] + u[o->imy] * q[o->ivy] + u[o->imz] * q[o->ivz]);
    ei = u[o->ien] - ek;
    q[o->ipr] = o->gammam1 * ei;
  } else {
    q[o->ibx] = u[o->ibx];
    q[o->iby] = u[o->iby];
    q[o->ibz] = u[o->ibz];
    q[o->ibp] = u[o->ibp];
    ek = 0.5 * (u[o->imx] * q[o->ivx] + u[o->imy] * q[o->ivy] + u[o->imz] * q[o->ivz]);
    em = 0.5 * ((q[o->ibx] * q[o->ibx] + q[o->iby] * q[o->iby]) + q[o->ibz] * q[o->ibz]);
    ei = u[o->ien] - (ek + em);
    q[o->ipr] = o->gammam1 * ei;
  }
}

static void fluxspeed(const ocfg* o, const double* q, const double* u,
                      double* f, double* cm, double* cp)
{
  const double g = o->c.gamma;
  if (o->srhd) { fluxspeed_srhd(o, q, u, f, cm, cp); return; }
  if (o->sr) { fluxspeed_sr(o, q, u, f, cm, cp); return; }
  if (o->iso && !o->mhd) {              /* fluxspeed_hd_iso 1484-1542 */
    f[o->idn] = u[o->imx];
    f[o->imx] = q[o->ivx] * u[o->imx];
    f[o->imy] = q[o->ivx] * u[o->imy];
    f[o->imz] = q[o->ivx] * u[o->imz];
    f[o->imx] = f[o->imx] + o->csnd2 * q[o->idn];
    *cm = q[o->ivx] - o->csnd;
    *cp = q[o->ivx] + o->csnd;
  } else if (o->iso) {                  /* fluxspeed_mhd_iso 2352-2435 */
    double bx2, by2, bz2, bb, pt, fa, fb, fc, cf;
    bx2 = q[o->ibx] * q[o->ibx];
    by2 = q[o->iby] * q[o->iby];
    bz2 = q[o->ibz] * q[o->ibz];
    bb  = bx2 + by2 + bz2;
    pt  = o->csnd2 * q[o->idn] + 0.5 * bb;
    f[o->idn] = u[o->imx];
    f[o->imx] = q[o->ivx] * u[o->imx] - bx2;
    f[o->imy] = q[o->ivx] * u[o->imy] - q[o->ibx] * q[o->iby];
    f[o->imz] = q[o->ivx] * u[o->imz] - q[o->ibx] * q[o->ibz];
    f[o->imx] = f[o->imx] + pt;
    f[o->ibx] = q[o->ibp];
    f[o->iby] = q[o->ivx] * q[o->iby] - q[o->ibx] * q[o->ivy];
    f[o->ibz] = q[o->ivx] * q[o->ibz] - q[o->ibx] * q[o->ivz];
    f[o->ibp] = o->cmax2 * q[o->ibx];
    fa = o->csnd2 * q[o->idn];
    fb = fa + bb;
    fc = fb * fb - 4.0 * fa * bx2;
    if (fc > 0.0) cf = sqrt(0.5 * (fb + sqrt(fc)) / q[o->idn]);
    else          cf = sqrt(0.5 * fb / q[o->idn]);
    *cm = q[o->ivx] - cf;
    *cp = q[o->ivx] + cf;
  } else if (o->c.eqsys == AMUN_EQSYS_HD) {
    double cs;
    f[o->idn] = u[o->imx];
    f[o->imx] = q[o->ivx] * u[o->imx];
    f[o->imy] = q[o->ivx] * u[o->imy];
    f[o->imz] = q[o->ivx] * u[o->imz];
    f[o->imx] = f[o->imx] + q[o->ipr];
    f[o->ien] = q[o->ivx] * (u[o->ien] + q[o->ipr]);
    cs  = sqrt(g * q[o->ipr] / q[o->idn]);
    *cm = q[o->ivx] - cs;
    *cp = q[o->ivx] + cs;
  } else {
    double bx2, by2, bz2, bb, vb, pt, fa, fb, fc, cf;
    bx2 = q[o->ibx] * q[o->ibx];
    by2 = q[o->iby] * q[o->iby];
    bz2 = q[o->ibz] * q[o->ibz];
    bb  = (bx2 + by2) + bz2;
    vb  = (q[o->ivx] * q[o->ibx] + q[o->ivy] * q[o->iby]) + q[o->ivz] * q[o->ibz];
    pt  = q[o->ipr] + 0.5 * bb;
    f[o->idn] = u[o->imx];
    f[o->imx] = q[o->ivx] * u[o->imx] - bx2;
    f[o->imy] = q[o->ivx] * u[o->imy] - q[o->ibx] * q[o->iby];
    f[o->imz] = q[o->ivx] * u[o->imz] - q[o->ibx] * q[o->ibz];
    f[o->imx] = f[o->imx] + pt;
    f[o->ibx] = q[o->ibp];
    f[o->iby] = q[o->ivx] * q[o->iby] - q[o->ibx] * q[o->ivy];
    f[o->ibz] = q[o->ivx] * q[o->ibz] - q[o->ibx] * q[o->ivz];
    f[o->ibp] = o->cmax2 * q[o->ibx];
    f[o->ien] = q[o->ivx] * (u[o->ien] + pt) - q[o->ibx] * vb;
    fa = g * q[o->ipr];
    fb = fa + bb;
    fc = fb * fb - 4.0 * fa * bx2;
    if (fc > 0.0) cf = sqrt(0.5 * (fb + sqrt(fc)) / q[o->idn]);
    else          cf = sqrt(0.5 * fb / q[o->idn]);
    *cm = q[o->ivx] - cf;
    *cp = q[o->ivx] + cf;
  }
}

/* maxspeed_hd_adi (equations.F90:1974-2044), maxspeed_mhd_adi (3060-3131);
 * maxspeed_srmhd_adi (5116-5163) returns the speed of light, 1 */
static double maxspeed_cell(const ocfg* o, const double* q)
{
  double vv, v, c, bb;
  if (o->srhd) return maxspeed_srhd(o, q);
  if (o->sr) return 1.0;
  vv = (q[o->ivx] * q[o->ivx] + q[o->ivy] * q[o->ivy]) + q[o->ivz] * q[o->ivz];
  v  = sqrt(vv);
  if (o->iso && !o->mhd) {              /* maxspeed_hd_iso 1546-1570 */
    c = o->csnd;
  } else if (o->iso) {                  /* maxspeed_mhd_iso 2438-2510 */
    bb = (q[o->ibx] * q[o->ibx] + q[o->iby] * q[o->iby]) + q[o->ibz] * q[o->ibz];
    c  = sqrt(o->csnd2 + bb / q[o->idn]);
  } else if (o->c.eqsys == AMUN_EQSYS_HD) {
    c = sqrt(o->c.gamma * q[o->ipr] / q[o->idn]);
  } else {
    bb = (q[o->ibx] * q[o->ibx] + q[o->iby] * q[o->iby]) + q[o->ibz] * q[o->ibz];
    c  = sqrt((o->c.gamma * q[o->ipr] + bb) / q[o->idn]);
  }
  return v + c;
}

/* ------------------------------------------------------------------------ */
/* Riemann solvers for one interface; ql, qr are modified (intent(inout))     */
/* ------------------------------------------------------------------------ */
#define NVMAX 9

/* HLL: riemann_hd_adi_hll (schemes.F90:1248-1341) and riemann_mhd_adi_hll
 * (3163-3280, with the GLM pre-step) */
static void riemann_hll(const ocfg* o, double* ql, double* qr, double* f)
{
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], wl[NVMAX], wr[NVMAX];
  double clm, clp, crm, crp, sl, sr, srml;
  int p, nv = o->nv;
  if (o->mhd) {
    double bx = 0.5 * ((qr[o->ibx] + ql[o->ibx]) - (qr[o->ibp] - ql[o->ibp]) / o->cmax);
    double bp = 0.5 * ((qr[o->ibp] + ql[o->ibp]) - (qr[o->ibx] - ql[o->ibx]) * o->cmax);
    ql[o->ibx] = bx; qr[o->ibx] = bx;
    ql[o->ibp] = bp; qr[o->ibp] = bp;
  }
  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &clm, &clp);
  fluxspeed(o, qr, ur, fr, &crm, &crp);
  sl = fmin_r(clm, crm);
  sr = fmax_r(clp, crp);
  if (sl >= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fl[p];
  } else if (sr <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fr[p];
  } else {
    srml = sr - sl;
    for (p = 0; p < nv; p++) {
      wl[p] = sl * ul[p] - fl[p];
      wr[p] = sr * ur[p] - fr[p];
      f[p]  = (sl * wr[p] - sr * wl[p]) / srml;
    }
  }
}

/* riemann_hd_adi_hllc, schemes.F90:1366-1513 */
static void riemann_hd_hllc(const ocfg* o, double* ql, double* qr, double* f)
{
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], wl[NVMAX], wr[NVMAX], ui[NVMAX];
  double clm, clp, crm, crp, sl, sr, sm, dn, pr, slmm, srmm;
  int p, nv = o->nv;
  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &clm, &clp);
  fluxspeed(o, qr, ur, fr, &crm, &crp);
  sl = fmin_r(clm, crm);
  sr = fmax_r(clp, crp);
  if (sl >= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fl[p];
  } else if (sr <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fr[p];
  } else {
    for (p = 0; p < nv; p++) {
      wl[p] = sl * ul[p] - fl[p];
      wr[p] = sr * ur[p] - fr[p];
    }
    dn = wr[o->idn] - wl[o->idn];
    sm = (wr[o->imx] - wl[o->imx]) / dn;
    pr = (wl[o->idn] * wr[o->imx] - wr[o->idn] * wl[o->imx]) / dn;
    if (sm > 0.0) {
      slmm = sl - sm;
      ui[o->idn] = wl[o->idn] / slmm;
      ui[o->imx] = ui[o->idn] * sm;
      ui[o->imy] = ui[o->idn] * ql[o->ivy];
      ui[o->imz] = ui[o->idn] * ql[o->ivz];
      ui[o->ien] = (wl[o->ien] + sm * pr) / slmm;
      for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
    } else if (sm < 0.0) {
      srmm = sr - sm;
      ui[o->idn] = wr[o->idn] / srmm;
      ui[o->imx] = ui[o->idn] * sm;
      ui[o->imy] = ui[o->idn] * qr[o->ivy];
      ui[o->imz] = ui[o->idn] * qr[o->ivz];
      ui[o->ien] = (wr[o->ien] + sm * pr) / srmm;
      for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
    } else {
      f[o->idn] = 0.0;
      f[o->imx] = -wl[o->imx];
      f[o->imy] = 0.0;
      f[o->imz] = 0.0;
      f[o->ien] = 0.0;
    }
  }
}

/* riemann_hd_adi_roe (schemes.F90:1541-1661) with esystem_roe_hd_adi
 * (equations.F90:2075-2196): Roe averages, the left (l[conserved][wave]) and
 * right (r[wave][conserved]) eigenvectors, and the flux 1/2 (fl + fr) minus
 * the wave contributions, summed from the left wave for a non-negative Roe
 * velocity and from the right one otherwise */
static void riemann_hd_roe(const ocfg* o, double* ql, double* qr, double* f)
{
  const int idn = o->idn, ivx = o->ivx, ivy = o->ivy, ivz = o->ivz, ien = o->ien;
  const int nv = o->nv;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], qi[NVMAX], ci[NVMAX], al[NVMAX];
  double li[NVMAX][NVMAX], ri[NVMAX][NVMAX];
  double sdl, sdr, sds, vv, vh, c2, na, cc, vc, ng, nd, nw, nh, nc, xx, cm_, cp_;
  int p, k;
  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &cm_, &cp_);
  fluxspeed(o, qr, ur, fr, &cm_, &cp_);
  sdl = sqrt(ql[idn]);
  sdr = sqrt(qr[idn]);
  sds = sdl + sdr;
  qi[idn] = sdl * sdr;
  qi[ivx] = (sdl * ql[ivx] + sdr * qr[ivx]) / sds;
  qi[ivy] = (sdl * ql[ivy] + sdr * qr[ivy]) / sds;
  qi[ivz] = (sdl * ql[ivz] + sdr * qr[ivz]) / sds;
  qi[o->ipr] = ((ul[ien] + ql[o->ipr]) / sdl + (ur[ien] + qr[o->ipr]) / sdr) / sds;
  /* esystem_roe_hd_adi: q(ien) is the Roe enthalpy */
  for (p = 0; p < nv; p++)
    for (k = 0; k < nv; k++) { li[p][k] = 0.0; ri[p][k] = 0.0; }
  vv = (qi[ivx] * qi[ivx] + qi[ivy] * qi[ivy]) + qi[ivz] * qi[ivz];
  vh = 0.5 * vv;
  c2 = o->gammam1 * (qi[ien] - vh);
  na = 0.5 / c2;
  cc = sqrt(c2);
  vc = qi[ivx] * cc;
  ng = na * o->gammam1;
  nd = 2.0 * ng;
  nw = na * vc;
  nh = (na * o->gammam1) * vh;
  nc = na * cc;
  ci[0] = qi[ivx] - cc;
  ci[1] = qi[ivx];
  ci[2] = qi[ivx];
  ci[3] = qi[ivx];
  ci[4] = qi[ivx] + cc;
  li[ivy][1] = 1.0;
  li[ivz][2] = 1.0;
  ri[0][idn] = 1.0; ri[1][ivy] = 1.0; ri[2][ivz] = 1.0; ri[3][idn] = 1.0; ri[4][idn] = 1.0;
  li[idn][0] = nh + nw;
  li[ivx][0] = -ng * qi[ivx] - nc;
  li[ivy][0] = -ng * qi[ivy];
  li[ivz][0] = -ng * qi[ivz];
  li[ien][0] = ng;
  li[idn][1] = -qi[ivy];
  li[idn][2] = -qi[ivz];
  li[idn][3] = 1.0 - ng * vv;
  li[ivx][3] = nd * qi[ivx];
  li[ivy][3] = nd * qi[ivy];
  li[ivz][3] = nd * qi[ivz];
  li[ien][3] = -nd;
  li[idn][4] = nh - nw;
  li[ivx][4] = -ng * qi[ivx] + nc;
  li[ivy][4] = -ng * qi[ivy];
  li[ivz][4] = -ng * qi[ivz];
  li[ien][4] = ng;
  ri[0][ivx] = qi[ivx] - cc; ri[0][ivy] = qi[ivy]; ri[0][ivz] = qi[ivz]; ri[0][ien] = qi[ien] - vc;
  ri[1][ien] = qi[ivy];
  ri[2][ien] = qi[ivz];
  ri[3][ivx] = qi[ivx]; ri[3][ivy] = qi[ivy]; ri[3][ivz] = qi[ivz]; ri[3][ien] = vh;
  ri[4][ivx] = qi[ivx] + cc; ri[4][ivy] = qi[ivy]; ri[4][ivz] = qi[ivz]; ri[4][ien] = qi[ien] + vc;
  if (ci[0] >= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fl[p];
  } else if (ci[nv - 1] <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fr[p];
  } else {
    for (k = 0; k < nv; k++) al[k] = 0.0;
    for (p = 0; p < nv; p++)
      for (k = 0; k < nv; k++) al[k] = al[k] + li[p][k] * (ur[p] - ul[p]);
    for (k = 0; k < nv; k++) f[k] = 0.5 * (fl[k] + fr[k]);
    if (qi[ivx] >= 0.0) {
      for (p = 0; p < nv; p++) {
        xx = -(0.5 * al[p] * fabs(ci[p]));
        for (k = 0; k < nv; k++) f[k] = f[k] + xx * ri[p][k];
      }
    } else {
      for (p = nv - 1; p >= 0; p--) {
        xx = -(0.5 * al[p] * fabs(ci[p]));
        for (k = 0; k < nv; k++) f[k] = f[k] + xx * ri[p][k];
      }
    }
  }
}

/* riemann_mhd_adi_hllc, schemes.F90:3305-3571 (Gurski / Li HLLC-MHD with
 * the GLM pre-step; its own Bx = 0 branch) */
static void riemann_mhd_hllc(const ocfg* o, double* ql, double* qr, double* f)
{
  const int idn = o->idn, imx = o->imx, imy = o->imy, imz = o->imz;
  const int ivy = o->ivy, ivz = o->ivz;
  const int ibx = o->ibx, iby = o->iby, ibz = o->ibz, ibp = o->ibp, ien = o->ien;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], wl[NVMAX], wr[NVMAX], ui[NVMAX];
  double clm, clp, crm, crp, sl, sr, sm, srml, slmm, srmm;
  double dn, bx, bp, b2, pt, vy, vz, by, bz, vb;
  int p, nv = o->nv;

  bx = 0.5 * ((qr[ibx] + ql[ibx]) - (qr[ibp] - ql[ibp]) / o->cmax);
  bp = 0.5 * ((qr[ibp] + ql[ibp]) - (qr[ibx] - ql[ibx]) * o->cmax);
  ql[ibx] = bx; qr[ibx] = bx;
  ql[ibp] = bp; qr[ibp] = bp;

  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &clm, &clp);
  fluxspeed(o, qr, ur, fr, &crm, &crp);
  sl = fmin_r(clm, crm);
  sr = fmax_r(clp, crp);
  if (sl >= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (sr <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  srml = sr - sl;
  for (p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
  }
  dn = wr[idn] - wl[idn];
  sm = (wr[imx] - wl[imx]) / dn;
  bx = ql[ibx];
  b2 = ql[ibx] * qr[ibx];
  if (b2 > 0.0) {
    pt = 0.5 * ((wl[idn] * sm - wl[imx]) + (wr[idn] * sm - wr[imx])) + b2;
    vy = (wr[imy] - wl[imy]) / dn;
    vz = (wr[imz] - wl[imz]) / dn;
    by = (wr[iby] - wl[iby]) / srml;
    bz = (wr[ibz] - wl[ibz]) / srml;
    vb = (sm * bx + vy * by) + vz * bz;
    if (sm > 0.0) {
      slmm = sl - sm;
      ui[idn] = wl[idn] / slmm;
      ui[imx] = ui[idn] * sm; ui[imy] = ui[idn] * vy; ui[imz] = ui[idn] * vz;
      ui[ibx] = bx; ui[iby] = by; ui[ibz] = bz; ui[ibp] = ql[ibp];
      ui[ien] = ((wl[ien] + sm * pt) - bx * vb) / slmm;
      for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
    } else if (sm < 0.0) {
      srmm = sr - sm;
      ui[idn] = wr[idn] / srmm;
      ui[imx] = ui[idn] * sm; ui[imy] = ui[idn] * vy; ui[imz] = ui[idn] * vz;
      ui[ibx] = bx; ui[iby] = by; ui[ibz] = bz; ui[ibp] = qr[ibp];
      ui[ien] = ((wr[ien] + sm * pt) - bx * vb) / srmm;
      for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
    } else {
      for (p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) / srml;
    }
  } else {
    pt = 0.5 * ((wl[idn] * sm - wl[imx]) + (wr[idn] * sm - wr[imx]));
    if (sm > 0.0) {
      slmm = sl - sm;
      ui[idn] = wl[idn] / slmm;
      ui[imx] = ui[idn] * sm; ui[imy] = ui[idn] * ql[ivy]; ui[imz] = ui[idn] * ql[ivz];
      ui[ibx] = 0.0; ui[iby] = wl[iby] / slmm; ui[ibz] = wl[ibz] / slmm; ui[ibp] = ql[ibp];
      ui[ien] = (wl[ien] + sm * pt) / slmm;
      for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
    } else if (sm < 0.0) {
      srmm = sr - sm;
      ui[idn] = wr[idn] / srmm;
      ui[imx] = ui[idn] * sm; ui[imy] = ui[idn] * qr[ivy]; ui[imz] = ui[idn] * qr[ivz];
      ui[ibx] = 0.0; ui[iby] = wr[iby] / srmm; ui[ibz] = wr[ibz] / srmm; ui[ibp] = qr[ibp];
      ui[ien] = (wr[ien] + sm * pt) / srmm;
      for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
    } else {
      f[idn] = 0.0;
      f[imx] = -0.5 * (wl[imx] + wr[imx]);
      f[imy] = 0.0;
      f[imz] = 0.0;
      f[ibx] = fl[ibx];
      f[iby] = 0.0;
      f[ibz] = 0.0;
      f[ibp] = fl[ibp];
      f[ien] = 0.0;
    }
  }
}

/* riemann_mhd_adi_hlld, schemes.F90:3596-4203 */
static void riemann_mhd_hlld(const ocfg* o, double* ql, double* qr, double* f)
{
  const int idn = o->idn, imx = o->imx, imy = o->imy, imz = o->imz;
  const int ibx = o->ibx, iby = o->iby, ibz = o->ibz, ibp = o->ibp, ien = o->ien;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], wl[NVMAX], wr[NVMAX];
  double wcl[NVMAX], wcr[NVMAX], ui[NVMAX];
  double clm, clp, crm, crp, sl, sr, sm, srml, slmm, srmm;
  double dn, bx, bp, b2, pt, vy, vz, by, bz, vb, dnl, dnr, cal, car, sml, smr;
  double dv, dvl, dvr;
  int p, nv = o->nv;

  bx = 0.5 * ((qr[ibx] + ql[ibx]) - (qr[ibp] - ql[ibp]) / o->cmax);
  bp = 0.5 * ((qr[ibp] + ql[ibp]) - (qr[ibx] - ql[ibx]) * o->cmax);
  ql[ibx] = bx; qr[ibx] = bx;
  ql[ibp] = bp; qr[ibp] = bp;

  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &clm, &clp);
  fluxspeed(o, qr, ur, fr, &crm, &crp);
  sl = fmin_r(clm, crm);
  sr = fmax_r(clp, crp);

  if (sl >= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (sr <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  srml = sr - sl;
  for (p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
  }
  dn   = wr[idn] - wl[idn];
  sm   = (wr[imx] - wl[imx]) / dn;
  slmm = sl - sm;
  srmm = sr - sm;
  dnl  = wl[idn] / slmm;
  dnr  = wr[idn] / srmm;
  bx   = ql[ibx];
  b2   = ql[ibx] * qr[ibx];
  pt   = (wl[idn] * wr[imx] - wr[idn] * wl[imx]) / dn + b2;

#define OUTER_LEFT()                                                           \
  vy = (slmm * wl[imy] - bx * wl[iby]) / dvl;                                  \
  vz = (slmm * wl[imz] - bx * wl[ibz]) / dvl;                                  \
  by = (wl[idn] * wl[iby] - bx * wl[imy]) / dvl;                               \
  bz = (wl[idn] * wl[ibz] - bx * wl[imz]) / dvl;                               \
  vb = (sm * bx + vy * by) + vz * bz;                                          \
  ui[idn] = dnl; ui[imx] = dnl * sm; ui[imy] = dnl * vy; ui[imz] = dnl * vz;   \
  ui[ibx] = bx; ui[iby] = by; ui[ibz] = bz; ui[ibp] = ul[ibp];                 \
  ui[ien] = ((wl[ien] + sm * pt) - bx * vb) / slmm;
#define OUTER_RIGHT()                                                          \
  vy = (srmm * wr[imy] - bx * wr[iby]) / dvr;                                  \
  vz = (srmm * wr[imz] - bx * wr[ibz]) / dvr;                                  \
  by = (wr[idn] * wr[iby] - bx * wr[imy]) / dvr;                               \
  bz = (wr[idn] * wr[ibz] - bx * wr[imz]) / dvr;                               \
  vb = (sm * bx + vy * by) + vz * bz;                                          \
  ui[idn] = dnr; ui[imx] = dnr * sm; ui[imy] = dnr * vy; ui[imz] = dnr * vz;   \
  ui[ibx] = bx; ui[iby] = by; ui[ibz] = bz; ui[ibp] = ur[ibp];                 \
  ui[ien] = ((wr[ien] + sm * pt) - bx * vb) / srmm;

  if (b2 > 1.0e-08 * fmax_r(dnl * (sl * sl), dnr * (sr * sr))) {
    cal = sqrt(b2 / dnl);
    car = sqrt(b2 / dnr);
    sml = sm - cal;
    smr = sm + car;
    dvl = slmm * wl[idn] - b2;
    dvr = srmm * wr[idn] - b2;
    if (dvl > 0.0 && dvr > 0.0) {
      if (sml >= 0.0) {
        OUTER_LEFT();
        for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
      } else if (smr <= 0.0) {
        OUTER_RIGHT();
        for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
      } else {
        OUTER_LEFT();
        for (p = 0; p < nv; p++) wcl[p] = (sml - sl) * ui[p] + wl[p];
        OUTER_RIGHT();
        for (p = 0; p < nv; p++) wcr[p] = (smr - sr) * ui[p] + wr[p];
        dv = car * dnr + cal * dnl;
        vy = (wcr[imy] - wcl[imy]) / dv;
        vz = (wcr[imz] - wcl[imz]) / dv;
        dv = car + cal;
        by = (wcr[iby] - wcl[iby]) / dv;
        bz = (wcr[ibz] - wcl[ibz]) / dv;
        vb = (sm * bx + vy * by) + vz * bz;
        if (sm > 0.0) {
          ui[idn] = dnl; ui[imx] = dnl * sm; ui[imy] = dnl * vy; ui[imz] = dnl * vz;
          ui[ibx] = bx; ui[iby] = by; ui[ibz] = bz; ui[ibp] = ul[ibp];
          ui[ien] = -((wcl[ien] + sm * pt) - bx * vb) / cal;
          for (p = 0; p < nv; p++) f[p] = sml * ui[p] - wcl[p];
        } else if (sm < 0.0) {
          ui[idn] = dnr; ui[imx] = dnr * sm; ui[imy] = dnr * vy; ui[imz] = dnr * vz;
          ui[ibx] = bx; ui[iby] = by; ui[ibz] = bz; ui[ibp] = ur[ibp];
          ui[ien] = ((wcr[ien] + sm * pt) - bx * vb) / car;
          for (p = 0; p < nv; p++) f[p] = smr * ui[p] - wcr[p];
        } else {
          for (p = 0; p < nv; p++) f[p] = (sml * wcr[p] - smr * wcl[p]) / (smr - sml);
        }
      }
    } else {
      if (dvl > 0.0) {
        OUTER_LEFT();
        if (sml >= 0.0) {
          for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
        } else {
          for (p = 0; p < nv; p++) wcl[p] = (sml - sl) * ui[p] + wl[p];
          dv = srmm * dnr + cal * dnl;
          vy = (wr[imy] - wcl[imy]) / dv;
          vz = (wr[imz] - wcl[imz]) / dv;
          dv = sr - sml;
          by = (wr[iby] - wcl[iby]) / dv;
          bz = (wr[ibz] - wcl[ibz]) / dv;
          vb = (sm * bx + vy * by) + vz * bz;
          ui[idn] = dnl; ui[imx] = dnl * sm; ui[imy] = dnl * vy; ui[imz] = dnl * vz;
          ui[ibx] = bx; ui[iby] = by; ui[ibz] = bz; ui[ibp] = ul[ibp];
          ui[ien] = -((wcl[ien] + sm * pt) - bx * vb) / cal;
          if (sm >= 0.0) {
            for (p = 0; p < nv; p++) f[p] = sml * ui[p] - wcl[p];
          } else {
            for (p = 0; p < nv; p++) wcr[p] = (sm - sml) * ui[p] + wcl[p];
            for (p = 0; p < nv; p++) f[p] = (sm * wr[p] - sr * wcr[p]) / srmm;
          }
        }
      } else if (dvr > 0.0) {
        OUTER_RIGHT();
        if (smr <= 0.0) {
          for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
        } else {
          for (p = 0; p < nv; p++) wcr[p] = (smr - sr) * ui[p] + wr[p];
          dv = slmm * dnl - car * dnr;
          vy = (wl[imy] - wcr[imy]) / dv;
          vz = (wl[imz] - wcr[imz]) / dv;
          dv = sl - smr;
          by = (wl[iby] - wcr[iby]) / dv;
          bz = (wl[ibz] - wcr[ibz]) / dv;
          vb = (sm * bx + vy * by) + vz * bz;
          ui[idn] = dnr; ui[imx] = dnr * sm; ui[imy] = dnr * vy; ui[imz] = dnr * vz;
          ui[ibx] = bx; ui[iby] = by; ui[ibz] = bz; ui[ibp] = ur[ibp];
          ui[ien] = ((wcr[ien] + sm * pt) - bx * vb) / car;
          if (sm <= 0.0) {
            for (p = 0; p < nv; p++) f[p] = smr * ui[p] - wcr[p];
          } else {
            for (p = 0; p < nv; p++) wcl[p] = (sm - smr) * ui[p] + wcr[p];
            for (p = 0; p < nv; p++) f[p] = (sm * wl[p] - sl * wcl[p]) / slmm;
          }
        }
      } else {
        for (p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) / srml;
      }
    }
  } else if (b2 > 0.0) {
    vy = (wr[imy] - wl[imy]) / dn;
    vz = (wr[imz] - wl[imz]) / dn;
    by = (wr[iby] - wl[iby]) / srml;
    bz = (wr[ibz] - wl[ibz]) / srml;
    vb = (sm * bx + vy * by) + vz * bz;
    if (sm > 0.0) {
      ui[idn] = dnl; ui[imx] = dnl * sm; ui[imy] = dnl * vy; ui[imz] = dnl * vz;
      ui[ibx] = bx; ui[iby] = by; ui[ibz] = bz; ui[ibp] = ul[ibp];
      ui[ien] = ((wl[ien] + sm * pt) - bx * vb) / slmm;
      for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
    } else if (sm < 0.0) {
      ui[idn] = dnr; ui[imx] = dnr * sm; ui[imy] = dnr * vy; ui[imz] = dnr * vz;
      ui[ibx] = bx; ui[iby] = by; ui[ibz] = bz; ui[ibp] = ur[ibp];
      ui[ien] = ((wr[ien] + sm * pt) - bx * vb) / srmm;
      for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
    } else {
      for (p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) / srml;
    }
  } else {
    if (sm > 0.0) {
      ui[idn] = dnl; ui[imx] = dnl * sm;
      ui[imy] = wl[imy] / slmm; ui[imz] = wl[imz] / slmm;
      ui[ibx] = 0.0; ui[iby] = wl[iby] / slmm; ui[ibz] = wl[ibz] / slmm;
      ui[ibp] = ul[ibp];
      ui[ien] = (wl[ien] + sm * pt) / slmm;
      for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
    } else if (sm < 0.0) {
      ui[idn] = dnr; ui[imx] = dnr * sm;
      ui[imy] = wr[imy] / srmm; ui[imz] = wr[imz] / srmm;
      ui[ibx] = 0.0; ui[iby] = wr[iby] / srmm; ui[ibz] = wr[ibz] / srmm;
      ui[ibp] = ur[ibp];
      ui[ien] = (wr[ien] + sm * pt) / srmm;
      for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
    } else {
      for (p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) / srml;
    }
  }
#undef OUTER_LEFT
#undef OUTER_RIGHT
}

/* esystem_roe_mhd_adi, equations.F90:3162-3454: the eigenvalues c[9] and
 * the left (l[var][wave]) and right (r[wave][var]) eigenvectors of the
 * adiabatic MHD Roe matrix at the averaged state q (q[ien] is the Roe
 * enthalpy; x and y the magnetic corrections).  Waves 5 and 9 and the
 * components ibx and ibp stay zero; nothing but c is set when c[0] >= 0 or
 * c[7] <= 0 (returns 0 then). */
static int esystem_roe_mhd(const ocfg* o, double x, double y, const double* q, double* c,
                           double (*l)[NVMAX], double (*r)[NVMAX])
{
  const int idn = o->idn, ivx = o->ivx, ivy = o->ivy, ivz = o->ivz, ipr = o->ipr;
  const int iby = o->iby, ibz = o->ibz, ibx = o->ibx, nv = o->nv;
  const double gammam1 = o->gammam1, gammam2 = o->c.gamma - 2.0;
  const double eps = 2.220446049250313e-16;
  double di, vsq, btsq, bt_starsq, casq, hp, twid_asq, ct2, tsum, tdif, cf2_cs2, cfsq, cf;
  double cssq, cs, ca, bt, bt_star, bet2, bet3, bet2_star, bet3_star, bet_starsq, vbet;
  double alpha_f, alpha_s, af_prime, as_prime, sqrtd, isqrtd, s, twid_a, qf, qs, afpbb, aspbb;
  double qa, qb, qc, qd, norm, cff, css, af, as, afpb, aspb, q2_star, q3_star, vqstr;
  int p, k;
  di = 1.0 / q[idn];
  casq = q[ibx] * q[ibx] * di;
  ca = sqrt(casq);
  vsq = (q[ivx] * q[ivx] + q[ivy] * q[ivy]) + q[ivz] * q[ivz];
  btsq = q[iby] * q[iby] + q[ibz] * q[ibz];
  bt_starsq = (gammam1 - gammam2 * y) * btsq;
  hp = q[ipr] - (casq + btsq * di);
  twid_asq = fmax_r(eps, gammam1 * (hp - 0.5 * vsq) - gammam2 * x);
  ct2 = bt_starsq * di;
  tsum = (casq + ct2) + twid_asq;
  tdif = (casq + ct2) - twid_asq;
  cf2_cs2 = sqrt(tdif * tdif + 4.0 * twid_asq * ct2);
  cfsq = 0.5 * (tsum + cf2_cs2);
  cf = sqrt(cfsq);
  cssq = twid_asq * casq / cfsq;
  cs = sqrt(cssq);
  c[0] = q[ivx] - cf; c[1] = q[ivx] - ca; c[2] = q[ivx] - cs; c[3] = q[ivx];
  c[4] = q[ivx]; c[5] = q[ivx] + cs; c[6] = q[ivx] + ca; c[7] = q[ivx] + cf; c[8] = c[7];
  if (c[0] >= 0.0 || c[7] <= 0.0) return 0;
  for (p = 0; p < nv; p++)
    for (k = 0; k < nv; k++) { l[p][k] = 0.0; r[p][k] = 0.0; }
  bt = sqrt(btsq);
  bt_star = sqrt(bt_starsq);
  if (bt == 0.0) { bet2 = 1.0; bet3 = 0.0; }
  else { bet2 = q[iby] / bt; bet3 = q[ibz] / bt; }
  bet2_star = bet2 / sqrt(gammam1 - gammam2 * y);
  bet3_star = bet3 / sqrt(gammam1 - gammam2 * y);
  bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
  vbet = q[ivy] * bet2_star + q[ivz] * bet3_star;
  if ((cfsq - cssq) == 0.0) { alpha_f = 1.0; alpha_s = 0.0; }
  else if ((twid_asq - cssq) <= 0.0) { alpha_f = 0.0; alpha_s = 1.0; }
  else if ((cfsq - twid_asq) <= 0.0) { alpha_f = 1.0; alpha_s = 0.0; }
  else {
    alpha_f = sqrt((twid_asq - cssq) / (cfsq - cssq));
    alpha_s = sqrt((cfsq - twid_asq) / (cfsq - cssq));
  }
  sqrtd = sqrt(q[idn]);
  isqrtd = 1.0 / sqrtd;
  s = copysign(1.0, q[ibx]);
  twid_a = sqrt(twid_asq);
  qf = cf * alpha_f * s;
  qs = cs * alpha_s * s;
  af_prime = twid_a * alpha_f * isqrtd;
  as_prime = twid_a * alpha_s * isqrtd;
  afpbb = af_prime * bt_star * bet_starsq;
  aspbb = as_prime * bt_star * bet_starsq;
  /* right eigenvectors r[wave][var] (evroe(2, wave, var)) */
  r[3][idn] = 1.0;
  r[0][idn] = alpha_f; r[2][idn] = alpha_s; r[5][idn] = alpha_s; r[7][idn] = alpha_f;
  r[0][ivx] = alpha_f * c[0]; r[2][ivx] = alpha_s * c[2]; r[3][ivx] = q[ivx];
  r[5][ivx] = alpha_s * c[5]; r[7][ivx] = alpha_f * c[7];
  qa = alpha_f * q[ivy]; qb = alpha_s * q[ivy]; qc = qs * bet2_star; qd = qf * bet2_star;
  r[0][ivy] = qa + qc; r[1][ivy] = -bet3; r[2][ivy] = qb - qd; r[3][ivy] = q[ivy];
  r[5][ivy] = qb + qd; r[6][ivy] = bet3; r[7][ivy] = qa - qc;
  qa = alpha_f * q[ivz]; qb = alpha_s * q[ivz]; qc = qs * bet3_star; qd = qf * bet3_star;
  r[0][ivz] = qa + qc; r[1][ivz] = bet2; r[2][ivz] = qb - qd; r[3][ivz] = q[ivz];
  r[5][ivz] = qb + qd; r[6][ivz] = -bet2; r[7][ivz] = qa - qc;
  r[0][ipr] = (alpha_f * (hp - q[ivx] * cf) + qs * vbet) + aspbb;
  r[1][ipr] = -(q[ivy] * bet3 - q[ivz] * bet2);
  r[2][ipr] = (alpha_s * (hp - q[ivx] * cs) - qf * vbet) - afpbb;
  r[3][ipr] = 0.5 * vsq + gammam2 * x / gammam1;
  r[5][ipr] = (alpha_s * (hp + q[ivx] * cs) + qf * vbet) - afpbb;
  r[6][ipr] = -r[1][ipr];
  r[7][ipr] = (alpha_f * (hp + q[ivx] * cf) - qs * vbet) + aspbb;
  r[0][iby] = as_prime * bet2_star; r[1][iby] = -(bet3 * s * isqrtd);
  r[2][iby] = -(af_prime * bet2_star);
  r[5][iby] = r[2][iby]; r[6][iby] = r[1][iby]; r[7][iby] = r[0][iby];
  r[0][ibz] = as_prime * bet3_star; r[1][ibz] = bet2 * s * isqrtd;
  r[2][ibz] = -(af_prime * bet3_star);
  r[5][ibz] = r[2][ibz]; r[6][ibz] = r[1][ibz]; r[7][ibz] = r[0][ibz];
  /* left eigenvectors l[var][wave] (evroe(1, var, wave)) */
  norm = 0.5 / twid_asq;
  cff = norm * alpha_f * cf;
  css = norm * alpha_s * cs;
  qf = qf * norm;
  qs = qs * norm;
  af = norm * af_prime * q[idn];
  as = norm * as_prime * q[idn];
  afpb = norm * af_prime * bt_star;
  aspb = norm * as_prime * bt_star;
  norm = norm * gammam1;
  alpha_f = alpha_f * norm;
  alpha_s = alpha_s * norm;
  q2_star = bet2_star / bet_starsq;
  q3_star = bet3_star / bet_starsq;
  vqstr = q[ivy] * q2_star + q[ivz] * q3_star;
  norm = 2.0 * norm;
  l[idn][0] = ((alpha_f * (vsq - hp) + cff * (cf + q[ivx])) - qs * vqstr) - aspb;
  l[ivx][0] = -(alpha_f * q[ivx]) - cff;
  l[ivy][0] = -(alpha_f * q[ivy]) + qs * q2_star;
  l[ivz][0] = -(alpha_f * q[ivz]) + qs * q3_star;
  l[ipr][0] = alpha_f;
  l[iby][0] = as * q2_star - alpha_f * q[iby];
  l[ibz][0] = as * q3_star - alpha_f * q[ibz];
  l[idn][1] = 0.5 * (q[ivy] * bet3 - q[ivz] * bet2);
  l[ivy][1] = -(0.5 * bet3);
  l[ivz][1] = 0.5 * bet2;
  l[iby][1] = -(0.5 * sqrtd * bet3 * s);
  l[ibz][1] = 0.5 * sqrtd * bet2 * s;
  l[idn][2] = ((alpha_s * (vsq - hp) + css * (cs + q[ivx])) + qf * vqstr) + afpb;
  l[ivx][2] = -(alpha_s * q[ivx]) - css;
  l[ivy][2] = -(alpha_s * q[ivy]) - qf * q2_star;
  l[ivz][2] = -(alpha_s * q[ivz]) - qf * q3_star;
  l[ipr][2] = alpha_s;
  l[iby][2] = -(af * q2_star) - alpha_s * q[iby];
  l[ibz][2] = -(af * q3_star) - alpha_s * q[ibz];
  l[idn][3] = 1.0 - norm * (0.5 * vsq - gammam2 * x / gammam1);
  l[ivx][3] = norm * q[ivx];
  l[ivy][3] = norm * q[ivy];
  l[ivz][3] = norm * q[ivz];
  l[ipr][3] = -norm;
  l[iby][3] = norm * q[iby];
  l[ibz][3] = norm * q[ibz];
  l[idn][5] = ((alpha_s * (vsq - hp) + css * (cs - q[ivx])) - qf * vqstr) + afpb;
  l[ivx][5] = -(alpha_s * q[ivx]) + css;
  l[ivy][5] = -(alpha_s * q[ivy]) + qf * q2_star;
  l[ivz][5] = -(alpha_s * q[ivz]) + qf * q3_star;
  l[ipr][5] = alpha_s;
  l[iby][5] = l[iby][2];
  l[ibz][5] = l[ibz][2];
  l[idn][6] = -l[idn][1];
  l[ivy][6] = -l[ivy][1];
  l[ivz][6] = -l[ivz][1];
  l[iby][6] = l[iby][1];
  l[ibz][6] = l[ibz][1];
  l[idn][7] = ((alpha_f * (vsq - hp) + cff * (cf - q[ivx])) + qs * vqstr) - aspb;
  l[ivx][7] = -(alpha_f * q[ivx]) + cff;
  l[ivy][7] = -(alpha_f * q[ivy]) - qs * q2_star;
  l[ivz][7] = -(alpha_f * q[ivz]) - qs * q3_star;
  l[ipr][7] = alpha_f;
  l[iby][7] = l[iby][0];
  l[ibz][7] = l[ibz][0];
  return 1;
}

/* riemann_mhd_adi_roe, schemes.F90:4230-4410: the GLM pre-step, Roe
 * averages (with the magnetic corrections x, y), the eigensystem above and
 * the flux 1/2 (fl + fr) minus the wave contributions, added from the left
 * wave for a non-negative Roe velocity and from the right one otherwise;
 * the Bx and psi fluxes are the left ones */
static void riemann_mhd_roe(const ocfg* o, double* ql, double* qr, double* f)
{
  const int idn = o->idn, ivx = o->ivx, ivy = o->ivy, ivz = o->ivz, ipr = o->ipr;
  const int ibx = o->ibx, iby = o->iby, ibz = o->ibz, ibp = o->ibp, ien = o->ien;
  const int nv = o->nv;
  const int fv[7] = {o->idn, o->imx, o->imy, o->imz, o->ien, o->iby, o->ibz};
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], qi[NVMAX], ci[NVMAX], al[NVMAX];
  double li[NVMAX][NVMAX], ri[NVMAX][NVMAX], sdl, sdr, sds, pml, pmr, xx, yy, bx, bp, cm_, cp_;
  int p, k, m;
  bx = 0.5 * ((qr[ibx] + ql[ibx]) - (qr[ibp] - ql[ibp]) / o->cmax);
  bp = 0.5 * ((qr[ibp] + ql[ibp]) - (qr[ibx] - ql[ibx]) * o->cmax);
  ql[ibx] = bx; qr[ibx] = bx;
  ql[ibp] = bp; qr[ibp] = bp;
  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &cm_, &cp_);
  fluxspeed(o, qr, ur, fr, &cm_, &cp_);
  sdl = sqrt(ql[idn]);
  sdr = sqrt(qr[idn]);
  sds = sdl + sdr;
  pml = 0.5 * ((ql[ibx] * ql[ibx] + ql[iby] * ql[iby]) + ql[ibz] * ql[ibz]);
  pmr = 0.5 * ((qr[ibx] * qr[ibx] + qr[iby] * qr[iby]) + qr[ibz] * qr[ibz]);
  for (p = 0; p < nv; p++) qi[p] = 0.0;
  qi[idn] = sdl * sdr;
  qi[ivx] = (sdl * ql[ivx] + sdr * qr[ivx]) / sds;
  qi[ivy] = (sdl * ql[ivy] + sdr * qr[ivy]) / sds;
  qi[ivz] = (sdl * ql[ivz] + sdr * qr[ivz]) / sds;
  qi[ipr] = (((ul[ien] + ql[ipr]) + pml) / sdl + ((ur[ien] + qr[ipr]) + pmr) / sdr) / sds;
  qi[ibx] = ql[ibx];
  qi[iby] = (sdr * ql[iby] + sdl * qr[iby]) / sds;
  qi[ibz] = (sdr * ql[ibz] + sdl * qr[ibz]) / sds;
  qi[ibp] = ql[ibp];
  xx = 0.5 * ((ql[iby] - qr[iby]) * (ql[iby] - qr[iby]) + (ql[ibz] - qr[ibz]) * (ql[ibz] - qr[ibz]))
     / (sds * sds);
  yy = 0.5 * (ql[idn] + qr[idn]) / qi[idn];
  if (!esystem_roe_mhd(o, xx, yy, qi, ci, li, ri)) {
    const double* fs = ci[0] >= 0.0 ? fl : fr;
    for (p = 0; p < nv; p++) f[p] = fs[p];
    return;
  }
  f[ibx] = fl[ibx];
  f[ibp] = fl[ibp];
  for (k = 0; k < nv; k++) al[k] = 0.0;
  for (p = 0; p < nv; p++)
    for (k = 0; k < nv; k++) al[k] = al[k] + li[p][k] * (ur[p] - ul[p]);
  for (m = 0; m < 7; m++) f[fv[m]] = 0.5 * (fl[fv[m]] + fr[fv[m]]);
  for (k = 0; k < nv; k++) {
    const int w = qi[ivx] >= 0.0 ? k : nv - 1 - k;
    xx = -(0.5 * al[w] * fabs(ci[w]));
    for (m = 0; m < 7; m++) f[fv[m]] = f[fv[m]] + xx * ri[w][fv[m]];
  }
}

/* riemann_srmhd_adi_hllc, schemes.F90:5382-5700 (Mignone & Bodo 2006): the
 * GLM pre-step, the HLL average state and flux, the contact speed as the
 * smaller root of the quadratic of the average state, and the star states
 * of the Bx != 0 and Bx = 0 cases; every degenerate case returns the HLL
 * flux */
static void riemann_srmhd_hllc(const ocfg* o, double* ql, double* qr, double* f)
{
  const int idn = o->idn, imx = o->imx, imy = o->imy, imz = o->imz, ien = o->ien;
  const int ibx = o->ibx, iby = o->iby, ibz = o->ibz, ibp = o->ibp, ivx = o->ivx;
  const int nv = o->nv;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], wl[NVMAX], wr[NVMAX];
  double uh[NVMAX], fh[NVMAX], us[NVMAX], a[3], x[2], vs[3];
  double clm, clp, crm, crp, sl, sr, srml, sm, bx, bp, bx2, uu, ff, uf, vv, vb, gi, pt, dv, fc;
  int p, nr;
  bx = 0.5 * ((qr[ibx] + ql[ibx]) - (qr[ibp] - ql[ibp]) / o->cmax);
  bp = 0.5 * ((qr[ibp] + ql[ibp]) - (qr[ibx] - ql[ibx]) * o->cmax);
  ql[ibx] = bx; qr[ibx] = bx;
  ql[ibp] = bp; qr[ibp] = bp;
  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &clm, &clp);
  fluxspeed(o, qr, ur, fr, &crm, &crp);
  sl = fmin_r(clm, crm);
  sr = fmax_r(clp, crp);
  if (sl >= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (sr <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  srml = sr - sl;
  for (p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
    uh[p] = (wr[p] - wl[p]) / srml;
    fh[p] = (sl * wr[p] - sr * wl[p]) / srml;
  }
  wl[ien] = wl[ien] + wl[idn];
  wr[ien] = wr[ien] + wr[idn];
  for (p = 0; p < nv; p++) f[p] = fh[p];     /* the HLL flux of every fallback */
  bx2 = ql[ibx] * qr[ibx];
  if (bx2 >= 1.0e-16) {
    uu = uh[iby] * uh[iby] + uh[ibz] * uh[ibz];
    ff = fh[iby] * fh[iby] + fh[ibz] * fh[ibz];
    uf = uh[iby] * fh[iby] + uh[ibz] * fh[ibz];
    a[0] = uh[imx] - uf;
    a[1] = (uu + ff) - ((fh[imx] + uh[ien]) + uh[idn]);
    a[2] = (fh[ien] + fh[idn]) - uf;
    nr = alg_quadratic(a, x);
    if (nr < 1) return;
    sm = x[0];
    if (sm <= sl || sm >= sr) return;
    us[ibx] = ql[ibx];
    us[iby] = uh[iby];
    us[ibz] = uh[ibz];
    us[ibp] = ql[ibp];
    vs[0] = sm;
    vs[1] = (us[iby] * sm - fh[iby]) / us[ibx];
    vs[2] = (us[ibz] * sm - fh[ibz]) / us[ibx];
    vv = (vs[0] * vs[0] + vs[1] * vs[1]) + vs[2] * vs[2];
    vb = (vs[0] * us[ibx] + vs[1] * us[iby]) + vs[2] * us[ibz];
    gi = 1.0 - vv;
    pt = (fh[imx] + gi * bx2) - ((fh[ien] + fh[idn]) - us[ibx] * vb) * sm;
    if (pt <= 0.0) return;
    if (sm > 0.0 || sm < 0.0) {
      const double s = sm > 0.0 ? sl : sr;
      const double* qq = sm > 0.0 ? ql : qr;
      const double* uq = sm > 0.0 ? ul : ur;
      const double* fq = sm > 0.0 ? fl : fr;
      dv = s - sm;
      fc = (s - qq[ivx]) / dv;
      us[idn] = fc * uq[idn];
      us[imy] = ((s * uq[imy] - fq[imy]) - us[ibx] * (gi * us[iby] + vs[1] * vb)) / dv;
      us[imz] = ((s * uq[imz] - fq[imz]) - us[ibx] * (gi * us[ibz] + vs[2] * vb)) / dv;
      us[ien] = ((((s * (uq[ien] + uq[idn])) - (fq[ien] + fq[idn])) + pt * sm) - us[ibx] * vb) / dv
              - us[idn];
      us[imx] = ((us[ien] + us[idn]) + pt) * sm - us[ibx] * vb;
      for (p = 0; p < nv; p++) f[p] = fq[p] + s * (us[p] - uq[p]);
    }
  } else {
    a[0] = uh[imx];
    a[1] = -((fh[imx] + uh[ien]) + uh[idn]);
    a[2] = fh[ien] + fh[idn];
    nr = alg_quadratic(a, x);
    if (nr < 1) return;
    sm = x[0];
    if (sm <= sl || sm >= sr) return;
    pt = fh[imx] - (fh[ien] + fh[idn]) * sm;
    if (pt <= 0.0) return;
    if (sm > 0.0 || sm < 0.0) {
      const double s = sm > 0.0 ? sl : sr;
      const double* w = sm > 0.0 ? wl : wr;
      const double* qq = sm > 0.0 ? ql : qr;
      const double* uq = sm > 0.0 ? ul : ur;
      const double* fq = sm > 0.0 ? fl : fr;
      dv = s - sm;
      us[idn] = w[idn] / dv;
      us[imy] = w[imy] / dv;
      us[imz] = w[imz] / dv;
      us[ien] = (w[ien] + pt * sm) / dv;
      us[imx] = (us[ien] + pt) * sm;
      us[ien] = us[ien] - us[idn];
      us[ibx] = 0.0;
      us[iby] = w[iby] / dv;
      us[ibz] = w[ibz] / dv;
      us[ibp] = qq[ibp];
      for (p = 0; p < nv; p++) f[p] = fq[p] + s * (us[p] - uq[p]);
    }
  }
}

static int riemann_check(const amun_cfg* c)
{
  if (c->riemann_solver == AMUN_RS_HLL) return 0;
  if (c->riemann_solver == AMUN_RS_HLLC &&
      (c->eqsys == AMUN_EQSYS_HD || c->eqsys == AMUN_EQSYS_MHD)) return 0;
  if (c->riemann_solver == AMUN_RS_ROE &&
      (c->eqsys == AMUN_EQSYS_HD || c->eqsys == AMUN_EQSYS_MHD)) return 0;
  if (c->riemann_solver == AMUN_RS_HLLD && c->eqsys == AMUN_EQSYS_MHD) return 0;
  /* HLLD-M is the isothermal MHD system's (schemes.F90:219-221) */
  if (c->riemann_solver == AMUN_RS_HLLDM && c->eqsys == AMUN_EQSYS_MHD &&
      c->eos == AMUN_EOS_ISO) return 0;
  /* SRMHD: "hllc" selects riemann_srmhd_adi_hllc, any other name HLL
   * (schemes.F90:479-498) */
  if (c->eqsys == AMUN_EQSYS_SRMHD || c->eqsys == AMUN_EQSYS_SRHD) return 0;
  return 300;
}

/* riemann_hd_iso_roe (schemes.F90:939-1057) with esystem_roe_hd_iso
 * (equations.F90:1655-1738): Roe averages of density and velocity, the
 * eigenvalues v -+ csnd and v, the upwinded sum over the waves in
 * increasing (v >= 0) or decreasing order */
static void riemann_hd_iso_roe(const ocfg* o, double* ql, double* qr, double* f)
{
  const int idn = o->idn, ivx = o->ivx, ivy = o->ivy, ivz = o->ivz, nv = o->nv;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], qi[NVMAX], ci[NVMAX], al[NVMAX];
  double li[NVMAX][NVMAX], ri[NVMAX][NVMAX];
  double sdl, sdr, sds, cm, cp, ch;
  int p, q;
  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &cm, &cp);
  fluxspeed(o, qr, ur, fr, &cm, &cp);
  sdl = sqrt(ql[idn]);
  sdr = sqrt(qr[idn]);
  sds = sdl + sdr;
  qi[idn] = sdl * sdr;
  qi[ivx] = (sdl * ql[ivx] + sdr * qr[ivx]) / sds;
  qi[ivy] = (sdl * ql[ivy] + sdr * qr[ivy]) / sds;
  qi[ivz] = (sdl * ql[ivz] + sdr * qr[ivz]) / sds;
  /* esystem_roe_hd_iso: l(wave, var), r(wave, var) */
  ch = 0.5 / o->csnd;
  for (p = 0; p < nv; p++)
    for (q = 0; q < nv; q++) li[p][q] = ri[p][q] = 0.0;
  ci[0] = qi[ivx] - o->csnd;
  ci[1] = qi[ivx];
  ci[2] = qi[ivx];
  ci[3] = qi[ivx] + o->csnd;
  li[ivx][0] = -ch;
  li[ivy][1] = 1.0;
  li[ivz][2] = 1.0;
  li[ivx][3] = ch;
  li[idn][0] = ch * ci[3];
  li[idn][1] = -qi[ivy];
  li[idn][2] = -qi[ivz];
  li[idn][3] = -ch * ci[0];
  ri[0][idn] = 1.0;
  ri[1][ivy] = 1.0;
  ri[2][ivz] = 1.0;
  ri[3][idn] = 1.0;
  ri[0][ivx] = ci[0];
  ri[0][ivy] = qi[ivy];
  ri[0][ivz] = qi[ivz];
  ri[3][ivx] = ci[3];
  ri[3][ivy] = qi[ivy];
  ri[3][ivz] = qi[ivz];
  if (ci[0] >= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fl[p];
  } else if (ci[nv - 1] <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fr[p];
  } else {
    /* al(1:nv) += li(p, 1:nv) (ur(p) - ul(p)), p = 1..nv: li is l(p, wave)
     * with the variable first, as evroe(1, var, wave) */
    for (q = 0; q < nv; q++) al[q] = 0.0;
    for (p = 0; p < nv; p++)
      for (q = 0; q < nv; q++) al[q] = al[q] + li[p][q] * (ur[p] - ul[p]);
    for (q = 0; q < nv; q++) f[q] = 0.5 * (fl[q] + fr[q]);
    if (qi[ivx] >= 0.0) {
      for (p = 0; p < nv; p++) {
        const double xx = -0.5 * al[p] * fabs(ci[p]);
        for (q = 0; q < nv; q++) f[q] = f[q] + xx * ri[p][q];
      }
    } else {
      for (p = nv - 1; p >= 0; p--) {
        const double xx = -0.5 * al[p] * fabs(ci[p]);
        for (q = 0; q < nv; q++) f[q] = f[q] + xx * ri[p][q];
      }
    }
  }
}

/* riemann_mhd_iso_hlld (schemes.F90:2007-2364): the GLM pre-step, the HLL
 * bounds, the contact speed sm from the HLL state's momentum, the Alfven
 * speeds sm -+ sqrt(Bx^2 / rho*), and the branch tree over Bx^2 (against
 * 1e-8 of the outer waves' scale, > 0, = 0) and the degeneracies */
static void riemann_mhd_iso_hlld(const ocfg* o, double* ql, double* qr, double* f)
{
  const int idn = o->idn, imx = o->imx, imy = o->imy, imz = o->imz;
  const int ibx = o->ibx, iby = o->iby, ibz = o->ibz, ibp = o->ibp, nv = o->nv;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], wl[NVMAX], wr[NVMAX];
  double wcl[NVMAX], wcr[NVMAX], ui[NVMAX];
  double clm, clp, crm, crp, sl, sr, sm, sml, smr, srml, slmm, srmm;
  double bx, bp, b2, dn, dnl, dnr, dvl, dvr;
  int p;
  bx = 0.5 * ((qr[ibx] + ql[ibx]) - (qr[ibp] - ql[ibp]) / o->cmax);
  bp = 0.5 * ((qr[ibp] + ql[ibp]) - (qr[ibx] - ql[ibx]) * o->cmax);
  ql[ibx] = bx; qr[ibx] = bx;
  ql[ibp] = bp; qr[ibp] = bp;
  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &clm, &clp);
  fluxspeed(o, qr, ur, fr, &crm, &crp);
  sl = fmin_r(clm, crm);
  sr = fmax_r(clp, crp);
  if (sl >= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (sr <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  srml = sr - sl;
  for (p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
  }
  dn = wr[idn] - wl[idn];
  sm = (wr[imx] - wl[imx]) / dn;
  bx = ql[ibx];
  b2 = ql[ibx] * qr[ibx];
  slmm = sl - sm;
  srmm = sr - sm;
  dnl = wl[idn] / slmm;
  dnr = wr[idn] / srmm;
#define STAR_L()                                                               \
  ui[idn] = dnl;                                                               \
  ui[imx] = dnl * sm;                                                          \
  ui[imy] = dnl * (slmm * wl[imy] - bx * wl[iby]) / dvl;                       \
  ui[imz] = dnl * (slmm * wl[imz] - bx * wl[ibz]) / dvl;                       \
  ui[ibx] = bx;                                                                \
  ui[iby] = (wl[idn] * wl[iby] - bx * wl[imy]) / dvl;                          \
  ui[ibz] = (wl[idn] * wl[ibz] - bx * wl[imz]) / dvl;                          \
  ui[ibp] = ul[ibp]
#define STAR_R()                                                               \
  ui[idn] = dnr;                                                               \
  ui[imx] = dnr * sm;                                                          \
  ui[imy] = dnr * (srmm * wr[imy] - bx * wr[iby]) / dvr;                       \
  ui[imz] = dnr * (srmm * wr[imz] - bx * wr[ibz]) / dvr;                       \
  ui[ibx] = bx;                                                                \
  ui[iby] = (wr[idn] * wr[iby] - bx * wr[imy]) / dvr;                          \
  ui[ibz] = (wr[idn] * wr[ibz] - bx * wr[imz]) / dvr;                          \
  ui[ibp] = ur[ibp]
  if (b2 > 1.0e-08 * fmax_r(dnl * (sl * sl), dnr * (sr * sr))) {
    sml = sm - sqrt(b2 / dnl);
    smr = sm + sqrt(b2 / dnr);
    dvl = slmm * wl[idn] - b2;
    dvr = srmm * wr[idn] - b2;
    if (fmin_r(dvl, dvr) > 0.0) {
      if (sml >= 0.0) {
        STAR_L();
        for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
      } else if (smr <= 0.0) {
        STAR_R();
        for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
      } else {
        STAR_L();
        for (p = 0; p < nv; p++) wcl[p] = (sml - sl) * ui[p] + wl[p];
        STAR_R();
        for (p = 0; p < nv; p++) wcr[p] = (smr - sr) * ui[p] + wr[p];
        for (p = 0; p < nv; p++) f[p] = (sml * wcr[p] - smr * wcl[p]) / (smr - sml);
      }
    } else {
      if (dvl > 0.0) {
        STAR_L();
        if (sml >= 0.0) {
          for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
        } else {
          for (p = 0; p < nv; p++) wcl[p] = (sml - sl) * ui[p] + wl[p];
          for (p = 0; p < nv; p++) f[p] = (sml * wr[p] - sr * wcl[p]) / (sr - sml);
        }
      } else if (dvr > 0.0) {
        STAR_R();
        if (smr <= 0.0) {
          for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
        } else {
          for (p = 0; p < nv; p++) wcr[p] = (smr - sr) * ui[p] + wr[p];
          for (p = 0; p < nv; p++) f[p] = (sl * wcr[p] - smr * wl[p]) / (smr - sl);
        }
      } else {
        for (p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) / srml;
      }
    }
  } else if (b2 > 0.0) {
    for (p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) / srml;
  } else {
    if (sm > 0.0) {
      ui[idn] = dnl;
      ui[imx] = dnl * sm;
      ui[imy] = wl[imy] / slmm;
      ui[imz] = wl[imz] / slmm;
      ui[ibx] = 0.0;
      ui[iby] = wl[iby] / slmm;
      ui[ibz] = wl[ibz] / slmm;
      ui[ibp] = ul[ibp];
      for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
    } else if (sm < 0.0) {
      ui[idn] = dnr;
      ui[imx] = dnr * sm;
      ui[imy] = wr[imy] / srmm;
      ui[imz] = wr[imz] / srmm;
      ui[ibx] = 0.0;
      ui[iby] = wr[iby] / srmm;
      ui[ibz] = wr[ibz] / srmm;
      ui[ibp] = ur[ibp];
      for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
    } else {
      for (p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) / srml;
    }
  }
#undef STAR_L
#undef STAR_R
}

/* esystem_roe_mhd_iso, equations.F90:2552-2791: the eigenvalues c[8] and
 * the left (l[var][wave]) and right (r[wave][var]) eigenvectors of the
 * isothermal MHD Roe matrix at the averaged state q (x and y the magnetic
 * corrections).  Waves 4 and 8 and the components ibx and ibp stay zero;
 * nothing but c is set when c[0] >= 0 or c[6] <= 0 (returns 0 then). */
static int esystem_roe_mhd_iso(const ocfg* o, double x, double y, const double* q, double* c,
                               double (*l)[NVMAX], double (*r)[NVMAX])
{
  const int idn = o->idn, ivx = o->ivx, ivy = o->ivy, ivz = o->ivz;
  const int iby = o->iby, ibz = o->ibz, ibx = o->ibx, nv = o->nv;
  double di, btsq, bt_starsq, casq, twid_csq, ct2, tsum, tdif, cf2_cs2, cfsq, cf, cssq, cs, ca;
  double bt, bt_star, bet2, bet3, bet2_star, bet3_star, bet_starsq, alpha_f, alpha_s;
  double sqrtd, s, twid_c, qf, qs, af_prime, as_prime, norm, cff, css, af, as, afpb, aspb;
  double q2_star, q3_star, vqstr;
  int p, k;
  di = 1.0 / q[idn];
  casq = q[ibx] * q[ibx] * di;
  ca = sqrt(casq);
  btsq = q[iby] * q[iby] + q[ibz] * q[ibz];
  bt_starsq = btsq * y;
  twid_csq = o->csnd2 + x;
  ct2 = bt_starsq * di;
  tsum = (casq + ct2) + twid_csq;
  tdif = (casq + ct2) - twid_csq;
  cf2_cs2 = sqrt(tdif * tdif + 4.0 * twid_csq * ct2);
  cfsq = 0.5 * (tsum + cf2_cs2);
  cf = sqrt(cfsq);
  cssq = twid_csq * casq / cfsq;
  cs = sqrt(cssq);
  c[0] = q[ivx] - cf; c[1] = q[ivx] - ca; c[2] = q[ivx] - cs; c[3] = q[ivx];
  c[4] = q[ivx] + cs; c[5] = q[ivx] + ca; c[6] = q[ivx] + cf; c[7] = c[6];
  if (c[0] >= 0.0 || c[6] <= 0.0) return 0;
  for (p = 0; p < nv; p++)
    for (k = 0; k < nv; k++) { l[p][k] = 0.0; r[p][k] = 0.0; }
  bt = sqrt(btsq);
  bt_star = sqrt(bt_starsq);
  if (bt == 0.0) { bet2 = 1.0; bet3 = 0.0; }
  else { bet2 = q[iby] / bt; bet3 = q[ibz] / bt; }
  bet2_star = bet2 / sqrt(y);
  bet3_star = bet3 / sqrt(y);
  bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
  if ((cfsq - cssq) == 0.0) { alpha_f = 1.0; alpha_s = 0.0; }
  else if ((twid_csq - cssq) <= 0.0) { alpha_f = 0.0; alpha_s = 1.0; }
  else if ((cfsq - twid_csq) <= 0.0) { alpha_f = 1.0; alpha_s = 0.0; }
  else {
    alpha_f = sqrt((twid_csq - cssq) / (cfsq - cssq));
    alpha_s = sqrt((cfsq - twid_csq) / (cfsq - cssq));
  }
  sqrtd = sqrt(q[idn]);
  s = copysign(1.0, q[ibx]);
  twid_c = sqrt(twid_csq);
  qf = cf * alpha_f * s;
  qs = cs * alpha_s * s;
  af_prime = twid_c * alpha_f / sqrtd;
  as_prime = twid_c * alpha_s / sqrtd;
  /* right eigenvectors r[wave][var] */
  r[0][idn] = alpha_f;
  r[0][ivx] = alpha_f * c[0];
  r[0][ivy] = alpha_f * q[ivy] + qs * bet2_star;
  r[0][ivz] = alpha_f * q[ivz] + qs * bet3_star;
  r[0][iby] = as_prime * bet2_star;
  r[0][ibz] = as_prime * bet3_star;
  r[1][ivy] = -bet3;
  r[1][ivz] = bet2;
  r[1][iby] = -(bet3 * s / sqrtd);
  r[1][ibz] = bet2 * s / sqrtd;
  r[2][idn] = alpha_s;
  r[2][ivx] = alpha_s * c[2];
  r[2][ivy] = alpha_s * q[ivy] - qf * bet2_star;
  r[2][ivz] = alpha_s * q[ivz] - qf * bet3_star;
  r[2][iby] = -(af_prime * bet2_star);
  r[2][ibz] = -(af_prime * bet3_star);
  r[4][idn] = alpha_s;
  r[4][ivx] = alpha_s * c[4];
  r[4][ivy] = alpha_s * q[ivy] + qf * bet2_star;
  r[4][ivz] = alpha_s * q[ivz] + qf * bet3_star;
  r[4][iby] = r[2][iby];
  r[4][ibz] = r[2][ibz];
  r[5][ivy] = bet3;
  r[5][ivz] = -bet2;
  r[5][iby] = r[1][iby];
  r[5][ibz] = r[1][ibz];
  r[6][idn] = alpha_f;
  r[6][ivx] = alpha_f * c[6];
  r[6][ivy] = alpha_f * q[ivy] - qs * bet2_star;
  r[6][ivz] = alpha_f * q[ivz] - qs * bet3_star;
  r[6][iby] = r[0][iby];
  r[6][ibz] = r[0][ibz];
  /* left eigenvectors l[var][wave] */
  norm = 0.5 / twid_csq;
  cff = norm * alpha_f * cf;
  css = norm * alpha_s * cs;
  qf = qf * norm;
  qs = qs * norm;
  af = norm * af_prime * q[idn];
  as = norm * as_prime * q[idn];
  afpb = norm * af_prime * bt_star;
  aspb = norm * as_prime * bt_star;
  q2_star = bet2_star / bet_starsq;
  q3_star = bet3_star / bet_starsq;
  vqstr = q[ivy] * q2_star + q[ivz] * q3_star;
  l[idn][0] = (cff * c[6] - qs * vqstr) - aspb;
  l[ivx][0] = -cff;
  l[ivy][0] = qs * q2_star;
  l[ivz][0] = qs * q3_star;
  l[iby][0] = as * q2_star;
  l[ibz][0] = as * q3_star;
  l[idn][1] = 0.5 * (q[ivy] * bet3 - q[ivz] * bet2);
  l[ivy][1] = -(0.5 * bet3);
  l[ivz][1] = 0.5 * bet2;
  l[iby][1] = -(0.5 * sqrtd * bet3 * s);
  l[ibz][1] = 0.5 * sqrtd * bet2 * s;
  l[idn][2] = (css * c[4] + qf * vqstr) + afpb;
  l[ivx][2] = -css;
  l[ivy][2] = -(qf * q2_star);
  l[ivz][2] = -(qf * q3_star);
  l[iby][2] = -(af * q2_star);
  l[ibz][2] = -(af * q3_star);
  l[idn][4] = (-(css * c[2]) - qf * vqstr) + afpb;
  l[ivx][4] = css;
  l[ivy][4] = -l[ivy][2];
  l[ivz][4] = -l[ivz][2];
  l[iby][4] = l[iby][2];
  l[ibz][4] = l[ibz][2];
  l[idn][5] = -l[idn][1];
  l[ivy][5] = -l[ivy][1];
  l[ivz][5] = -l[ivz][1];
  l[iby][5] = l[iby][1];
  l[ibz][5] = l[ibz][1];
  l[idn][6] = (-(cff * c[0]) + qs * vqstr) - aspb;
  l[ivx][6] = cff;
  l[ivy][6] = -l[ivy][0];
  l[ivz][6] = -l[ivz][0];
  l[iby][6] = l[iby][0];
  l[ibz][6] = l[ibz][0];
  return 1;
}

/* riemann_mhd_iso_roe, schemes.F90:2778-2906: the GLM pre-step, Roe
 * averages with the magnetic corrections x, y, the eigensystem above and
 * the flux 1/2 (fl + fr) minus the wave contributions, added from the left
 * wave for a non-negative Roe velocity and from the right one otherwise;
 * the Bx and psi fluxes are the left ones */
static void riemann_mhd_iso_roe(const ocfg* o, double* ql, double* qr, double* f)
{
  const int idn = o->idn, ivx = o->ivx, ivy = o->ivy, ivz = o->ivz;
  const int ibx = o->ibx, iby = o->iby, ibz = o->ibz, ibp = o->ibp, nv = o->nv;
  const int fv[6] = {o->idn, o->imx, o->imy, o->imz, o->iby, o->ibz};
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], qi[NVMAX], ci[NVMAX], al[NVMAX];
  double li[NVMAX][NVMAX], ri[NVMAX][NVMAX], sdl, sdr, sds, xx, yy, bx, bp, cm_, cp_;
  int p, k, m;
  bx = 0.5 * ((qr[ibx] + ql[ibx]) - (qr[ibp] - ql[ibp]) / o->cmax);
  bp = 0.5 * ((qr[ibp] + ql[ibp]) - (qr[ibx] - ql[ibx]) * o->cmax);
  ql[ibx] = bx; qr[ibx] = bx;
  ql[ibp] = bp; qr[ibp] = bp;
  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &cm_, &cp_);
  fluxspeed(o, qr, ur, fr, &cm_, &cp_);
  sdl = sqrt(ql[idn]);
  sdr = sqrt(qr[idn]);
  sds = sdl + sdr;
  for (p = 0; p < nv; p++) qi[p] = 0.0;
  qi[idn] = sdl * sdr;
  qi[ivx] = (sdl * ql[ivx] + sdr * qr[ivx]) / sds;
  qi[ivy] = (sdl * ql[ivy] + sdr * qr[ivy]) / sds;
  qi[ivz] = (sdl * ql[ivz] + sdr * qr[ivz]) / sds;
  qi[ibx] = ql[ibx];
  qi[iby] = (sdr * ql[iby] + sdl * qr[iby]) / sds;
  qi[ibz] = (sdr * ql[ibz] + sdl * qr[ibz]) / sds;
  qi[ibp] = ql[ibp];
  xx = 0.5 * ((ql[iby] - qr[iby]) * (ql[iby] - qr[iby]) + (ql[ibz] - qr[ibz]) * (ql[ibz] - qr[ibz]))
     / (sds * sds);
  yy = 0.5 * (ql[idn] + qr[idn]) / qi[idn];
  if (!esystem_roe_mhd_iso(o, xx, yy, qi, ci, li, ri)) {
    const double* fs = ci[0] >= 0.0 ? fl : fr;
    for (p = 0; p < nv; p++) f[p] = fs[p];
    return;
  }
  f[ibx] = fl[ibx];
  f[ibp] = fl[ibp];
  for (k = 0; k < nv; k++) al[k] = 0.0;
  for (p = 0; p < nv; p++)
    for (k = 0; k < nv; k++) al[k] = al[k] + li[p][k] * (ur[p] - ul[p]);
  for (m = 0; m < 6; m++) f[fv[m]] = 0.5 * (fl[fv[m]] + fr[fv[m]]);
  for (k = 0; k < nv; k++) {
    const int w = qi[ivx] >= 0.0 ? k : nv - 1 - k;
    xx = -(0.5 * al[w] * fabs(ci[w]));
    for (m = 0; m < 6; m++) f[fv[m]] = f[fv[m]] + xx * ri[w][fv[m]];
  }
}

/* riemann_mhd_iso_hlldm, schemes.F90:2388-2777: HLLD with one star density
 * for both sides, the HLL state's (wr - wl)(idn) / (sr - sl) (Mignone's
 * variant): Alfven speeds sm -+ sqrt(Bx^2 / rho*), the threshold against
 * 1e-8 max(sl^2, sr^2) and the same branch tree as riemann_mhd_iso_hlld */
static void riemann_mhd_iso_hlldm(const ocfg* o, double* ql, double* qr, double* f)
{
  const int idn = o->idn, imx = o->imx, imy = o->imy, imz = o->imz;
  const int ibx = o->ibx, iby = o->iby, ibz = o->ibz, ibp = o->ibp, nv = o->nv;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], wl[NVMAX], wr[NVMAX];
  double wcl[NVMAX], wcr[NVMAX], ui[NVMAX];
  double clm, clp, crm, crp, sl, sr, sm, sml, smr, srml, slmm, srmm;
  double bx, bp, b2, dn, dvl, dvr, ca, ca2;
  int p;
  bx = 0.5 * ((qr[ibx] + ql[ibx]) - (qr[ibp] - ql[ibp]) / o->cmax);
  bp = 0.5 * ((qr[ibp] + ql[ibp]) - (qr[ibx] - ql[ibx]) * o->cmax);
  ql[ibx] = bx; qr[ibx] = bx;
  ql[ibp] = bp; qr[ibp] = bp;
  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &clm, &clp);
  fluxspeed(o, qr, ur, fr, &crm, &crp);
  sl = fmin_r(clm, crm);
  sr = fmax_r(clp, crp);
  if (sl >= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (sr <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  srml = sr - sl;
  for (p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
  }
  dn = wr[idn] - wl[idn];
  sm = (wr[imx] - wl[imx]) / dn;
  bx = ql[ibx];
  b2 = ql[ibx] * qr[ibx];
  slmm = sl - sm;
  srmm = sr - sm;
  dn = dn / srml;
  ca2 = b2 / dn;
#define STAR_L()                                                               \
  ui[idn] = dn;                                                                \
  ui[imx] = dn * sm;                                                           \
  ui[imy] = dn * (slmm * wl[imy] - bx * wl[iby]) / dvl;                        \
  ui[imz] = dn * (slmm * wl[imz] - bx * wl[ibz]) / dvl;                        \
  ui[ibx] = bx;                                                                \
  ui[iby] = (slmm * dn * wl[iby] - bx * wl[imy]) / dvl;                        \
  ui[ibz] = (slmm * dn * wl[ibz] - bx * wl[imz]) / dvl;                        \
  ui[ibp] = ul[ibp]
#define STAR_R()                                                               \
  ui[idn] = dn;                                                                \
  ui[imx] = dn * sm;                                                           \
  ui[imy] = dn * (srmm * wr[imy] - bx * wr[iby]) / dvr;                        \
  ui[imz] = dn * (srmm * wr[imz] - bx * wr[ibz]) / dvr;                        \
  ui[ibx] = bx;                                                                \
  ui[iby] = (srmm * dn * wr[iby] - bx * wr[imy]) / dvr;                        \
  ui[ibz] = (srmm * dn * wr[ibz] - bx * wr[imz]) / dvr;                        \
  ui[ibp] = ur[ibp]
  if (ca2 > 1.0e-08 * fmax_r(sl * sl, sr * sr)) {
    ca = sqrt(ca2);
    sml = sm - ca;
    smr = sm + ca;
    dvl = slmm * slmm * dn - b2;
    dvr = srmm * srmm * dn - b2;
    if (fmin_r(dvl, dvr) > 0.0) {
      if (sml >= 0.0) {
        STAR_L();
        for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
      } else if (smr <= 0.0) {
        STAR_R();
        for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
      } else {
        STAR_L();
        for (p = 0; p < nv; p++) wcl[p] = (sml - sl) * ui[p] + wl[p];
        STAR_R();
        for (p = 0; p < nv; p++) wcr[p] = (smr - sr) * ui[p] + wr[p];
        for (p = 0; p < nv; p++) f[p] = (sml * wcr[p] - smr * wcl[p]) / (smr - sml);
      }
    } else if (dvl > 0.0) {
      STAR_L();
      if (sml >= 0.0) {
        for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
      } else {
        for (p = 0; p < nv; p++) wcl[p] = (sml - sl) * ui[p] + wl[p];
        for (p = 0; p < nv; p++) f[p] = (sml * wr[p] - sr * wcl[p]) / (sr - sml);
      }
    } else if (dvr > 0.0) {
      STAR_R();
      if (smr <= 0.0) {
        for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
      } else {
        for (p = 0; p < nv; p++) wcr[p] = (smr - sr) * ui[p] + wr[p];
        for (p = 0; p < nv; p++) f[p] = (sl * wcr[p] - smr * wl[p]) / (smr - sl);
      }
    } else {
      for (p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) / srml;
    }
  } else if (b2 > 0.0) {
    for (p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) / srml;
  } else if (sm > 0.0) {
    ui[idn] = dn;
    ui[imx] = dn * sm;
    ui[imy] = wl[imy] / slmm;
    ui[imz] = wl[imz] / slmm;
    ui[ibx] = 0.0;
    ui[iby] = wl[iby] / slmm;
    ui[ibz] = wl[ibz] / slmm;
    ui[ibp] = ul[ibp];
    for (p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
  } else if (sm < 0.0) {
    ui[idn] = dn;
    ui[imx] = dn * sm;
    ui[imy] = wr[imy] / srmm;
    ui[imz] = wr[imz] / srmm;
    ui[ibx] = 0.0;
    ui[iby] = wr[iby] / srmm;
    ui[ibz] = wr[ibz] / srmm;
    ui[ibp] = ur[ibp];
    for (p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
  } else {
    for (p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) / srml;
  }
#undef STAR_L
#undef STAR_R
}

/* riemann_srhd_adi_hllc, schemes.F90:4775-4968 (Mignone & Bodo 2005): the
 * HLL average state and flux, the contact speed as the smaller root of
 * F_E x^2 - (E + F_m) x + m = 0 of that state, the star state on the side
 * of the contact; no root, a root outside (sl, sr) or a non-positive star
 * pressure return the HLL flux */
static void riemann_srhd_hllc(const ocfg* o, double* ql, double* qr, double* f)
{
  const int idn = o->idn, imx = o->imx, imy = o->imy, imz = o->imz, ien = o->ien;
  const int nv = o->nv;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], wl[NVMAX], wr[NVMAX];
  double uh[NVMAX], fh[NVMAX], us[NVMAX], a[3], x[2];
  double clm, clp, crm, crp, sl, sr, srml, sm, pr, dv;
  int p, nr;
  prim2cons(o, ql, ul);
  prim2cons(o, qr, ur);
  fluxspeed(o, ql, ul, fl, &clm, &clp);
  fluxspeed(o, qr, ur, fr, &crm, &crp);
  sl = fmin_r(clm, crm);
  sr = fmax_r(clp, crp);
  if (sl >= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (sr <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  srml = sr - sl;
  for (p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
  }
  for (p = 0; p < nv; p++) {
    uh[p] = (wr[p] - wl[p]) / srml;
    fh[p] = (sl * wr[p] - sr * wl[p]) / srml;
  }
  wl[ien] = wl[ien] + wl[idn];
  wr[ien] = wr[ien] + wr[idn];
  a[0] = uh[imx];
  a[1] = -((fh[imx] + uh[ien]) + uh[idn]);
  a[2] = fh[ien] + fh[idn];
  nr = alg_quadratic(a, x);
  if (nr < 1) {
    for (p = 0; p < nv; p++) f[p] = fh[p];
    return;
  }
  sm = x[0];
  if (sm <= sl || sm >= sr) {
    for (p = 0; p < nv; p++) f[p] = fh[p];
    return;
  }
  pr = fh[imx] - (fh[ien] + fh[idn]) * sm;
  if (pr <= 0.0) {
    for (p = 0; p < nv; p++) f[p] = fh[p];
    return;
  }
  if (sm > 0.0) {
    dv = sl - sm;
    us[idn] = wl[idn] / dv;
    us[imy] = wl[imy] / dv;
    us[imz] = wl[imz] / dv;
    us[ien] = (wl[ien] + pr * sm) / dv;
    us[imx] = (us[ien] + pr) * sm;
    us[ien] = us[ien] - us[idn];
    for (p = 0; p < nv; p++) f[p] = fl[p] + sl * (us[p] - ul[p]);
  } else if (sm < 0.0) {
    dv = sr - sm;
    us[idn] = wr[idn] / dv;
    us[imy] = wr[imy] / dv;
    us[imz] = wr[imz] / dv;
    us[ien] = (wr[ien] + pr * sm) / dv;
    us[imx] = (us[ien] + pr) * sm;
    us[ien] = us[ien] - us[idn];
    for (p = 0; p < nv; p++) f[p] = fr[p] + sr * (us[p] - ur[p]);
  } else {
    f[idn] = 0.0;
    f[imx] = pr;
    f[imy] = 0.0;
    f[imz] = 0.0;
    f[ien] = 0.0;
  }
}

static void riemann(const ocfg* o, double* ql, double* qr, double* f)
{
  if (o->iso) {
    /* schemes.F90:163-310: HD iso takes "roe", MHD iso "hlld", every other
     * name the HLL solver */
    if (!o->mhd && o->c.riemann_solver == AMUN_RS_ROE) riemann_hd_iso_roe(o, ql, qr, f);
    else if (o->mhd && o->c.riemann_solver == AMUN_RS_HLLD) riemann_mhd_iso_hlld(o, ql, qr, f);
    else if (o->mhd && o->c.riemann_solver == AMUN_RS_HLLDM) riemann_mhd_iso_hlldm(o, ql, qr, f);
    else if (o->mhd && o->c.riemann_solver == AMUN_RS_ROE) riemann_mhd_iso_roe(o, ql, qr, f);
    else riemann_hll(o, ql, qr, f);
    return;
  }
  if (o->srhd) {
    /* schemes.F90:400-430: "hllc" selects riemann_srhd_adi_hllc, every
     * other name HLL */
    if (o->c.riemann_solver == AMUN_RS_HLLC) riemann_srhd_hllc(o, ql, qr, f);
    else riemann_hll(o, ql, qr, f);               /* riemann_srhd_adi_hll */
    return;
  }
  if (o->sr) {
    if (o->c.riemann_solver == AMUN_RS_HLLC) riemann_srmhd_hllc(o, ql, qr, f);
    else riemann_hll(o, ql, qr, f);               /* riemann_srmhd_adi_hll */
    return;
  }
  switch (o->c.riemann_solver) {
  case AMUN_RS_HLLC:
    if (o->mhd) riemann_mhd_hllc(o, ql, qr, f);
    else riemann_hd_hllc(o, ql, qr, f);
    break;
  case AMUN_RS_HLLD: riemann_mhd_hlld(o, ql, qr, f); break;
  case AMUN_RS_ROE:
    if (o->mhd) riemann_mhd_roe(o, ql, qr, f);
    else riemann_hd_roe(o, ql, qr, f);
    break;
  default:           riemann_hll(o, ql, qr, f); break;
  }
}

/* ------------------------------------------------------------------------ */
/* update_flux for one block (schemes.F90:1080-1223 / 2970-3138)              */
/* f layout: Fortran f(NDIMS,nv,im,jm,km)                                     */
/* ------------------------------------------------------------------------ */
#define FI(o, d, p, i, j, k) (((((size_t)(k) * (o)->jm + (j)) * (o)->im + (i)) * (o)->nv + (p)) * (o)->nd + (d))

static void update_flux(ocfg* o, const double* q, double* f, double* qs)
{
  const int nv = o->nv, nd = o->nd, im = o->im, jm = o->jm, km = o->km;
  const size_t ncell = (size_t)im * jm * km;
  int perm[3][NVMAX];   /* rotated position -> original variable */
  int d, p, i, j, k;
  double ql[NVMAX], qr[NVMAX], fx[NVMAX];

  /* component rotation of schemes.F90:3063-3071 / 3100-3108 */
  for (d = 0; d < 3; d++)
    for (p = 0; p < nv; p++) perm[d][p] = p;
  perm[1][o->ivx] = o->ivy; perm[1][o->ivy] = o->ivz; perm[1][o->ivz] = o->ivx;
  perm[2][o->ivx] = o->ivz; perm[2][o->ivy] = o->ivx; perm[2][o->ivz] = o->ivy;
  if (o->mhd) {
    perm[1][o->ibx] = o->iby; perm[1][o->iby] = o->ibz; perm[1][o->ibz] = o->ibx;
    perm[2][o->ibx] = o->ibz; perm[2][o->iby] = o->ibx; perm[2][o->ibz] = o->iby;
  }

  memset(f, 0, ncell * nv * nd * sizeof(double));
  if (o->sr && o->c.states_4vec) {
    /* state_variables = 4vec (schemes.F90:4481-4517, 5040-5076): rho W
     * and the four-velocity W v are reconstructed, the face states turned
     * back into rho and v with W = sqrt(1 + |W v|^2) */
    static double* qq = NULL;
    static size_t nqq = 0;
    size_t n, m;
    if (nqq < ncell * nv) {
      free(qq);
      nqq = ncell * nv;
      qq = (double*)malloc(nqq * sizeof(double));
    }
    for (n = 0; n < ncell; n++) {
      const double* qc = q + n * nv;
      double* qo = qq + n * nv;
      const double vm = sqrt(1.0 - ((qc[o->ivx] * qc[o->ivx] + qc[o->ivy] * qc[o->ivy]) +
                                    qc[o->ivz] * qc[o->ivz]));
      for (p = 0; p < nv; p++) qo[p] = qc[p];
      qo[o->idn] = qc[o->idn] / vm;
      qo[o->ivx] = qc[o->ivx] / vm;
      qo[o->ivy] = qc[o->ivy] / vm;
      qo[o->ivz] = qc[o->ivz] / vm;
    }
    reconstruct_interfaces(o, qq, qs);
    for (m = 0; m < (size_t)nd * 2 * ncell; m++) {
      double* qf = qs + m * nv;
      const double vm = sqrt(1.0 + ((qf[o->ivx] * qf[o->ivx] + qf[o->ivy] * qf[o->ivy]) +
                                    qf[o->ivz] * qf[o->ivz]));
      qf[o->idn] = qf[o->idn] / vm;
      qf[o->ivx] = qf[o->ivx] / vm;
      qf[o->ivy] = qf[o->ivy] / vm;
      qf[o->ivz] = qf[o->ivz] / vm;
    }
  } else {
    reconstruct_interfaces(o, q, qs);
  }

#define QSV(d, s, i, j, k, p) qs[((size_t)((d) * 2 + (s)) * ncell + ((size_t)(k) * jm + (j)) * im + (i)) * nv + (p)]
  for (d = 0; d < nd; d++) {
    int n1lo, n1hi, n2lo, n2hi, a, b, c, len;
    if (d == 0) { n1lo = o->jbl; n1hi = o->jeu; n2lo = o->kbl; n2hi = o->keu; len = im; }
    else if (d == 1) { n1lo = o->ibl; n1hi = o->ieu; n2lo = o->kbl; n2hi = o->keu; len = jm; }
    else { n1lo = o->ibl; n1hi = o->ieu; n2lo = o->jbl; n2hi = o->jeu; len = km; }
    for (b = n2lo; b <= n2hi; b++)
      for (a = n1lo; a <= n1hi; a++)
        for (c = 0; c < len; c++) {
          if (d == 0) { i = c; j = a; k = b; }
          else if (d == 1) { i = a; j = c; k = b; }
          else { i = a; j = b; k = c; }
          for (p = 0; p < nv; p++) {
            ql[p] = QSV(d, 0, i, j, k, perm[d][p]);
            qr[p] = QSV(d, 1, i, j, k, perm[d][p]);
          }
          riemann(o, ql, qr, fx);
          for (p = 0; p < nv; p++) f[FI(o, d, perm[d][p], i, j, k)] = fx[p];
        }
  }
#undef QSV
}

/* ------------------------------------------------------------------------ */
/* block grid with ghost exchange                                             */
/* ------------------------------------------------------------------------ */
#define OMAXLEV 12
typedef struct {
  ocfg o;
  int ir, jr, kr, nblk;
  double *u0, *u1, *q, *f, *du, *qs;    /* u0,u1,q: nblk blocks */
  double time, dt, dtn;
  long step;
  /* multi-level meshes (leaf blocks only) */
  int amr, nlev;                         /* levels 1..nlev may hold blocks  */
  int *lev, *bx, *by, *bz;               /* level, position in blocks there */
  int *lut[OMAXLEV + 1];                 /* position -> block id or -1      */
  double lvd[OMAXLEV + 1][6];            /* adx, ady, adz, adxi, adyi, adzi */
  unsigned char* gflag;                  /* ghost cell filled by prolongation */
  double* fall;                          /* fluxes of every block           */
  int skip_fix;                          /* tests: leave the coarse fluxes  */
} ogrid;

static size_t bsize(const ocfg* o) { return (size_t)o->im * o->jm * o->km * o->nv; }

static int nbr_block(const ogrid* g, int b, int di, int dj, int dk)
{
  int bx = b % g->ir, by = (b / g->ir) % g->jr, bz = b / (g->ir * g->jr);
  bx = (bx + di + g->ir) % g->ir;
  by = (by + dj + g->jr) % g->jr;
  bz = (bz + dk + g->kr) % g->kr;
  return (bz * g->jr + by) * g->ir + bx;
}

/* block_boundary_specific (boundaries.F90:5024-5637) for the open, outflow
 * and reflective types: the ghost layers of axis n on side s (0 low, 1 high)
 * over the block's whole transverse range, ghosts included */
static void block_boundary_specific(const ocfg* o, double* qb, int n, int s, int type)
{
  const int nv = o->nv, ng = o->ng;
  const int vn = n == 0 ? o->ivx : (n == 1 ? o->ivy : o->ivz);
  const int bn = o->mhd ? (n == 0 ? o->ibx : (n == 1 ? o->iby : o->ibz)) : -1;
  const int lo = ng, hi = ng + o->nc - 1;            /* ib, ie (0-based) */
  const int m1 = n == 0 ? o->jm : o->im, m2 = n == 2 ? o->jm : o->km;
  int a, b, c, p;
  for (b = 0; b < m2; b++)
    for (a = 0; a < m1; a++)
      for (c = 1; c <= ng; c++) {
        /* target ghost t and source cell r along axis n */
        int t = s == 0 ? lo - c : hi + c, r;
        double* qt;
        const double* qr;
        if (type == AMUN_BND_REFLECTIVE) r = s == 0 ? lo - 1 + c : hi + 1 - c;
        else r = s == 0 ? lo : hi;
#define AT(x) (n == 0 ? &qb[QI(o, 0, x, a, b)] : (n == 1 ? &qb[QI(o, 0, a, x, b)] : &qb[QI(o, 0, a, b, x)]))
        qt = AT(t);
        qr = AT(r);
#undef AT
        for (p = 0; p < nv; p++) qt[p] = qr[p];
        if (type == AMUN_BND_OUTFLOW)
          qt[vn] = s == 0 ? fmin_r(0.0, qr[vn]) : fmax_r(0.0, qr[vn]);
        if (type == AMUN_BND_REFLECTIVE) {
          qt[vn] = -qr[vn];
          if (bn >= 0) qt[bn] = -qr[bn];
        }
      }
}

/* boundary_variables for a single-level mesh: every ghost cell of q is the
 * neighbour's interior value (face, edge and corner copies) where the
 * neighbour exists (periodic axes wrap), then boundaries_specific fills the
 * ghost layers at the non-periodic domain sides (per block, x, y, z in
 * turn), then update_ghost_cells converts the ghost q to u */
static void boundary_variables(ogrid* g, double* u)
{
  ocfg* o = &g->o;
  const size_t bs = bsize(o);
  const int* bt = o->c.bnd_type;
  int b, i, j, k;
  for (b = 0; b < g->nblk; b++) {
    double* qb = g->q + b * bs;
    for (k = 0; k < o->km; k++)
      for (j = 0; j < o->jm; j++)
        for (i = 0; i < o->im; i++) {
          int gi = i - o->ng, gj = j - o->ng, gk = (o->nd == 3) ? k - o->ng : 0;
          int di = gi < 0 ? -1 : (gi >= o->nc ? 1 : 0);
          int dj = gj < 0 ? -1 : (gj >= o->nc ? 1 : 0);
          int dk = (o->nd == 3) ? (gk < 0 ? -1 : (gk >= o->nc ? 1 : 0)) : 0;
          int nb, li, lj, lk;
          if (di == 0 && dj == 0 && dk == 0) continue;
          {
            /* no neighbour across a non-periodic domain side */
            const int px = b % g->ir + di, py = (b / g->ir) % g->jr + dj;
            const int pz = b / (g->ir * g->jr) + dk;
            if ((bt[0] && (px < 0 || px >= g->ir)) || (bt[2] && (py < 0 || py >= g->jr)) ||
                (o->nd == 3 && bt[4] && (pz < 0 || pz >= g->kr)))
              continue;
          }
          nb = nbr_block(g, b, di, dj, dk);
          li = gi - di * o->nc + o->ng;
          lj = gj - dj * o->nc + o->ng;
          lk = (o->nd == 3) ? gk - dk * o->nc + o->ng : 0;
          memcpy(&qb[QI(o, 0, i, j, k)], &g->q[nb * bs + QI(o, 0, li, lj, lk)],
                 o->nv * sizeof(double));
          prim2cons(o, &qb[QI(o, 0, i, j, k)], &u[b * bs + QI(o, 0, i, j, k)]);
        }
  }
  if (!(bt[0] || bt[2] || bt[4])) return;
  for (b = 0; b < g->nblk; b++) {
    const int pos[3] = {b % g->ir, (b / g->ir) % g->jr, b / (g->ir * g->jr)};
    const int nbk[3] = {g->ir, g->jr, g->kr};
    double* qb = g->q + b * bs;
    int n;
    for (n = 0; n < o->nd; n++) {
      if (!bt[2 * n]) continue;
      if (pos[n] == 0) block_boundary_specific(o, qb, n, 0, bt[2 * n]);
      if (pos[n] == nbk[n] - 1) block_boundary_specific(o, qb, n, 1, bt[2 * n + 1]);
    }
    for (k = 0; k < o->km; k++)
      for (j = 0; j < o->jm; j++)
        for (i = 0; i < o->im; i++) {
          int gi = i - o->ng, gj = j - o->ng, gk = (o->nd == 3) ? k - o->ng : 0;
          if (gi >= 0 && gi < o->nc && gj >= 0 && gj < o->nc && gk >= 0 && (o->nd == 2 || gk < o->nc))
            continue;
          prim2cons(o, &qb[QI(o, 0, i, j, k)], &u[b * bs + QI(o, 0, i, j, k)]);
        }
  }
}

/* ---- multi-level meshes ----------------------------------------------- */
/* Every ghost cell takes the value the reference's boundary_variables leaves
 * there (boundaries.F90:372-479).  Under the 2:1 level balance of AMUN's
 * meshes the region a ghost cell lies in is covered by exactly one leaf, and
 * each of the reference's copy, restriction and prolongation regions fills
 * the cells of its neighbour's coverage:
 *   same level  -> copy of the neighbour's cell (boundaries_*_copy);
 *   finer leaf  -> average of the 2^d covering cells (block_*_restrict,
 *                  boundaries.F90:5637, 5919, 6256; the summation order is
 *                  theirs);
 *   coarser leaf-> the parent cell plus limited slopes (block_*_prolong,
 *                  5762, 6071, 6351) from the parent's neighbours, which the
 *                  copies and restrictions have filled before.
 * The regions overlap in places (the edge prolongation covers corner cells,
 * boundaries.F90:3445-3470 via edges_gp, coordinates.F90:428-436); every
 * overlapping write of a cell carries the same value, so the order of the
 * reference's passes matters only between restriction and prolongation. */
static long wrapc(long x, long n) { x %= n; return x < 0 ? x + n : x; }

static long lev_cells(const ogrid* g, int L, int d)
{
  const int r = d == 0 ? g->ir : (d == 1 ? g->jr : g->kr);
  return (long)r * g->o.nc << (L - 1);
}

/* leaf block covering cell (x, y, z) of level L (already wrapped), or -1 */
static int owner(const ogrid* g, int L, long x, long y, long z)
{
  const int nc = g->o.nc;
  const long nbx = (long)g->ir << (L - 1), nby = (long)g->jr << (L - 1);
  if (L < 1 || L > g->nlev) return -1;
  return g->lut[L][((z / nc) * nby + y / nc) * nbx + x / nc];
}

/* padded local index of global cell (x, y, z) in block b */
static size_t local_cell(const ogrid* g, int b, long x, long y, long z)
{
  const ocfg* o = &g->o;
  const int i = (int)(x - (long)g->bx[b] * o->nc) + o->ng;
  const int j = (int)(y - (long)g->by[b] * o->nc) + o->ng;
  const int k = (o->nd == 3) ? (int)(z - (long)g->bz[b] * o->nc) + o->ng : 0;
  return QI(o, 0, i, j, k);
}

static int boundary_variables_amr(ogrid* g, double* u)
{
  ocfg* o = &g->o;
  const size_t bs = bsize(o), ncp = (size_t)o->im * o->jm * o->km;
  const int nv = o->nv, nd = o->nd;
  int b, i, j, k, p, pass;
  for (pass = 0; pass < 2; pass++)
    for (b = 0; b < g->nblk; b++) {
      const int L = g->lev[b];
      const long nx = lev_cells(g, L, 0), ny = lev_cells(g, L, 1);
      const long nz = nd == 3 ? lev_cells(g, L, 2) : 1;
      double* qb = g->q + b * bs;
      for (k = 0; k < o->km; k++)
        for (j = 0; j < o->jm; j++)
          for (i = 0; i < o->im; i++) {
            const int gi = i - o->ng, gj = j - o->ng, gk = nd == 3 ? k - o->ng : 0;
            long x, y, z;
            int nb;
            const size_t lc = QI(o, 0, i, j, k);
            if (gi >= 0 && gi < o->nc && gj >= 0 && gj < o->nc && gk >= 0 && gk < o->nc)
              continue;
            x = wrapc((long)g->bx[b] * o->nc + gi, nx);
            y = wrapc((long)g->by[b] * o->nc + gj, ny);
            z = nd == 3 ? wrapc((long)g->bz[b] * o->nc + gk, nz) : 0;
            if (pass == 0) {
              g->gflag[b * ncp + lc / nv] = 0;
              if ((nb = owner(g, L, x, y, z)) >= 0) {
                memcpy(&qb[lc], &g->q[nb * bs + local_cell(g, nb, x, y, z)],
                       nv * sizeof(double));
              } else if ((nb = owner(g, L + 1, 2 * x, 2 * y, 2 * z)) >= 0) {
                /* restriction: il/ip, jl/jp, kl/kp = the lower/upper child */
                const double* F = g->q + nb * bs;
                const size_t lll = local_cell(g, nb, 2 * x, 2 * y, 2 * z);
                const size_t di = nv, dj = (size_t)o->im * nv, dk = (size_t)o->im * o->jm * nv;
                for (p = 0; p < nv; p++) {
                  const double* a = F + lll + p;
                  if (nd == 2)
                    qb[lc + p] = 2.50e-01 * ((a[0] + a[di + dj]) + (a[dj] + a[di]));
                  else
                    qb[lc + p] = 1.25e-01 * (((a[0] + a[di + dj + dk]) + (a[dk] + a[di + dj]))
                                           + ((a[dj + dk] + a[di]) + (a[dj] + a[di + dk])));
                }
              } else if (owner(g, L - 1, x / 2, y / 2, z / 2) >= 0) {
                g->gflag[b * ncp + lc / nv] = 1;
              } else {
                return 700;   /* hole in the mesh or level jump above one */
              }
            } else if (g->gflag[b * ncp + lc / nv]) {
              /* prolongation from the parent cell of coarse leaf C */
              const long cx = x / 2, cy = y / 2, cz = z / 2;
              const int C = owner(g, L - 1, cx, cy, cz);
              const double* Q = g->q + C * bs;
              const size_t c0 = local_cell(g, C, cx, cy, cz);
              const size_t st[3] = {(size_t)nv, (size_t)o->im * nv, (size_t)o->im * o->jm * nv};
              const int sx = (int)(x & 1), sy = (int)(y & 1), sz = (int)(z & 1);
              int d;
              for (d = 0; d < nd; d++)
                if (g->gflag[C * ncp + (c0 - st[d]) / nv] || g->gflag[C * ncp + (c0 + st[d]) / nv])
                  return 701;  /* the slope would read a prolonged cell */
              for (p = 0; p < nv; p++) {
                const double qc = Q[c0 + p];
                double dq[3] = {0.0, 0.0, 0.0}, v;
                for (d = 0; d < nd; d++) {
                  const double dql = qc - Q[c0 - st[d] + p], dqr = Q[c0 + st[d] + p] - qc;
                  dq[d] = lim(o->c.prol_limiter, 0.5, dql, dqr);
                }
                if (p == o->idn || p == o->ipr) {
                  if (qc > 0.0) {
                    int it;
                    for (it = 0; it < OMAXHALV; it++) {
                      double sa = fabs(dq[0]) + fabs(dq[1]);
                      if (nd == 3) sa = sa + fabs(dq[2]);
                      if (!(qc <= sa)) break;
                      for (d = 0; d < nd; d++) dq[d] = 0.5 * dq[d];
                    }
                  } else {
                    o->npos_events++;
                    dq[0] = dq[1] = dq[2] = 0.0;
                  }
                }
                for (d = 0; d < nd; d++) dq[d] = 0.5 * dq[d];
                if (nd == 2) {
                  const double dq1 = dq[0] + dq[1], dq2 = dq[0] - dq[1];
                  const int s = sx + 2 * sy;
                  v = s == 0 ? qc - dq1 : (s == 1 ? qc + dq2 : (s == 2 ? qc - dq2 : qc + dq1));
                } else {
                  const double dq1 = dq[0] + dq[1] + dq[2], dq2 = dq[0] - dq[1] - dq[2];
                  const double dq3 = dq[0] - dq[1] + dq[2], dq4 = dq[0] + dq[1] - dq[2];
                  switch (sx + 2 * sy + 4 * sz) {
                  case 0: v = qc - dq1; break;
                  case 1: v = qc + dq2; break;
                  case 2: v = qc - dq3; break;
                  case 3: v = qc + dq4; break;
                  case 4: v = qc - dq4; break;
                  case 5: v = qc + dq3; break;
                  case 6: v = qc - dq2; break;
                  default: v = qc + dq1; break;
                  }
                }
                qb[lc + p] = v;
              }
            }
          }
    }
  /* update_ghost_cells: u of every ghost cell */
  for (b = 0; b < g->nblk; b++)
    for (k = 0; k < o->km; k++)
      for (j = 0; j < o->jm; j++)
        for (i = 0; i < o->im; i++) {
          const int gi = i - o->ng, gj = j - o->ng, gk = nd == 3 ? k - o->ng : 0;
          if (gi >= 0 && gi < o->nc && gj >= 0 && gj < o->nc && gk >= 0 && gk < o->nc)
            continue;
          prim2cons(o, &g->q[b * bs + QI(o, 0, i, j, k)], &u[b * bs + QI(o, 0, i, j, k)]);
        }
  return 0;
}

/* boundary_fluxes (boundaries.F90:494-1060) with block_update_flux
 * (6539-6614): the flux through a coarse block's face with finer
 * neighbours becomes the average of the fine fluxes through it, 2 (2D) or 4
 * (3D) per coarse face, summed in the reference's order */
static void boundary_fluxes(ogrid* g)
{
  ocfg* o = &g->o;
  const size_t bs = bsize(o);
  const int nd = o->nd, nv = o->nv, nc = o->nc;
  int b, n, side, a1, a2, p;
  for (b = 0; b < g->nblk; b++) {
    const int L = g->lev[b];
    long org[3];
    if (L >= g->nlev) continue;
    org[0] = (long)g->bx[b] * nc; org[1] = (long)g->by[b] * nc;
    org[2] = nd == 3 ? (long)g->bz[b] * nc : 0;
    for (n = 0; n < nd; n++)
      for (side = 0; side < 2; side++)
        for (a2 = 0; a2 < (nd == 3 ? nc : 1); a2++)
          for (a1 = 0; a1 < nc; a1++) {
            /* coarse cell across the face, and this block's face cell */
            long c[3], fc[3];
            int t1 = n == 0 ? 1 : 0, t2 = n == 2 ? 1 : 2, F, d;
            size_t fo[4];
            double* fb;
            c[n] = side == 0 ? org[n] - 1 : org[n] + nc;
            c[t1] = org[t1] + a1;
            if (nd == 3) c[t2] = org[t2] + a2; else c[2] = 0;
            for (d = 0; d < nd; d++) c[d] = wrapc(c[d], lev_cells(g, L, d));
            if (owner(g, L, c[0], c[1], c[2]) >= 0) continue;
            F = owner(g, L + 1, 2 * c[0], 2 * c[1], 2 * c[2]);
            if (F < 0) continue;
            /* the fine cells next to the face: the upper cell layer of a
             * lower neighbour (flux at its ie), the lower layer of an upper
             * neighbour (flux at ibl = ib - 1) */
            for (d = 0; d < 3; d++) fc[d] = 2 * c[d];
            if (side == 0) fc[n] = 2 * c[n] + 1;
            for (d = 0; d < 4; d++) {
              long e[3] = {fc[0], fc[1], fc[2]};
              size_t cell;
              e[t1] += d & 1;
              if (nd == 3) e[t2] += d >> 1;
              cell = local_cell(g, F, e[0], e[1], e[2]) / nv;
              if (side == 1) cell -= (n == 0 ? 1 : (n == 1 ? o->im : (size_t)o->im * o->jm));
              fo[d] = cell;
              if (nd == 2 && d == 1) break;
            }
            {
              /* this block's face: f at ibl (lower) or ie (upper) along n */
              int ii[3] = {o->ng, o->ng, nd == 3 ? o->ng : 0};
              size_t cb;
              ii[n] = side == 0 ? o->ng - 1 : o->ng + nc - 1;
              ii[t1] += a1;
              if (nd == 3) ii[t2] += a2;
              cb = QI(o, 0, ii[0], ii[1], ii[2]) / nv;
              fb = g->fall + b * bs * nd;
              for (p = 0; p < nv; p++) {
                const double* ff = g->fall + F * bs * nd;
                double v;
#define FF(cell) ff[((cell) * nv + p) * nd + n]
                if (nd == 2)
                  v = (FF(fo[0]) + FF(fo[1])) / 2.0;
                else
                  v = ((FF(fo[0]) + FF(fo[3])) + (FF(fo[2]) + FF(fo[1]))) / 4.0;
#undef FF
                fb[(cb * nv + p) * nd + n] = v;
              }
            }
          }
  }
}

/* update_primitive_variables (equations.F90:1088-1148): interior cons2prim,
 * the ghost layers are then replaced by the exchange */
static void update_variables(ogrid* g, double* u)
{
  ocfg* o = &g->o;
  const size_t bs = bsize(o);
  int b, i, j, k;
  for (b = 0; b < g->nblk; b++)
    for (k = o->kb; k <= o->ke; k++)
      for (j = o->jb; j <= o->je; j++)
        for (i = o->ib; i <= o->ie; i++)
          cons2prim(o, &u[b * bs + QI(o, 0, i, j, k)], &g->q[b * bs + QI(o, 0, i, j, k)]);
  if (g->amr) {
    if (!g->o.amr_err) g->o.amr_err = boundary_variables_amr(g, u);
  } else {
    boundary_variables(g, u);
  }
}

/* update_increment, evolution.F90:1898-1977 (cell sizes of the block's
 * level) */
static void update_increment_dx(const ocfg* o, const double* f, double* du, double adxi,
                                double adyi, double adzi)
{
  int p, i, j, k;
  memset(du, 0, bsize(o) * sizeof(double));
  for (k = o->kbl; k <= o->keu; k++)
    for (j = o->jbl; j <= o->jeu; j++)
      for (i = o->ibl; i <= o->ieu; i++)
        for (p = 0; p < o->nv; p++) {
          double v;
          v = -adxi * (f[FI(o, 0, p, i, j, k)] - f[FI(o, 0, p, i - 1, j, k)])
              - adyi * (f[FI(o, 1, p, i, j, k)] - f[FI(o, 1, p, i, j - 1, k)]);
          if (o->nd == 3)
            v = v - adzi * (f[FI(o, 2, p, i, j, k)] - f[FI(o, 2, p, i, j, k - 1)]);
          du[QI(o, p, i, j, k)] = v;
        }
}

static void update_increment(const ocfg* o, const double* f, double* du)
{
  update_increment_dx(o, f, du, o->adxi, o->adyi, o->adzi);
}

/* new_time_step, evolution.F90:379-493: dt = cfl dx_min / max(cmax + 2 max(nu, eta) / dx_min,
 * eps) */
static void new_time_step(ogrid* g, double dtnext)
{
  ocfg* o = &g->o;
  const size_t bs = bsize(o);
  const double eps = 2.2250738585072014e-308;   /* tiny(1d0) */
  double dxmin;
  int b, i, j, k;
  o->cmax = eps;
  for (b = 0; b < g->nblk; b++) {
    double ms = 0.0;
    for (k = o->kb; k <= o->ke; k++)
      for (j = o->jb; j <= o->je; j++)
        for (i = o->ib; i <= o->ie; i++)
          ms = fmax_r(ms, maxspeed_cell(o, &g->q[b * bs + QI(o, 0, i, j, k)]));
    o->cmax = fmax_r(o->cmax, ms);
  }
  o->cmax2 = o->cmax * o->cmax;
  if (g->amr) {                 /* dx_min of the highest level in use */
    int mlev = 1;
    for (b = 0; b < g->nblk; b++) mlev = g->lev[b] > mlev ? g->lev[b] : mlev;
    dxmin = fmin_r(g->lvd[mlev][0], g->lvd[mlev][1]);
    if (o->nd == 3) dxmin = fmin_r(dxmin, g->lvd[mlev][2]);
  } else {
    dxmin = fmin_r(o->adx, o->ady);
    if (o->nd == 3) dxmin = fmin_r(dxmin, o->adz);
  }
  g->dtn = o->cflx * dxmin /
           fmax_r(o->cmax + 2.0 * fmax_r(o->c.viscosity, o->c.resistivity) / dxmin, eps);
  if (dtnext > 0.0) g->dt = fmin_r(g->dtn, dtnext);
}

/* the gravitational source terms of update_sources (sources.F90:288-335)
 * for the uniform acceleration (0, gacc, 0) of the Rayleigh-Taylor problem
 * (gravity.F90:283-330): du(m) += rho g, du(E) += rho g . v, in the
 * reference's order (the zero components included) */
static void gravity_sources(const ocfg* o, const double* q, double* du)
{
  int i, j, k;
  if (!o->c.gravity) return;
  for (k = o->kb; k <= o->ke; k++)
    for (j = o->jb; j <= o->je; j++)
      for (i = o->ib; i <= o->ie; i++) {
        const size_t c = QI(o, 0, i, j, k);
        const double* qc = q + c;
        double* dc = du + c;
        const double gx = qc[o->idn] * 0.0, gy = qc[o->idn] * o->c.gacc;
        const double gz = qc[o->idn] * 0.0;
        dc[o->imx] = dc[o->imx] + gx;
        dc[o->imy] = dc[o->imy] + gy;
        if (o->nd == 3) dc[o->imz] = dc[o->imz] + gz;
        if (o->ien >= 0) {
          if (o->nd == 3)
            dc[o->ien] = ((dc[o->ien] + gx * qc[o->ivx]) + gy * qc[o->ivy]) + gz * qc[o->ivz];
          else
            dc[o->ien] = (dc[o->ien] + gx * qc[o->ivx]) + gy * qc[o->ivy];
        }
      }
}

/* derivative_1st (operators.F90:538-615) of variable v at cell c along d:
 * the central difference of the interior cells */
static double d1c(const double* q, size_t c, size_t st, int v, double h)
{
  return 0.5 * (q[c + st + v] - q[c - st + v]) / h;
}

/* the viscous stress of update_sources (sources.F90:340-384) at cell c:
 * velocity gradients g[a][d] = d v_a / d x_d (zero beyond ndims, as
 * gradient leaves them), gc = nu rho, fc = 2 gc,
 *   t(a,a) = fc (g(a,a) - div v / 3),  t(a,b) = gc (g(a,b) + g(b,a)) */
static void visc_tau(const ocfg* o, const double* q, size_t c, const size_t st[3],
                     const double h[3], double t[3][3])
{
  double g[3][3], gc, fc, divv;
  int a, d;
  for (a = 0; a < 3; a++)
    for (d = 0; d < 3; d++) g[a][d] = d < o->nd ? d1c(q, c, st[d], o->ivx + a, h[d]) : 0.0;
  gc = o->c.viscosity * q[c + o->idn];
  fc = 2.0 * gc;
  divv = ((g[0][0] + g[1][1]) + g[2][2]) / 3.0;
  t[0][0] = fc * (g[0][0] - divv);
  t[1][1] = fc * (g[1][1] - divv);
  t[2][2] = fc * (g[2][2] - divv);
  t[0][1] = gc * (g[0][1] + g[1][0]);
  t[0][2] = gc * (g[0][2] + g[2][0]);
  t[1][2] = gc * (g[1][2] + g[2][1]);
  t[1][0] = t[0][1];
  t[2][0] = t[0][2];
  t[2][1] = t[1][2];
}

/* the viscous source terms of update_sources (sources.F90:340-467) on the
 * interior cells of one block: du(m_a) += div t(a,:), and with an energy
 * equation du(E) += div (t v); the divergences are derivative_1st's central
 * differences summed over the directions in order (operators.F90:181-243).
 * Every system, when viscosity > 0. */
static void viscous_sources(const ocfg* o, const double* q, double* du, const double h[3])
{
  const int nd = o->nd, nv = o->nv;
  const size_t st[3] = {(size_t)nv, (size_t)o->im * nv, (size_t)o->im * o->jm * nv};
  int i, j, k, d, a, b;
  if (!(o->c.viscosity > 0.0)) return;
  for (k = o->kb; k <= o->ke; k++)
    for (j = o->jb; j <= o->je; j++)
      for (i = o->ib; i <= o->ie; i++) {
        const size_t c = QI(o, 0, i, j, k);
        double tp[3][3][3], tm[3][3][3], gp[3], gm[3], s;
        for (d = 0; d < nd; d++) {
          visc_tau(o, q, c + st[d], st, h, tp[d]);
          visc_tau(o, q, c - st[d], st, h, tm[d]);
        }
        for (a = 0; a < 3; a++) {
          s = 0.0;
          for (d = 0; d < nd; d++) s = s + 0.5 * (tp[d][a][d] - tm[d][a][d]) / h[d];
          du[c + o->imx + a] = du[c + o->imx + a] + s;
        }
        if (o->ien >= 0) {
          s = 0.0;
          for (d = 0; d < nd; d++) {
            const double* vp = q + c + st[d] + o->ivx;
            const double* vm = q + c - st[d] + o->ivx;
            b = d;
            gp[b] = (vp[0] * tp[d][b][0] + vp[1] * tp[d][b][1]) + vp[2] * tp[d][b][2];
            gm[b] = (vm[0] * tm[d][b][0] + vm[1] * tm[d][b][1]) + vm[2] * tm[d][b][2];
            s = s + 0.5 * (gp[b] - gm[b]) / h[d];
          }
          du[c + o->ien] = du[c + o->ien] + s;
        }
      }
}

/* the resistive source terms of update_sources (sources.F90:543-598) on
 * the interior cells of one block, for the systems with a magnetic field:
 * du(B) += eta lap B (derivative_2nd's ((u(i+1) + u(i-1)) - 2 u(i)) / dh^2
 * summed over the directions, operators.F90:453-711), and with an energy
 * equation du(E) += B . (eta lap B), then du(E) += eta |curl B|^2 (curl
 * from derivative_1st, operators.F90:328-434) */
static void resistive_sources(const ocfg* o, const double* q, double* du, const double h[3])
{
  const int nd = o->nd, nv = o->nv, bx = o->ibx;
  const size_t st[3] = {(size_t)nv, (size_t)o->im * nv, (size_t)o->im * o->jm * nv};
  const double eta = o->c.resistivity;
  int i, j, k, d, a;
  if (!(eta > 0.0) || o->ibx < 0) return;
  for (k = o->kb; k <= o->ke; k++)
    for (j = o->jb; j <= o->je; j++)
      for (i = o->ib; i <= o->ie; i++) {
        const size_t c = QI(o, 0, i, j, k);
        const double* qc = q + c;
        double el[3], cv[3], db;
        for (a = 0; a < 3; a++) {
          double lap = 0.0;
          for (d = 0; d < nd; d++) {
            const double w = ((q[c + st[d] + bx + a] + q[c - st[d] + bx + a]) - 2.0 * qc[bx + a]) /
                             (h[d] * h[d]);
            lap = d == 0 ? w : lap + w;
          }
          el[a] = eta * lap;
          du[c + bx + a] = du[c + bx + a] + el[a];
        }
        if (o->ien >= 0) {
          du[c + o->ien] = du[c + o->ien] +
                           ((qc[bx] * el[0] + qc[bx + 1] * el[1]) + qc[bx + 2] * el[2]);
          if (nd == 3) {
            cv[0] = d1c(q, c, st[1], bx + 2, h[1]) - d1c(q, c, st[2], bx + 1, h[2]);
            cv[1] = d1c(q, c, st[2], bx, h[2]) - d1c(q, c, st[0], bx + 2, h[0]);
          } else {
            cv[0] = d1c(q, c, st[1], bx + 2, h[1]);
            cv[1] = -d1c(q, c, st[0], bx + 2, h[0]);
          }
          cv[2] = d1c(q, c, st[0], bx + 1, h[0]) - d1c(q, c, st[1], bx, h[1]);
          db = (cv[0] * cv[0] + cv[1] * cv[1]) + cv[2] * cv[2];
          du[c + o->ien] = du[c + o->ien] + eta * db;
        }
      }
}

/* The GLM source terms of update_sources (sources.F90:479-539) on the
 * interior cells of one block, after update_increment: with db = div B and
 * grad psi from derivative_1st's central differences (operators.F90:181-308,
 * 0.5 (u(i+1) - u(i-1)) / dh, summed over the directions in order),
 *   EGLM (1): du(m) -= db B,  du(E) -= B . grad psi;
 *   HEGLM (2), in addition: du(B) -= db v,  du(E) -= db (v . B).
 * Only for systems with a magnetic field (ibx > 0 in the reference). */
static void glm_sources(const ocfg* o, const double* q, double* du, const double h[3])
{
  const int nd = o->nd, nv = o->nv;
  const size_t st[3] = {(size_t)nv, (size_t)o->im * nv, (size_t)o->im * o->jm * nv};
  int i, j, k, d;
  if (o->c.glm_source_terms <= 0 || o->ibx < 0) return;
  for (k = o->kb; k <= o->ke; k++)
    for (j = o->jb; j <= o->je; j++)
      for (i = o->ib; i <= o->ie; i++) {
        const size_t c = QI(o, 0, i, j, k);
        const double* qc = q + c;
        double* dc = du + c;
        double db = 0.0, gp[3] = {0.0, 0.0, 0.0};
        for (d = 0; d < nd; d++)
          db = db + 0.5 * (q[c + st[d] + o->ibx + d] - q[c - st[d] + o->ibx + d]) / h[d];
        dc[o->imx] = dc[o->imx] - db * qc[o->ibx];
        dc[o->imy] = dc[o->imy] - db * qc[o->iby];
        dc[o->imz] = dc[o->imz] - db * qc[o->ibz];
        if (o->ien >= 0 && o->ibp >= 0) {
          for (d = 0; d < nd; d++)
            gp[d] = 0.5 * (q[c + st[d] + o->ibp] - q[c - st[d] + o->ibp]) / h[d];
          dc[o->ien] = dc[o->ien] -
                       ((qc[o->ibx] * gp[0] + qc[o->iby] * gp[1]) + qc[o->ibz] * gp[2]);
        }
        if (o->c.glm_source_terms > 1) {
          dc[o->ibx] = dc[o->ibx] - db * qc[o->ivx];
          dc[o->iby] = dc[o->iby] - db * qc[o->ivy];
          dc[o->ibz] = dc[o->ibz] - db * qc[o->ivz];
          if (o->ien >= 0)
            dc[o->ien] = dc[o->ien] - db * ((qc[o->ivx] * qc[o->ibx] + qc[o->ivy] * qc[o->iby]) +
                                            qc[o->ivz] * qc[o->ibz]);
        }
      }
}

/* one stage: fluxes of u (through q), increment, combine into dst (which
 * may be a0 or a1: every element is read before it is written) */
static void stage2(ogrid* g, int kind, double ds, double* dst, const double* a0,
                   const double* a1, double ca, double cb)
{
  ocfg* o = &g->o;
  const size_t bs = bsize(o);
  int b;
  size_t n;
  if (g->amr) {                 /* update_fluxes: every block, then the correction */
    for (b = 0; b < g->nblk; b++) {
      const double* lv = g->lvd[g->lev[b]];
      o->h[0] = lv[0]; o->h[1] = lv[1]; o->h[2] = lv[2];
      update_flux(o, g->q + b * bs, g->fall + b * bs * o->nd, g->qs);
    }
    if (!g->skip_fix) boundary_fluxes(g);
  }
  for (b = 0; b < g->nblk; b++) {
    if (g->amr) {
      const double* lv = g->lvd[g->lev[b]];
      update_increment_dx(o, g->fall + b * bs * o->nd, g->du, lv[3], lv[4], lv[5]);
      gravity_sources(o, g->q + b * bs, g->du);
      viscous_sources(o, g->q + b * bs, g->du, lv);
      glm_sources(o, g->q + b * bs, g->du, lv);
      resistive_sources(o, g->q + b * bs, g->du, lv);
    } else {
      const double h[3] = {o->adx, o->ady, o->adz};
      update_flux(o, g->q + b * bs, g->f, g->qs);
      update_increment(o, g->f, g->du);
      gravity_sources(o, g->q + b * bs, g->du);
      viscous_sources(o, g->q + b * bs, g->du, h);
      glm_sources(o, g->q + b * bs, g->du, h);
      resistive_sources(o, g->q + b * bs, g->du, h);
    }
    for (n = 0; n < bs; n++) {
      double v;
      if (kind == 0)       /* u1 = u0 + ds du */
        v = a0[b * bs + n] + ds * g->du[n];
      else if (kind == 1)  /* rk3 2nd: 3/4 u0 + 1/4 u1 + ds du */
        v = ((3.0 / 4.0) * a0[b * bs + n] + (1.0 / 4.0) * a1[b * bs + n]) + ds * g->du[n];
      else if (kind == 2)  /* rk3 3rd: 1/3 u0 + 2/3 u1 + ds du */
        v = ((1.0 / 3.0) * a0[b * bs + n] + (2.0 / 3.0) * a1[b * bs + n]) + ds * g->du[n];
      else if (kind == 3)  /* rk2 2nd: 1/2 (u0 + u1 + dt du) */
        v = 0.5 * ((a0[b * bs + n] + a1[b * bs + n]) + ds * g->du[n]);
      else if (kind == 4)  /* ssprk: u1 + ds du */
        v = a1[b * bs + n] + ds * g->du[n];
      else if (kind == 5)  /* ssprk(m,2) last: fl u0 + fr (u1 + ds du) */
        v = ca * a0[b * bs + n] + cb * (a1[b * bs + n] + ds * g->du[n]);
      else if (kind == 6)  /* a u0 + b u1 + ds du */
        v = (ca * a0[b * bs + n] + cb * a1[b * bs + n]) + ds * g->du[n];
      else                 /* a u1 + b u0 + ds du */
        v = (ca * a1[b * bs + n] + cb * a0[b * bs + n]) + ds * g->du[n];
      dst[b * bs + n] = v;
    }
  }
}

static void stage(ogrid* g, int kind, double ds, double* dst, const double* a0,
                  const double* a1)
{
  stage2(g, kind, ds, dst, a0, a1, 0.0, 0.0);
}

static void psi_decay(ogrid* g, double* u)
{
  ocfg* o = &g->o;
  size_t c, ncell = (size_t)g->nblk * o->im * o->jm * o->km;
  if (o->ibp < 0) return;
  for (c = 0; c < ncell; c++) u[c * o->nv + o->ibp] = o->decay * u[c * o->nv + o->ibp];
}

static void advance(ogrid* g, double dtnext)
{
  ocfg* o = &g->o;
  new_time_step(g, dtnext);
  if (o->c.time_advance == AMUN_TA_EULER) {       /* evolve_euler 517 */
    stage(g, 0, g->dt, g->u0, g->u0, NULL);
    psi_decay(g, g->u0);
    update_variables(g, g->u0);
  } else if (o->c.time_advance == AMUN_TA_RK2) {  /* evolve_rk2 628 */
    stage(g, 0, g->dt, g->u1, g->u0, NULL);
    update_variables(g, g->u1);
    stage(g, 3, g->dt, g->u0, g->u0, g->u1);
    psi_decay(g, g->u0);
    update_variables(g, g->u0);
  } else if (o->c.time_advance == AMUN_TA_SSPRK_M2) {   /* evolve_ssprk2 791 */
    const double ft = 1.0 / (o->stages - 1), fl = 1.0 / o->stages, fr = 1.0 - fl;
    const double ds = ft * g->dt;
    int n;
    memcpy(g->u1, g->u0, bsize(o) * g->nblk * sizeof(double));
    for (n = 1; n <= o->stages - 1; n++) {
      stage(g, 4, ds, g->u1, g->u0, g->u1);
      update_variables(g, g->u1);
    }
    stage2(g, 5, ds, g->u0, g->u0, g->u1, fl, fr);
    psi_decay(g, g->u0);
    update_variables(g, g->u0);
  } else if (o->c.time_advance == AMUN_TA_SSPRK43) {    /* evolve_ssprk34 1228 */
    const double b1 = 1.0 / 2.0, b3 = 1.0 / 6.0, a31 = 2.0 / 3.0, a33 = 1.0 / 3.0;
    stage(g, 0, b1 * g->dt, g->u1, g->u0, NULL);
    update_variables(g, g->u1);
    stage(g, 4, b1 * g->dt, g->u1, g->u0, g->u1);
    update_variables(g, g->u1);
    stage2(g, 6, b3 * g->dt, g->u1, g->u0, g->u1, a31, a33);
    update_variables(g, g->u1);
    stage(g, 4, b1 * g->dt, g->u0, g->u0, g->u1);
    psi_decay(g, g->u0);
    update_variables(g, g->u0);
  } else if (o->c.time_advance == AMUN_TA_SSPRK53) {    /* evolve_ssprk35 1492 */
    const double b1 = 3.77268915331368e-01, b3 = 2.42995220537396e-01;
    const double b4 = 2.38458932846290e-01, b5 = 2.87632146308408e-01;
    const double a31 = 3.55909775063327e-01, a33 = 6.44090224936673e-01;
    const double a41 = 3.67933791638137e-01, a44 = 6.32066208361863e-01;
    const double a53 = 2.37593836598569e-01, a55 = 7.62406163401431e-01;
    stage(g, 0, b1 * g->dt, g->u1, g->u0, NULL);
    update_variables(g, g->u1);
    stage(g, 4, b1 * g->dt, g->u1, g->u0, g->u1);
    update_variables(g, g->u1);
    stage2(g, 6, b3 * g->dt, g->u1, g->u0, g->u1, a31, a33);
    update_variables(g, g->u1);
    stage2(g, 6, b4 * g->dt, g->u0, g->u0, g->u1, a41, a44);
    update_variables(g, g->u0);
    stage2(g, 7, b5 * g->dt, g->u0, g->u0, g->u1, a53, a55);
    psi_decay(g, g->u0);
    update_variables(g, g->u0);
  } else {                                         /* evolve_rk3 1005 */
    stage(g, 0, g->dt, g->u1, g->u0, NULL);
    update_variables(g, g->u1);
    stage(g, 1, (1.0 / 4.0) * g->dt, g->u1, g->u0, g->u1);
    update_variables(g, g->u1);
    stage(g, 2, (2.0 / 3.0) * g->dt, g->u0, g->u0, g->u1);
    psi_decay(g, g->u0);
    update_variables(g, g->u0);
  }
}

/* ------------------------------------------------------------------------ */
/* exported API (ctypes)                                                      */
/* ------------------------------------------------------------------------ */
int orc_create(const amun_cfg* c, void** handle)
{
  ogrid* g;
  int r;
  size_t bs, ncell;
  *handle = NULL;
  g = (ogrid*)calloc(1, sizeof(ogrid));
  if (!g) return 500;
  r = oc_setup(&g->o, c);
  if (!r) r = riemann_check(c);
  if (!r && c->reconstruction != AMUN_REC_TVD && c->reconstruction != AMUN_REC_WENO5Z &&
      c->reconstruction != AMUN_REC_WENO5YC && c->reconstruction != AMUN_REC_WENO5NS &&
      c->reconstruction != AMUN_REC_MP5 && c->reconstruction != AMUN_REC_MP7 &&
      c->reconstruction != AMUN_REC_GP && c->reconstruction != AMUN_REC_WENO3 &&
      c->reconstruction != AMUN_REC_LIMO3 && c->reconstruction != AMUN_REC_PPM &&
      c->reconstruction != AMUN_REC_CRWENO5Z && c->reconstruction != AMUN_REC_CRWENO5YC &&
      c->reconstruction != AMUN_REC_CRWENO5NS && c->reconstruction != AMUN_REC_CRMP5 &&
      c->reconstruction != AMUN_REC_CRMP5LD && c->reconstruction != AMUN_REC_CRMP7 &&
      c->reconstruction != AMUN_REC_MGP) r = 200;
  if (!r && (c->time_advance < AMUN_TA_EULER || c->time_advance > AMUN_TA_SSPRK53)) r = 400;
  if (!r) {
    int n;
    for (n = 0; n < 3; n++) {
      const int a = c->bnd_type[2 * n], b = c->bnd_type[2 * n + 1];
      if (a < 0 || a > AMUN_BND_REFLECTIVE || b < 0 || b > AMUN_BND_REFLECTIVE) r = 600;
      if ((a == AMUN_BND_PERIODIC) != (b == AMUN_BND_PERIODIC)) r = 600;
    }
  }
  if (r) { free(g); return r; }
  g->ir = c->xblocks; g->jr = c->yblocks; g->kr = (c->ndims == 3) ? c->zblocks : 1;
  g->nblk = g->ir * g->jr * g->kr;
  bs = bsize(&g->o);
  ncell = (size_t)g->o.im * g->o.jm * g->o.km;
  g->u0 = (double*)calloc(bs * g->nblk, sizeof(double));
  g->u1 = (double*)calloc(bs * g->nblk, sizeof(double));
  g->q  = (double*)calloc(bs * g->nblk, sizeof(double));
  g->f  = (double*)calloc(bs * g->o.nd, sizeof(double));
  g->du = (double*)calloc(bs, sizeof(double));
  g->qs = (double*)calloc(ncell * g->o.nv * 2 * g->o.nd, sizeof(double));
  g->dt = 1.0; g->dtn = 1.0;
  *handle = g;
  return 0;
}

void orc_destroy(void* h)
{
  ogrid* g = (ogrid*)h;
  int L;
  if (!g) return;
  free(g->u0); free(g->u1); free(g->q); free(g->f); free(g->du); free(g->qs);
  free(g->lev); free(g->bx); free(g->by); free(g->bz); free(g->gflag); free(g->fall);
  for (L = 0; L <= OMAXLEV; L++) free(g->lut[L]);
  free(g);
}

/* A multi-level mesh of nblk leaf blocks: meta[4 b + 0..3] = level and the
 * block's first interior cell (i0, j0, k0) in cells of its level; the base
 * level has the cfg's xblocks x yblocks (x zblocks) blocks, periodic.  Cell
 * sizes per level as coordinates.F90:330-356. */
int orc_create_amr(const amun_cfg* c, int nblk, const int32_t* meta, void** handle)
{
  ogrid* g;
  int r, b, L;
  size_t bs, ncell;
  const double xlen = c->xmax - c->xmin, ylen = c->ymax - c->ymin, zlen = c->zmax - c->zmin;
  *handle = NULL;
  if (nblk < 1) return 500;
  r = orc_create(c, (void**)&g);
  if (r) return r;
  if (c->nghosts % 2 || c->ncells < 2 * c->nghosts) { orc_destroy(g); return 600; }
  if (c->bnd_type[0] || c->bnd_type[2] || c->bnd_type[4]) { orc_destroy(g); return 600; }
  free(g->u0); free(g->u1); free(g->q);
  g->nblk = nblk;
  g->amr = 1;
  bs = bsize(&g->o);
  ncell = (size_t)g->o.im * g->o.jm * g->o.km;
  g->u0 = (double*)calloc(bs * nblk, sizeof(double));
  g->u1 = (double*)calloc(bs * nblk, sizeof(double));
  g->q  = (double*)calloc(bs * nblk, sizeof(double));
  g->fall = (double*)calloc(bs * g->o.nd * nblk, sizeof(double));
  g->gflag = (unsigned char*)calloc(ncell * nblk, 1);
  g->lev = (int*)calloc(nblk, sizeof(int));
  g->bx = (int*)calloc(nblk, sizeof(int));
  g->by = (int*)calloc(nblk, sizeof(int));
  g->bz = (int*)calloc(nblk, sizeof(int));
  g->nlev = 1;
  for (b = 0; b < nblk; b++) {
    g->lev[b] = meta[4 * b];
    if (g->lev[b] < 1 || g->lev[b] > OMAXLEV) { orc_destroy(g); return 500; }
    if (meta[4 * b + 1] % c->ncells || meta[4 * b + 2] % c->ncells ||
        meta[4 * b + 3] % c->ncells) { orc_destroy(g); return 500; }
    g->bx[b] = meta[4 * b + 1] / c->ncells;
    g->by[b] = meta[4 * b + 2] / c->ncells;
    g->bz[b] = meta[4 * b + 3] / c->ncells;
    if (g->lev[b] > g->nlev) g->nlev = g->lev[b];
  }
  for (L = 1; L <= g->nlev; L++) {
    const long nbx = (long)g->ir << (L - 1), nby = (long)g->jr << (L - 1);
    const long nbz = (g->o.nd == 3) ? (long)g->kr << (L - 1) : 1;
    const int ff = 1 << (L - 1), ni = g->o.nc * ff;
    long n;
    g->lut[L] = (int*)malloc(nbx * nby * nbz * sizeof(int));
    for (n = 0; n < nbx * nby * nbz; n++) g->lut[L][n] = -1;
    g->lvd[L][0] = xlen / (g->ir * ni);
    g->lvd[L][1] = ylen / (g->jr * ni);
    g->lvd[L][2] = (g->o.nd == 3) ? zlen / (g->kr * ni) : 1.0;
    g->lvd[L][3] = 1.0 / g->lvd[L][0];
    g->lvd[L][4] = 1.0 / g->lvd[L][1];
    g->lvd[L][5] = (g->o.nd == 3) ? 1.0 / g->lvd[L][2] : 1.0;
  }
  for (b = 0; b < nblk; b++) {
    L = g->lev[b];
    {
      const long nbx = (long)g->ir << (L - 1), nby = (long)g->jr << (L - 1);
      const long nbz = (g->o.nd == 3) ? (long)g->kr << (L - 1) : 1;
      const long id = ((long)g->bz[b] * nby + g->by[b]) * nbx + g->bx[b];
      if (g->bx[b] >= nbx || g->by[b] >= nby || g->bz[b] >= nbz || g->lut[L][id] >= 0) {
        orc_destroy(g);
        return 500;
      }
      g->lut[L][id] = b;
    }
  }
  *handle = g;
  return 0;
}

/* interiors of every block, u[b][k][j][i][v] (v fastest), then q, ghost fill
 * and the first time step estimate as orc_set_state */
int orc_set_blocks(void* h, const double* u, double dtnext)
{
  ogrid* g = (ogrid*)h;
  ocfg* o = &g->o;
  const size_t bs = bsize(o);
  const int kn = o->nd == 3 ? o->nc : 1;
  int b, i, j, k;
  for (b = 0; b < g->nblk; b++)
    for (k = 0; k < kn; k++)
      for (j = 0; j < o->nc; j++)
        for (i = 0; i < o->nc; i++)
          memcpy(&g->u0[b * bs + QI(o, 0, i + o->ib, j + o->jb, k + o->kb)],
                 &u[((((size_t)b * kn + k) * o->nc + j) * o->nc + i) * o->nv],
                 o->nv * sizeof(double));
  o->amr_err = 0;
  update_variables(g, g->u0);
  if (o->amr_err) return o->amr_err;
  new_time_step(g, dtnext);
  return 0;
}

/* Tests only: with skip != 0 the coarse faces keep their own fluxes (the
 * step without boundary_fluxes), to show what the correction changes. */
void orc_amr_skip_correction(void* h, int skip) { ((ogrid*)h)->skip_fix = skip; }

/* interiors of u0 (and, if q != NULL, every block's q with its ghosts) */
int orc_get_blocks(void* h, double* u, double* q)
{
  ogrid* g = (ogrid*)h;
  ocfg* o = &g->o;
  const size_t bs = bsize(o);
  const int kn = o->nd == 3 ? o->nc : 1;
  int b, i, j, k;
  for (b = 0; b < g->nblk; b++)
    for (k = 0; k < kn; k++)
      for (j = 0; j < o->nc; j++)
        for (i = 0; i < o->nc; i++)
          memcpy(&u[((((size_t)b * kn + k) * o->nc + j) * o->nc + i) * o->nv],
                 &g->u0[b * bs + QI(o, 0, i + o->ib, j + o->jb, k + o->kb)],
                 o->nv * sizeof(double));
  if (q) memcpy(q, g->q, bs * g->nblk * sizeof(double));
  return o->amr_err;
}

/* global array U[nv][nz][ny][nx] <-> block interiors */
static void xfer_global(ogrid* g, double* blk, double* glob, int to_blocks)
{
  ocfg* o = &g->o;
  const size_t bs = bsize(o);
  const size_t nx = (size_t)g->ir * o->nc, ny = (size_t)g->jr * o->nc;
  const size_t nz = (o->nd == 3) ? (size_t)g->kr * o->nc : 1;
  int b, i, j, k, p;
  for (b = 0; b < g->nblk; b++) {
    int bx = b % g->ir, by = (b / g->ir) % g->jr, bz = b / (g->ir * g->jr);
    for (k = o->kb; k <= o->ke; k++)
      for (j = o->jb; j <= o->je; j++)
        for (i = o->ib; i <= o->ie; i++)
          for (p = 0; p < o->nv; p++) {
            size_t gx = (size_t)bx * o->nc + (i - o->ib);
            size_t gy = (size_t)by * o->nc + (j - o->jb);
            size_t gz = (o->nd == 3) ? (size_t)bz * o->nc + (k - o->kb) : 0;
            size_t gi = (((size_t)p * nz + gz) * ny + gy) * nx + gx;
            if (to_blocks) blk[b * bs + QI(o, p, i, j, k)] = glob[gi];
            else glob[gi] = blk[b * bs + QI(o, p, i, j, k)];
          }
  }
}

/* Set the state as the reference driver does after problem set-up:
 * interior u0 given, q = cons2prim(u0), ghost exchange, first dt estimate. */
int orc_set_state(void* h, const double* uglob, double dtnext)
{
  ogrid* g = (ogrid*)h;
  xfer_global(g, g->u0, (double*)uglob, 1);
  update_variables(g, g->u0);
  new_time_step(g, dtnext);
  return 0;
}

int orc_get_state(void* h, double* uglob, double* qglob)
{
  ogrid* g = (ogrid*)h;
  xfer_global(g, g->u0, uglob, 0);
  if (qglob) xfer_global(g, g->q, qglob, 0);
  return 0;
}

/* the statistics of store_integrals (integrals.F90:254-485) of the current
 * primitive state: volume-weighted mean, minimum and maximum of density,
 * pressure, |v|, |B|, psi, the sonic Mach number sqrt(rho)|v|/sqrt(gamma p)
 * (|v| / csnd for the isothermal systems, whose pressure entries stay 0)
 * and the Alfvenic one sqrt(rho)|v|/max(eps, |B|), in that order, out[21];
 * per-block sums times the block's cell volume, added in block order */
int orc_statistics(void* h, double* out)
{
  ogrid* g = (ogrid*)h;
  const ocfg* o = &g->o;
  const size_t bs = bsize(o);
  const double eps = 2.220446049250313e-16, big = 1.7976931348623157e308;
  double av[7], mn[7], mx[7], vol;
  int b, i, j, k, m;
  for (m = 0; m < 7; m++) { av[m] = 0.0; mn[m] = big; mx[m] = -big; }
  if (!o->mhd) { mn[3] = mx[3] = mn[4] = mx[4] = mn[6] = mx[6] = 0.0; }
  if (o->iso) mn[1] = mx[1] = 0.0;     /* no pressure (integrals.F90:320-323) */
  for (b = 0; b < g->nblk; b++) {
    const double* qb = g->q + b * bs;
    double dvol, s[7] = {0, 0, 0, 0, 0, 0, 0}, lo[7], hi[7];
    if (g->amr) {
      const double* lv = g->lvd[g->lev[b]];
      dvol = (lv[0] * lv[1]) * lv[2];
    } else {
      dvol = (o->adx * o->ady) * o->adz;
    }
    for (m = 0; m < 7; m++) { lo[m] = big; hi[m] = -big; }
    for (k = o->kb; k <= o->ke; k++)
      for (j = o->jb; j <= o->je; j++)
        for (i = o->ib; i <= o->ie; i++) {
          const double* q = qb + QI(o, 0, i, j, k);
          double t[7], vel, mag = 0.0, sqd;
          vel = sqrt((q[o->ivx] * q[o->ivx] + q[o->ivy] * q[o->ivy]) + q[o->ivz] * q[o->ivz]);
          if (o->mhd)
            mag = sqrt((q[o->ibx] * q[o->ibx] + q[o->iby] * q[o->iby]) + q[o->ibz] * q[o->ibz]);
          sqd = sqrt(q[o->idn]);
          t[0] = q[o->idn];
          t[1] = o->iso ? 0.0 : q[o->ipr];
          t[2] = vel;
          t[3] = mag;
          t[4] = o->mhd ? q[o->ibp] : 0.0;
          /* the isothermal Mach number is |v| / csnd (integrals.F90:418-423) */
          t[5] = o->iso ? vel / o->csnd : sqd * vel / sqrt(o->c.gamma * q[o->ipr]);
          t[6] = sqd * vel / fmax_r(eps, mag);
          for (m = 0; m < 7; m++) {
            if (!o->mhd && (m == 3 || m == 4 || m == 6)) continue;
            if (o->iso && m == 1) continue;
            s[m] += t[m];
            lo[m] = fmin_r(lo[m], t[m]);
            hi[m] = fmax_r(hi[m], t[m]);
          }
        }
    for (m = 0; m < 7; m++) {
      if (!o->mhd && (m == 3 || m == 4 || m == 6)) continue;
      if (o->iso && m == 1) continue;
      av[m] = av[m] + s[m] * dvol;
      mn[m] = fmin_r(mn[m], lo[m]);
      mx[m] = fmax_r(mx[m], hi[m]);
    }
  }
  {
    const double xlen = o->c.xmax - o->c.xmin, ylen = o->c.ymax - o->c.ymin;
    const double zlen = o->nd == 3 ? o->c.zmax - o->c.zmin : 1.0;
    vol = 1.0 / ((xlen * ylen) * zlen);
  }
  for (m = 0; m < 7; m++) {
    out[3 * m] = av[m] * vol;
    out[3 * m + 1] = mn[m];
    out[3 * m + 2] = mx[m];
  }
  return 0;
}

/* nsteps of advance(dtnext); trace (may be NULL) gets (dt, cmax) per step */
int orc_run(void* h, int nsteps, double dtnext, double* trace)
{
  ogrid* g = (ogrid*)h;
  int s;
  for (s = 0; s < nsteps; s++) {
    if (g->o.amr_err) return g->o.amr_err;
    advance(g, dtnext);
    if (trace) { trace[2 * s] = g->dt; trace[2 * s + 1] = g->o.cmax; }
    g->time += g->dt;
    g->step += 1;
  }
  return 0;
}

long orc_positivity_events(void* h) { return ((ogrid*)h)->o.npos_events; }

/* second_derivative_error (refinement.F90:359-462) of primitive variable p
 * of one block with filled ghosts: the largest normalised second difference
 * over the cells ibl..ieu, in the reference's mix of real(4) and real(8)
 * (differences and magnitudes rounded to real(4), the quotient in real(8)
 * with the real(4) sums promoted, the result and the running maximum in
 * real(4)) */
static float ref_second_derivative_error(const ocfg* o, const double* q, int p, double epsref)
{
  const double eps = 2.220446049250313e-16;
  float error = 0.0f;
  int i, j, k, d;
  for (k = o->kbl; k <= o->keu; k++)
    for (j = o->jbl; j <= o->jeu; j++)
      for (i = o->ibl; i <= o->ieu; i++) {
        const double qc = q[QI(o, p, i, j, k)];
        for (d = 0; d < o->nd; d++) {
          const int di = d == 0, dj = d == 1, dk = d == 2;
          const double qp = q[QI(o, p, i + di, j + dj, k + dk)];
          const double qm = q[QI(o, p, i - di, j - dj, k - dk)];
          const float fr = (float)(qp - qc);
          const float fl = (float)(qm - qc);
          const float fc = (float)((fabs(qp) + fabs(qm)) + 2.0 * fabs(qc));
          const float fsum = fabsf(fr) + fabsf(fl);
          const float f = (float)((double)fabsf(fr + fl) /
                                  (((double)fsum + epsref * (double)fc) + eps));
          error = error > f ? error : f;
        }
      }
  return error;
}

/* the largest |curl u| of the vector at variables i0..i0+2 (velocity or
 * magnetic field) over the cells ibl..ieu (vorticity_magnitude /
 * current_density_magnitude, refinement.F90:481-640, with curl of
 * operators.F90:328-434: central differences with unit spacing, the 2D form
 * without z derivatives); the squared magnitude in real(8), its running
 * maximum in real(4) */
static float ref_curl_max(const ocfg* o, const double* q, int i0)
{
  float wmax = 0.0f;
  int i, j, k;
#define D1(c, di, dj, dk) (0.5 * (q[QI(o, i0 + (c), i + (di), j + (dj), k + (dk))] - \
                               q[QI(o, i0 + (c), i - (di), j - (dj), k - (dk))]) / 1.0)
  for (k = o->kbl; k <= o->keu; k++)
    for (j = o->jbl; j <= o->jeu; j++)
      for (i = o->ibl; i <= o->ieu; i++) {
        double wx, wy, wz;
        if (o->nd == 3) {
          wx = D1(2, 0, 1, 0) - D1(1, 0, 0, 1);
          wy = D1(0, 0, 0, 1) - D1(2, 1, 0, 0);
          wz = D1(1, 1, 0, 0) - D1(0, 0, 1, 0);
        } else {
          wx = D1(2, 0, 1, 0);
          wy = -D1(2, 1, 0, 0);
          wz = D1(1, 1, 0, 0) - D1(0, 0, 1, 0);
        }
        {
          const float v = (float)((wx * wx + wy * wy) + wz * wz);
          wmax = wmax > v ? wmax : v;
        }
      }
#undef D1
  return sqrtf(wmax);
}

/* check_refinement_criterion (refinement.F90:257-317) of every block:
 * crit[b] = -1 (may derefine), 0 (keep) or 1 (refine) from the largest
 * second-derivative error of the primitive variables in varmask, and the
 * vorticity / current-density magnitudes when asked for; ind[3 b ..] (may be
 * NULL) the three indicator values (the error is the largest over the
 * variables).  par = crefmin, crefmax, epsref, vortmin, vortmax, jabsmin,
 * jabsmax; flags = varmask (bit p: primitive variable p), vort, jabs. */
int orc_refinement(void* h, const double* par, const int* flags, int* crit, float* ind)
{
  ogrid* g = (ogrid*)h;
  const ocfg* o = &g->o;
  const size_t bs = bsize(o);
  int b, p;
  for (b = 0; b < g->nblk; b++) {
    const double* qb = g->q + (size_t)b * bs;
    int c = -1;
    float emax = 0.0f, w = 0.0f, jc = 0.0f;
    for (p = 0; p < o->nv; p++)
      if (flags[0] & (1 << p)) {
        const float cref = ref_second_derivative_error(o, qb, p, par[2]);
        if ((double)cref > par[0]) c = c > 0 ? c : 0;
        if ((double)cref > par[1]) c = 1;
        emax = emax > cref ? emax : cref;
      }
    if (flags[1]) {
      w = ref_curl_max(o, qb, o->ivx);
      if ((double)w > par[3]) c = c > 0 ? c : 0;
      if ((double)w > par[4]) c = 1;
    }
    if (flags[2] && o->mhd) {
      jc = ref_curl_max(o, qb, o->ibx);
      if ((double)jc > par[5]) c = c > 0 ? c : 0;
      if ((double)jc > par[6]) c = 1;
    }
    crit[b] = c;
    if (ind) {
      ind[3 * b] = emax;
      ind[3 * b + 1] = w;
      ind[3 * b + 2] = jc;
    }
  }
  return 0;
}

/* prolong_block (mesh.F90:1031-1257) for one child cell and variable: the
 * parent's cell value U[c] plus its limited slopes over the face neighbours
 * U[c -/+ st[d]] (prolongation limiter at 1/2), halved until the parent's
 * density or energy exceeds their absolute sum (set to zero with an event
 * when that value is not positive), halved once more and combined for
 * child position s = sx + 2 sy + 4 sz (sub-cell bit 0 = lower half). */
static double prolong_block_cell(ocfg* o, const double* U, size_t c, const size_t st[3], int p,
                                 int s)
{
  const int nd = o->nd;
  const double uc = U[c + p];
  double du[3] = {0.0, 0.0, 0.0};
  int d;
  for (d = 0; d < nd; d++)
    du[d] = lim(o->c.prol_limiter, 0.5, uc - U[c - st[d] + p], U[c + st[d] + p] - uc);
  if (p == o->idn || p == o->ien) {
    if (uc > 0.0) {
      int it;
      for (it = 0; it < OMAXHALV; it++) {
        double sa = fabs(du[0]) + fabs(du[1]);
        if (nd == 3) sa = sa + fabs(du[2]);
        if (!(uc <= sa)) break;
        for (d = 0; d < nd; d++) du[d] = 0.5 * du[d];
      }
    } else {
      o->npos_events++;
      du[0] = du[1] = du[2] = 0.0;
    }
  }
  for (d = 0; d < nd; d++) du[d] = 0.5 * du[d];
  if (nd == 2) {
    const double du1 = du[0] + du[1], du2 = du[0] - du[1];
    return s == 0 ? uc - du1 : (s == 1 ? uc + du2 : (s == 2 ? uc - du2 : uc + du1));
  } else {
    const double du1 = du[0] + du[1] + du[2], du2 = du[0] - du[1] - du[2];
    const double du3 = du[0] - du[1] + du[2], du4 = du[0] + du[1] - du[2];
    switch (s) {
    case 0: return uc - du1;
    case 1: return uc + du2;
    case 2: return uc - du3;
    case 3: return uc + du4;
    case 4: return uc - du4;
    case 5: return uc + du3;
    case 6: return uc - du2;
    default: return uc + du1;
    }
  }
}

/* The data movement of update_mesh (mesh.F90:756-833) onto a new leaf set:
 * every block of meta either exists in the old mesh (copied), is a child of
 * an old leaf (prolong_block, mesh.F90:1031-1257, from the parent's
 * conserved state including one layer of its ghost cells, which hold
 * prim2cons of the filled primitive ghosts), or is the parent of 2^d old
 * leaves (restrict_block, mesh.F90:1258-1454: the mean of the 2^d covering
 * cells in the reference's summation order).  Only the interiors are moved;
 * the new grid then fills its ghosts (update_variables after update_mesh,
 * evolution.F90:321-340) and keeps the old grid's time, step and dt.
 * Returns 702 when a block of meta has no such source. */
int orc_amr_remesh(const amun_cfg* c, void* hold, int nblk, const int32_t* meta, void** hnew)
{
  ogrid* g = (ogrid*)hold;
  ogrid* n;
  ocfg* o;
  const int nc = g->o.nc, nd = g->o.nd, half = nc / 2;
  size_t bs, st[3];
  int r, b, i, j, k, p;
  *hnew = NULL;
  if (!g->amr) return 600;
  r = orc_create_amr(c, nblk, meta, (void**)&n);
  if (r) return r;
  o = &n->o;
  bs = bsize(o);
  st[0] = (size_t)o->nv;
  st[1] = (size_t)o->im * o->nv;
  st[2] = (size_t)o->im * o->jm * o->nv;
  for (b = 0; b < nblk; b++) {
    const int L = n->lev[b];
    const long x0 = (long)n->bx[b] * nc, y0 = (long)n->by[b] * nc;
    const long z0 = nd == 3 ? (long)n->bz[b] * nc : 0;
    const int same = owner(g, L, x0, y0, z0);
    double* dst = n->u0 + (size_t)b * bs;
    if (same >= 0) {
      memcpy(dst, g->u0 + (size_t)same * bs, bs * sizeof(double));
      continue;
    }
    {
      /* the old leaf covering this block one level down */
      const int P = owner(g, L - 1, x0 / 2, y0 / 2, z0 / 2);
      if (P >= 0) {
        const double* U = g->u0 + (size_t)P * bs;
        const int px = n->bx[b] & 1, py = n->by[b] & 1, pz = nd == 3 ? (n->bz[b] & 1) : 0;
        for (k = 0; k < (nd == 3 ? nc : 1); k++)
          for (j = 0; j < nc; j++)
            for (i = 0; i < nc; i++) {
              const int ci = (i + px * nc) / 2 + o->ib, cj = (j + py * nc) / 2 + o->jb;
              const int ck = nd == 3 ? (k + pz * nc) / 2 + o->kb : 0;
              const int s = (i & 1) + 2 * (j & 1) + (nd == 3 ? 4 * (k & 1) : 0);
              const size_t cc = QI(o, 0, ci, cj, ck);
              const size_t dc = QI(o, 0, i + o->ib, j + o->jb, k + o->kb);
              for (p = 0; p < o->nv; p++)
                dst[dc + p] = prolong_block_cell(&g->o, U, cc, st, p, s);
            }
        continue;
      }
    }
    {
      /* the 2^d old leaves one level up */
      int kid[8], q;
      for (q = 0; q < (1 << nd); q++) {
        kid[q] = owner(g, L + 1, 2 * x0 + (q & 1) * nc, 2 * y0 + ((q >> 1) & 1) * nc,
                       2 * z0 + ((q >> 2) & 1) * nc);
        if (kid[q] < 0) { orc_destroy(n); return 702; }
      }
      for (k = 0; k < (nd == 3 ? nc : 1); k++)
        for (j = 0; j < nc; j++)
          for (i = 0; i < nc; i++) {
            const int px = i >= half, py = j >= half, pz = nd == 3 && k >= half;
            const double* U = g->u0 + (size_t)kid[px + 2 * py + 4 * pz] * bs;
            const int fi = 2 * (i - px * half) + o->ib, fj = 2 * (j - py * half) + o->jb;
            const int fk = nd == 3 ? 2 * (k - pz * half) + o->kb : 0;
            const size_t l = QI(o, 0, fi, fj, fk);
            const size_t dc = QI(o, 0, i + o->ib, j + o->jb, k + o->kb);
            const size_t a = st[0], bb = st[1], e = st[2];
            for (p = 0; p < o->nv; p++) {
              if (nd == 2)
                dst[dc + p] = 2.50e-01 * ((U[l + p] + U[l + a + bb + p]) +
                                          (U[l + bb + p] + U[l + a + p]));
              else
                dst[dc + p] = 1.25e-01 * (((U[l + p] + U[l + a + bb + e + p]) +
                                           (U[l + e + p] + U[l + a + bb + p])) +
                                          ((U[l + bb + e + p] + U[l + a + p]) +
                                           (U[l + bb + p] + U[l + a + e + p])));
            }
          }
    }
  }
  n->time = g->time;
  n->step = g->step;
  n->dt = g->dt;
  n->dtn = g->dtn;
  o->npos_events = g->o.npos_events;
  o->amr_err = 0;
  update_variables(n, n->u0);
  if (o->amr_err) {
    r = o->amr_err;
    orc_destroy(n);
    return r;
  }
  *hnew = n;
  return 0;
}

/* the MGP fallback record so far: count and position hash */
void orc_mgp_fallbacks(void* h, unsigned long long* out)
{
  out[0] = ((ogrid*)h)->o.mgp_fallbacks;
  out[1] = ((ogrid*)h)->o.mgp_hash;
}

/* messages of the relativistic primitive solver so far: bracketing failed,
 * non-positive pressure, no convergence */
void orc_c2p_events(void* h, long* n)
{
  ogrid* g = (ogrid*)h;
  n[0] = g->o.c2p_fail;
  n[1] = g->o.c2p_pneg;
  n[2] = g->o.c2p_noconv;
}

/* unit level: the Gaussian-process weights cgp(1..ngp) of prepare_gp */
int orc_gp_weights(const amun_cfg* c, double* cgp)
{
  ocfg o;
  amun_cfg cc = *c;
  int i, r;
  cc.reconstruction = AMUN_REC_GP;
  r = oc_setup(&o, &cc);
  if (r) return r;
  for (i = 0; i < cc.ngp; i++) cgp[i] = o.cgp[i];
  return 0;
}

/* unit level: the root finders of algebra.F90 on count coefficient sets
 * a(5) (increasing powers).  Per set, out receives 19 values:
 * [nr, x1, x2] quadratic(a(1:3)), [nr, x1, x2] quadratic_normalized(a(1:2)),
 * [nr, x1..x3] cubic(a(1:4)), [nr, x1..x3] cubic_normalized(a(1:3)),
 * [nr, x1..x4] quartic(a(1:5)); unused roots as the routines leave them */
void orc_algebra(int count, const double* a, double* out)
{
  int i, k;
  for (i = 0; i < count; i++) {
    const double* c = a + (size_t)i * 5;
    double* o = out + (size_t)i * 19;
    double x[4];
    x[0] = x[1] = x[2] = x[3] = 0.0;
    o[0] = alg_quadratic(c, x);
    o[1] = x[0]; o[2] = x[1];
    x[0] = x[1] = x[2] = x[3] = 0.0;
    o[3] = alg_quadratic_normalized(c, x);
    o[4] = x[0]; o[5] = x[1];
    x[0] = x[1] = x[2] = x[3] = 0.0;
    o[6] = alg_cubic(c, x);
    for (k = 0; k < 3; k++) o[7 + k] = x[k];
    x[0] = x[1] = x[2] = x[3] = 0.0;
    o[10] = alg_cubic_normalized(c, x);
    for (k = 0; k < 3; k++) o[11 + k] = x[k];
    x[0] = x[1] = x[2] = x[3] = 0.0;
    o[14] = alg_quartic(c, x);
    for (k = 0; k < 4; k++) o[15 + k] = x[k];
  }
}

/* unit level: cons2prim on n conserved states u(nv,n); q(nv,n) is updated in
 * place (entries a failing cell does not write keep their values); ev (may be
 * NULL) receives the three solver message counts */
int orc_cons2prim(const amun_cfg* c, int n, const double* u, double* q, long* ev)
{
  ocfg o;
  int i, r = oc_setup(&o, c);
  if (r) return r;
  for (i = 0; i < n; i++) cons2prim(&o, u + (size_t)i * o.nv, q + (size_t)i * o.nv);
  if (ev) { ev[0] = o.c2p_fail; ev[1] = o.c2p_pneg; ev[2] = o.c2p_noconv; }
  return 0;
}

/* unit level: update_flux on one block q(nv,im,jm,km) with a given cmax */
int orc_update_flux(const amun_cfg* c, const double* q, const double* dx,
                    double cmax, double* f)
{
  ocfg o;
  double* qs;
  int r = oc_setup(&o, c);
  if (!r) r = riemann_check(c);
  if (r) return r;
  o.h[0] = dx[0]; o.h[1] = dx[1]; o.h[2] = o.nd == 3 ? dx[2] : 1.0;
  o.cmax = cmax;
  o.cmax2 = cmax * cmax;
  qs = (double*)malloc((size_t)o.im * o.jm * o.km * o.nv * 2 * o.nd * sizeof(double));
  update_flux(&o, q, f, qs);
  free(qs);
  return 0;
}

/* unit level: reconstruct() on npencil pencils of length n (WENO5-Z, GP) */
int orc_reconstruct(const amun_cfg* c, int n, int npencil, double h, const double* fp,
                    double* fl, double* fr)
{
  ocfg o;
  double* w;
  int p, r = oc_setup(&o, c);
  if (r) return r;
  w = (double*)malloc(3 * (size_t)(n + 4) * sizeof(double));
  for (p = 0; p < npencil; p++)
    recon_dir(&o, n, h, fp + (size_t)p * n, fl + (size_t)p * n, fr + (size_t)p * n, w, w + n + 4,
              w + 2 * (n + 4));
  free(w);
  return 0;
}

/* unit level: prim2cons, cons2prim, fluxspeed on n states q(nv,n) */
int orc_eqs(const amun_cfg* c, int n, double cmax, const double* q, double* u,
            double* q2, double* f, double* cm, double* cp)
{
  ocfg o;
  int i, r = oc_setup(&o, c);
  if (r) return r;
  o.cmax = cmax;
  o.cmax2 = cmax * cmax;
  for (i = 0; i < n; i++) {
    prim2cons(&o, q + (size_t)i * o.nv, u + (size_t)i * o.nv);
    cons2prim(&o, u + (size_t)i * o.nv, q2 + (size_t)i * o.nv);
    fluxspeed(&o, q + (size_t)i * o.nv, u + (size_t)i * o.nv, f + (size_t)i * o.nv, cm + i, cp + i);
  }
  return 0;
}

/* unit level: one Riemann problem per state pair, ql/qr (nv,n) primitive */
int orc_riemann(const amun_cfg* c, int n, double cmax, const double* ql,
                const double* qr, double* f)
{
  ocfg o;
  int i, p, r = oc_setup(&o, c);
  double a[NVMAX], b[NVMAX];
  if (!r) r = riemann_check(c);
  if (r) return r;
  o.cmax = cmax;
  o.cmax2 = cmax * cmax;
  for (i = 0; i < n; i++) {
    for (p = 0; p < o.nv; p++) { a[p] = ql[(size_t)i * o.nv + p]; b[p] = qr[(size_t)i * o.nv + p]; }
    riemann(&o, a, b, f + (size_t)i * o.nv);
  }
  return 0;
}
