// kernels.h -- device code of the MI355X (gfx950) AMUN per-block update: data
// layout, the sweep kernels and their helpers, and the table of kernel
// instantiations a configuration selects.  Shared by amun_dev.hip (host side
// of the C-ABI) and the instantiation units inst_*.hip, which are compiled in
// parallel.
//
// Device data (block pool, structure of arrays, interior cells only):
//   U0, U1   conserved state, double [nblk][nv][nc^nd]   (u0/u1 of blocks.F90)
//   Q[2]     primitive state ping-pong, double, same layout      (q)
//   D        flux-divergence accumulator of the current stage     (du)
//   nbr      int [nblk][27] block index of every same-level neighbour
//            (face, edge and corner), periodic wrap included
// Ghost zones are never materialised on the hot path: a stencil that leaves
// its block reads the neighbour's interior through nbr, which is exactly the
// value the reference copies into its ghost layers (boundaries.F90:372-479,
// same-level face/edge/corner copies).
//
// One RK stage = one sweep kernel per direction.  A sweep thread owns one
// pencil of a block and walks it with a sliding window: each cell is
// reconstructed once (both faces), each face gets one Riemann solve, and the
// flux difference goes straight into the cell's increment.  The x sweep
// stores -dxi*dF, the y sweep subtracts dyi*dF, and the last sweep completes
// du, applies the stage combine of the integrator, the psi decay, recomputes
// the primitive variables of the new state and (last stage) reduces the
// maximum speed for the next CFL step -- i.e. update_fluxes, update_increment,
// the stage update and update_variables of evolution.F90:1005-2094 fused into
// nd kernels.  dt, cmax and the simulation time stay on the device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <string>
#include <vector>

#pragma once

#include "../../include/amun_dev.h"
#include "physics.h"
#include "srmhd.h"

using namespace amun;

namespace amun_dev {

inline thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(1000 + (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr double kTiny = 2.2250738585072014e-308;   // tiny(1d0)
constexpr int kTrace = 1 << 16;                      // dt trace entries (all runs)

// ---------------------------------------------------------------------------
// device-side step state
// ---------------------------------------------------------------------------
struct DevState {
  unsigned long long npos;          // non-positive density/pressure events
  unsigned long long c2p[3];        // SR primitive solver: failed, p <= 0, no convergence
  unsigned long long tridiag;       // compact schemes: pencils with a zero Thomas pivot
};

// step state of one run (one independent grid of a batch; a single run is a
// batch of one): what new_time_step and the driver keep per simulation
struct RunState {
  unsigned long long cmax_bits[2];  // cmax of step s lives in slot s&1
  double dt;                        // last dt (kept when dtnext <= 0)
  double time;
  long long step;
  double dxmin;                     // smallest cell size (evolution.F90:467-476)
  double dxi[3];                    // inverse cell sizes
  double h[3];                      // cell sizes (the reconstructions' h)
  double fix_dt, fix_cmax;          // > 0: dt and cmax held at these values
};

// constants of one run, passed by value to every kernel
struct RunConst {
  int nd, nc, nblk, ncell;
  int nruns;                        // independent grids in the block pool
  int limiter;
  int prol_limiter;                 // prolongation of multi-level ghost fills
  double gamma, gammam1, gammam1i;
  double csnd, csnd2;               // isothermal sound speed and its square
  double nu, eta;                   // viscosity and resistivity (sources.F90:52-58)
  double gacc;                      // uniform gravity (0, gacc, 0) when grav
  int grav;
  int v4;                           // relativistic: four-velocity reconstruction
  double cfl, eps, decay;
  double cgp[5];                    // Gaussian-process weights (prepare_gp)
  double kappa, kbeta;              // MP limiter of the GP reconstruction
  double rad, ppm_const;            // LimO3 radius, PPM curvature ratio
  int bnd[6];                       // AMUN_BND_* per domain side (xl, xu, ...)
  int ng;                           // ghost layers of the reference's pencils
  void* fsl;                        // compact schemes: face states of the
  void* fsr;                        //   current sweep direction
  const double* ugp;                // MGP weights [dir][side][stencil]
  int mgp, dgp;                     // MGP half width and stencil size ngp^nd
  SrConst sc;                       // relativistic primitive solver
};

__device__ __forceinline__ double cmax_of(const RunState* R, int slot) {
  if (R->fix_cmax > 0.0) return R->fix_cmax;
  return fmax(__longlong_as_double((long long)R->cmax_bits[slot]), kTiny);
}

// dt of new_time_step (evolution.F90:478-483), identical in every thread
__device__ __forceinline__ double dt_of(const RunConst& rc, const RunState* R,
                                       int slot, double dtnext) {
  if (R->fix_dt > 0.0) return R->fix_dt;
  double cmax = cmax_of(R, slot);
  double dtn = rc.cfl * R->dxmin / fmax(cmax + 2.0 * fmax(rc.nu, rc.eta) / R->dxmin, kTiny);
  return dtnext > 0.0 ? fmin(dtn, dtnext) : R->dt;
}

// Position of cell (i,j,k) inside a block's variable plane.  Cells are stored
// in 4x4 x-y bricks: one 128-byte line holds x = 4i'..4i'+3 of four
// consecutive y, so the x, y and z sweeps (lanes across the two transverse
// directions, walking the third) each touch few lines per load and reuse
// them within four steps.  Requires nc % 4 == 0.
// FP32 arrays (element size 4) of 3D runs use 4x4x2 bricks instead, so that
// a line (32 floats) again spans four steps of every sweep direction.
template <int ES>
__host__ __device__ __forceinline__ int lidx(int nc, int nd, int i, int j, int k) {
  if (nd == 2) return (((i >> 2) * nc + j) << 2) | (i & 3);
  if (ES == 8) return ((((i >> 2) * nc + k) * nc + j) << 2) | (i & 3);
  return (((((i >> 2) * (nc >> 1) + (k >> 1)) * nc + j) * 2 + (k & 1)) << 2) | (i & 3);
}

// position of (i,j,k), possibly outside the block by at most nc, in the pool
// (array of element size ES)
template <int ES>
__device__ __forceinline__ size_t cell_index(const int* __restrict__ nbr,
                                             const RunConst& rc, int b, int i,
                                             int j, int k) {
  const int nc = rc.nc;
  int di = (i < 0) ? -1 : (i >= nc ? 1 : 0);
  int dj = (j < 0) ? -1 : (j >= nc ? 1 : 0);
  int dk = (k < 0) ? -1 : (k >= nc ? 1 : 0);
  if (di | dj | dk) {
    b = nbr[b * 27 + (dk + 1) * 9 + (dj + 1) * 3 + (di + 1)];
    i -= di * nc;
    j -= dj * nc;
    k -= dk * nc;
  }
  // packed: block in the high word, cell in the low word
  return ((size_t)b << 32) | (size_t)lidx<ES>(nc, rc.nd, i, j, k);
}

template <int NV>
__device__ __forceinline__ size_t var_index(const RunConst& rc, size_t packed, int v) {
  size_t b = packed >> 32, c = packed & 0xffffffffull;
  return (b * NV + v) * (size_t)rc.ncell + c;
}

// component rotation of the directional solvers (schemes.F90:3063-3071,
// 3100-3108): rotated slot p takes original variable rot<P,DIR>(p)
template <class P, int DIR>
__device__ __forceinline__ constexpr int rot(int p) {
  if constexpr (DIR == 0) {
    return p;
  } else if constexpr (DIR == 1) {
    if (p == P::ivx) return P::ivy;
    if (p == P::ivy) return P::ivz;
    if (p == P::ivz) return P::ivx;
    if constexpr (P::mhd) {
      if (p == P::ibx) return P::iby;
      if (p == P::iby) return P::ibz;
      if (p == P::ibz) return P::ibx;
    }
    return p;
  } else {
    if (p == P::ivx) return P::ivz;
    if (p == P::ivy) return P::ivx;
    if (p == P::ivz) return P::ivy;
    if constexpr (P::mhd) {
      if (p == P::ibx) return P::ibz;
      if (p == P::iby) return P::ibx;
      if (p == P::ibz) return P::iby;
    }
    return p;
  }
}

enum {
  REC_CRWENO5Z = AMUN_REC_CRWENO5Z, REC_CRWENO5YC = AMUN_REC_CRWENO5YC,
  REC_CRWENO5NS = AMUN_REC_CRWENO5NS, REC_CRMP5 = AMUN_REC_CRMP5,
  REC_CRMP5LD = AMUN_REC_CRMP5LD, REC_CRMP7 = AMUN_REC_CRMP7, REC_MGP = AMUN_REC_MGP,
  REC_TVD = AMUN_REC_TVD, REC_WENO3 = AMUN_REC_WENO3, REC_LIMO3 = AMUN_REC_LIMO3,
  REC_PPM = AMUN_REC_PPM, REC_WENO5Z = AMUN_REC_WENO5Z, REC_WENO5YC = AMUN_REC_WENO5YC,
  REC_WENO5NS = AMUN_REC_WENO5NS, REC_MP5 = AMUN_REC_MP5, REC_MP7 = AMUN_REC_MP7,
  REC_GP = AMUN_REC_GP
};
enum { RS_HLL = AMUN_RS_HLL, RS_HLLC = AMUN_RS_HLLC, RS_HLLD = AMUN_RS_HLLD, RS_ROE = AMUN_RS_ROE,
       RS_HLLDM = AMUN_RS_HLLDM };
// kernel-table keys of the isothermal systems: the equation system + 16
enum { EQ_ISO = 16, EQ_HD_ISO = AMUN_EQSYS_HD + EQ_ISO, EQ_MHD_ISO = AMUN_EQSYS_MHD + EQ_ISO };

template <class P, int RS, typename T>
__device__ __forceinline__ void riemann(const EosConst<T>& e, const SrConst& sc, T cmax,
                                        T cmax2, T* ql, T* qr, T* f) {
  if constexpr (P::iso && !P::mhd && RS == RS_ROE)
    riemann_roe_hd_iso<P, T>(e, cmax, cmax2, ql, qr, f);
  else if constexpr (P::iso && P::mhd && RS == RS_HLLD)
    riemann_hlld_iso<P, false, T>(e, cmax, cmax2, ql, qr, f);
  else if constexpr (P::iso && P::mhd && RS == RS_HLLDM)
    riemann_hlld_iso<P, true, T>(e, cmax, cmax2, ql, qr, f);
  else if constexpr (P::iso && P::mhd && RS == RS_ROE)
    riemann_roe_mhd_iso<P, T>(e, cmax, cmax2, ql, qr, f);
  else if constexpr (P::iso) riemann_hll<P, T>(e, cmax, cmax2, ql, qr, f);
  else if constexpr (P::sr && RS == RS_HLLC) riemann_sr_hllc<P, T>(sc, cmax, cmax2, ql, qr, f);
  else if constexpr (P::sr) riemann_sr_hll<P, T>(sc, cmax, cmax2, ql, qr, f);
  else if constexpr (RS == RS_HLLD) riemann_hlld<P, T>(e, cmax, cmax2, ql, qr, f);
  else if constexpr (RS == RS_HLLC && P::mhd) riemann_hllc_mhd<P, T>(e, cmax, cmax2, ql, qr, f);
  else if constexpr (RS == RS_ROE && P::mhd) riemann_roe_mhd<P, T>(e, cmax, cmax2, ql, qr, f);
  else if constexpr (RS == RS_ROE) riemann_roe_hd<P, T>(e, cmax, cmax2, ql, qr, f);
  else if constexpr (RS == RS_HLLC) riemann_hllc<P, T>(e, cmax, cmax2, ql, qr, f);
  else riemann_hll<P, T>(e, cmax, cmax2, ql, qr, f);
}

// lanes of the per-lane LDS slots of the chunked (x and y) sweeps
constexpr int kLdsLanes = 256;

// maximum speed of one sweep workgroup for the next CFL step, and its run
// (run < 0: the workgroup spanned several runs and used atomics instead)
struct SpeedPart {
  double smax;
  int run, pad;
};

struct SweepArgs {
  const void* Q;                   // primitive state of the stage input (T)
  const double* U0;                // state at the start of the step
  const double* U1;                // intermediate state (stage input)
  double* Uout;                    // stage output (may alias U0 or U1)
  void* Qout;                      // primitive state of Uout (T)
  void* D;                         // increment accumulator (T)
  float* C;                        // Kahan compensation of U0 (mixed mode)
  const int* __restrict__ nbr;
  DevState* st;                    // event counters
  RunState* rs;                    // per-run step state
  const int* __restrict__ brun;    // run of every block
  SpeedPart* spart;                // per-workgroup maximum speeds (last stage)
  int slot;                        // cmax slot of this step
  int swz;                         // XCD-aware workgroup order (autotuned)
  int kind;                        // stage combine, see combine()
  int last;                        // last stage of the step
  int comp;                        // the stage writes the step state U0
  int first;                       // first sweep of the step
  double fac;                      // ds = fac * dt
  double ca, cb;                   // state weights of combine kinds 5-7
  double dtnext;
  // multi-level meshes (null on single-level grids): level - 1 of every
  // block (cell sizes scale by 2^-(level-1)), the flux-correction slot of
  // every block face (block * 6 + 2 dir + side, -1 if none) and the
  // averaged fine fluxes of the slots
  const int* __restrict__ blev;
  const int* __restrict__ ffix;
  const double* __restrict__ FB;
  // GLM source terms (null when off): the eight products per cell of
  // k_glm_src, and the kind (1 EGLM, 2 HEGLM)
  const void* src;
  int glm;
  // viscous and resistive source terms (null when off): the nine products
  // per cell of k_visc_src, and which are on (1 viscosity, 2 resistivity)
  const void* vsrc;
  int vis;
  // the primitive state of the stage input: Q itself, or the primitives
  // behind the four-velocity states of Q (the recovery keeps them for a
  // failed cell; gravity reads them)
  const void* Qold;
  // multi-dimensional limiting (null when off): the slopes of k_mlp for the
  // interior cells (S) and the cells beyond the pencil ends (G)
  const void* mlps;
  const void* mlpg;
};

// stage combines of evolve_euler/rk2/rk3 (evolution.F90:570, 700, 740, 1080,
// 1130, 1177) and of the SSPRK schemes (evolution.F90:917, 956, 1303, 1344,
// 1395, 1442, 1667, 1714, 1766)
__device__ __forceinline__ double combine(int kind, double u0, double u1, double ds,
                                          double du, double ca, double cb) {
  switch (kind) {
  case 0: return u0 + ds * du;
  case 1: return ((3.0 / 4.0) * u0 + (1.0 / 4.0) * u1) + ds * du;
  case 2: return ((1.0 / 3.0) * u0 + (2.0 / 3.0) * u1) + ds * du;
  case 3: return 0.5 * ((u0 + u1) + ds * du);
  case 4: return u1 + ds * du;
  case 5: return ca * u0 + cb * (u1 + ds * du);
  case 6: return (ca * u0 + cb * u1) + ds * du;
  default: return (ca * u1 + cb * u0) + ds * du;
  }
}

// the same combines as increments of the compensated step state u0 - cp
// (every scheme's state weights sum to one)
__device__ __forceinline__ double combine_delta(int kind, double u0, double u1, double cp,
                                                double dd, double ca, double cb) {
  const double w = (u1 - u0) + cp;
  switch (kind) {
  case 0: return dd;
  case 1: return 0.25 * w + dd;
  case 2: return (2.0 / 3.0) * w + dd;
  case 3: return 0.5 * (w + dd);
  case 4: return w + dd;
  case 5: return cb * (w + dd);
  case 6: return cb * w + dd;
  default: return ca * w + dd;
  }
}

// bound of the positivity halving loops (more than the 2098 halvings that
// take the largest double below the smallest positive one)
constexpr int kMaxHalvings = 2200;

// reconstruction of one cell along DIR from its primitive window:
// flc = state at the upper face, frc = state at the lower face
template <class P, int REC, int ND, int DIR, typename T>
struct Recon;

// WENO5-Z window of five cells (interfaces_dir -> reconstruct_weno5z)
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_WENO5Z, ND, DIR, T> {
  static constexpr int R = 2;    // window half width
  static constexpr int H = 3;    // ghost depth needed (faces of cells -1, nc)
  T w[P::nv][5];
  __device__ __forceinline__ void load(int slotm, const T* q) {
#pragma unroll
    for (int v = 0; v < P::nv; v++) w[v][slotm] = q[v];
  }
  __device__ __forceinline__ void shift() {
#pragma unroll
    for (int v = 0; v < P::nv; v++)
#pragma unroll
      for (int m = 0; m < 4; m++) w[v][m] = w[v][m + 1];
  }
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc, const T*,
                                       const int*, int, int, int, int, DevState*, double) {
    const T eps = T(rc.eps);
#pragma unroll
    for (int v = 0; v < P::nv; v++) weno5z_cell<T>(w[v], eps, flc[v], frc[v]);
  }
};

// WENO5-YC and WENO5-NS (reconstruct_weno5yc / _weno5ns): the window of
// WENO5-Z with their own weights
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_WENO5YC, ND, DIR, T> : Recon<P, REC_WENO5Z, ND, DIR, T> {
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc, const T*,
                                       const int*, int, int, int, int, DevState*, double) {
    const T eps = T(rc.eps);
#pragma unroll
    for (int v = 0; v < P::nv; v++) weno5z_cell<T, true>(this->w[v], eps, flc[v], frc[v]);
  }
};

template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_WENO5NS, ND, DIR, T> : Recon<P, REC_WENO5Z, ND, DIR, T> {
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc, const T*,
                                       const int*, int, int, int, int, DevState*, double) {
    const T eps = T(rc.eps);
#pragma unroll
    for (int v = 0; v < P::nv; v++) weno5ns_cell<T>(this->w[v], eps, flc[v], frc[v]);
  }
};

// MP5 (reconstruct_mp5): the five-cell window, polynomial face values and
// mp_limiting
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_MP5, ND, DIR, T> : Recon<P, REC_WENO5Z, ND, DIR, T> {
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc, const T*,
                                       const int*, int, int, int, int, DevState*, double) {
    const T kappa = T(rc.kappa), kbeta = T(rc.kbeta), eps = T(rc.eps);
#pragma unroll
    for (int v = 0; v < P::nv; v++) mp_cell<T, 2>(this->w[v], kappa, kbeta, eps, flc[v], frc[v]);
  }
};

// MP7 (reconstruct_mp7): a seven-cell window
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_MP7, ND, DIR, T> {
  static constexpr int R = 3;
  static constexpr int H = 4;
  T w[P::nv][7];
  __device__ __forceinline__ void load(int slotm, const T* q) {
#pragma unroll
    for (int v = 0; v < P::nv; v++) w[v][slotm] = q[v];
  }
  __device__ __forceinline__ void shift() {
#pragma unroll
    for (int v = 0; v < P::nv; v++)
#pragma unroll
      for (int m = 0; m < 6; m++) w[v][m] = w[v][m + 1];
  }
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc, const T*,
                                       const int*, int, int, int, int, DevState*, double) {
    const T kappa = T(rc.kappa), kbeta = T(rc.kbeta), eps = T(rc.eps);
#pragma unroll
    for (int v = 0; v < P::nv; v++) mp_cell<T, 3>(w[v], kappa, kbeta, eps, flc[v], frc[v]);
  }
};

// Gaussian process (interfaces_dir -> reconstruct_gp, mp_limiting): the
// five-cell window of WENO5-Z
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_GP, ND, DIR, T> : Recon<P, REC_WENO5Z, ND, DIR, T> {
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc, const T*,
                                       const int*, int, int, int, int, DevState*, double) {
    const T cgp[5] = {T(rc.cgp[0]), T(rc.cgp[1]), T(rc.cgp[2]), T(rc.cgp[3]), T(rc.cgp[4])};
    const T kappa = T(rc.kappa), kbeta = T(rc.kbeta), eps = T(rc.eps);
#pragma unroll
    for (int v = 0; v < P::nv; v++) gp_cell<T>(this->w[v], cgp, kappa, kbeta, eps, flc[v], frc[v]);
  }
};

// WENO3 and LimO3 (interfaces_dir -> reconstruct_weno3 / _limo3): a
// three-cell window; both weigh by the cell size h of the sweep direction
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_WENO3, ND, DIR, T> {
  static constexpr int R = 1;
  static constexpr int H = 2;
  T w[P::nv][3];
  __device__ __forceinline__ void load(int slotm, const T* q) {
#pragma unroll
    for (int v = 0; v < P::nv; v++) w[v][slotm] = q[v];
  }
  __device__ __forceinline__ void shift() {
#pragma unroll
    for (int v = 0; v < P::nv; v++) {
      w[v][0] = w[v][1];
      w[v][1] = w[v][2];
    }
  }
  __device__ __forceinline__ void cell(const RunConst&, T* flc, T* frc, const T*,
                                       const int*, int, int, int, int, DevState*, double h) {
    const T h2 = T(h * h);
#pragma unroll
    for (int v = 0; v < P::nv; v++) weno3_cell<T>(w[v], h2, flc[v], frc[v]);
  }
};

template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_LIMO3, ND, DIR, T> : Recon<P, REC_WENO3, ND, DIR, T> {
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc, const T*,
                                       const int*, int, int, int, int, DevState*, double h) {
    const double rdx = rc.rad * h;
    const T eps = T(rc.eps), rdx2 = T(rdx * rdx);
#pragma unroll
    for (int v = 0; v < P::nv; v++) limo3_cell<T>(this->w[v], eps, rdx2, flc[v], frc[v]);
  }
};

// PPM (interfaces_dir -> reconstruct_ppm): the seven-cell window of MP7
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_PPM, ND, DIR, T> : Recon<P, REC_MP7, ND, DIR, T> {
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc, const T*,
                                       const int*, int, int, int, int, DevState*, double) {
    const T pc = T(rc.ppm_const);
#pragma unroll
    for (int v = 0; v < P::nv; v++) ppm_cell<T>(this->w[v], pc, flc[v], frc[v]);
  }
};

// compact WENO (CRWENO5-Z/-YC/-NS) and compact MP (CRMP5, CRMP5LD, CRMP7):
// the face states of the whole pencil come
// from a tridiagonal solve (k_compact, run before the sweep of each
// direction into rc.fsl / rc.fsr, [block][var][cell -1 .. nc][pencil])
constexpr bool is_compact(int rec) {
  return rec == REC_CRWENO5Z || rec == REC_CRWENO5YC || rec == REC_CRWENO5NS ||
         rec == REC_CRMP5 || rec == REC_CRMP5LD || rec == REC_CRMP7;
}

template <class P, int REC, int ND, int DIR, typename T>
struct ReconCompact {
  static constexpr int R = 0;
  static constexpr int H = 2;
  __device__ __forceinline__ void load(int, const T*) {}
  __device__ __forceinline__ void shift() {}
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc, const T*,
                                       const int*, int b, int i, int j, int k, DevState*,
                                       double) {
    const int nc = rc.nc, npen = ND == 3 ? nc * nc : nc;
    const int c = DIR == 0 ? i : (DIR == 1 ? j : k);
    const int t1 = DIR == 0 ? j : i, t2 = DIR == 2 ? j : k;
    const int p = t1 + (ND == 3 ? nc * t2 : 0);
    const T* L = static_cast<const T*>(rc.fsl);
    const T* Rr = static_cast<const T*>(rc.fsr);
#pragma unroll
    for (int v = 0; v < P::nv; v++) {
      const size_t o = ((size_t)(b * P::nv + v) * (nc + 2) + (c + 1)) * npen + p;
      flc[v] = L[o];
      frc[v] = Rr[o];
    }
  }
};
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_CRWENO5Z, ND, DIR, T> : ReconCompact<P, REC_CRWENO5Z, ND, DIR, T> {};
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_CRWENO5YC, ND, DIR, T> : ReconCompact<P, REC_CRWENO5YC, ND, DIR, T> {};
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_CRWENO5NS, ND, DIR, T> : ReconCompact<P, REC_CRWENO5NS, ND, DIR, T> {};
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_CRMP5, ND, DIR, T> : ReconCompact<P, REC_CRMP5, ND, DIR, T> {};
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_CRMP5LD, ND, DIR, T> : ReconCompact<P, REC_CRMP5LD, ND, DIR, T> {};
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_CRMP7, ND, DIR, T> : ReconCompact<P, REC_CRMP7, ND, DIR, T> {};

// the longest pencil (ncells + 2 nghosts) of the compact schemes
constexpr int kMaxPencil = 96;

// reconstruct_crweno5z / _yc / _ns (interpolations.F90:2477-3646) and
// tridiag (algebra.F90:1115-1174) on one pencil f[0..n) of the reference's
// layout (ng ghosts each side): the nonlinear weights, the explicit WENO5
// rows within ng cells of the ends, the compact rows between them, the
// Thomas algorithm with the rows built on the fly.  Writes the states at the
// faces of the cells -1 .. nc: ol[c + 1] upper, or[c + 1] lower.
template <int REC, typename T>
__device__ bool crweno_pencil(int n, int ng, const T* f, T eps, T* ol, T* orr) {
  const T c1 = T(13.0 / 12.0), c2 = T(0.25), cl = T(1.0 / 9.0), cc = T(5.0 / 9.0),
          cr = T(1.0 / 3.0), dl = T(0.1), dc = T(0.6), dr = T(0.3), dq = T(0.5), xi = T(0.4);
  const T a11 = T(2.0 / 6.0), a12 = T(-7.0 / 6.0), a13 = T(11.0 / 6.0);
  const T a21 = T(-1.0 / 6.0), a22 = T(5.0 / 6.0), a23 = T(2.0 / 6.0);
  const T a31 = T(2.0 / 6.0), a32 = T(5.0 / 6.0), a33 = T(-1.0 / 6.0);
  auto dfp = [&](int i) { return i < n - 1 ? f[i + 1] - f[i] : f[n - 1] - f[n - 2]; };
  auto dfm = [&](int i) { return i > 0 ? f[i] - f[i - 1] : f[1] - f[0]; };
  auto df2 = [&](int i) {
    const T d = dfp(i) - dfm(i);
    return REC == REC_CRWENO5NS ? T(0.5) * tabs(d) : c1 * (d * d);
  };
  // weights (al, ac, ar) of cell i for the left (s = 0) or right state
  auto weights = [&](int i, int s, T& al, T& ac, T& ar) {
    if constexpr (REC == REC_CRWENO5NS) {
      const T df = s == 0 ? tabs(dfp(i)) : tabs(dfm(i)), lq = xi * df;
      const T bl = s == 0 ? df2(i - 1) + xi * tabs(T(2) * dfm(i) - dfm(i - 1)) : df2(i - 1) + lq;
      const T bc = df2(i) + lq;
      const T br = s == 0 ? df2(i + 1) + lq : df2(i + 1) + xi * tabs(T(2) * dfp(i) - dfp(i + 1));
      const T l3 = (df * df) * df, q3 = l3 / (T(1) + l3);
      const T zt = T(0.5) * ((bl - br) * (bl - br) + q3 * q3);
      al = T(1) + zt / ((bl + eps) * (bl + eps));
      ac = T(1) + zt / ((bc + eps) * (bc + eps));
      ar = T(1) + zt / ((br + eps) * (br + eps));
    } else {
      const T t1 = T(3) * dfm(i) - dfm(i - 1), t2 = dfp(i) + dfm(i), t3 = T(3) * dfp(i) - dfp(i + 1);
      const T bl = df2(i - 1) + c2 * (t1 * t1), bc = df2(i) + c2 * (t2 * t2);
      const T br = df2(i + 1) + c2 * (t3 * t3);
      T tt;
      if constexpr (REC == REC_CRWENO5YC) {
        const int im2 = i - 2 > 0 ? i - 2 : 0, ip2 = i + 2 < n - 1 ? i + 2 : n - 1;
        const T d4 = (T(6) * f[i] + (f[im2] + f[ip2])) - T(4) * (f[i - 1] + f[i + 1]);
        tt = d4 * d4;
      } else {
        tt = tabs(br - bl);
      }
      al = T(1) + tt / (bl + eps);
      ac = T(1) + tt / (bc + eps);
      ar = T(1) + tt / (br + eps);
    }
  };
  // row i of system s: sub-diagonal a, diagonal b, super-diagonal c, rhs r
  auto row = [&](int i, int s, T& a, T& b, T& c, T& r) {
    if (i == 0) { a = T(0); b = T(1); c = T(0); r = s == 0 ? T(0.5) * (f[0] + f[1]) : f[0]; return; }
    if (i == n - 1) {
      a = T(0); b = T(1); c = T(0); r = s == 0 ? f[n - 1] : T(0.5) * (f[n - 2] + f[n - 1]);
      return;
    }
    T al, ac, ar;
    weights(i, s, al, ac, ar);
    const bool expl = s == 0 ? (i <= ng - 1 || i >= n - ng - 1) : (i <= ng || i >= n - ng);
    if (expl) {
      const int im2 = i - 2 > 0 ? i - 2 : 0, ip2 = i + 2 < n - 1 ? i + 2 : n - 1;
      T wl = s == 0 ? dl * al : dl * ar, wc = dc * ac, wr = s == 0 ? dr * ar : dr * al;
      const T ww = (wl + wr) + wc;
      wl = wl / ww; wr = wr / ww; wc = T(1) - (wl + wr);
      T ql, qc, qr;
      if (s == 0) {
        ql = (a11 * f[im2] + a12 * f[i - 1]) + a13 * f[i];
        qc = (a21 * f[i - 1] + a22 * f[i]) + a23 * f[i + 1];
        qr = (a31 * f[i] + a32 * f[i + 1]) + a33 * f[ip2];
      } else {
        ql = (a11 * f[ip2] + a12 * f[i + 1]) + a13 * f[i];
        qc = (a21 * f[i + 1] + a22 * f[i]) + a23 * f[i - 1];
        qr = (a31 * f[i] + a32 * f[i - 1]) + a33 * f[im2];
      }
      a = T(0); b = T(1); c = T(0);
      r = (wl * ql + wr * qr) + wc * qc;
      return;
    }
    T wl = s == 0 ? cl * al : cl * ar, wc = cc * ac, wr = s == 0 ? cr * ar : cr * al;
    const T ww = (wl + wr) + wc;
    wl = wl / ww; wr = wr / ww; wc = T(1) - (wl + wr);
    b = wl + T(2) * (wc + wr);
    if (s == 0) {
      a = T(2) * wl + wc; c = wr;
      r = ((wl * f[i - 1] + (T(5) * (wl + wc) + wr) * f[i]) + (wc + T(5) * wr) * f[i + 1]) * dq;
    } else {
      a = wr; c = T(2) * wl + wc;
      r = ((wl * f[i + 1] + (T(5) * (wl + wc) + wr) * f[i]) + (wc + T(5) * wr) * f[i - 1]) * dq;
    }
  };
  T g[kMaxPencil], x[kMaxPencil];
  for (int s = 0; s < 2; s++) {
    T a, b, c, r, cprev;
    row(0, s, a, b, c, r);
    T t = b;
    x[0] = r / t;
    cprev = c;
    for (int i = 1; i < n; i++) {
      row(i, s, a, b, c, r);
      g[i] = cprev / t;
      t = b - a * g[i];
      // a zero pivot: tridiag prints "Could not find solution!" and leaves
      // the rest of the solution undefined (algebra.F90:1145-1152)
      if (t == T(0)) return false;
      x[i] = (r - a * x[i - 1]) / t;
      cprev = c;
    }
    for (int m = n - 2; m >= 0; m--) x[m] = x[m] - g[m + 1] * x[m + 1];
    // cells -1 .. nc of the block: pencil cells l = ng - 1 .. n - ng; the
    // left system solves the upper-face state of cell l, the right one (as
    // fr(1:n-1) = u(2:n)) the lower-face state of cell l at row l as well
    for (int cidx = 0; cidx < n - 2 * ng + 2; cidx++) {
      if (s == 0) ol[cidx] = x[ng - 1 + cidx];
      else orr[cidx] = x[ng - 1 + cidx];
    }
  }
  return true;
}

// reconstruct_crmp5 / _crmp5ld / _crmp7 (interpolations.F90:3930-4427) on
// one pencil f[0..n), n ghosts ng >= 2 each side: a tridiagonal system of
// constant rows (identity within ng cells of the ends, where the right-hand
// side is the explicit 2-, 3-, 5- or 7-point interpolation) solved by the
// Thomas algorithm, then the Suresh-Huynh limiter of each face value; the
// right states solve the same system on the mirrored pencil.  Writes ol / or
// as crweno_pencil does.
template <int REC, typename T>
__device__ bool crmp_pencil(int n, int ng, const T* f, T kappa, T kbeta, T eps, T* ol,
                            T* orr) {
  constexpr int V = REC == REC_CRMP5 ? 0 : (REC == REC_CRMP5LD ? 1 : 2);
  const T d0 = V == 0 ? T(0.3) : (V == 1 ? T(0.25) : T(2.0 / 7.0));
  const T d1 = V == 2 ? T(4.0 / 7.0) : T(0.6);
  const T d2 = V == 0 ? T(0.1) : (V == 1 ? T(0.15) : T(1.0 / 7.0));
  const T e5[5] = {T(2.0 / 60.0), T(-13.0 / 60.0), T(47.0 / 60.0), T(27.0 / 60.0), T(-3.0 / 60.0)};
  const T e3[3] = {T(-1.0 / 6.0), T(5.0 / 6.0), T(2.0 / 6.0)};
  const T e7[7] = {T(-3.0 / 420.0), T(25.0 / 420.0), T(-101.0 / 420.0), T(319.0 / 420.0),
                   T(214.0 / 420.0), T(-38.0 / 420.0), T(4.0 / 420.0)};
  T g[kMaxPencil], x[kMaxPencil];
  for (int s = 0; s < 2; s++) {
    // pencil of system s: f itself (s = 0) or mirrored (s = 1)
    auto F = [&](int m) { return s == 0 ? f[m] : f[n - 1 - m]; };
    auto sum = [&](const T* c, int k, int i0) {
      T a = c[0] * F(i0);
      for (int m = 1; m < k; m++) a = a + c[m] * F(i0 + m);
      return a;
    };
    auto rhs = [&](int i) -> T {
      if (i == 0) return T(0.5) * F(0) + T(0.5) * F(1);
      if (i == 1) return sum(e3, 3, 0);
      if (i == n - 2) return sum(e3, 3, n - 3);
      if (i == n - 1) return F(n - 1);
      if (V < 2) {
        if (i <= ng - 1 || i >= n - ng - 1) return sum(e5, 5, i - 2);
        if (V == 0) return (T(1.0 / 30.0) * F(i - 1) + T(19.0 / 30.0) * F(i)) + T(10.0 / 30.0) * F(i + 1);
        return ((T(3.0 / 120.0) * F(i - 1) + T(67.0 / 120.0) * F(i)) + T(49.0 / 120.0) * F(i + 1)) +
               T(1.0 / 120.0) * F(i + 2);
      }
      if (i == 2 || i == n - 3) return sum(e5, 5, i - 2);
      if (i <= ng - 1 || i >= n - ng - 1) return sum(e7, 7, i - 3);
      return (((T(-1.0 / 420.0) * F(i - 2) + T(19.0 / 420.0) * F(i - 1)) + T(239.0 / 420.0) * F(i)) +
              T(159.0 / 420.0) * F(i + 1)) + T(4.0 / 420.0) * F(i + 2);
    };
    // rows ng .. n - ng - 2 carry (d0, d1, d2), the others are identity
    auto inner = [&](int i) { return i >= ng && i <= n - ng - 2; };
    T t = T(1), cprev = T(0);
    x[0] = rhs(0);
    for (int i = 1; i < n; i++) {
      const T a = inner(i) ? d0 : T(0), b = inner(i) ? d1 : T(1);
      g[i] = cprev / t;
      t = b - a * g[i];
      if (t == T(0)) return false;   // as in crweno_pencil
      x[i] = (rhs(i) - a * x[i - 1]) / t;
      cprev = inner(i) ? d2 : T(0);
    }
    for (int m = n - 2; m >= 0; m--) x[m] = x[m] - g[m + 1] * x[m + 1];
    // the limited face values the block needs: system 0 at pencil cells
    // l = ng - 1 .. n - ng (upper faces), system 1 at the mirrored rows
    // n - 1 - l (lower faces, fr(1:n-1) = u(n-1:1:-1))
    for (int cidx = 0; cidx < n - 2 * ng + 2; cidx++) {
      const int q = s == 0 ? ng - 1 + cidx : n - ng - cidx;
      T w[5];
#pragma unroll
      for (int m = 0; m < 5; m++) {
        const int r = q - 2 + m;
        w[m] = F(r < 0 ? 0 : (r > n - 1 ? n - 1 : r));
      }
      const T u = mp_limit<T>(w, x[q], kappa, kbeta, eps);
      if (s == 0) ol[cidx] = u;
      else orr[cidx] = u;
    }
  }
  return true;
}

// the face states of one direction for the compact schemes: one thread per
// (block, pencil, variable) gathers the pencil with ng ghosts each side
template <class P, int REC, int ND, int DIR, typename T>
__global__ __launch_bounds__(64) void k_compact(const void* Qv, const int* __restrict__ nbr,
                                                RunConst rc, DevState* st) {
  constexpr int NV = P::nv;
  const int nc = rc.nc, ng = rc.ng, n = nc + 2 * ng, npen = ND == 3 ? nc * nc : nc;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)rc.nblk * npen * NV) return;
  const int v = (int)(g % NV), p = (int)((g / NV) % npen), b = (int)(g / ((long long)NV * npen));
  const int t1 = p % nc, t2 = ND == 3 ? p / nc : 0;
  const T* Q = static_cast<const T*>(Qv);
  T f[kMaxPencil], ol[kMaxPencil], orr[kMaxPencil];
  for (int m = 0; m < n; m++) {
    const int c = m - ng;       // the sweep's coordinates of pencil (t1, t2)
    const int i = DIR == 0 ? c : t1;
    const int j = DIR == 0 ? t1 : (DIR == 1 ? c : t2);
    const int k = DIR == 2 ? c : t2;
    f[m] = Q[var_index<NV>(rc, cell_index<sizeof(T)>(nbr, rc, b, i, j, k), v)];
  }
  bool ok;
  if constexpr (REC == REC_CRMP5 || REC == REC_CRMP5LD || REC == REC_CRMP7)
    ok = crmp_pencil<REC, T>(n, ng, f, T(rc.kappa), T(rc.kbeta), T(rc.eps), ol, orr);
  else
    ok = crweno_pencil<REC, T>(n, ng, f, T(rc.eps), ol, orr);
  if (!ok) {
    // a zero pivot in the pencil's solve: counted (amun_dev_tridiag_events)
    // and the pencil's face states set to its cell values (first order),
    // where the reference goes on with an undefined solution
    atomicAdd(&st->tridiag, 1ull);
    for (int c = 0; c < nc + 2; c++) ol[c] = orr[c] = f[ng - 1 + c];
  }
  T* L = static_cast<T*>(rc.fsl);
  T* Rr = static_cast<T*>(rc.fsr);
  for (int c = 0; c < nc + 2; c++) {
    const size_t o = ((size_t)(b * NV + v) * (nc + 2) + c) * npen + p;
    L[o] = ol[c];
    Rr[o] = orr[c];
  }
}

// TVD (interfaces_tvd, interpolations.F90:787-833): three-cell window along
// DIR; density and pressure also need the transverse slopes for the joint
// positivity halving
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_TVD, ND, DIR, T> {
  static constexpr int R = 1;
  static constexpr int H = 2;
  T w[P::nv][3];
  __device__ __forceinline__ void load(int slotm, const T* q) {
#pragma unroll
    for (int v = 0; v < P::nv; v++) w[v][slotm] = q[v];
  }
  __device__ __forceinline__ void shift() {
#pragma unroll
    for (int v = 0; v < P::nv; v++) {
      w[v][0] = w[v][1];
      w[v][1] = w[v][2];
    }
  }
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc,
                                       const T* __restrict__ Q,
                                       const int* __restrict__ nbr, int b, int i, int j,
                                       int k, DevState* st, double) {
#pragma unroll
    for (int v = 0; v < P::nv; v++) {
      const T qc = w[v][1];
      T dq = limiter<T>(rc.limiter, T(0.5), qc - w[v][0], w[v][2] - qc);
      if (v == P::idn || v == P::ipr) {
        T dqt[3] = {T(0), T(0), T(0)};
        dqt[DIR] = dq;
#pragma unroll
        for (int d = 0; d < ND; d++) {
          if (d == DIR) continue;
          int im = i - (d == 0), jm = j - (d == 1), km = k - (d == 2);
          int ip = i + (d == 0), jp = j + (d == 1), kp = k + (d == 2);
          T qm = Q[var_index<P::nv>(rc, cell_index<sizeof(T)>(nbr, rc, b, im, jm, km), v)];
          T qp = Q[var_index<P::nv>(rc, cell_index<sizeof(T)>(nbr, rc, b, ip, jp, kp), v)];
          dqt[d] = limiter<T>(rc.limiter, T(0.5), qc - qm, qp - qc);
        }
        if (qc > T(0)) {
          // a finite slope falls below qc within kMaxHalvings halvings; the
          // bound only stops an infinite one (the reference loops forever)
          for (int it = 0; it < kMaxHalvings; it++) {
            T sa = tabs(dqt[0]) + tabs(dqt[1]);
            if (ND == 3) sa += tabs(dqt[2]);
            if (!(qc <= sa)) break;
            dqt[0] *= T(0.5); dqt[1] *= T(0.5); dqt[2] *= T(0.5);
          }
        } else {
          atomicAdd(&st->npos, 1ull);
          dqt[0] = dqt[1] = dqt[2] = T(0);
        }
        dq = dqt[DIR];
      }
      flc[v] = qc + dq;
      frc[v] = qc - dq;
    }
  }
};

// Multi-dimensional Gaussian process (interfaces_mgp, interpolations.F90:
// 978-1144, mlp_limiting off): the 2 ND face values of the cell from the
// weights ugp over its ngp^ND neighbourhood (read from the pool, i fastest,
// as the reference's reshape); where any face value leaves the range of its
// two cells by more than eps, all of them fall back to the limited TVD slopes
// with the positivity halving.  The sweep along DIR keeps that direction's
// pair.
template <class P, int ND, int DIR, typename T>
struct Recon<P, REC_MGP, ND, DIR, T> {
  static constexpr int R = 0;
  static constexpr int H = 2;
  __device__ __forceinline__ void load(int, const T*) {}
  __device__ __forceinline__ void shift() {}
  __device__ __forceinline__ void cell(const RunConst& rc, T* flc, T* frc,
                                       const T* __restrict__ Q,
                                       const int* __restrict__ nbr, int b, int i, int j,
                                       int k, DevState* st, double) {
    const int mg = rc.mgp, dgp = rc.dgp;
    const double* __restrict__ ug = rc.ugp;
    const T eps = T(rc.eps);
    for (int v = 0; v < P::nv; v++) {
      auto at = [&](int a, int bb, int c) {
        return Q[var_index<P::nv>(rc, cell_index<sizeof(T)>(nbr, rc, b, i + a, j + bb, k + c), v)];
      };
      const T qc = at(0, 0, 0);
      T fv[2 * ND];
#pragma unroll
      for (int s = 0; s < 2 * ND; s++) fv[s] = T(0);
      int m = 0;
      for (int c = (ND == 3 ? -mg : 0); c <= (ND == 3 ? mg : 0); c++)
        for (int bb = -mg; bb <= mg; bb++)
          for (int a = -mg; a <= mg; a++, m++) {
            const T u = at(a, bb, c) - qc;
#pragma unroll
            for (int s = 0; s < 2 * ND; s++) fv[s] = fv[s] + T(ug[s * dgp + m]) * u;
          }
      T qn[2 * ND];   // the upper and lower neighbours along each direction
#pragma unroll
      for (int d = 0; d < ND; d++) {
        qn[2 * d] = at(d == 0, d == 1, d == 2);
        qn[2 * d + 1] = at(-(d == 0), -(d == 1), -(d == 2));
      }
      bool flag = false;
#pragma unroll
      for (int s = 0; s < 2 * ND; s++) {
        fv[s] = fv[s] + qc;
        flag = flag || ((fv[s] - qn[s]) * (fv[s] - qc) > eps);
      }
      if (flag) {
        T dq[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int d = 0; d < ND; d++)
          dq[d] = limiter<T>(rc.limiter, T(0.5), qc - qn[2 * d + 1], qn[2 * d] - qc);
        if (v == P::idn || v == P::ipr) {
          if (qc > T(0)) {
            for (int it = 0; it < kMaxHalvings; it++) {
              T sa = tabs(dq[0]) + tabs(dq[1]);
              if (ND == 3) sa += tabs(dq[2]);
              if (!(qc <= sa)) break;
              dq[0] *= T(0.5); dq[1] *= T(0.5); dq[2] *= T(0.5);
            }
          } else {
            atomicAdd(&st->npos, 1ull);
            dq[0] = dq[1] = dq[2] = T(0);
          }
        }
        fv[2 * DIR] = qc + dq[DIR];
        fv[2 * DIR + 1] = qc - dq[DIR];
      }
      flc[v] = fv[2 * DIR];
      frc[v] = fv[2 * DIR + 1];
    }
  }
};

// Counts relativistic primitive-solver events (the reference prints each to
// stderr and goes on).
__device__ __forceinline__ void c2p_count(DevState* st, int cs) {
  if (cs & (C2P_FAIL | C2P_LOWER)) atomicAdd(&st->c2p[0], 1ull);
  if (cs & C2P_PNEG) atomicAdd(&st->c2p[1], 1ull);
  if (cs & C2P_NOCONV) atomicAdd(&st->c2p[2], 1ull);
}

// A cell whose bracketing failed keeps its previous primitive values (read
// from Qold at offset base, variable stride ncell); p = 0 is already set.
template <class P, typename T>
__device__ __forceinline__ void c2p_keep_old(int cs, double* q, const T* __restrict__ Qold,
                                             size_t base, int ncell) {
  if (cs & C2P_FAIL) {
#pragma unroll
    for (int v = 0; v < P::nv; v++)
      if (v != P::ipr) q[v] = double(Qold[base + (size_t)v * ncell]);
  }
}

// Workgroup index of this workgroup.  The dispatcher deals workgroups out to
// the eight XCDs round robin (workgroup i runs on XCD i mod 8), and each XCD
// has its own L2.  With swz set, the index is remapped so that every XCD gets
// one contiguous range of the pencil space (neighbouring pencils, hence the
// same bricks and the same neighbour blocks, share an L2); the map is a
// bijection of [0, gridDim.x).
__device__ __forceinline__ int workgroup_index(int swz) {
  const int i = blockIdx.x;
  if (!swz) return i;
  const int g = gridDim.x, q = g >> 3, r = g & 7, x = i & 7;
  return x * q + (x < r ? x : r) + (i >> 3);
}

// Maximum speed of a wave into its run's accumulator: one atomic per wave
// when every lane belongs to the same run (the common case), one per lane
// otherwise (a wave straddling two grids of a batch).  Positive doubles
// order as their bit patterns.
__device__ __forceinline__ void wave_max_speed(RunState* rs, int run, int slot, double smax) {
  const int r0 = __shfl(run, 0);
  if (__all(run == r0)) {
    for (int off = 32; off > 0; off >>= 1) smax = fmax(smax, __shfl_xor(smax, off));
    if ((threadIdx.x & 63) == 0 && smax > 0.0)
      atomicMax(&rs[r0].cmax_bits[slot], (unsigned long long)__double_as_longlong(smax));
  } else if (smax > 0.0) {
    atomicMax(&rs[run].cmax_bits[slot], (unsigned long long)__double_as_longlong(smax));
  }
}

// Maximum speed of a sweep workgroup (at most kMaxWaves waves) into
// part[blockIdx.x], reduced over the workgroups by k_reduce_speed after the
// sweep.  Device-scope atomics on one address from every wave of a 256^3
// sweep (16K of them) serialise at the memory side; a workgroup whose lanes
// belong to more than one run (a batch) still uses them.  Every thread of
// the workgroup must call this.
constexpr int kMaxWaves = 16;
__device__ __forceinline__ void wg_max_speed(RunState* rs, SpeedPart* part, int run, int slot,
                                             double smax) {
  __shared__ double wm[kMaxWaves];
  __shared__ int wr[kMaxWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  const int r0 = __shfl(run, 0);
  const bool uni = __all(run == r0);
  if (uni) {
    for (int off = 32; off > 0; off >>= 1) smax = fmax(smax, __shfl_xor(smax, off));
  } else if (smax > 0.0) {
    atomicMax(&rs[run].cmax_bits[slot], (unsigned long long)__double_as_longlong(smax));
  }
  if (lane == 0) {
    wm[w] = uni ? smax : 0.0;
    wr[w] = uni ? r0 : -1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wm[0];
    bool same = wr[0] >= 0;
    for (int i = 1; i < nw; i++) {
      same = same && wr[i] == wr[0];
      m = fmax(m, wm[i]);
    }
    if (same) {
      part[blockIdx.x].smax = m;
      part[blockIdx.x].run = wr[0];
    } else {
      part[blockIdx.x].smax = 0.0;
      part[blockIdx.x].run = -1;
      for (int i = 0; i < nw; i++)
        if (wr[i] >= 0 && wm[i] > 0.0)
          atomicMax(&rs[wr[i]].cmax_bits[slot], (unsigned long long)__double_as_longlong(wm[i]));
    }
  }
}

// The per-workgroup maxima of the last sweep into the runs' cmax slots (one
// workgroup of 1024 threads; per-run maxima in LDS when the runs fit)
constexpr int kReduceRuns = 2048;
static __global__ __launch_bounds__(1024) void k_reduce_speed(const SpeedPart* __restrict__ part,
                                                              int n, RunState* rs, int slot,
                                                              int nruns) {
  __shared__ unsigned long long lm[kReduceRuns];
  if (nruns <= kReduceRuns) {
    for (int r = threadIdx.x; r < nruns; r += blockDim.x) lm[r] = 0ull;
    __syncthreads();
    if (nruns == 1) {
      double m = 0.0;
      for (int i = threadIdx.x; i < n; i += blockDim.x)
        if (part[i].run == 0) m = fmax(m, part[i].smax);
      for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
      if ((threadIdx.x & 63) == 0 && m > 0.0)
        atomicMax(&lm[0], (unsigned long long)__double_as_longlong(m));
    } else {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const SpeedPart p = part[i];
        if (p.run >= 0 && p.smax > 0.0)
          atomicMax(&lm[p.run], (unsigned long long)__double_as_longlong(p.smax));
      }
    }
    __syncthreads();
    for (int r = threadIdx.x; r < nruns; r += blockDim.x)
      if (lm[r]) atomicMax(&rs[r].cmax_bits[slot], lm[r]);
  } else {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const SpeedPart p = part[i];
      if (p.run >= 0 && p.smax > 0.0)
        atomicMax(&rs[p.run].cmax_bits[slot], (unsigned long long)__double_as_longlong(p.smax));
    }
  }
}

// ---------------------------------------------------------------------------
// the sweep kernel
// ---------------------------------------------------------------------------
// FP64 sweeps hold more than 256 registers per lane and run one wave per
// SIMD (spilling to reach two is several times slower); FP32 sweeps fit two.
template <typename T>
constexpr int sweep_min_waves() { return sizeof(T) == 4 ? 2 : 1; }

// The FP64 x and y sweeps also move the increment through LDS in chunks; the
// FP32 ones read and write it per cell, which leaves LDS for two workgroups
// per CU.
#include "mlp.h"

template <typename T>
constexpr bool sweep_chunk_d() { return sizeof(T) == 8; }

// the sweep along DIR of one workgroup: its lanes walk the pencils
// workgroup_index(swz) * blockDim.x + threadIdx.x along DIR.  Every thread
// of the workgroup must call it (the last sweep reduces the maximum speed
// over the workgroup).
template <class P, int REC, int RS, int ND, int DIR, typename T>
__device__ __forceinline__ void sweep_body(const SweepArgs& a, const RunConst& rc) {
  constexpr int NV = P::nv;
  constexpr bool kLast = (DIR == ND - 1);
  using RC = Recon<P, REC, ND, DIR, T>;
  // chunked pencil I/O along x and y (see loadq below)
  constexpr bool kChunk = DIR < 2;
  constexpr bool kDin = kChunk && DIR > 0 && sweep_chunk_d<T>();    // increment read in chunks
  constexpr bool kDout = kChunk && !kLast && sweep_chunk_d<T>();    // increment written in chunks
  // per-lane LDS slots of the chunked pencil I/O: sq holds four cells of the
  // primitive window, sd four cells of the increment (4 NV values per lane
  // each), slot-major with a fixed stride of kLdsLanes lanes (consecutive
  // lanes hit consecutive banks; a compile-time stride keeps the slot
  // addresses out of the register file).  The chunked sweeps launch
  // workgroups of at most kLdsLanes lanes.
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int wg = kLdsLanes;
  T* const sq = reinterpret_cast<T*>(smem);
  T* const sd = sq + (kChunk ? 4 * NV * wg : 0);
  const int nc = rc.nc;
  const int npen = (ND == 3) ? nc * nc : nc;
  const long long gid = (long long)workgroup_index(a.swz) * blockDim.x + threadIdx.x;
  const int b = (int)(gid / npen);
  const int t = (int)(gid % npen);
  const bool active = b < rc.nblk;

  if (a.first && gid < rc.nruns) {
    // reset the maximum-speed accumulators of the next step (new_time_step:
    // cmax = tiny before the reduction), one thread per run
    a.rs[gid].cmax_bits[a.slot ^ 1] = (unsigned long long)__double_as_longlong(kTiny);
  }

  const int run = active ? a.brun[b] : 0;
  const RunState* R = a.rs + run;
  const double cmaxd = cmax_of(R, a.slot);
  const T cmax = T(cmaxd), cmax2 = T(cmaxd * cmaxd);
  const double dt = dt_of(rc, R, a.slot, a.dtnext);
  const double ds = a.fac * dt;
  const EosConst<T> e{T(rc.gamma), T(rc.gammam1), T(rc.gammam1i), T(rc.csnd), T(rc.csnd2)};
  const EosConst<double> ed{rc.gamma, rc.gammam1, rc.gammam1i, rc.csnd, rc.csnd2};
  const T* __restrict__ Qp = static_cast<const T*>(a.Q);
  T* __restrict__ Dp = static_cast<T*>(a.D);
  T* __restrict__ Qo = static_cast<T*>(a.Qout);
  double smax = 0.0;

  if (active) {
    // transverse coordinates of the pencil
    int c1 = t % nc, c2 = (ND == 3) ? t / nc : 0;
    auto coords = [&](int c, int& i, int& j, int& k) {
      if constexpr (DIR == 0) { i = c; j = c1; k = c2; }
      else if constexpr (DIR == 1) { i = c1; j = c; k = c2; }
      else { i = c1; j = c2; k = c; }
    };
    // x and y sweeps move their pencil data in chunks of four cells: a
    // brick line spans four cells along x and along y, so loading (and
    // storing) four cells at once uses whole lines instead of a quarter of a
    // line per walk step.  The chunks live in per-lane LDS slots.
    const int tid = threadIdx.x;
    auto loadq = [&](int c, T* q) {
      if constexpr (kChunk) {
        const int first = -1 - RC::R;
        if (c == first || (c & 3) == 0) {
          const int c0 = c & ~3;
#pragma unroll
          for (int m = 0; m < 4; m++) {
            int i, j, k;
            coords(c0 + m, i, j, k);
            const size_t pk = cell_index<sizeof(T)>(a.nbr, rc, b, i, j, k);
#pragma unroll
            for (int v = 0; v < NV; v++) sq[(m * NV + v) * wg + tid] = Qp[var_index<NV>(rc, pk, v)];
          }
        }
#pragma unroll
        for (int v = 0; v < NV; v++) q[v] = sq[((c & 3) * NV + v) * wg + tid];
      } else {
        int i, j, k;
        coords(c, i, j, k);
        size_t pk = cell_index<sizeof(T)>(a.nbr, rc, b, i, j, k);
#pragma unroll
        for (int v = 0; v < NV; v++) q[v] = Qp[var_index<NV>(rc, pk, v)];
      }
    };
    // global offsets (without the variable term) of the four cells of the
    // increment chunk that starts at cell c0
    auto dchunk_base = [&](int c0, size_t* base) {
#pragma unroll
      for (int m = 0; m < 4; m++) {
        int i, j, k;
        coords(c0 + m, i, j, k);
        base[m] = (size_t)b * NV * (size_t)rc.ncell + (size_t)lidx<sizeof(T)>(nc, ND, i, j, k);
      }
    };

    // cell size along DIR at this block's level (coordinates.F90:330-341)
    const double hdir = a.blev ? R->h[DIR] / double(1 << a.blev[b]) : R->h[DIR];
    RC rec;
    T qtmp[NV];
    // prime the window for cell c = -1: cells -1-R .. -1+R-1
#pragma unroll
    for (int m = 0; m < 2 * RC::R; m++) {
      loadq(-1 - RC::R + m, qtmp);
      rec.load(m, qtmp);
    }
    T flp[NV], fprev[NV];
    // unchunked sweeps: the window's next cell is requested one iteration
    // ahead of its use
    T qnext[NV];
    if constexpr (!kChunk) loadq(RC::R - 1, qnext);
    for (int c = -1; c <= nc; c++) {
      // operands of the cell finished in this iteration (c-1), loaded before
      // the reconstruction and Riemann solve so that their latency overlaps
      // the arithmetic
      T dpre[NV];
      double u0pre[NV], u1pre[NV];
      float cpre[NV];
      if (c >= 1) {
        int i, j, k;
        coords(c - 1, i, j, k);
        const size_t bb = (size_t)b * NV * (size_t)rc.ncell;
        const size_t bt = bb + (size_t)lidx<sizeof(T)>(nc, ND, i, j, k);
        const size_t bd = bb + (size_t)lidx<8>(nc, ND, i, j, k);
        const size_t bf = bb + (size_t)lidx<4>(nc, ND, i, j, k);
        if constexpr (kDin) {
          if (((c - 1) & 3) == 0) {
            size_t base[4];
            dchunk_base(c - 1, base);
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
              for (int v = 0; v < NV; v++)
                sd[(m * NV + v) * wg + tid] = Dp[base[m] + (size_t)v * rc.ncell];
          }
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const size_t off = (size_t)v * rc.ncell;
          if constexpr (DIR > 0 && !kDin) dpre[v] = Dp[bt + off];
          if constexpr (kLast) {
            u0pre[v] = a.U0[bd + off];
            if (a.kind != 0) u1pre[v] = a.U1[bd + off];
            // the compensation enters the updates of the step state; the
            // transient stages ignore it (it is below the FP64 rounding of U0)
            if constexpr (sizeof(T) == 4) cpre[v] = a.comp ? a.C[bf + off] : 0.0f;
          }
        }
      }
      if (c > -1) rec.shift();
      if constexpr (kChunk) {
        loadq(c + RC::R, qtmp);
        rec.load(2 * RC::R, qtmp);
      } else {
        rec.load(2 * RC::R, qnext);
        if (c < nc) loadq(c + RC::R + 1, qnext);
      }
      T flc[NV], frc[NV];
      int ci, cj, ck;
      coords(c, ci, cj, ck);
      rec.cell(rc, flc, frc, Qp, a.nbr, b, ci, cj, ck, a.st, hdir);
      if (a.mlps)
        mlp_apply<P, ND, DIR, T>(rc, Qp, static_cast<const T*>(a.mlps),
                                 static_cast<const T*>(a.mlpg), a.nbr, b, c, t, ci, cj, ck,
                                 c == -1 && rc.ng == 2, flc, frc);
      if constexpr (P::sr) {
        if (rc.v4) {
          states_from_4vec<P, T>(flc);
          states_from_4vec<P, T>(frc);
        }
      }
      if (c >= 0) {
        // face c-1/2: left state from cell c-1, right state from cell c
        T ql[NV], qr[NV], fr[NV], f[NV];
#pragma unroll
        for (int p = 0; p < NV; p++) {
          ql[p] = flp[rot<P, DIR>(p)];
          qr[p] = frc[rot<P, DIR>(p)];
        }
        riemann<P, RS, T>(e, rc.sc, cmax, cmax2, ql, qr, fr);
#pragma unroll
        for (int p = 0; p < NV; p++) f[rot<P, DIR>(p)] = fr[p];
        if (a.ffix && (c == 0 || c == nc)) {
          // a face with finer neighbours: their averaged fluxes
          // (boundary_fluxes, boundaries.F90:494-1060)
          const int fs = a.ffix[b * 6 + 2 * DIR + (c == nc)];
          if (fs >= 0) {
            const int nf = (ND == 3) ? nc * nc : nc;
#pragma unroll
            for (int p = 0; p < NV; p++) f[p] = T(a.FB[((size_t)fs * NV + p) * nf + t]);
          }
        }
        if (c >= 1) {
          // cell c-1 has both faces
          int i, j, k;
          coords(c - 1, i, j, k);
          const size_t bb = (size_t)b * NV * (size_t)rc.ncell;
          const size_t bt = bb + (size_t)lidx<sizeof(T)>(nc, ND, i, j, k);
          const size_t bd = bb + (size_t)lidx<8>(nc, ND, i, j, k);
          const size_t bf = bb + (size_t)lidx<4>(nc, ND, i, j, k);
          const T dxi = T(a.blev ? R->dxi[DIR] * double(1 << a.blev[b]) : R->dxi[DIR]);
          double unew[NV];
#pragma unroll
          for (int v = 0; v < NV; v++) {
            const size_t off = (size_t)v * rc.ncell;
            // update_increment order: -dxi*dfx - dyi*dfy - dzi*dfz
            const T dfx = f[v] - fprev[v];
            T du;
            if constexpr (DIR == 0) du = -dxi * dfx;
            else if constexpr (kDin) du = sd[(((c - 1) & 3) * NV + v) * wg + tid] - dxi * dfx;
            else du = dpre[v] - dxi * dfx;
            if constexpr (kLast) {
              // gravity (update_sources, sources.F90:288-335): du(m) += rho g,
              // du(E) += rho g . v with g = (0, gacc, 0), zero terms included
              if (rc.grav) {
#pragma clang fp contract(off)
                const T* Qg = static_cast<const T*>(a.Qold);
                const size_t n1 = (size_t)rc.ncell;
                const T rho = Qg[bt + (size_t)P::idn * n1];
                const T gx = rho * T(0), gy = rho * T(rc.gacc), gz = rho * T(0);
                if (v == P::imx) du = du + gx;
                if (v == P::imy) du = du + gy;
                if (ND == 3 && v == P::imz) du = du + gz;
                if (v == P::ien) {
                  du = (du + gx * Qg[bt + (size_t)P::ivx * n1]) + gy * Qg[bt + (size_t)P::ivy * n1];
                  if (ND == 3) du = du + gz * Qg[bt + (size_t)P::ivz * n1];
                }
              }
              // viscous source terms after the completed increment
              // (update_sources, sources.F90:340-467; products of k_visc_src)
              if (a.vsrc && (a.vis & 1)) {
#pragma clang fp contract(off)
                const T* V = static_cast<const T*>(a.vsrc) + (size_t)b * 9 * rc.ncell + (bt - bb);
                const size_t n1 = (size_t)rc.ncell;
                if (v == P::imx) du = du + V[0];
                if (v == P::imy) du = du + V[n1];
                if (v == P::imz) du = du + V[2 * n1];
                if (v == P::ien) du = du + V[3 * n1];
              }
            }
            if constexpr (kLast && P::mhd) {
              // GLM source terms after the completed increment
              // (update_sources, sources.F90:479-539; products from k_glm_src)
              if (a.src) {
#pragma clang fp contract(off)
                const T* Sp = static_cast<const T*>(a.src) + (size_t)b * 8 * rc.ncell + (bt - bb);
                const size_t n1 = (size_t)rc.ncell;
                if (v == P::imx) du = du - Sp[0];
                if (v == P::imy) du = du - Sp[n1];
                if (v == P::imz) du = du - Sp[2 * n1];
                if (v == P::ien) du = du - Sp[3 * n1];
                if (a.glm > 1) {
                  if (v == P::ibx) du = du - Sp[4 * n1];
                  if (v == P::iby) du = du - Sp[5 * n1];
                  if (v == P::ibz) du = du - Sp[6 * n1];
                  if (v == P::ien) du = du - Sp[7 * n1];
                }
              }
              // resistive source terms after the GLM ones (sources.F90:543-598)
              if (a.vsrc && (a.vis & 2)) {
#pragma clang fp contract(off)
                const T* V = static_cast<const T*>(a.vsrc) + (size_t)b * 9 * rc.ncell + (bt - bb);
                const size_t n1 = (size_t)rc.ncell;
                if (v == P::ibx) du = du + V[4 * n1];
                if (v == P::iby) du = du + V[5 * n1];
                if (v == P::ibz) du = du + V[6 * n1];
                if (v == P::ien) du = (du + V[7 * n1]) + V[8 * n1];
              }
            }
            if constexpr (!kLast) {
              if constexpr (kDout) sd[(((c - 1) & 3) * NV + v) * wg + tid] = du;
              else Dp[bt + off] = du;
            } else if constexpr (sizeof(T) == 8) {
              double un = combine(a.kind, u0pre[v], u1pre[v], ds, du, a.ca, a.cb);
              if (a.last && v == P::ibp) un = rc.decay * un;
              unew[v] = un;
            } else {
              // mixed precision: the FP32 increment is added to the FP64
              // state in increment form with Kahan compensation.  The true
              // step state is U0 - C; a stage that writes the step state
              // forms U0_true + delta and keeps the rounding error of that
              // sum in C (the stages into U1 use U0 alone).
              const double u0 = u0pre[v];
              const double cp = double(cpre[v]);
              const double dd = ds * double(du);
              const double y =
                  combine_delta(a.kind, u0, u1pre[v], cp, dd, a.ca, a.cb) - cp;
              double un = u0 + y;
              if (a.comp) {
                double cn = (un - u0) - y;
                if (a.last && v == P::ibp) { un = rc.decay * un; cn = rc.decay * cn; }
                a.C[bf + off] = float(cn);
              }
              unew[v] = un;
            }
          }
          if constexpr (kLast) {
            double qn[NV];
            if constexpr (P::sr) {
              const int cs = cons2prim_sr<P>(rc.sc, unew, qn);
              if (cs) {
                c2p_keep_old<P, T>(cs, qn, static_cast<const T*>(a.Qold), bt, rc.ncell);
                c2p_count(a.st, cs);
              }
            } else {
              cons2prim<P, double>(ed, unew, qn);
            }
#pragma unroll
            for (int v = 0; v < NV; v++) {
              const size_t off = (size_t)v * rc.ncell;
              a.Uout[bd + off] = unew[v];
              Qo[bt + off] = T(qn[v]);
            }
            if (a.last) smax = fmax(smax, maxspeed<P, double>(ed, qn));
          }
        }
        if constexpr (kDout) {
          if (c >= 1 && ((c - 1) & 3) == 3) {
            size_t base[4];
            dchunk_base(c - 4, base);
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
              for (int v = 0; v < NV; v++)
                Dp[base[m] + (size_t)v * rc.ncell] = sd[(m * NV + v) * wg + tid];
          }
        }
#pragma unroll
        for (int p = 0; p < NV; p++) fprev[p] = f[p];
      }
#pragma unroll
      for (int p = 0; p < NV; p++) flp[p] = flc[p];
    }
  }

  if constexpr (kLast) {
    if (a.last) wg_max_speed(a.rs, a.spart, run, a.slot ^ 1, smax);
  }
}

template <class P, int REC, int RS, int ND, int DIR, typename T>
__global__ __launch_bounds__(256, sweep_min_waves<T>()) void k_sweep(SweepArgs a, RunConst rc) {
  sweep_body<P, REC, RS, ND, DIR, T>(a, rc);
}

// One launch per stage: when a workgroup's pencils are whole blocks in
// every direction (npen divides the workgroup size), the sweeps of one
// block depend only on each other through its own increment D (each
// direction reads the stage input Q, including the neighbours' halo, which
// no sweep of the stage writes), so one workgroup runs its blocks' x, y (and
// z) sweeps in turn with a workgroup barrier between them; D stays in the
// caches between the directions.
template <class P, int REC, int RS, int ND, typename T>
__global__ __launch_bounds__(256, sweep_min_waves<T>()) void k_stage(SweepArgs a, RunConst rc) {
  sweep_body<P, REC, RS, ND, 0, T>(a, rc);
  __syncthreads();
  a.first = 0;
  sweep_body<P, REC, RS, ND, 1, T>(a, rc);
  if constexpr (ND == 3) {
    __syncthreads();
    sweep_body<P, REC, RS, ND, 2, T>(a, rc);
  }
}

// ---------------------------------------------------------------------------
// refinement criterion of the leaf blocks (check_refinement_criterion,
// refinement.F90:257-640), one workgroup per block
// ---------------------------------------------------------------------------
struct RefineArgs {
  double crefmin, crefmax, epsref, vortmin, vortmax, jabsmin, jabsmax;
  int varmask, vort, jabs;
};

// one cell's normalised second difference along each direction (the
// reference's real(4) differences and magnitudes, real(8) quotient,
// real(4) result); contraction off so that the real(8) sums round as there
#pragma clang fp contract(off)
__device__ __forceinline__ float sde_term(double qp, double qm, double qc, double epsref) {
  const float fr = float(qp - qc);
  const float fl = float(qm - qc);
  const float fc = float((fabs(qp) + fabs(qm)) + 2.0 * fabs(qc));
  const float fsum = fabsf(fr) + fabsf(fl);
  return float(double(fabsf(fr + fl)) / ((double(fsum) + epsref * double(fc)) + 2.220446049250313e-16));
}

// |curl|^2 of the vector at variables i0..i0+2 with unit spacing (curl,
// operators.F90:328-434), summed in real(8)
template <int NV, int ND, typename T>
__device__ __forceinline__ double curl2(const T* __restrict__ Q, const int* __restrict__ nbr,
                                        const RunConst& rc, int b, int i, int j, int k, int i0) {
  auto at = [&](int c, int di, int dj, int dk) {
    return double(Q[var_index<NV>(rc, cell_index<sizeof(T)>(nbr, rc, b, i + di, j + dj, k + dk),
                                  i0 + c)]);
  };
  auto d1 = [&](int c, int di, int dj, int dk) {
    return 0.5 * (at(c, di, dj, dk) - at(c, -di, -dj, -dk)) / 1.0;
  };
  double wx, wy, wz;
  if constexpr (ND == 3) {
    wx = d1(2, 0, 1, 0) - d1(1, 0, 0, 1);
    wy = d1(0, 0, 0, 1) - d1(2, 1, 0, 0);
    wz = d1(1, 1, 0, 0) - d1(0, 0, 1, 0);
  } else {
    wx = d1(2, 0, 1, 0);
    wy = -d1(2, 1, 0, 0);
    wz = d1(1, 1, 0, 0) - d1(0, 0, 1, 0);
  }
  return (wx * wx + wy * wy) + wz * wz;
}
#pragma clang fp contract(on)

// per block: crit[b] = -1 / 0 / 1 and ind[3 b ..] = the largest
// second-derivative error over the selected variables, the vorticity and
// the current-density magnitude, over the cells one layer beyond the
// interior (ibl .. ieu), from the primitive state Q (ghosts filled)
template <class P, typename T>
__global__ __launch_bounds__(256) void k_refine(const void* Qv, const int* __restrict__ nbr,
                                                RunConst rc, RefineArgs ra, int* crit,
                                                float* ind) {
  constexpr int NV = P::nv;
  const T* __restrict__ Q = static_cast<const T*>(Qv);
  const int b = blockIdx.x;
  const int nc = rc.nc, m = nc + 2, nd = rc.nd;
  const int ncell = nd == 3 ? m * m * m : m * m;
  float err[NV], w2 = 0.0f, j2 = 0.0f;
#pragma unroll
  for (int v = 0; v < NV; v++) err[v] = 0.0f;
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const int i = c % m - 1, j = (c / m) % m - 1, k = nd == 3 ? c / (m * m) - 1 : 0;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      if (!(ra.varmask & (1 << v))) continue;
      auto at = [&](int di, int dj, int dk) {
        return double(Q[var_index<NV>(rc, cell_index<sizeof(T)>(nbr, rc, b, i + di, j + dj, k + dk), v)]);
      };
      const double qc = at(0, 0, 0);
      float e = err[v];
      e = fmaxf(e, sde_term(at(1, 0, 0), at(-1, 0, 0), qc, ra.epsref));
      e = fmaxf(e, sde_term(at(0, 1, 0), at(0, -1, 0), qc, ra.epsref));
      if (nd == 3) e = fmaxf(e, sde_term(at(0, 0, 1), at(0, 0, -1), qc, ra.epsref));
      err[v] = e;
    }
    if (ra.vort) {
      const double s = nd == 3 ? curl2<NV, 3, T>(Q, nbr, rc, b, i, j, k, P::ivx)
                               : curl2<NV, 2, T>(Q, nbr, rc, b, i, j, k, P::ivx);
      w2 = fmaxf(w2, float(s));
    }
    if constexpr (P::mhd) {
      if (ra.jabs) {
        const double s = nd == 3 ? curl2<NV, 3, T>(Q, nbr, rc, b, i, j, k, P::ibx)
                                 : curl2<NV, 2, T>(Q, nbr, rc, b, i, j, k, P::ibx);
        j2 = fmaxf(j2, float(s));
      }
    }
  }
  // workgroup maxima (order-free: maxima of floats)
  __shared__ float red[NV + 2][256];
#pragma unroll
  for (int v = 0; v < NV; v++) red[v][threadIdx.x] = err[v];
  red[NV][threadIdx.x] = w2;
  red[NV + 1][threadIdx.x] = j2;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int v = 0; v < NV + 2; v++)
        red[v][threadIdx.x] = fmaxf(red[v][threadIdx.x], red[v][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int cr = -1;
    float emax = 0.0f;
    for (int v = 0; v < NV; v++) {
      if (!(ra.varmask & (1 << v))) continue;
      const float cref = red[v][0];
      if (double(cref) > ra.crefmin) cr = cr > 0 ? cr : 0;
      if (double(cref) > ra.crefmax) cr = 1;
      emax = fmaxf(emax, cref);
    }
    const float wv = ra.vort ? sqrtf(red[NV][0]) : 0.0f;
    if (ra.vort) {
      if (double(wv) > ra.vortmin) cr = cr > 0 ? cr : 0;
      if (double(wv) > ra.vortmax) cr = 1;
    }
    const float jv = (P::mhd && ra.jabs) ? sqrtf(red[NV + 1][0]) : 0.0f;
    if (P::mhd && ra.jabs) {
      if (double(jv) > ra.jabsmin) cr = cr > 0 ? cr : 0;
      if (double(jv) > ra.jabsmax) cr = 1;
    }
    crit[b] = cr;
    ind[3 * b] = emax;
    ind[3 * b + 1] = wv;
    ind[3 * b + 2] = jv;
  }
}

// Q = cons2prim(U) on every interior cell, and the maximum speed into a slot
// (update_primitive_variables + maxspeed); used after uploads
template <class P, typename T>
__global__ __launch_bounds__(256) void k_prim(const double* __restrict__ U,
                                              void* Qv, RunConst rc, DevState* st,
                                              RunState* rs, const int* __restrict__ brun,
                                              int slot, int b0, int nb) {
  T* __restrict__ Q = static_cast<T*>(Qv);
  constexpr int NV = P::nv;
  const long long n = (long long)nb * rc.ncell;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const EosConst<double> e{rc.gamma, rc.gammam1, rc.gammam1i, rc.csnd, rc.csnd2};
  double smax = 0.0;
  int run = 0;
  if (gid < n) {
    const long long b = b0 + gid / rc.ncell, c = gid % rc.ncell;
    run = brun[b];
    const int i = (int)(c % rc.nc), j = (int)((c / rc.nc) % rc.nc), k = (int)(c / ((long long)rc.nc * rc.nc));
    const long long cu = lidx<8>(rc.nc, rc.nd, i, j, k), cq = lidx<sizeof(T)>(rc.nc, rc.nd, i, j, k);
    double u[NV], q[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = U[(b * NV + v) * rc.ncell + cu];
    if constexpr (P::sr) {
      const int cs = cons2prim_sr<P>(rc.sc, u, q);
      if (cs) {
        c2p_keep_old<P, T>(cs, q, Q, (size_t)b * NV * rc.ncell + cq, rc.ncell);
        c2p_count(st, cs);
      }
    } else {
      cons2prim<P, double>(e, u, q);
    }
#pragma unroll
    for (int v = 0; v < NV; v++) Q[(b * NV + v) * rc.ncell + cq] = T(q[v]);
    smax = maxspeed<P, double>(e, q);
  }
  wave_max_speed(rs, run, slot, smax);
}

static __global__ void k_zero_f32(float* x, long long n) {
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x)
    x[g] = 0.0f;
}

static __global__ void k_reset_slot(RunState* rs, int slot, int r0, int nr) {
  const int r = r0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r < r0 + nr) rs[r].cmax_bits[slot] = (unsigned long long)__double_as_longlong(kTiny);
}

// end of a step, one thread per run: dt into the trace ring
// trace[(step mod tsteps) * nruns + run], time and step advanced
static __global__ void k_end_step(RunState* rs, double* trace, int tsteps, RunConst rc, int slot,
                           double dtnext) {
  const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= rc.nruns) return;
  RunState* R = rs + r;
  const double dt = dt_of(rc, R, slot, dtnext);
  trace[(size_t)(R->step % tsteps) * rc.nruns + r] = dt;
  R->dt = dt;
  R->time += dt;
  R->step += 1;
}

// dt of the current slot without advancing (new_time_step alone), per run:
// out[2 r] = dt, out[2 r + 1] = cmax
static __global__ void k_peek_dt(RunState* rs, RunConst rc, int slot, double dtnext, double* out) {
  const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= rc.nruns) return;
  RunState* R = rs + r;
  const double dt = dt_of(rc, R, slot, dtnext);
  R->dt = dt;
  out[2 * r] = dt;
  out[2 * r + 1] = cmax_of(R, slot);
}

// volume integrals (store_integrals, integrals.F90:331-358): per-workgroup
// partial sums, then a fixed-order final sum (deterministic)
template <class P>
__global__ __launch_bounds__(256) void k_integrals(const double* __restrict__ U,
                                                   RunConst rc, double* part, int b0,
                                                   int nb, const int* __restrict__ blev) {
  constexpr int NV = P::nv;
  __shared__ double sh[8][256];
  U += (size_t)b0 * NV * rc.ncell;
  const long long n = (long long)nb * rc.ncell;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x) {
    long long b = g / rc.ncell, c = g % rc.ncell;
    // cells of level l hold 2^(-nd (l - 1)) of a base-level cell's volume
    const double w = blev ? ldexp(1.0, -rc.nd * blev[b0 + b]) : 1.0;
    double u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = w * U[(b * NV + v) * rc.ncell + c];
    acc[0] += u[P::idn];
    acc[1] += u[P::imx];
    acc[2] += u[P::imy];
    acc[3] += u[P::imz];
    // the isothermal systems carry no energy (integrals.F90:355, 457)
    if constexpr (!P::iso) acc[4] += u[P::ien];
    double ek = (u[P::imx] * u[P::imx] + u[P::imy] * u[P::imy] + u[P::imz] * u[P::imz]) / u[P::idn];
    acc[5] += ek;
    double em = 0.0;
    if constexpr (P::mhd) em = u[P::ibx] * u[P::ibx] + u[P::iby] * u[P::iby] + u[P::ibz] * u[P::ibz];
    acc[6] += em;
    if constexpr (!P::iso) acc[7] += u[P::ien] - 0.5 * ek - 0.5 * em;
  }
#pragma unroll
  for (int m = 0; m < 8; m++) sh[m][threadIdx.x] = acc[m];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int m = 0; m < 8; m++) sh[m][threadIdx.x] += sh[m][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int m = 0; m < 8; m++) part[blockIdx.x * 8 + m] = sh[m][0];
}

// statistics of store_integrals (integrals.F90:331-431) from the primitive
// state: per workgroup, the volume-weighted sums, minima and maxima of
// density, pressure, |v|, |B|, psi, sqrt(rho)|v|/sqrt(gamma p) and
// sqrt(rho)|v|/max(eps, |B|); part[wg][3 m + {0, 1, 2}]
template <class P, typename TQ>
__global__ __launch_bounds__(256) void k_stats(const void* Qv, RunConst rc, double* part,
                                               int b0, int nb, const int* __restrict__ blev) {
  constexpr int NV = P::nv;
  constexpr double big = 1.7976931348623157e308, eps = 2.220446049250313e-16;
  __shared__ double sh[21][256];
  const TQ* __restrict__ Q = static_cast<const TQ*>(Qv) + (size_t)b0 * NV * rc.ncell;
  const long long n = (long long)nb * rc.ncell;
  double s[7], lo[7], hi[7];
#pragma unroll
  for (int m = 0; m < 7; m++) { s[m] = 0.0; lo[m] = big; hi[m] = -big; }
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x) {
    const long long b = g / rc.ncell, c = g % rc.ncell;
    const double w = blev ? ldexp(1.0, -rc.nd * blev[b0 + b]) : 1.0;
    double q[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) q[v] = double(Q[(b * NV + v) * rc.ncell + c]);
    const double vel = sqrt((q[P::ivx] * q[P::ivx] + q[P::ivy] * q[P::ivy]) + q[P::ivz] * q[P::ivz]);
    double mag = 0.0, psi = 0.0;
    if constexpr (P::mhd) {
      mag = sqrt((q[P::ibx] * q[P::ibx] + q[P::iby] * q[P::iby]) + q[P::ibz] * q[P::ibz]);
      psi = q[P::ibp];
    }
    const double sv = sqrt(q[P::idn]) * vel;
    // isothermal: no pressure, the sonic Mach number |v| / csnd
    // (integrals.F90:383-423)
    double pr = 0.0, ms;
    if constexpr (P::iso) {
      ms = vel / rc.csnd;
    } else {
      pr = q[P::ipr];
      ms = sv / sqrt(rc.gamma * q[P::ipr]);
    }
    const double t[7] = {q[P::idn], pr, vel, mag, psi, ms, sv / tmax(eps, mag)};
#pragma unroll
    for (int m = 0; m < 7; m++) {
      s[m] += w * t[m];
      lo[m] = tmin(lo[m], t[m]);
      hi[m] = tmax(hi[m], t[m]);
    }
  }
#pragma unroll
  for (int m = 0; m < 7; m++) {
    sh[3 * m][threadIdx.x] = s[m];
    sh[3 * m + 1][threadIdx.x] = lo[m];
    sh[3 * m + 2][threadIdx.x] = hi[m];
  }
  __syncthreads();
  for (int r = 128; r > 0; r >>= 1) {
    if ((int)threadIdx.x < r)
#pragma unroll
      for (int m = 0; m < 7; m++) {
        sh[3 * m][threadIdx.x] += sh[3 * m][threadIdx.x + r];
        sh[3 * m + 1][threadIdx.x] = tmin(sh[3 * m + 1][threadIdx.x], sh[3 * m + 1][threadIdx.x + r]);
        sh[3 * m + 2][threadIdx.x] = tmax(sh[3 * m + 2][threadIdx.x], sh[3 * m + 2][threadIdx.x + r]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int m = 0; m < 21; m++) part[blockIdx.x * 21 + m] = sh[m][0];
}

// pack interior + ghost layers of every block into the reference host layout
// u(nv,im,jm,km) (boundary_variables: ghost cells are neighbour copies)
template <int NV, typename TX>
__global__ void k_pack_blocks(const void* Xv, double* __restrict__ out,
                              const int* __restrict__ nbr, RunConst rc, int ng) {
  const TX* __restrict__ X = static_cast<const TX*>(Xv);
  const int nc = rc.nc;
  const int im = nc + 2 * ng, km = (rc.nd == 3) ? im : 1;
  const long long per = (long long)im * im * km;
  const long long n = per * rc.nblk;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x) {
    int b = (int)(g / per);
    long long r = g % per;
    int i = (int)(r % im), j = (int)((r / im) % im), k = (int)(r / ((long long)im * im));
    int li = i - ng, lj = j - ng, lk = (rc.nd == 3) ? k - ng : 0;
    size_t pk = cell_index<sizeof(TX)>(nbr, rc, b, li, lj, lk);
#pragma unroll
    for (int v = 0; v < NV; v++) out[(size_t)g * NV + v] = double(X[var_index<NV>(rc, pk, v)]);
  }
}

// ---------------------------------------------------------------------------
// unit-level kernels (parity tests): the pointwise routines of the sweeps on
// vectors of states, in the arithmetic type of the configuration, one thread
// per state (the riemann_iface / prim2cons / cons2prim / fluxspeed seams of
// schemes.F90:66-71 and equations.F90:70-108)
// ---------------------------------------------------------------------------
template <class P, int RS, typename T>
__global__ void k_unit_riemann(int n, double cmaxd, const double* __restrict__ ql,
                               const double* __restrict__ qr, double* __restrict__ f,
                               RunConst rc) {
  constexpr int NV = P::nv;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const EosConst<T> e{T(rc.gamma), T(rc.gammam1), T(rc.gammam1i), T(rc.csnd), T(rc.csnd2)};
  T a[NV], b[NV], g[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    a[v] = T(ql[(size_t)i * NV + v]);
    b[v] = T(qr[(size_t)i * NV + v]);
  }
  riemann<P, RS, T>(e, rc.sc, T(cmaxd), T(cmaxd * cmaxd), a, b, g);
#pragma unroll
  for (int v = 0; v < NV; v++) f[(size_t)i * NV + v] = double(g[v]);
}

// the 1D reconstruction of the sweeps (reconstruct_states of interfaces_dir,
// interpolations.F90:867-959) on pencils: one thread per pencil slides the
// window of the sweep along it; cell i writes fl(i) (its upper face) and
// fr(i-1) (its lower face) for the cells whose window lies in the pencil
template <class P, int REC, typename T>
__global__ void k_unit_recon(int n, int npen, double h, const double* __restrict__ fp,
                             double* __restrict__ fl, double* __restrict__ fr, RunConst rc) {
  using RC = Recon<P, REC, 3, 0, T>;
  constexpr int NV = P::nv;
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= npen) return;
  const double* f = fp + (size_t)p * n;
  RC rec;
  T q[NV];
  for (int m = 0; m < 2 * RC::R; m++) {
#pragma unroll
    for (int v = 0; v < NV; v++) q[v] = T(f[m]);
    rec.load(m, q);
  }
  for (int i = RC::R; i < n - RC::R; i++) {
    if (i > RC::R) rec.shift();
#pragma unroll
    for (int v = 0; v < NV; v++) q[v] = T(f[i + RC::R]);
    rec.load(2 * RC::R, q);
    T a[NV], b[NV];
    rec.cell(rc, a, b, nullptr, nullptr, 0, 0, 0, 0, nullptr, h);
    fl[(size_t)p * n + i] = double(a[0]);
    fr[(size_t)p * n + i - 1] = double(b[0]);
  }
}

// q = cons2prim(u) in FP64 on conserved vectors (cons2prim_iface); q holds
// the previous primitives on entry (kept by a failed relativistic recovery)
template <class P>
__global__ void k_unit_c2p(int n, const double* __restrict__ u, double* __restrict__ q,
                           RunConst rc, DevState* st) {
  constexpr int NV = P::nv;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  double ud[NV], qd[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    ud[v] = u[(size_t)i * NV + v];
    qd[v] = q[(size_t)i * NV + v];
  }
  if constexpr (P::sr) {
    const int cs = cons2prim_sr<P>(rc.sc, ud, qd);
    if (cs) c2p_count(st, cs);
  } else {
    const EosConst<double> ed{rc.gamma, rc.gammam1, rc.gammam1i, rc.csnd, rc.csnd2};
    cons2prim<P, double>(ed, ud, qd);
  }
#pragma unroll
  for (int v = 0; v < NV; v++) q[(size_t)i * NV + v] = qd[v];
}

// u = prim2cons(q), f/cm/cp = fluxspeed(q, u) in T; q2 = cons2prim(u) in
// double (the kernels recover primitives in FP64 in every mode; q2 enters
// holding the previous primitives, kept by a failed relativistic recovery)
template <class P, typename T>
__global__ void k_unit_eqs(int n, double cmaxd, const double* __restrict__ q,
                           double* __restrict__ u, double* __restrict__ q2,
                           double* __restrict__ f, double* __restrict__ cm,
                           double* __restrict__ cp, RunConst rc, DevState* st) {
  constexpr int NV = P::nv;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const EosConst<T> e{T(rc.gamma), T(rc.gammam1), T(rc.gammam1i), T(rc.csnd), T(rc.csnd2)};
  const EosConst<double> ed{rc.gamma, rc.gammam1, rc.gammam1i, rc.csnd, rc.csnd2};
  T qq[NV], uu[NV], ff[NV], m, pp;
#pragma unroll
  for (int v = 0; v < NV; v++) qq[v] = T(q[(size_t)i * NV + v]);
  const T cmax2 = T(cmaxd * cmaxd);
  if constexpr (P::sr) {
    prim2cons_sr<P>(T(rc.sc.rgammaxi), qq, uu);
    fluxspeed_sr<P>(rc.sc, cmax2, qq, uu, ff, m, pp);
  } else {
    prim2cons<P>(e, qq, uu);
    fluxspeed<P>(e, cmax2, qq, uu, ff, m, pp);
  }
  double ud[NV], qd[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    ud[v] = double(uu[v]);
    qd[v] = q2[(size_t)i * NV + v];
  }
  if constexpr (P::sr) {
    const int cs = cons2prim_sr<P>(rc.sc, ud, qd);
    if (cs) c2p_count(st, cs);
  } else {
    cons2prim<P, double>(ed, ud, qd);
  }
#pragma unroll
  for (int v = 0; v < NV; v++) {
    u[(size_t)i * NV + v] = ud[v];
    q2[(size_t)i * NV + v] = qd[v];
    f[(size_t)i * NV + v] = double(ff[v]);
  }
  cm[i] = double(m);
  cp[i] = double(pp);
}

typedef void (*SweepFn)(SweepArgs, RunConst);

}  // namespace amun_dev

#include "adjoint.h"
#include "amr.h"

using namespace amun_dev;

// kernel instantiations selected for one configuration
struct Dispatch {
  SweepFn sweep[3] = {nullptr, nullptr, nullptr};
  SweepFn stage = nullptr;   // all directions of a stage in one launch (k_stage)
  void (*compact[3])(const void*, const int*, RunConst, DevState*) = {nullptr, nullptr, nullptr};
  void (*prim)(const double*, void*, RunConst, DevState*, RunState*, const int*, int, int,
               int) = nullptr;
  void (*integ)(const double*, RunConst, double*, int, int, const int*) = nullptr;
  void (*stats)(const void*, RunConst, double*, int, int, const int*) = nullptr;
  void (*pack_u)(const void*, double*, const int*, RunConst, int) = nullptr;
  void (*pack_q)(const void*, double*, const int*, RunConst, int) = nullptr;
  void (*refine)(const void*, const int*, RunConst, RefineArgs, int*, float*) = nullptr;
  void (*unit_riemann)(int, double, const double*, const double*, double*, RunConst) = nullptr;
  void (*unit_recon)(int, int, double, const double*, double*, double*, RunConst) = nullptr;
  void (*unit_c2p)(int, const double*, double*, RunConst, DevState*) = nullptr;
  void (*unit_eqs)(int, double, const double*, double*, double*, double*, double*, double*,
                   RunConst, DevState*) = nullptr;
  // tangent-linear / adjoint kernels (FP64, TVD reconstruction, HD and MHD)
  bool has_adj = false;
  void (*adj_prim)(const double*, double*, RunConst) = nullptr;
  void (*adj_face[3])(const double*, const double*, double*, double*, double*, const int*,
                      RunConst, double, double) = {nullptr, nullptr, nullptr};
  void (*tlm_face[3])(const double*, const double*, const double*, double, double*, const int*,
                      RunConst, double) = {nullptr, nullptr, nullptr};
  void (*adj_cell)(const double*, const double*, const double*, const double*, const double*,
                   const double*, const double*, const double*, const double*, const double*,
                   const double*, double*, double*, const int*, RunConst) = nullptr;
  void (*tlm_cell)(const double*, const double*, const double*, double*, const int*, RunConst,
                   double, double, double) = nullptr;
  // multi-level meshes: averaged fine fluxes of coarse faces, per direction
  void (*fineflux[3])(const int4*, int, const void*, const int*, RunConst, const RunState*, int,
                      double*, DevState*, const int*, const void*,
                      const void*) = {nullptr, nullptr, nullptr};
};


namespace amun_dev {

template <class P, int REC, int RS, int ND, typename T>
void fill(Dispatch& d) {
  if constexpr (is_compact(REC)) {
    d.compact[0] = k_compact<P, REC, ND, 0, T>;
    d.compact[1] = k_compact<P, REC, ND, 1, T>;
    if constexpr (ND == 3) d.compact[2] = k_compact<P, REC, ND, 2, T>;
  }
  d.sweep[0] = k_sweep<P, REC, RS, ND, 0, T>;
  d.sweep[1] = k_sweep<P, REC, RS, ND, 1, T>;
  if constexpr (ND == 3) d.sweep[2] = k_sweep<P, REC, RS, ND, 2, T>;
  // the fused stage for the reconstructions of the production workloads
  if constexpr ((REC == REC_TVD || REC == REC_WENO5Z) && !P::sr)
    d.stage = k_stage<P, REC, RS, ND, T>;
  d.prim = k_prim<P, T>;
  d.integ = k_integrals<P>;
  d.stats = k_stats<P, T>;
  d.pack_u = k_pack_blocks<P::nv, double>;
  d.pack_q = k_pack_blocks<P::nv, T>;
  d.refine = k_refine<P, T>;
  d.unit_riemann = k_unit_riemann<P, RS, T>;
  d.unit_c2p = k_unit_c2p<P>;
  // TVD reads the transverse slopes of the block: no pencil form
  if constexpr (REC != REC_TVD && REC != REC_MGP && !is_compact(REC)) d.unit_recon = k_unit_recon<P, REC, T>;
  d.unit_eqs = k_unit_eqs<P, T>;
  d.fineflux[0] = k_amr_fineflux<P, REC, RS, ND, 0, T>;
  d.fineflux[1] = k_amr_fineflux<P, REC, RS, ND, 1, T>;
  if constexpr (ND == 3) d.fineflux[2] = k_amr_fineflux<P, REC, RS, ND, 2, T>;
  if constexpr (REC == REC_TVD && !P::sr && !P::iso) {
    // the face Jacobians in the arithmetic of the forward fluxes (T)
    d.has_adj = true;
    d.adj_prim = k_adj_prim<P>;
    d.adj_face[0] = k_adj_face<P, RS, ND, 0, T>;
    d.adj_face[1] = k_adj_face<P, RS, ND, 1, T>;
    d.tlm_face[0] = k_tlm_face<P, RS, ND, 0, T>;
    d.tlm_face[1] = k_tlm_face<P, RS, ND, 1, T>;
    if constexpr (ND == 3) {
      d.adj_face[2] = k_adj_face<P, RS, ND, 2, T>;
      d.tlm_face[2] = k_tlm_face<P, RS, ND, 2, T>;
    }
    d.adj_cell = k_adj_cell<P, ND>;
    d.tlm_cell = k_tlm_cell<P, ND>;
  }
}

template <class P, int RS, int REC, typename T>
bool pick_nd(Dispatch& d, int nd) {
  if (nd == 2) { fill<P, REC, RS, 2, T>(d); return true; }
  if (nd == 3) { fill<P, REC, RS, 3, T>(d); return true; }
  return false;
}

// kernel table of one reconstruction, equation system and precision; each
// combination is its own instantiation unit (inst_unit.hip compiled once per
// entry of inst_list.h), so the units build in parallel
template <int REC, int EQ, typename T>
bool pick_eq(Dispatch& d, const amun_cfg& c) {
  if constexpr (EQ == AMUN_EQSYS_HD) {
    if (c.riemann_solver == AMUN_RS_HLL) return pick_nd<HD, RS_HLL, REC, T>(d, c.ndims);
    if (c.riemann_solver == AMUN_RS_HLLC) return pick_nd<HD, RS_HLLC, REC, T>(d, c.ndims);
    if (c.riemann_solver == AMUN_RS_ROE) return pick_nd<HD, RS_ROE, REC, T>(d, c.ndims);
  } else if constexpr (EQ == AMUN_EQSYS_MHD) {
    if (c.riemann_solver == AMUN_RS_HLL) return pick_nd<MHD, RS_HLL, REC, T>(d, c.ndims);
    if (c.riemann_solver == AMUN_RS_HLLD) return pick_nd<MHD, RS_HLLD, REC, T>(d, c.ndims);
    if (c.riemann_solver == AMUN_RS_HLLC) return pick_nd<MHD, RS_HLLC, REC, T>(d, c.ndims);
    if (c.riemann_solver == AMUN_RS_ROE) return pick_nd<MHD, RS_ROE, REC, T>(d, c.ndims);
  } else if constexpr (EQ == AMUN_EQSYS_SRMHD) {
    // "hllc" selects riemann_srmhd_adi_hllc, every other name HLL
    // (schemes.F90:479-498)
    if (c.riemann_solver == AMUN_RS_HLLC) return pick_nd<SRMHD, RS_HLLC, REC, T>(d, c.ndims);
    return pick_nd<SRMHD, RS_HLL, REC, T>(d, c.ndims);
  } else if constexpr (EQ == AMUN_EQSYS_SRHD) {
    // "hllc" selects riemann_srhd_adi_hllc, every other name HLL
    // (schemes.F90:400-430)
    if (c.riemann_solver == AMUN_RS_HLLC) return pick_nd<SRHD, RS_HLLC, REC, T>(d, c.ndims);
    return pick_nd<SRHD, RS_HLL, REC, T>(d, c.ndims);
  } else if constexpr (EQ == EQ_HD_ISO) {
    // "roe" selects riemann_hd_iso_roe, every other name HLL
    // (schemes.F90:163-176)
    if (c.riemann_solver == AMUN_RS_ROE) return pick_nd<HDI, RS_ROE, REC, T>(d, c.ndims);
    return pick_nd<HDI, RS_HLL, REC, T>(d, c.ndims);
  } else if constexpr (EQ == EQ_MHD_ISO) {
    // "hlld", "hlldm", "roe", every other name HLL (schemes.F90:201-225)
    if (c.riemann_solver == AMUN_RS_HLLD) return pick_nd<MHDI, RS_HLLD, REC, T>(d, c.ndims);
    if (c.riemann_solver == AMUN_RS_HLLDM) return pick_nd<MHDI, RS_HLLDM, REC, T>(d, c.ndims);
    if (c.riemann_solver == AMUN_RS_ROE) return pick_nd<MHDI, RS_ROE, REC, T>(d, c.ndims);
    return pick_nd<MHDI, RS_HLL, REC, T>(d, c.ndims);
  }
  return false;
}

#define AMUN_PICK_DECL(name, rec, eq, prec) bool name(Dispatch& d, const amun_cfg& c);
#include "inst_list.h"
AMUN_INST_LIST(AMUN_PICK_DECL)
#undef AMUN_PICK_DECL

}  // namespace amun_dev
