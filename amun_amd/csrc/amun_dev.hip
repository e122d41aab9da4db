// amun_dev.hip -- MI355X (gfx950) implementation of the AMUN per-block update
// behind the C-ABI of include/amun_dev.h.
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
#include <cstring>
#include <string>
#include <vector>

#include "../../include/amun_dev.h"
#include "physics.h"
#include "srmhd.h"

using namespace amun;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
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
constexpr int kTrace = 1 << 16;                      // dt trace ring

// ---------------------------------------------------------------------------
// device-side step state
// ---------------------------------------------------------------------------
struct DevState {
  unsigned long long cmax_bits[2];  // cmax of step s lives in slot s&1
  double dt;                        // last dt (kept when dtnext <= 0)
  double time;
  long long step;
  unsigned long long npos;          // non-positive density/pressure events
  unsigned long long c2p[3];        // SR primitive solver: failed, p <= 0, no convergence
  double trace[kTrace];             // dt of every step (ring)
};

// constants of one run, passed by value to every kernel
struct RunConst {
  int nd, nc, nblk, ncell;
  int limiter;
  double gamma, gammam1, gammam1i;
  double cfl, dxmin, eps, decay;
  double dxi[3];
  SrConst sc;                       // relativistic primitive solver
};

__device__ __forceinline__ double cmax_of(const DevState* st, int slot) {
  return fmax(__longlong_as_double((long long)st->cmax_bits[slot]), kTiny);
}

// dt of new_time_step (evolution.F90:478-483), identical in every thread
__device__ __forceinline__ double dt_of(const RunConst& rc, const DevState* st,
                                       int slot, double dtnext) {
  double cmax = cmax_of(st, slot);
  double dtn = rc.cfl * rc.dxmin / fmax(cmax + 2.0 * 0.0 / rc.dxmin, kTiny);
  return dtnext > 0.0 ? fmin(dtn, dtnext) : st->dt;
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

enum { REC_TVD = AMUN_REC_TVD, REC_WENO5Z = AMUN_REC_WENO5Z };
enum { RS_HLL = AMUN_RS_HLL, RS_HLLC = AMUN_RS_HLLC, RS_HLLD = AMUN_RS_HLLD };

template <class P, int RS, typename T>
__device__ __forceinline__ void riemann(const EosConst<T>& e, const SrConst& sc, T cmax,
                                        T cmax2, T* ql, T* qr, T* f) {
  if constexpr (P::sr) riemann_sr_hll<P, T>(sc, cmax, cmax2, ql, qr, f);
  else if constexpr (RS == RS_HLLD) riemann_hlld<P, T>(e, cmax, cmax2, ql, qr, f);
  else if constexpr (RS == RS_HLLC) riemann_hllc<P, T>(e, cmax, cmax2, ql, qr, f);
  else riemann_hll<P, T>(e, cmax, cmax2, ql, qr, f);
}

struct SweepArgs {
  const void* Q;                   // primitive state of the stage input (T)
  const double* U0;                // state at the start of the step
  const double* U1;                // intermediate state (stage input)
  double* Uout;                    // stage output (may alias U0 or U1)
  void* Qout;                      // primitive state of Uout (T)
  void* D;                         // increment accumulator (T)
  float* C;                        // Kahan compensation of U0 (mixed mode)
  const int* __restrict__ nbr;
  DevState* st;
  int slot;                        // cmax slot of this step
  int kind;                        // stage combine, see combine()
  int last;                        // last stage of the step
  int first;                       // first sweep of the step
  double fac;                      // ds = fac * dt
  double dtnext;
};

// stage combines of evolve_euler/rk2/rk3 (evolution.F90:570, 700, 740, 1080,
// 1130, 1177)
__device__ __forceinline__ double combine(int kind, double u0, double u1, double ds,
                                          double du) {
  switch (kind) {
  case 0: return u0 + ds * du;
  case 1: return ((3.0 / 4.0) * u0 + (1.0 / 4.0) * u1) + ds * du;
  case 2: return ((1.0 / 3.0) * u0 + (2.0 / 3.0) * u1) + ds * du;
  default: return 0.5 * ((u0 + u1) + ds * du);
  }
}

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
                                       const int*, int, int, int, int, DevState*) {
    const T eps = T(rc.eps);
#pragma unroll
    for (int v = 0; v < P::nv; v++) weno5z_cell<T>(w[v], eps, flc[v], frc[v]);
  }
};

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
                                       int k, DevState* st) {
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
          for (;;) {
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

// Counts relativistic primitive-solver events (the reference prints each to
// stderr and goes on).
__device__ __forceinline__ void c2p_count(DevState* st, int cs) {
  if (cs & C2P_FAIL) atomicAdd(&st->c2p[0], 1ull);
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

// ---------------------------------------------------------------------------
// the sweep kernel
// ---------------------------------------------------------------------------
template <class P, int REC, int RS, int ND, int DIR, typename T>
__global__ __launch_bounds__(256) void k_sweep(SweepArgs a, RunConst rc) {
  constexpr int NV = P::nv;
  constexpr bool kLast = (DIR == ND - 1);
  using RC = Recon<P, REC, ND, DIR, T>;
  // chunked pencil I/O along x and y (see loadq below)
  constexpr bool kChunk = DIR < 2;
  constexpr bool kDin = kChunk && DIR > 0;       // increment read in chunks
  constexpr bool kDout = kChunk && !kLast;       // increment written in chunks
  __shared__ T sq[kChunk ? 4 * NV * 256 : 1];
  __shared__ T sd[(kDin || kDout) ? 4 * NV * 256 : 1];
  const int nc = rc.nc;
  const int npen = (ND == 3) ? nc * nc : nc;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = (int)(gid / npen);
  const int t = (int)(gid % npen);
  const bool active = b < rc.nblk;

  if (a.first && gid == 0) {
    // reset the maximum-speed accumulator of the next step (new_time_step:
    // cmax = tiny before the reduction)
    a.st->cmax_bits[a.slot ^ 1] = (unsigned long long)__double_as_longlong(kTiny);
  }

  const double cmaxd = cmax_of(a.st, a.slot);
  const T cmax = T(cmaxd), cmax2 = T(cmaxd * cmaxd);
  const double dt = dt_of(rc, a.st, a.slot, a.dtnext);
  const double ds = a.fac * dt;
  const EosConst<T> e{T(rc.gamma), T(rc.gammam1), T(rc.gammam1i)};
  const EosConst<double> ed{rc.gamma, rc.gammam1, rc.gammam1i};
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
            for (int v = 0; v < NV; v++) sq[(m * NV + v) * 256 + tid] = Qp[var_index<NV>(rc, pk, v)];
          }
        }
#pragma unroll
        for (int v = 0; v < NV; v++) q[v] = sq[((c & 3) * NV + v) * 256 + tid];
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

    RC rec;
    T qtmp[NV];
    // prime the window for cell c = -1: cells -1-R .. -1+R-1
#pragma unroll
    for (int m = 0; m < 2 * RC::R; m++) {
      loadq(-1 - RC::R + m, qtmp);
      rec.load(m, qtmp);
    }
    T flp[NV], fprev[NV];
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
                sd[(m * NV + v) * 256 + tid] = Dp[base[m] + (size_t)v * rc.ncell];
          }
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const size_t off = (size_t)v * rc.ncell;
          if constexpr (DIR > 0 && !kDin) dpre[v] = Dp[bt + off];
          if constexpr (kLast) {
            u0pre[v] = a.U0[bd + off];
            if (a.kind != 0) u1pre[v] = a.U1[bd + off];
            // the compensation enters the step update; the transient
            // stages ignore it (it is below the FP64 rounding of U0)
            if constexpr (sizeof(T) == 4) cpre[v] = a.last ? a.C[bf + off] : 0.0f;
          }
        }
      }
      if (c > -1) rec.shift();
      loadq(c + RC::R, qtmp);
      rec.load(2 * RC::R, qtmp);
      T flc[NV], frc[NV];
      int ci, cj, ck;
      coords(c, ci, cj, ck);
      rec.cell(rc, flc, frc, Qp, a.nbr, b, ci, cj, ck, a.st);
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
        if (c >= 1) {
          // cell c-1 has both faces
          int i, j, k;
          coords(c - 1, i, j, k);
          const size_t bb = (size_t)b * NV * (size_t)rc.ncell;
          const size_t bt = bb + (size_t)lidx<sizeof(T)>(nc, ND, i, j, k);
          const size_t bd = bb + (size_t)lidx<8>(nc, ND, i, j, k);
          const size_t bf = bb + (size_t)lidx<4>(nc, ND, i, j, k);
          const T dxi = T(rc.dxi[DIR]);
          double unew[NV];
#pragma unroll
          for (int v = 0; v < NV; v++) {
            const size_t off = (size_t)v * rc.ncell;
            // update_increment order: -dxi*dfx - dyi*dfy - dzi*dfz
            const T dfx = f[v] - fprev[v];
            T du;
            if constexpr (DIR == 0) du = -dxi * dfx;
            else if constexpr (kDin) du = sd[(((c - 1) & 3) * NV + v) * 256 + tid] - dxi * dfx;
            else du = dpre[v] - dxi * dfx;
            if constexpr (!kLast) {
              if constexpr (kDout) sd[(((c - 1) & 3) * NV + v) * 256 + tid] = du;
              else Dp[bt + off] = du;
            } else if constexpr (sizeof(T) == 8) {
              double un = combine(a.kind, u0pre[v], u1pre[v], ds, du);
              if (a.last && v == P::ibp) un = rc.decay * un;
              unew[v] = un;
            } else {
              // mixed precision: the FP32 increment is added to the FP64
              // state in increment form with Kahan compensation.  The true
              // step state is U0 - C; the last stage forms U0_true + delta
              // and keeps the rounding error of that sum in C (the
              // intermediate stages use U0 alone).
              const double u0 = u0pre[v];
              const double cp = double(cpre[v]);
              const double dd = ds * double(du);
              double delta;
              switch (a.kind) {
              case 0: delta = dd; break;
              case 1: delta = 0.25 * ((u1pre[v] - u0) + cp) + dd; break;
              case 2: delta = (2.0 / 3.0) * ((u1pre[v] - u0) + cp) + dd; break;
              default: delta = 0.5 * (((u1pre[v] - u0) + cp) + dd); break;
              }
              const double y = delta - cp;
              double un = u0 + y;
              if (a.last) {
                double cn = (un - u0) - y;
                if (v == P::ibp) { un = rc.decay * un; cn = rc.decay * cn; }
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
                c2p_keep_old<P, T>(cs, qn, Qp, bt, rc.ncell);
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
                Dp[base[m] + (size_t)v * rc.ncell] = sd[(m * NV + v) * 256 + tid];
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
    if (a.last) {
      // wave-level max, one atomic per wave (positive doubles order as
      // their bit patterns)
      for (int off = 32; off > 0; off >>= 1) smax = fmax(smax, __shfl_xor(smax, off));
      if ((threadIdx.x & 63) == 0 && smax > 0.0)
        atomicMax(&a.st->cmax_bits[a.slot ^ 1],
                  (unsigned long long)__double_as_longlong(smax));
    }
  }
}

// Q = cons2prim(U) on every interior cell, and the maximum speed into a slot
// (update_primitive_variables + maxspeed); used after uploads
template <class P, typename T>
__global__ __launch_bounds__(256) void k_prim(const double* __restrict__ U,
                                              void* Qv, RunConst rc,
                                              DevState* st, int slot) {
  T* __restrict__ Q = static_cast<T*>(Qv);
  constexpr int NV = P::nv;
  const long long n = (long long)rc.nblk * rc.ncell;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const EosConst<double> e{rc.gamma, rc.gammam1, rc.gammam1i};
  double smax = 0.0;
  if (gid < n) {
    const long long b = gid / rc.ncell, c = gid % rc.ncell;
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
  for (int off = 32; off > 0; off >>= 1) smax = fmax(smax, __shfl_xor(smax, off));
  if ((threadIdx.x & 63) == 0 && smax > 0.0)
    atomicMax(&st->cmax_bits[slot], (unsigned long long)__double_as_longlong(smax));
}

__global__ void k_zero_f32(float* x, long long n) {
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x)
    x[g] = 0.0f;
}

__global__ void k_reset_slot(DevState* st, int slot) {
  st->cmax_bits[slot] = (unsigned long long)__double_as_longlong(kTiny);
}

__global__ void k_end_step(DevState* st, RunConst rc, int slot, double dtnext) {
  double dt = dt_of(rc, st, slot, dtnext);
  st->trace[st->step & (kTrace - 1)] = dt;
  st->dt = dt;
  st->time += dt;
  st->step += 1;
}

// dt of the current slot without advancing (new_time_step alone)
__global__ void k_peek_dt(DevState* st, RunConst rc, int slot, double dtnext,
                          double* out) {
  double dt = dt_of(rc, st, slot, dtnext);
  st->dt = dt;
  out[0] = dt;
  out[1] = cmax_of(st, slot);
}

// volume integrals (store_integrals, integrals.F90:331-358): per-workgroup
// partial sums, then a fixed-order final sum (deterministic)
template <class P>
__global__ __launch_bounds__(256) void k_integrals(const double* __restrict__ U,
                                                   RunConst rc, double* part) {
  constexpr int NV = P::nv;
  __shared__ double sh[8][256];
  const long long n = (long long)rc.nblk * rc.ncell;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x) {
    long long b = g / rc.ncell, c = g % rc.ncell;
    double u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = U[(b * NV + v) * rc.ncell + c];
    acc[0] += u[P::idn];
    acc[1] += u[P::imx];
    acc[2] += u[P::imy];
    acc[3] += u[P::imz];
    acc[4] += u[P::ien];
    double ek = (u[P::imx] * u[P::imx] + u[P::imy] * u[P::imy] + u[P::imz] * u[P::imz]) / u[P::idn];
    acc[5] += ek;
    double em = 0.0;
    if constexpr (P::mhd) em = u[P::ibx] * u[P::ibx] + u[P::iby] * u[P::iby] + u[P::ibz] * u[P::ibz];
    acc[6] += em;
    acc[7] += u[P::ien] - 0.5 * ek - 0.5 * em;
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

typedef void (*SweepFn)(SweepArgs, RunConst);

}  // namespace

// kernel instantiations selected for one configuration
struct Dispatch {
  SweepFn sweep[3] = {nullptr, nullptr, nullptr};
  void (*prim)(const double*, void*, RunConst, DevState*, int) = nullptr;
  void (*integ)(const double*, RunConst, double*) = nullptr;
  void (*pack_u)(const void*, double*, const int*, RunConst, int) = nullptr;
  void (*pack_q)(const void*, double*, const int*, RunConst, int) = nullptr;
};

// ===========================================================================
// context
// ===========================================================================
struct amun_ctx {
  amun_cfg cfg;
  RunConst rc;
  int nv;
  size_t nstate;            // doubles per state array
  double *U0 = nullptr, *U1 = nullptr;
  void *Q[2] = {nullptr, nullptr}, *D = nullptr;   // element size qsize
  float* C = nullptr;       // Kahan compensation of U0 (mixed mode)
  size_t qsize = 8;
  int* nbr = nullptr;
  DevState* st = nullptr;
  double* scratch = nullptr;   // reductions / small transfers
  int qcur = 0;                // Q[qcur] holds the primitive state of U0
  long long nsteps = 0;        // steps launched (slot = nsteps & 1)
  hipStream_t stream = nullptr;
  std::vector<int32_t> coords;
  Dispatch disp;
};

namespace {


template <class P, int REC, int RS, int ND, typename T>
void fill(Dispatch& d) {
  d.sweep[0] = k_sweep<P, REC, RS, ND, 0, T>;
  d.sweep[1] = k_sweep<P, REC, RS, ND, 1, T>;
  if constexpr (ND == 3) d.sweep[2] = k_sweep<P, REC, RS, ND, 2, T>;
  d.prim = k_prim<P, T>;
  d.integ = k_integrals<P>;
  d.pack_u = k_pack_blocks<P::nv, double>;
  d.pack_q = k_pack_blocks<P::nv, T>;
}

template <class P, int RS, typename T>
bool pick_nd_rec(Dispatch& d, int nd, int rec) {
  if (nd == 2 && rec == REC_TVD) { fill<P, REC_TVD, RS, 2, T>(d); return true; }
  if (nd == 3 && rec == REC_TVD) { fill<P, REC_TVD, RS, 3, T>(d); return true; }
  if (nd == 2 && rec == REC_WENO5Z) { fill<P, REC_WENO5Z, RS, 2, T>(d); return true; }
  if (nd == 3 && rec == REC_WENO5Z) { fill<P, REC_WENO5Z, RS, 3, T>(d); return true; }
  return false;
}

template <typename T>
bool pick_t(Dispatch& d, const amun_cfg& c) {
  if (c.eqsys == AMUN_EQSYS_HD) {
    if (c.riemann_solver == AMUN_RS_HLL) return pick_nd_rec<HD, RS_HLL, T>(d, c.ndims, c.reconstruction);
    if (c.riemann_solver == AMUN_RS_HLLC) return pick_nd_rec<HD, RS_HLLC, T>(d, c.ndims, c.reconstruction);
  } else if (c.eqsys == AMUN_EQSYS_MHD) {
    if (c.riemann_solver == AMUN_RS_HLL) return pick_nd_rec<MHD, RS_HLL, T>(d, c.ndims, c.reconstruction);
    if (c.riemann_solver == AMUN_RS_HLLD) return pick_nd_rec<MHD, RS_HLLD, T>(d, c.ndims, c.reconstruction);
  } else if (c.eqsys == AMUN_EQSYS_SRMHD) {
    // every solver name but "hllc" selects HLL (schemes.F90:479-498)
    return pick_nd_rec<SRMHD, RS_HLL, T>(d, c.ndims, c.reconstruction);
  }
  return false;
}

bool pick(Dispatch& d, const amun_cfg& c) {
  return c.precision == AMUN_PREC_MIXED ? pick_t<float>(d, c) : pick_t<double>(d, c);
}

int validate(const amun_cfg& c, int* nv) {
  if (c.ndims != 2 && c.ndims != 3) return fail(500, "ndims must be 2 or 3");
  if (c.eqsys == AMUN_EQSYS_HD) *nv = 5;
  else if (c.eqsys == AMUN_EQSYS_MHD || c.eqsys == AMUN_EQSYS_SRMHD) *nv = 9;
  else return fail(100, "unsupported equation_system");
  if (c.eos != AMUN_EOS_ADI) return fail(110, "unsupported equation_of_state");
  if (c.reconstruction != AMUN_REC_TVD && c.reconstruction != AMUN_REC_WENO5Z)
    return fail(200, "unsupported reconstruction");
  if (c.limiter < 0 || c.limiter > 5) return fail(200, "unsupported limiter");
  if (c.eqsys == AMUN_EQSYS_SRMHD) {
    if (c.riemann_solver == AMUN_RS_HLLC)
      return fail(600, "the relativistic HLLC solver is not supported on the device path");
    if (!(c.tolerance > 0.0) || c.nr_maxit < 1 || c.nr_extra < 0 || !(c.pmin >= 0.0))
      return fail(500, "tolerance, nr_maxit, nr_extra or pmin out of range");
  } else {
    bool rs_ok = c.riemann_solver == AMUN_RS_HLL ||
                 (c.riemann_solver == AMUN_RS_HLLC && c.eqsys == AMUN_EQSYS_HD) ||
                 (c.riemann_solver == AMUN_RS_HLLD && c.eqsys == AMUN_EQSYS_MHD);
    if (!rs_ok) return fail(300, "riemann_solver does not match the equation system");
  }
  if (c.time_advance < AMUN_TA_EULER || c.time_advance > AMUN_TA_RK3)
    return fail(400, "unsupported time_advance");
  if (c.precision != AMUN_PREC_FP64 && c.precision != AMUN_PREC_MIXED)
    return fail(500, "unsupported precision");
  if (c.fix_positivity || c.clip_extrema)
    return fail(600, "fix_positivity / clip_extrema are not supported on the device path");
  if (c.ncells < 4 || c.ncells > 64 || c.ncells % 4)
    return fail(500, "ncells must be a multiple of 4 in [4, 64] (4x4 brick layout)");
  int need = (c.reconstruction == AMUN_REC_WENO5Z) ? 4 : 2;
  if (c.nghosts < need)
    return fail(600, "nghosts too small for the reconstruction stencil (WENO5-Z needs 4, TVD 2)");
  if (c.ncells < 3) return fail(500, "ncells too small for the stencil");
  if (c.xblocks < 1 || c.yblocks < 1 || (c.ndims == 3 && c.zblocks < 1))
    return fail(500, "block counts must be positive");
  if (!(c.gamma > 1.0) || !(c.cfl > 0.0)) return fail(500, "gamma must exceed 1 and cfl be positive");
  return 0;
}

Dispatch& dispatch_of(amun_ctx* ctx) { return ctx->disp; }

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
  c->cfl = 0.5;
  c->eps = 2.220446049250313e-16;
  c->alpha = 2.0;
  c->xmin = c->ymin = c->zmin = 0.0;
  c->xmax = c->ymax = c->zmax = 1.0;
  c->tolerance = 1.0e-10;
  c->pmin = 1.0e-16;
  c->nr_maxit = 100;
  c->nr_extra = 2;
}

int amun_dev_init(const amun_cfg* cfg, amun_ctx** out) {
  if (!cfg || !out) return fail(500, "null argument");
  *out = nullptr;
  int nv = 0;
  int r = validate(*cfg, &nv);
  if (r) return r;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail(1000 + (int)hipErrorNoDevice, "no HIP device");

  amun_ctx* c = new amun_ctx();
  c->cfg = *cfg;
  c->cfg.nv = nv;
  c->nv = nv;
  const int nd = cfg->ndims, nc = cfg->ncells;
  const int ir = cfg->xblocks, jr = cfg->yblocks, kr = (nd == 3) ? cfg->zblocks : 1;
  RunConst& rc = c->rc;
  rc.nd = nd;
  rc.nc = nc;
  rc.nblk = ir * jr * kr;
  rc.ncell = (nd == 3) ? nc * nc * nc : nc * nc;
  rc.limiter = cfg->limiter;
  rc.gamma = cfg->gamma;
  rc.gammam1 = cfg->gamma - 1.0;
  rc.gammam1i = 1.0 / rc.gammam1;
  rc.cfl = cfg->cfl;
  rc.eps = cfg->eps;
  rc.decay = std::exp(-cfg->alpha * cfg->cfl);
  rc.sc.gamma = cfg->gamma;
  rc.sc.gammaxi = rc.gammam1 / cfg->gamma;       // equations.F90:843
  rc.sc.rgammaxi = 1.0 / rc.sc.gammaxi;
  rc.sc.tol = cfg->tolerance;
  rc.sc.pmin = cfg->pmin;
  rc.sc.nrmax = cfg->nr_maxit;
  rc.sc.nrext = cfg->nr_extra;
  // cell sizes of coordinates.F90:330-348 (one level)
  double adx = (cfg->xmax - cfg->xmin) / (ir * nc);
  double ady = (cfg->ymax - cfg->ymin) / (jr * nc);
  double adz = (nd == 3) ? (cfg->zmax - cfg->zmin) / (kr * nc) : 1.0;
  rc.dxi[0] = 1.0 / adx;
  rc.dxi[1] = 1.0 / ady;
  rc.dxi[2] = (nd == 3) ? 1.0 / adz : 1.0;
  rc.dxmin = std::fmin(adx, ady);
  if (nd == 3) rc.dxmin = std::fmin(rc.dxmin, adz);

  // periodic base-level neighbour table (setup_domain_default wiring)
  std::vector<int32_t> nbr((size_t)rc.nblk * 27, -1);
  c->coords.resize((size_t)rc.nblk * 3);
  for (int bz = 0; bz < kr; bz++)
    for (int by = 0; by < jr; by++)
      for (int bx = 0; bx < ir; bx++) {
        int b = (bz * jr + by) * ir + bx;
        c->coords[b * 3 + 0] = bx;
        c->coords[b * 3 + 1] = by;
        c->coords[b * 3 + 2] = bz;
        for (int dk = -1; dk <= 1; dk++)
          for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
              if (nd == 2 && dk != 0) continue;
              int nx = (bx + di + ir) % ir, ny = (by + dj + jr) % jr, nz = (bz + dk + kr) % kr;
              nbr[(size_t)b * 27 + (dk + 1) * 9 + (dj + 1) * 3 + (di + 1)] = (nz * jr + ny) * ir + nx;
            }
      }

  c->nstate = (size_t)rc.nblk * nv * rc.ncell;
  const size_t bytes = c->nstate * sizeof(double);
  c->qsize = (cfg->precision == AMUN_PREC_MIXED) ? sizeof(float) : sizeof(double);
  const size_t qbytes = c->nstate * c->qsize;
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
  ALLOC(c->st, sizeof(DevState));
  ALLOC(c->scratch, 4096 * 8 * sizeof(double));
#undef ALLOC
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    amun_dev_finalize(c);
    return fail(1000 + (int)e, "hipStreamCreate");
  }
  e = hipMemcpy(c->nbr, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    DevState h;
    std::memset(&h, 0, sizeof(h));
    h.cmax_bits[0] = h.cmax_bits[1] = 0;
    std::memcpy(&h.cmax_bits[0], &kTiny, 8);
    std::memcpy(&h.cmax_bits[1], &kTiny, 8);
    h.dt = 1.0;
    e = hipMemcpy(c->st, &h, sizeof(DevState), hipMemcpyHostToDevice);
  }
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
  *out = c;
  return 0;
}

void amun_dev_finalize(amun_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  hipFree(c->U0);
  hipFree(c->U1);
  hipFree(c->Q[0]);
  hipFree(c->Q[1]);
  hipFree(c->D);
  hipFree(c->C);
  hipFree(c->nbr);
  hipFree(c->st);
  hipFree(c->scratch);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

static int after_upload(amun_ctx* c) {
  Dispatch& d = dispatch_of(c);
  const int slot = (int)(c->nsteps & 1);
  hipLaunchKernelGGL(k_reset_slot, dim3(1), dim3(1), 0, c->stream, c->st, slot);
  long long n = (long long)c->rc.nblk * c->rc.ncell;
  // in place over the current primitive array (update_primitive_variables
  // overwrites pdata%q)
  hipLaunchKernelGGL(d.prim, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     c->U0, c->Q[c->qcur], c->rc, c->st, slot);
  if (c->C)
    hipLaunchKernelGGL(k_zero_f32, dim3(2048), dim3(256), 0, c->stream, c->C,
                       (long long)c->nstate);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int amun_dev_upload_global(amun_ctx* c, const double* u) {
  if (!c || !u) return fail(500, "null argument");
  const RunConst& rc = c->rc;
  const int nc = rc.nc, nv = c->nv;
  const int ir = c->cfg.xblocks, jr = c->cfg.yblocks;
  const size_t nx = (size_t)ir * nc, ny = (size_t)jr * nc;
  const size_t nz = (rc.nd == 3) ? (size_t)c->cfg.zblocks * nc : 1;
  const int kn = (rc.nd == 3) ? nc : 1;
  std::vector<double> h(c->nstate);
  for (int b = 0; b < rc.nblk; b++) {
    size_t bx = c->coords[b * 3], by = c->coords[b * 3 + 1], bz = c->coords[b * 3 + 2];
    for (int v = 0; v < nv; v++)
      for (int k = 0; k < kn; k++)
        for (int j = 0; j < nc; j++) {
          const double* src = u + ((v * nz + bz * kn + k) * ny + by * nc + j) * nx + bx * nc;
          double* dst = &h[((size_t)b * nv + v) * rc.ncell];
          for (int i = 0; i < nc; i++) dst[lidx<8>(nc, rc.nd, i, j, k)] = src[i];
        }
  }
  HIP_TRY(hipMemcpy(c->U0, h.data(), c->nstate * sizeof(double), hipMemcpyHostToDevice));
  return after_upload(c);
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
  HIP_TRY(hipMemcpy(c->U0, h.data(), c->nstate * sizeof(double), hipMemcpyHostToDevice));
  return after_upload(c);
}

int amun_dev_download_global(amun_ctx* c, double* u, double* q) {
  if (!c || !u) return fail(500, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  const RunConst& rc = c->rc;
  const int nc = rc.nc, nv = c->nv;
  const int ir = c->cfg.xblocks, jr = c->cfg.yblocks;
  const size_t nx = (size_t)ir * nc, ny = (size_t)jr * nc;
  const size_t nz = (rc.nd == 3) ? (size_t)c->cfg.zblocks * nc : 1;
  const int kn = (rc.nd == 3) ? nc : 1;
  std::vector<double> h(c->nstate);
  for (int pass = 0; pass < 2; pass++) {
    double* dst_arr = pass == 0 ? u : q;
    if (!dst_arr) continue;
    if (pass == 0 || c->qsize == sizeof(double)) {
      HIP_TRY(hipMemcpy(h.data(), pass == 0 ? (void*)c->U0 : c->Q[c->qcur],
                        c->nstate * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      std::vector<float> hf(c->nstate);
      HIP_TRY(hipMemcpy(hf.data(), c->Q[c->qcur], c->nstate * sizeof(float),
                        hipMemcpyDeviceToHost));
      for (size_t n = 0; n < c->nstate; n++) h[n] = hf[n];
    }
    for (int b = 0; b < rc.nblk; b++) {
      size_t bx = c->coords[b * 3], by = c->coords[b * 3 + 1], bz = c->coords[b * 3 + 2];
      for (int v = 0; v < nv; v++)
        for (int k = 0; k < kn; k++)
          for (int j = 0; j < nc; j++) {
            double* dst = dst_arr + ((v * nz + bz * kn + k) * ny + by * nc + j) * nx + bx * nc;
            const double* src = &h[((size_t)b * nv + v) * rc.ncell];
            if (pass == 0 || c->qsize == sizeof(double))
              for (int i = 0; i < nc; i++) dst[i] = src[lidx<8>(nc, rc.nd, i, j, k)];
            else
              for (int i = 0; i < nc; i++) dst[i] = src[lidx<4>(nc, rc.nd, i, j, k)];
          }
    }
  }
  return 0;
}

int amun_dev_download_blocks(amun_ctx* c, double* u, double* q) {
  if (!c || !u) return fail(500, "null argument");
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

// launch one whole step asynchronously; with ev != nullptr an event is
// recorded after every kernel (ev must hold 3*nd+2 events)
static int launch_step(amun_ctx* c, double dtnext, hipEvent_t* ev = nullptr) {
  Dispatch& d = dispatch_of(c);
  const RunConst& rc = c->rc;
  const int slot = (int)(c->nsteps & 1);
  const int npen = (rc.nd == 3) ? rc.nc * rc.nc : rc.nc;
  const long long nthreads = (long long)rc.nblk * npen;
  const dim3 grid((unsigned)((nthreads + 255) / 256)), block(256);

  struct Stage { int kind; double fac; int src_is_u0; double* out; };
  std::vector<Stage> stages;
  switch (c->cfg.time_advance) {
  case AMUN_TA_EULER:
    stages = {{0, 1.0, 1, c->U0}};
    break;
  case AMUN_TA_RK2:
    stages = {{0, 1.0, 1, c->U1}, {3, 1.0, 0, c->U0}};
    break;
  default:
    stages = {{0, 1.0, 1, c->U1}, {1, 1.0 / 4.0, 0, c->U1}, {2, 2.0 / 3.0, 0, c->U0}};
  }
  for (size_t s = 0; s < stages.size(); s++) {
    SweepArgs a;
    a.Q = c->Q[c->qcur];
    a.U0 = c->U0;
    a.U1 = c->U1;
    a.Uout = stages[s].out;
    a.Qout = c->Q[c->qcur ^ 1];
    a.D = c->D;
    a.C = c->C;
    a.nbr = c->nbr;
    a.st = c->st;
    a.slot = slot;
    a.kind = stages[s].kind;
    a.last = (s + 1 == stages.size());
    a.fac = stages[s].fac;
    a.dtnext = dtnext;
    for (int dir = 0; dir < rc.nd; dir++) {
      a.first = (s == 0 && dir == 0);
      hipLaunchKernelGGL(d.sweep[dir], grid, block, 0, c->stream, a, rc);
      if (ev) HIP_TRY(hipEventRecord(*ev++, c->stream));
    }
    c->qcur ^= 1;
  }
  hipLaunchKernelGGL(k_end_step, dim3(1), dim3(1), 0, c->stream, c->st, rc, slot, dtnext);
  if (ev) HIP_TRY(hipEventRecord(*ev++, c->stream));
  HIP_TRY(hipGetLastError());
  c->nsteps++;
  return 0;
}

static int kernels_per_step(const amun_ctx* c) {
  int ns = c->cfg.time_advance == AMUN_TA_EULER ? 1 : (c->cfg.time_advance == AMUN_TA_RK2 ? 2 : 3);
  return ns * c->rc.nd + 1;
}

int amun_dev_new_time_step(amun_ctx* c, double dtnext, double* dt, double* cmax) {
  if (!c) return fail(500, "null argument");
  const int slot = (int)(c->nsteps & 1);
  hipLaunchKernelGGL(k_peek_dt, dim3(1), dim3(1), 0, c->stream, c->st, c->rc, slot, dtnext,
                     c->scratch);
  HIP_TRY(hipGetLastError());
  double h[2];
  HIP_TRY(hipMemcpyAsync(h, c->scratch, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (dt) *dt = h[0];
  if (cmax) *cmax = h[1];
  return 0;
}

int amun_dev_step(amun_ctx* c, double dtnext, double* dt, double* cmax) {
  if (!c) return fail(500, "null argument");
  const int slot = (int)(c->nsteps & 1);
  hipLaunchKernelGGL(k_peek_dt, dim3(1), dim3(1), 0, c->stream, c->st, c->rc, slot, dtnext,
                     c->scratch);
  int r = launch_step(c, dtnext);
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
  if (nsteps > kTrace) return fail(500, "at most 65536 steps per call");
  long long s0 = c->nsteps;
  for (int s = 0; s < nsteps; s++) {
    int r = launch_step(c, dtnext);
    if (r) return r;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (dts && nsteps > 0) {
    std::vector<double> tr(kTrace);
    HIP_TRY(hipMemcpy(tr.data(), (const char*)c->st + offsetof(DevState, trace),
                      kTrace * sizeof(double), hipMemcpyDeviceToHost));
    for (int s = 0; s < nsteps; s++) dts[s] = tr[(s0 + s) & (kTrace - 1)];
  }
  return 0;
}

int amun_dev_timed_run(amun_ctx* c, int32_t nsteps, double dtnext, double* elapsed_ms,
                       double* kernel_ms, int32_t nkernel_max, int32_t* nkernel) {
  if (!c || nsteps < 1 || !elapsed_ms) return fail(500, "bad argument");
  const int nk = kernels_per_step(c);
  if (nkernel) *nkernel = nk;
  const bool per_kernel = kernel_ms && nkernel_max >= nk;
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

int amun_dev_sync(amun_ctx* c) {
  if (!c) return fail(500, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int amun_dev_integrals(amun_ctx* c, double out[8]) {
  if (!c || !out) return fail(500, "null argument");
  Dispatch& d = dispatch_of(c);
  const int nwg = 1024;
  hipLaunchKernelGGL(d.integ, dim3(nwg), dim3(256), 0, c->stream, c->U0, c->rc, c->scratch);
  HIP_TRY(hipGetLastError());
  std::vector<double> part((size_t)nwg * 8);
  HIP_TRY(hipMemcpyAsync(part.data(), c->scratch, part.size() * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int w = 0; w < nwg; w++)
    for (int m = 0; m < 8; m++) acc[m] += part[(size_t)w * 8 + m];
  double dvol = 1.0 / (c->rc.dxi[0] * c->rc.dxi[1] * (c->rc.nd == 3 ? c->rc.dxi[2] : 1.0));
  out[0] = acc[0] * dvol;
  out[1] = acc[1] * dvol;
  out[2] = acc[2] * dvol;
  out[3] = acc[3] * dvol;
  out[4] = acc[4] * dvol;
  out[5] = acc[5] * 0.5 * dvol;
  out[6] = acc[6] * 0.5 * dvol;
  out[7] = acc[7] * dvol;
  return 0;
}

int amun_dev_get_time(amun_ctx* c, double* time, int64_t* step) {
  if (!c) return fail(500, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  DevState h;
  HIP_TRY(hipMemcpy(&h, c->st, offsetof(DevState, trace), hipMemcpyDeviceToHost));
  if (time) *time = h.time;
  if (step) *step = h.step;
  return 0;
}

int amun_dev_nblocks(amun_ctx* c, int32_t* nblk) {
  if (!c || !nblk) return fail(500, "null argument");
  *nblk = c->rc.nblk;
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
  HIP_TRY(hipMemcpy(&h, c->st, offsetof(DevState, trace), hipMemcpyDeviceToHost));
  *n = h.npos;
  return 0;
}

int amun_dev_c2p_events(amun_ctx* c, uint64_t n[3]) {
  if (!c || !n) return fail(500, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  DevState h;
  HIP_TRY(hipMemcpy(&h, c->st, offsetof(DevState, trace), hipMemcpyDeviceToHost));
  n[0] = h.c2p[0];
  n[1] = h.c2p[1];
  n[2] = h.c2p[2];
  return 0;
}

void* amun_dev_stream(amun_ctx* c) { return c ? (void*)c->stream : nullptr; }

}  // extern "C"
