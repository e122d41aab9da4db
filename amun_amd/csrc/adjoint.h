// adjoint.h -- tangent-linear and adjoint (reverse-mode) kernels of one SSPRK
// stage of the TVD finite-volume update, for the gradient of a scalar
// objective of the final state with respect to the initial state and gamma.
//
// The stage is Y = a U0 + b X + c dt L(X) with L(X)_i = -sum_d dxi_d (F_d(i+1/2)
// - F_d(i-1/2)), F = Riemann(qL, qR; gamma), q = cons2prim(X), qL/qR the TVD
// face states (interfaces_tvd, interpolations.F90:729-848).  Linearised with
// dt and the GLM cmax of every step held at their forward values (both are
// maxima over the grid, evolution.F90:379-493, and are not differentiated),
// and with the piecewise-constant parts frozen: the limiter branch and the
// positivity halving count of every slope.
//
// Local Jacobians: cons2prim and the limiters are written out by hand; the
// face flux Jacobian comes from evaluating the solver templates of physics.h
// on dual numbers (dual.h), NCH input columns per pass.
//
// Arrays are FP64 in the block-pool layout of the forward solver
// ([block][variable][cell], lidx<8> bricks).  Included by amun_dev.hip.
#pragma once

#include "dual.h"

namespace amun_dev {

constexpr int kAdjNch = 4;          // Jacobian columns per dual pass

// cell of a thread: block b and (i, j, k) inside it
__device__ __forceinline__ bool adj_cell(const RunConst& rc, int& b, int& i, int& j, int& k) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)rc.nblk * rc.ncell) return false;
  b = (int)(gid / rc.ncell);
  const int c = (int)(gid % rc.ncell), nc = rc.nc;
  i = c % nc;
  j = (c / nc) % nc;
  k = (rc.nd == 3) ? c / (nc * nc) : 0;
  return true;
}

template <int NV>
__device__ __forceinline__ void load_vec(const double* __restrict__ A, const RunConst& rc,
                                         size_t pk, double* x) {
#pragma unroll
  for (int v = 0; v < NV; v++) x[v] = A[var_index<NV>(rc, pk, v)];
}

// primitive state of cell (b; i+di, j+dj, k+dk) of the FP64 array Q
template <int NV>
__device__ __forceinline__ void load_q(const double* __restrict__ Q, const int* __restrict__ nbr,
                                       const RunConst& rc, int b, int i, int j, int k, double* q) {
  load_vec<NV>(Q, rc, cell_index<8>(nbr, rc, b, i, j, k), q);
}

// limiter value and its partial derivatives in a and b (x = 1/2)
__device__ __forceinline__ double lim_d(int kind, double a, double b, double& da, double& db) {
  Dual<2> A(a), B(b);
  A.d[0] = 1.0;
  B.d[1] = 1.0;
  Dual<2> r = limiter<Dual<2>>(kind, Dual<2>(0.5), A, B);
  da = r.d[0];
  db = r.d[1];
  return r.v;
}

// positivity halving factor of a density or pressure slope at a cell
// (interfaces_tvd, interpolations.F90:805-822): 0.5^k with k the number of
// halvings of the joint slope sum, 0 for a non-positive value; 1 otherwise
template <class P, int ND>
__device__ __forceinline__ double halving(const double* __restrict__ Q, const int* __restrict__ nbr,
                                          const RunConst& rc, int b, int i, int j, int k, int v) {
  if (v != P::idn && v != P::ipr) return 1.0;
  const double qc = Q[var_index<P::nv>(rc, cell_index<8>(nbr, rc, b, i, j, k), v)];
  if (!(qc > 0.0)) return 0.0;
  double dq[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int d = 0; d < ND; d++) {
    const double qm = Q[var_index<P::nv>(rc, cell_index<8>(nbr, rc, b, i - (d == 0), j - (d == 1), k - (d == 2)), v)];
    const double qp = Q[var_index<P::nv>(rc, cell_index<8>(nbr, rc, b, i + (d == 0), j + (d == 1), k + (d == 2)), v)];
    dq[d] = limiter<double>(rc.limiter, 0.5, qc - qm, qp - qc);
  }
  double f = 1.0;
  for (int it = 0; it < kMaxHalvings; it++) {
    double sa = fabs(dq[0]) + fabs(dq[1]);
    if (ND == 3) sa += fabs(dq[2]);
    if (!(qc <= sa)) break;
    dq[0] *= 0.5; dq[1] *= 0.5; dq[2] *= 0.5;
    f *= 0.5;
  }
  return f;
}

// slope of variable v at cell c (offset o along DIR from (i,j,k)) and its
// partials with respect to q(c - e), q(c), q(c + e)
template <class P, int ND, int DIR>
__device__ __forceinline__ double slope_d(const double* __restrict__ Q, const int* __restrict__ nbr,
                                          const RunConst& rc, int b, int i, int j, int k, int o,
                                          int v, double& dm, double& d0, double& dp) {
  const int ci = i + (DIR == 0) * o, cj = j + (DIR == 1) * o, ck = k + (DIR == 2) * o;
  const double qm = Q[var_index<P::nv>(rc, cell_index<8>(nbr, rc, b, ci - (DIR == 0), cj - (DIR == 1), ck - (DIR == 2)), v)];
  const double q0 = Q[var_index<P::nv>(rc, cell_index<8>(nbr, rc, b, ci, cj, ck), v)];
  const double qp = Q[var_index<P::nv>(rc, cell_index<8>(nbr, rc, b, ci + (DIR == 0), cj + (DIR == 1), ck + (DIR == 2)), v)];
  double da, db;
  const double s = lim_d(rc.limiter, q0 - qm, qp - q0, da, db);
  const double h = halving<P, ND>(Q, nbr, rc, b, ci, cj, ck, v);
  dm = -h * da;
  d0 = h * (da - db);
  dp = h * db;
  return h * s;
}

// ---------------------------------------------------------------------------
// cons2prim and its linearisation (equations.F90:1828-1876, 2886-2939)
// ---------------------------------------------------------------------------
template <class P>
__global__ void k_adj_prim(const double* __restrict__ X, double* __restrict__ Q, RunConst rc) {
  int b, i, j, k;
  if (!adj_cell(rc, b, i, j, k)) return;
  constexpr int NV = P::nv;
  const size_t base = (size_t)b * NV * rc.ncell + lidx<8>(rc.nc, rc.nd, i, j, k);
  const EosConst<double> e{rc.gamma, rc.gammam1, rc.gammam1i};
  double u[NV], q[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) u[v] = X[base + (size_t)v * rc.ncell];
  cons2prim<P, double>(e, u, q);
#pragma unroll
  for (int v = 0; v < NV; v++) Q[base + (size_t)v * rc.ncell] = q[v];
}

// dq = (dQ/dU) du + (dQ/dgamma) dg at one cell
template <class P>
__device__ __forceinline__ void c2p_tangent(const RunConst& rc, const double* u, const double* q,
                                            const double* du, double dg, double* dq) {
  const double rho = u[P::idn], gm1 = rc.gammam1;
  dq[P::idn] = du[P::idn];
  double mdm = 0.0, v2 = 0.0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int iv = P::ivx + c;
    dq[iv] = (du[P::imx + c] - q[iv] * du[P::idn]) / rho;
    mdm += q[iv] * du[P::imx + c];
    v2 += q[iv] * q[iv];
  }
  double bdb = 0.0, eint = u[P::ien] - 0.5 * rho * v2;
  if constexpr (P::mhd) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      dq[P::ibx + c] = du[P::ibx + c];
      bdb += u[P::ibx + c] * du[P::ibx + c];
      eint -= 0.5 * u[P::ibx + c] * u[P::ibx + c];
    }
    dq[P::ibp] = du[P::ibp];
  }
  dq[P::ipr] = gm1 * (du[P::ien] - mdm + 0.5 * v2 * du[P::idn] - bdb) + eint * dg;
}

// lu = (dQ/dU)^T lq at one cell; returns the gamma part lq . dQ/dgamma
template <class P>
__device__ __forceinline__ double c2p_adjoint(const RunConst& rc, const double* u, const double* q,
                                              const double* lq, double* lu) {
  const double rho = u[P::idn], gm1 = rc.gammam1, lp = lq[P::ipr];
  double v2 = 0.0, lvv = 0.0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    v2 += q[P::ivx + c] * q[P::ivx + c];
    lvv += lq[P::ivx + c] * q[P::ivx + c];
    lu[P::imx + c] = lq[P::ivx + c] / rho - lp * gm1 * q[P::ivx + c];
  }
  lu[P::idn] = lq[P::idn] - lvv / rho + lp * gm1 * 0.5 * v2;
  double eint = u[P::ien] - 0.5 * rho * v2;
  if constexpr (P::mhd) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      lu[P::ibx + c] = lq[P::ibx + c] - lp * gm1 * u[P::ibx + c];
      eint -= 0.5 * u[P::ibx + c] * u[P::ibx + c];
    }
    lu[P::ibp] = lq[P::ibp];
  }
  lu[P::ien] = lp * gm1;
  return lp * eint;
}

// ---------------------------------------------------------------------------
// face kernels: thread = cell i, face i+1/2 along DIR
// ---------------------------------------------------------------------------
// face states of face i+1/2 (left from cell i, right from cell i+1)
template <class P, int ND, int DIR>
__device__ __forceinline__ void face_states(const double* __restrict__ Q, const int* __restrict__ nbr,
                                            const RunConst& rc, int b, int i, int j, int k,
                                            double* ql, double* qr) {
#pragma unroll
  for (int v = 0; v < P::nv; v++) {
    double dm, d0, dp;
    const double sl = slope_d<P, ND, DIR>(Q, nbr, rc, b, i, j, k, 0, v, dm, d0, dp);
    const double sr = slope_d<P, ND, DIR>(Q, nbr, rc, b, i, j, k, 1, v, dm, d0, dp);
    const double q0 = Q[var_index<P::nv>(rc, cell_index<8>(nbr, rc, b, i, j, k), v)];
    const double q1 = Q[var_index<P::nv>(rc, cell_index<8>(nbr, rc, b, i + (DIR == 0), j + (DIR == 1), k + (DIR == 2)), v)];
    ql[v] = q0 + sl;
    qr[v] = q1 - sr;
  }
}

// tangent of the face flux: DF(i) = dF(i+1/2) for tangent dX, dgamma
// TF: arithmetic of the face flux and its Jacobian (double, or float in the
// FP32-flux mode, whose forward evaluates the fluxes in float); everything
// else -- primitives, slopes, the stage combines -- stays FP64
template <class P, int RS, int ND, int DIR, typename TF>
__global__ void k_tlm_face(const double* __restrict__ X, const double* __restrict__ Q,
                           const double* __restrict__ dX, double dgam, double* __restrict__ DF,
                           const int* __restrict__ nbr, RunConst rc, double cmax) {
  int b, i, j, k;
  if (!adj_cell(rc, b, i, j, k)) return;
  constexpr int NV = P::nv;
  // tangent primitives of cells i-1 .. i+2 along DIR
  double dq[4][NV];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int o = m - 1;
    const size_t pk = cell_index<8>(nbr, rc, b, i + (DIR == 0) * o, j + (DIR == 1) * o, k + (DIR == 2) * o);
    double u[NV], q[NV], du[NV];
    load_vec<NV>(X, rc, pk, u);
    load_vec<NV>(Q, rc, pk, q);
    load_vec<NV>(dX, rc, pk, du);
    c2p_tangent<P>(rc, u, q, du, dgam, dq[m]);
  }
  using D1 = Dual<1, TF>;
  D1 ql[NV], qr[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    double dm, d0, dp;
    const double sl = slope_d<P, ND, DIR>(Q, nbr, rc, b, i, j, k, 0, v, dm, d0, dp);
    const double dsl = dm * dq[0][v] + d0 * dq[1][v] + dp * dq[2][v];
    const double sr = slope_d<P, ND, DIR>(Q, nbr, rc, b, i, j, k, 1, v, dm, d0, dp);
    const double dsr = dm * dq[1][v] + d0 * dq[2][v] + dp * dq[3][v];
    const double q0 = Q[var_index<NV>(rc, cell_index<8>(nbr, rc, b, i, j, k), v)];
    const double q1 = Q[var_index<NV>(rc, cell_index<8>(nbr, rc, b, i + (DIR == 0), j + (DIR == 1), k + (DIR == 2)), v)];
    ql[v].v = TF(q0 + sl);
    ql[v].d[0] = TF(dq[1][v] + dsl);
    qr[v].v = TF(q1 - sr);
    qr[v].d[0] = TF(dq[2][v] - dsr);
  }
  D1 g(rc.gamma);
  g.d[0] = TF(dgam);
  const EosConst<D1> e{g, g - 1.0, rcp(g - 1.0)};
  D1 a[NV], c[NV], f[NV];
#pragma unroll
  for (int p = 0; p < NV; p++) {
    a[p] = ql[rot<P, DIR>(p)];
    c[p] = qr[rot<P, DIR>(p)];
  }
  riemann<P, RS, D1>(e, rc.sc, D1(cmax), D1(cmax * cmax), a, c, f);
  const size_t base = (size_t)b * NV * rc.ncell + lidx<8>(rc.nc, rc.nd, i, j, k);
#pragma unroll
  for (int p = 0; p < NV; p++) DF[base + (size_t)rot<P, DIR>(p) * rc.ncell] = double(f[p].d[0]);
}

// adjoint of the face flux: from MU (= c dt lambda_Y) the face flux adjoint
// lF = dxi (MU(i+1) - MU(i)), then AL(i) = (dF/dqL)^T lF, AR(i) = (dF/dqR)^T
// lF, GF(i) = (dF/dgamma) . lF
template <class P, int RS, int ND, int DIR, typename TF>
__global__ void k_adj_face(const double* __restrict__ Q, const double* __restrict__ MU,
                           double* __restrict__ AL, double* __restrict__ AR,
                           double* __restrict__ GF, const int* __restrict__ nbr, RunConst rc,
                           double dxi, double cmax) {
  int b, i, j, k;
  if (!adj_cell(rc, b, i, j, k)) return;
  constexpr int NV = P::nv;
  double ql[NV], qr[NV], lf[NV];
  face_states<P, ND, DIR>(Q, nbr, rc, b, i, j, k, ql, qr);
  {
    double m0[NV], m1[NV];
    load_vec<NV>(MU, rc, cell_index<8>(nbr, rc, b, i, j, k), m0);
    load_vec<NV>(MU, rc, cell_index<8>(nbr, rc, b, i + (DIR == 0), j + (DIR == 1), k + (DIR == 2)), m1);
#pragma unroll
    for (int v = 0; v < NV; v++) lf[v] = dxi * (m1[v] - m0[v]);
  }
  // inputs in the solver's rotated frame: 0..NV-1 left, NV..2NV-1 right,
  // 2NV gamma
  constexpr int NIN = 2 * NV + 1;
  double gin[NIN];
  for (int c0 = 0; c0 < NIN; c0 += kAdjNch) {
    using DN = Dual<kAdjNch, TF>;
    DN a[NV], c[NV], f[NV];
#pragma unroll
    for (int p = 0; p < NV; p++) {
      a[p] = DN(ql[rot<P, DIR>(p)]);
      c[p] = DN(qr[rot<P, DIR>(p)]);
    }
    DN g(rc.gamma);
#pragma unroll
    for (int s = 0; s < kAdjNch; s++) {
      const int n = c0 + s;
      if (n < NV) a[n].d[s] = TF(1);
      else if (n < 2 * NV) c[n - NV].d[s] = TF(1);
      else if (n == 2 * NV) g.d[s] = TF(1);
    }
    const EosConst<DN> e{g, g - 1.0, rcp(g - 1.0)};
    riemann<P, RS, DN>(e, rc.sc, DN(cmax), DN(cmax * cmax), a, c, f);
#pragma unroll
    for (int s = 0; s < kAdjNch; s++) {
      const int n = c0 + s;
      if (n >= NIN) break;
      double acc = 0.0;
#pragma unroll
      for (int p = 0; p < NV; p++) acc += lf[rot<P, DIR>(p)] * double(f[p].d[s]);
      gin[n] = acc;
    }
  }
  const size_t base = (size_t)b * NV * rc.ncell + lidx<8>(rc.nc, rc.nd, i, j, k);
#pragma unroll
  for (int p = 0; p < NV; p++) {
    AL[base + (size_t)rot<P, DIR>(p) * rc.ncell] = gin[p];
    AR[base + (size_t)rot<P, DIR>(p) * rc.ncell] = gin[NV + p];
  }
  GF[(size_t)b * rc.ncell + lidx<8>(rc.nc, rc.nd, i, j, k)] = gin[2 * NV];
}

// ---------------------------------------------------------------------------
// cell kernels
// ---------------------------------------------------------------------------
// tangent of L: DL(i) = -sum_d dxi_d (DF_d(i) - DF_d(i - e_d))
template <class P, int ND>
__global__ void k_tlm_cell(const double* __restrict__ DF0, const double* __restrict__ DF1,
                           const double* __restrict__ DF2, double* __restrict__ DL,
                           const int* __restrict__ nbr, RunConst rc, double dxi0, double dxi1,
                           double dxi2) {
  int b, i, j, k;
  if (!adj_cell(rc, b, i, j, k)) return;
  constexpr int NV = P::nv;
  const size_t pc = cell_index<8>(nbr, rc, b, i, j, k);
  const double* DF[3] = {DF0, DF1, DF2};
  const double dxi[3] = {dxi0, dxi1, dxi2};
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) acc[v] = 0.0;
#pragma unroll
  for (int d = 0; d < ND; d++) {
    const size_t pm = cell_index<8>(nbr, rc, b, i - (d == 0), j - (d == 1), k - (d == 2));
#pragma unroll
    for (int v = 0; v < NV; v++)
      acc[v] -= dxi[d] * (DF[d][var_index<NV>(rc, pc, v)] - DF[d][var_index<NV>(rc, pm, v)]);
  }
#pragma unroll
  for (int v = 0; v < NV; v++) DL[var_index<NV>(rc, pc, v)] = acc[v];
}

// adjoint of L at a cell: the face-state adjoints of the faces around it,
// through the TVD slopes of the cell and of its neighbours, then through
// cons2prim; LT(i) = (dL/dX)^T MU at i, GC(i) = the gamma part of the cell
// (its cons2prim term plus the flux term of its upper faces)
template <class P, int ND>
__global__ void k_adj_cell(const double* __restrict__ X, const double* __restrict__ Q,
                           const double* __restrict__ AL0, const double* __restrict__ AR0,
                           const double* __restrict__ AL1, const double* __restrict__ AR1,
                           const double* __restrict__ AL2, const double* __restrict__ AR2,
                           const double* __restrict__ GF0, const double* __restrict__ GF1,
                           const double* __restrict__ GF2, double* __restrict__ LT,
                           double* __restrict__ GC, const int* __restrict__ nbr, RunConst rc) {
  int b, i, j, k;
  if (!adj_cell(rc, b, i, j, k)) return;
  constexpr int NV = P::nv;
  const double* AL[3] = {AL0, AL1, AL2};
  const double* AR[3] = {AR0, AR1, AR2};
  const size_t pc = cell_index<8>(nbr, rc, b, i, j, k);
  double lq[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) lq[v] = 0.0;
  double gsum = 0.0;
  auto per_dir = [&](auto dirc) {
    constexpr int D = decltype(dirc)::value;
    const int ei = (D == 0), ej = (D == 1), ek = (D == 2);
    // lambda of the slope of cells i-1, i, i+1: AL(c) - AR(c - e)
    size_t pcell[3], pface[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const int o = m - 1;
      pcell[m] = cell_index<8>(nbr, rc, b, i + ei * o, j + ej * o, k + ek * o);
      pface[m] = cell_index<8>(nbr, rc, b, i + ei * (o - 1), j + ej * (o - 1), k + ek * (o - 1));
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
      // direct terms: qL(i+1/2) = q_i + s_i, qR(i-1/2) = q_i - s_i
      lq[v] += AL[D][var_index<NV>(rc, pcell[1], v)] + AR[D][var_index<NV>(rc, pface[1], v)];
      double dm, d0, dp;
      // slope of cell i-1: d/dq_i is its "plus" partial
      slope_d<P, ND, D>(Q, nbr, rc, b, i, j, k, -1, v, dm, d0, dp);
      lq[v] += (AL[D][var_index<NV>(rc, pcell[0], v)] - AR[D][var_index<NV>(rc, pface[0], v)]) * dp;
      slope_d<P, ND, D>(Q, nbr, rc, b, i, j, k, 0, v, dm, d0, dp);
      lq[v] += (AL[D][var_index<NV>(rc, pcell[1], v)] - AR[D][var_index<NV>(rc, pface[1], v)]) * d0;
      slope_d<P, ND, D>(Q, nbr, rc, b, i, j, k, 1, v, dm, d0, dp);
      lq[v] += (AL[D][var_index<NV>(rc, pcell[2], v)] - AR[D][var_index<NV>(rc, pface[2], v)]) * dm;
    }
  };
  per_dir(std::integral_constant<int, 0>());
  per_dir(std::integral_constant<int, 1>());
  if constexpr (ND == 3) per_dir(std::integral_constant<int, 2>());
  const size_t pg = (size_t)b * rc.ncell + lidx<8>(rc.nc, rc.nd, i, j, k);
  gsum += GF0[pg] + GF1[pg];
  if constexpr (ND == 3) gsum += GF2[pg];
  double u[NV], q[NV], lu[NV];
  load_vec<NV>(X, rc, pc, u);
  load_vec<NV>(Q, rc, pc, q);
  gsum += c2p_adjoint<P>(rc, u, q, lq, lu);
#pragma unroll
  for (int v = 0; v < NV; v++) LT[var_index<NV>(rc, pc, v)] = lu[v];
  GC[pg] = gsum;
}

// y = alpha x + beta y (beta = 0: y = alpha x), over n doubles
static __global__ void k_axpby(long long n, double alpha, const double* __restrict__ x, double beta,
                        double* __restrict__ y) {
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x)
    y[g] = (beta == 0.0) ? alpha * x[g] : alpha * x[g] + beta * y[g];
}

// scale variable v of every cell by s (psi decay and its adjoint)
static __global__ void k_scale_var(double* __restrict__ Y, int nv, int v, double s, RunConst rc) {
  const long long n = (long long)rc.nblk * rc.ncell;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x) {
    const long long b = g / rc.ncell, c = g % rc.ncell;
    Y[(b * nv + v) * rc.ncell + c] *= s;
  }
}

// fixed-order partial sums of a per-cell array added to part[workgroup]
// (deterministic: the same cells always go to the same partial, in order)
static __global__ __launch_bounds__(256) void k_sum_acc(const double* __restrict__ x, long long n,
                                                 double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (long long)gridDim.x * blockDim.x)
    acc += x[g];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] += sh[0];
}

// dt and cmax of the coming step into the tape
static __global__ void k_tape(const RunState* rs, RunConst rc, int slot, double dtnext, double* out) {
  const RunState* R = rs;
  out[0] = dt_of(rc, R, slot, dtnext);
  out[1] = cmax_of(R, slot);
}

static __global__ void k_set_fixed(RunState* rs, double dt, double cmax) {
  rs->fix_dt = dt;
  rs->fix_cmax = cmax;
}

}  // namespace amun_dev
