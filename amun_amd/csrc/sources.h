// sources.h -- the GLM, viscous and resistive source terms of update_sources
// (sources.F90:340-598).
//
// With glm_source_terms = eglm (1) or heglm (2) the reference adds, after
// update_increment and before the stage combine, for every system with a
// magnetic field:
//   EGLM:  du(m) -= (div B) B,   du(E) -= B . grad psi
//   HEGLM: in addition du(B) -= (div B) v,   du(E) -= (div B) (v . B)
// where div B and grad psi come from derivative_1st's central differences
// 0.5 (u(i+1) - u(i-1)) / dh (operators.F90:181-308), summed over the
// directions in order, at the cell size of the block's level.  The terms
// read the stage input's primitive state one cell around each cell in every
// direction, which the pencil walks of the sweeps do not hold, so a kernel
// computes the eight products per cell beforehand and the last sweep of the
// stage subtracts them from the completed increment in the reference's
// order (k_sweep).  Products and differences are rounded as the reference
// rounds them (no contraction).
#pragma once

namespace amun_dev {

// S[b][s][cell] (lidx<sizeof(T)> cells): s = 0..2 db B, 3 B . grad psi,
// 4..6 db v, 7 db (v . B); one thread per interior cell
template <class P, typename T>
static __global__ void __launch_bounds__(256)
k_glm_src(const T* __restrict__ Q, const int* __restrict__ nbr, RunConst rc,
          const RunState* __restrict__ rs, const int* __restrict__ brun,
          const int* __restrict__ blev, T* __restrict__ S) {
#pragma clang fp contract(off)
  constexpr int NV = P::nv;
  const int nc = rc.nc, nd = rc.nd, ncell = rc.ncell;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)rc.nblk * ncell) return;
  const int b = (int)(gid / ncell), s = (int)(gid % ncell);
  // the cell of storage index s (the plane's brick order of element size T)
  int i, j, k;
  {
    const int n = nc, q = s >> 2;
    if (nd == 2) {
      j = q % n;
      i = ((q / n) << 2) | (s & 3);
      k = 0;
    } else if (sizeof(T) == 8) {
      j = q % n;
      k = (q / n) % n;
      i = ((q / (n * n)) << 2) | (s & 3);
    } else {
      // lidx<4>: ((((i>>2) (nc/2) + k>>1) nc + j) 2 + (k&1)) 4 + (i&3)
      const int kl = q & 1, r = q >> 1;
      j = r % n;
      const int r2 = r / n;
      k = ((r2 % (n >> 1)) << 1) | kl;
      i = ((r2 / (n >> 1)) << 2) | (s & 3);
    }
  }
  const RunState* R = rs + brun[b];
  const double sc = blev ? 1.0 / double(1 << blev[b]) : 1.0;
  const size_t here = (size_t)b * NV * ncell + s;
  auto at = [&](size_t packed, int v) {
    return Q[var_index<NV>(rc, packed, v)];
  };
  T db = T(0), gp[3] = {T(0), T(0), T(0)};
  for (int d = 0; d < nd; d++) {
    const T h = T(R->h[d] * sc);
    const int di = d == 0, dj = d == 1, dk = d == 2;
    const size_t pp = cell_index<sizeof(T)>(nbr, rc, b, i + di, j + dj, k + dk);
    const size_t pm = cell_index<sizeof(T)>(nbr, rc, b, i - di, j - dj, k - dk);
    db = db + T(0.5) * (at(pp, P::ibx + d) - at(pm, P::ibx + d)) / h;
    gp[d] = T(0.5) * (at(pp, P::ibp) - at(pm, P::ibp)) / h;
  }
  T q[NV];
  for (int v = 0; v < NV; v++) q[v] = Q[here + (size_t)v * ncell];
  T* o = S + (size_t)b * 8 * ncell + s;
  o[0 * (size_t)ncell] = db * q[P::ibx];
  o[1 * (size_t)ncell] = db * q[P::iby];
  o[2 * (size_t)ncell] = db * q[P::ibz];
  o[3 * (size_t)ncell] = (q[P::ibx] * gp[0] + q[P::iby] * gp[1]) + q[P::ibz] * gp[2];
  o[4 * (size_t)ncell] = db * q[P::ivx];
  o[5 * (size_t)ncell] = db * q[P::ivy];
  o[6 * (size_t)ncell] = db * q[P::ivz];
  o[7 * (size_t)ncell] =
      db * ((q[P::ivx] * q[P::ibx] + q[P::ivy] * q[P::iby]) + q[P::ivz] * q[P::ibz]);
}


// the cell (i, j, k) of storage index s of a block plane (element size T)
template <typename T>
__device__ __forceinline__ void src_cell(const RunConst& rc, int s, int& i, int& j, int& k) {
  const int n = rc.nc, q = s >> 2;
  if (rc.nd == 2) {
    j = q % n;
    i = ((q / n) << 2) | (s & 3);
    k = 0;
  } else if (sizeof(T) == 8) {
    j = q % n;
    k = (q / n) % n;
    i = ((q / (n * n)) << 2) | (s & 3);
  } else {
    const int kl = q & 1, r = q >> 1;
    j = r % n;
    const int r2 = r / n;
    k = ((r2 % (n >> 1)) << 1) | kl;
    i = ((r2 / (n >> 1)) << 2) | (s & 3);
  }
}

// the viscous stress at cell (i, j, k) of block b (sources.F90:340-384):
// velocity gradients g[a][d] = d v_a / d x_d from derivative_1st's central
// differences (zero beyond ndims), gc = nu rho, fc = 2 gc,
//   t(a,a) = fc (g(a,a) - div v / 3),  t(a,b) = gc (g(a,b) + g(b,a))
template <class P, int ND, typename T>
__device__ __forceinline__ void visc_tau(const T* __restrict__ Q, const int* __restrict__ nbr,
                                         const RunConst& rc, int b, int i, int j, int k,
                                         const T* h, T nu, T t[3][3]) {
#pragma clang fp contract(off)
  constexpr int NV = P::nv;
  T g[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int d = 0; d < 3; d++) g[a][d] = T(0);
#pragma unroll
  for (int d = 0; d < ND; d++) {
    const int di = d == 0, dj = d == 1, dk = d == 2;
    const size_t pp = cell_index<sizeof(T)>(nbr, rc, b, i + di, j + dj, k + dk);
    const size_t pm = cell_index<sizeof(T)>(nbr, rc, b, i - di, j - dj, k - dk);
#pragma unroll
    for (int a = 0; a < 3; a++)
      g[a][d] = T(0.5) * (Q[var_index<NV>(rc, pp, P::ivx + a)] -
                          Q[var_index<NV>(rc, pm, P::ivx + a)]) / h[d];
  }
  const T rho = Q[var_index<NV>(rc, cell_index<sizeof(T)>(nbr, rc, b, i, j, k), P::idn)];
  const T gc = nu * rho, fc = T(2) * gc;
  const T divv = ((g[0][0] + g[1][1]) + g[2][2]) / T(3);
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

// The viscous and resistive source terms of update_sources
// (sources.F90:340-467, 543-598) of the stage input Q, per interior cell, as
// V[b][s][cell]:
//   vis & 1: s = 0..2 div t(a,:) (momentum), 3 div (t v) (energy);
//   vis & 2: s = 4..6 eta lap B, 7 B . (eta lap B), 8 eta |curl B|^2
// with derivative_1st's central differences 0.5 (u(i+1) - u(i-1)) / dh and
// derivative_2nd's ((u(i+1) + u(i-1)) - 2 u(i)) / dh^2 summed over the
// directions in order (operators.F90:181-711) at the cell size of the
// block's level; the stress of the six face neighbours is formed from the
// velocities one cell further out.  Rounded as the reference rounds (no
// contraction).  The last sweep adds them to the completed increment.
template <class P, int ND, typename T>
static __global__ void __launch_bounds__(256)
k_visc_src(const T* __restrict__ Q, const int* __restrict__ nbr, RunConst rc,
           const RunState* __restrict__ rs, const int* __restrict__ brun,
           const int* __restrict__ blev, int vis, T* __restrict__ V) {
#pragma clang fp contract(off)
  constexpr int NV = P::nv;
  const int ncell = rc.ncell;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)rc.nblk * ncell) return;
  const int b = (int)(gid / ncell), s = (int)(gid % ncell);
  int i, j, k;
  src_cell<T>(rc, s, i, j, k);
  const RunState* R = rs + brun[b];
  const double sc = blev ? 1.0 / double(1 << blev[b]) : 1.0;
  T h[3];
#pragma unroll
  for (int d = 0; d < 3; d++) h[d] = T(R->h[d] * sc);
  T* o = V + (size_t)b * 9 * ncell + s;
  auto at = [&](int di, int dj, int dk, int v) {
    return Q[var_index<NV>(rc, cell_index<sizeof(T)>(nbr, rc, b, i + di, j + dj, k + dk), v)];
  };
  if (vis & 1) {
    const T nu = T(rc.nu);
    T sm[3] = {T(0), T(0), T(0)}, se = T(0);
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const int di = d == 0, dj = d == 1, dk = d == 2;
      T tp[3][3], tm[3][3];
      visc_tau<P, ND, T>(Q, nbr, rc, b, i + di, j + dj, k + dk, h, nu, tp);
      visc_tau<P, ND, T>(Q, nbr, rc, b, i - di, j - dj, k - dk, h, nu, tm);
#pragma unroll
      for (int a = 0; a < 3; a++) sm[a] = sm[a] + T(0.5) * (tp[a][d] - tm[a][d]) / h[d];
      if constexpr (!P::iso) {
        const T vpx = at(di, dj, dk, P::ivx), vpy = at(di, dj, dk, P::ivy);
        const T vpz = at(di, dj, dk, P::ivz);
        const T vmx = at(-di, -dj, -dk, P::ivx), vmy = at(-di, -dj, -dk, P::ivy);
        const T vmz = at(-di, -dj, -dk, P::ivz);
        const T gp = (vpx * tp[d][0] + vpy * tp[d][1]) + vpz * tp[d][2];
        const T gm = (vmx * tm[d][0] + vmy * tm[d][1]) + vmz * tm[d][2];
        se = se + T(0.5) * (gp - gm) / h[d];
      }
    }
    o[0] = sm[0];
    o[(size_t)ncell] = sm[1];
    o[2 * (size_t)ncell] = sm[2];
    o[3 * (size_t)ncell] = se;
  }
  if constexpr (P::mhd) {
    if (vis & 2) {
      const T eta = T(rc.eta);
      T el[3];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const T bc = at(0, 0, 0, P::ibx + a);
        T lap = T(0);
#pragma unroll
        for (int d = 0; d < ND; d++) {
          const int di = d == 0, dj = d == 1, dk = d == 2;
          const T w = ((at(di, dj, dk, P::ibx + a) + at(-di, -dj, -dk, P::ibx + a)) - T(2) * bc) /
                      (h[d] * h[d]);
          lap = d == 0 ? w : lap + w;
        }
        el[a] = eta * lap;
        o[(size_t)(4 + a) * ncell] = el[a];
      }
      if constexpr (!P::iso) {
        auto d1 = [&](int d, int v) {
          const int di = d == 0, dj = d == 1, dk = d == 2;
          return T(0.5) * (at(di, dj, dk, v) - at(-di, -dj, -dk, v)) / h[d];
        };
        const T bx = at(0, 0, 0, P::ibx), by = at(0, 0, 0, P::iby), bz = at(0, 0, 0, P::ibz);
        T c0, c1;
        if constexpr (ND == 3) {
          c0 = d1(1, P::ibz) - d1(2, P::iby);
          c1 = d1(2, P::ibx) - d1(0, P::ibz);
        } else {
          c0 = d1(1, P::ibz);
          c1 = -d1(0, P::ibz);
        }
        const T c2 = d1(0, P::iby) - d1(1, P::ibx);
        o[7 * (size_t)ncell] = (bx * el[0] + by * el[1]) + bz * el[2];
        o[8 * (size_t)ncell] = eta * ((c0 * c0 + c1 * c1) + c2 * c2);
      }
    }
  }
}

}  // namespace amun_dev
