// sources.h -- the GLM source terms of update_sources (sources.F90:479-539).
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

}  // namespace amun_dev
