// mlp.h -- the multi-dimensional limiting process (mlp_limiting,
// interpolations.F90:1171-1321).
//
// With mlp_limiting on, the reference passes the face states of every
// reconstruction through a per-cell process over the cell's 3^d box:
// minmod slopes in every direction; where their absolute sum exceeds the
// distance of the cell value to the box's extremes, slopes from the central
// differences, scaled to that distance and redistributed between the
// directions (the 2D and the cyclic 3D formulas), with the minmod slopes'
// signs; then the cell's two face states of direction d are reset to
// q +- dq(d) when either lies further from q than |dq(d)|.
//
// The process reads all directions, so k_mlp evaluates the slopes dq(d)
// beforehand from the stage input: for every interior cell of every real
// block (S) and, block by block, for the cells just beyond each face that
// a pencil walk reconstructs (G: the cell before the first and after the
// last of every pencil, through the block's own neighbour table, so that
// neighbour, virtual and ghost blocks all serve).  The sweeps then only
// compare and reset (mlp_apply).
//
// The reference runs the process over the whole block array, indices
// clamped at its edges, in increasing i, j, k.  The outermost cell of the
// lower side thus has a box cut to two cells along the direction, and its
// reset writes the left face state of the next cell, which that cell's test
// reads next.  With nghosts = 2 that next cell is the first ghost cell,
// whose right face state is used, so G also holds the slope of the
// outermost cell and mlp_apply replays its reset first.  Its right face
// state is q itself in the reference (the ghost-zone copy of interfaces_tvd
// and interfaces_mgp); with nghosts = 2 the process is therefore built for
// these two reconstructions only.
//
// Arithmetic in the reference's rounding (no contraction).
//
// Included by kernels.h inside namespace amun_dev, after cell_index().
#pragma once

// dq(d) after the process, for variable v of the cell whose 3^ND box
// (packed pool positions, (dk, dj, di) order) is box
template <class P, int ND, typename T>
__device__ __forceinline__ void mlp_slopes(const RunConst& rc, const T* __restrict__ Q,
                                           const size_t* box, int v, T* dq) {
#pragma clang fp contract(off)
  constexpr int NV = P::nv;
  constexpr int NB = ND == 3 ? 27 : 9;
  constexpr int C = ND == 3 ? 13 : 4;
  constexpr int dm[3] = {C - 1, C - 3, C - 9}, dp[3] = {C + 1, C + 3, C + 9};
  auto at = [&](int m) { return Q[var_index<NV>(rc, box[m], v)]; };
  const T qc = at(C);
  dq[0] = dq[1] = dq[2] = T(0);
  for (int d = 0; d < ND; d++)
    dq[d] = limiter<T>(LIM_MINMOD, T(0.5), qc - at(dm[d]), at(dp[d]) - qc);
  T qmn = qc, qmx = qc;
  for (int m = 0; m < NB; m++) {
    const T w = at(m);
    qmn = w < qmn ? w : qmn;
    qmx = w > qmx ? w : qmx;
  }
  const T a1 = qmx - qc, a2 = qc - qmn;
  const T dqc = a1 < a2 ? a1 : a2;
  T sa = fabs(dq[0]) + fabs(dq[1]);
  if (ND == 3) sa = sa + fabs(dq[2]);
  if (!(sa > dqc)) return;
  T dqm[3] = {T(0), T(0), T(0)}, ap[3] = {T(0), T(0), T(0)};
  for (int d = 0; d < ND; d++) dqm[d] = fabs(at(dp[d]) - at(dm[d]));
  T sm = fabs(dqm[0]) + fabs(dqm[1]);
  if (ND == 3) sm = sm + fabs(dqm[2]);
  T fc = dqc / (sm > T(1.0e-16) ? sm : T(1.0e-16));
  for (int d = 0; d < ND; d++) ap[d] = fc * fabs(dqm[d]);
  if constexpr (ND == 3) {
    T hh[3], uu[3];
    for (int n = 0; n < 3; n++) {
      const int n1 = (n + 1) % 3, n2 = (n + 2) % 3;
      const T h0 = ap[n] - fabs(dq[n]);
      hh[n] = h0 > T(0) ? h0 : T(0);
      uu[n] = ap[n] - hh[n];
      uu[n1] = ap[n1] + T(0.5) * hh[n];
      uu[n2] = ap[n2] + T(0.5) * hh[n];
      fc = hh[n] / (hh[n] + T(1.0e-16));
      const T e1 = ap[n1] - fabs(dq[n1]), e2 = ap[n2] - fabs(dq[n2]);
      const T fl = fc * ((e1 > T(0) ? e1 : T(0)) - (e2 > T(0) ? e2 : T(0)));
      ap[n] = uu[n];
      ap[n1] = uu[n1] - fl;
      ap[n2] = uu[n2] + fl;
    }
  } else {
    const T e1 = ap[0] - fabs(dq[0]), e2 = ap[1] - fabs(dq[1]);
    const T fl = (e1 > T(0) ? e1 : T(0)) - (e2 > T(0) ? e2 : T(0));
    ap[0] = ap[0] - fl;
    ap[1] = ap[1] + fl;
  }
  for (int d = 0; d < ND; d++) dq[d] = copysign(fabs(ap[d]), dq[d]);
}

// packed positions of the box of cell (i, j, k) of block b; with clampdir
// >= 0 the lower neighbour along that axis is the cell itself (the
// reference's clamped index at the block array's lower edge)
template <int ND, typename T>
__device__ __forceinline__ void mlp_box(const int* __restrict__ nbr, const RunConst& rc, int b,
                                        int i, int j, int k, int clampdir, size_t* box) {
  int m = 0;
  for (int dk = (ND == 3 ? -1 : 0); dk <= (ND == 3 ? 1 : 0); dk++)
    for (int dj = -1; dj <= 1; dj++)
      for (int di = -1; di <= 1; di++) {
        const int oi = (clampdir == 0 && di < 0) ? 0 : di;
        const int oj = (clampdir == 1 && dj < 0) ? 0 : dj;
        const int ok = (clampdir == 2 && dk < 0) ? 0 : dk;
        box[m++] = cell_index<sizeof(T)>(nbr, rc, b, i + oi, j + oj, k + ok);
      }
}

// G slopes per block, variable and direction: the cells -1 and nc and the
// clamped outermost cell -2 of every pencil
template <int ND>
__host__ __device__ __forceinline__ int mlp_gsize(int nc) {
  return 3 * (ND == 3 ? nc * nc : nc);
}

// S[((b NV + v) ND + d) ncell + cell] (cell: storage index of the pool's
// element size) for the interior cells; G[((b NV + v) ND + d) gsize + side
// npen + pencil], side 0: cell -1, 1: cell nc, 2: cell -2 with its clamped
// box.  Threads: per block ncell interior cells, then ND * gsize boundary
// cells.
template <class P, int ND, typename T>
static __global__ void __launch_bounds__(256)
k_mlp(const T* __restrict__ Q, const int* __restrict__ nbr, RunConst rc, T* __restrict__ S,
      T* __restrict__ G) {
  constexpr int NV = P::nv;
  constexpr int NB = ND == 3 ? 27 : 9;
  const int nc = rc.nc, ncell = rc.ncell;
  const int npen = ND == 3 ? nc * nc : nc, gs = mlp_gsize<ND>(nc);
  const int per = ncell + ND * gs;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)rc.nblk * per) return;
  const int b = (int)(gid / per), s = (int)(gid % per);
  size_t box[NB];
  T dq[3];
  if (s < ncell) {
    int i, j, k;
    const int q = s >> 2;
    if (ND == 2) {
      j = q % nc;
      i = ((q / nc) << 2) | (s & 3);
      k = 0;
    } else if (sizeof(T) == 8) {
      j = q % nc;
      k = (q / nc) % nc;
      i = ((q / (nc * nc)) << 2) | (s & 3);
    } else {
      const int kl = q & 1, r = q >> 1;
      j = r % nc;
      const int r2 = r / nc;
      k = ((r2 % (nc >> 1)) << 1) | kl;
      i = ((r2 / (nc >> 1)) << 2) | (s & 3);
    }
    mlp_box<ND, T>(nbr, rc, b, i, j, k, -1, box);
    for (int v = 0; v < NV; v++) {
      mlp_slopes<P, ND, T>(rc, Q, box, v, dq);
      for (int d = 0; d < ND; d++) S[(((size_t)b * NV + v) * ND + d) * ncell + s] = dq[d];
    }
    return;
  }
  const int g = s - ncell;
  const int d = g / gs, r = g % gs, side = r / npen, p = r % npen;
  const int c = side == 0 ? -1 : (side == 1 ? nc : -2);
  const int t1 = p % nc, t2 = ND == 3 ? p / nc : 0;
  int i, j, k;
  if (d == 0) { i = c; j = t1; k = t2; }
  else if (d == 1) { i = t1; j = c; k = t2; }
  else { i = t1; j = t2; k = c; }
  mlp_box<ND, T>(nbr, rc, b, i, j, k, side == 2 ? d : -1, box);
  for (int v = 0; v < NV; v++) {
    mlp_slopes<P, ND, T>(rc, Q, box, v, dq);
    G[(((size_t)b * NV + v) * ND + d) * gs + side * npen + p] = dq[d];
  }
}

// the reset of the face states along DIR of cell c (-1 .. nc) of pencil p
// of block b (fl: the state at the cell's right face, fr: at its left
// face); lowedge (cell -1 with nghosts = 2) replays the clamped outermost
// cell's reset first
template <class P, int ND, int DIR, typename T>
__device__ __forceinline__ void mlp_apply(const RunConst& rc, const T* __restrict__ Q,
                                          const T* __restrict__ S, const T* __restrict__ G,
                                          const int* __restrict__ nbr, int b, int c, int p,
                                          int i, int j, int k, bool lowedge, T* fl, T* fr) {
#pragma clang fp contract(off)
  constexpr int NV = P::nv;
  const int nc = rc.nc, ncell = rc.ncell;
  const int npen = ND == 3 ? nc * nc : nc, gs = mlp_gsize<ND>(nc);
  const size_t pk = cell_index<sizeof(T)>(nbr, rc, b, i, j, k);
  const size_t bb = pk >> 32, cc = pk & 0xffffffffull;
  size_t pk2 = 0;
  if (lowedge)
    pk2 = cell_index<sizeof(T)>(nbr, rc, b, i - (DIR == 0), j - (DIR == 1), k - (DIR == 2));
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const T qc = Q[(bb * NV + v) * (size_t)ncell + cc];
    const size_t gb = (((size_t)b * NV + v) * ND + DIR) * gs;
    T dq;
    if (c < 0) dq = G[gb + p];
    else if (c >= nc) dq = G[gb + npen + p];
    else dq = S[(((size_t)b * NV + v) * ND + DIR) * ncell + cc];
    if (lowedge) {
      // the outermost cell: its right face state is q (dl = 0); its left
      // face neighbour's state is this cell's fr
      const T q2 = Q[((pk2 >> 32) * NV + v) * (size_t)ncell + (pk2 & 0xffffffffull)];
      const T dq2 = G[gb + 2 * npen + p];
      if (fabs(fr[v] - q2) > fabs(dq2)) fr[v] = q2 - dq2;
    }
    const T dl = fabs(fl[v] - qc), dr = fabs(fr[v] - qc);
    if ((dl > dr ? dl : dr) > fabs(dq)) {
      fl[v] = qc + dq;
      fr[v] = qc - dq;
    }
  }
}
