// amr.h -- device side of multi-level (AMR) block meshes.
//
// The sweeps read every cell outside a block through the block's 27-entry
// neighbour table (cell_index).  On a multi-level mesh a neighbour position
// that holds no leaf of the block's own level points to a "virtual" block
// instead: a block of the same level at that position, stored in the same
// primitive pools behind the real blocks, whose cells within nghosts of the
// referring blocks hold what the reference's boundary_variables writes into
// the ghost zones there (boundaries.F90:372-479):
//   - the region is refined: the average of the 2^d covering cells of the
//     finer leaf (block_face/edge/corner_restrict, boundaries.F90:5637,
//     5919, 6256, in their summation order);
//   - the region lies inside a coarser leaf: the parent cell plus limited
//     slopes (block_face/edge/corner_prolong, 5762, 6071, 6351), whose
//     stencil reads the coarse leaf's neighbours through its own table
//     (real blocks or restriction-filled virtual blocks).
// The fill runs before the sweeps of every stage: all restrictions first,
// then all prolongations, as the reference orders them.  The host builds the
// task lists (one entry per virtual cell) and checks that no prolongation
// stencil reads a prolonged cell (2:1 level balance guarantees it).
//
// The flux correction of boundary_fluxes (boundaries.F90:494-1060 with
// block_update_flux 6539-6614) is applied inside the sweeps: a coarse block
// whose face has finer neighbours takes, at that face, the average of the
// fine faces' fluxes, which k_amr_fineflux computes beforehand from the same
// primitive state with the sweep's own reconstruction and Riemann code.
#pragma once

namespace amun_dev {

// task of one virtual cell: x = destination block, y = destination cell
// (i | j << 8 | k << 16), z = source block, w = source cell (i | j << 8 |
// k << 16) with the child position (sx << 24 | sy << 25 | sz << 26) of a
// prolongation
__device__ __forceinline__ void unpack_cell(int p, int& i, int& j, int& k) {
  i = p & 0xff;
  j = (p >> 8) & 0xff;
  k = (p >> 16) & 0xff;
}

template <typename T>
static __global__ void k_amr_restrict(const int4* __restrict__ task, int ntask,
                                      T* __restrict__ Q, RunConst rc, int nv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ntask) return;
  const int4 t = task[n];
  int i, j, k, fi, fj, fk;
  unpack_cell(t.y, i, j, k);
  unpack_cell(t.w, fi, fj, fk);
  const int nc = rc.nc, nd = rc.nd;
  constexpr int ES = sizeof(T);
  const size_t dst = (size_t)t.x * nv * rc.ncell + lidx<ES>(nc, nd, i, j, k);
  const size_t src = (size_t)t.z * nv * rc.ncell;
  // the 2^d children: l = lower, p = upper cell along each axis
  const size_t lll = src + lidx<ES>(nc, nd, fi, fj, fk);
  const size_t pll = src + lidx<ES>(nc, nd, fi + 1, fj, fk);
  const size_t lpl = src + lidx<ES>(nc, nd, fi, fj + 1, fk);
  const size_t ppl = src + lidx<ES>(nc, nd, fi + 1, fj + 1, fk);
  for (int v = 0; v < nv; v++) {
    const size_t o = (size_t)v * rc.ncell;
    double r;
    if (nd == 2) {
      r = 2.50e-01 * ((double(Q[lll + o]) + double(Q[ppl + o])) +
                      (double(Q[lpl + o]) + double(Q[pll + o])));
    } else {
      const size_t llp = src + lidx<ES>(nc, nd, fi, fj, fk + 1);
      const size_t plp = src + lidx<ES>(nc, nd, fi + 1, fj, fk + 1);
      const size_t lpp = src + lidx<ES>(nc, nd, fi, fj + 1, fk + 1);
      const size_t ppp = src + lidx<ES>(nc, nd, fi + 1, fj + 1, fk + 1);
      r = 1.25e-01 * (((double(Q[lll + o]) + double(Q[ppp + o])) +
                       (double(Q[llp + o]) + double(Q[ppl + o]))) +
                      ((double(Q[lpp + o]) + double(Q[pll + o])) +
                       (double(Q[lpl + o]) + double(Q[plp + o]))));
    }
    Q[dst + o] = T(r);
  }
}

template <typename T>
static __global__ void k_amr_prolong(const int4* __restrict__ task, int ntask,
                                     T* __restrict__ Q, const int* __restrict__ nbr,
                                     RunConst rc, int nv, int ipr, DevState* st) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ntask) return;
  const int4 t = task[n];
  int i, j, k, ci, cj, ck;
  unpack_cell(t.y, i, j, k);
  unpack_cell(t.w, ci, cj, ck);
  const int sx = (t.w >> 24) & 1, sy = (t.w >> 25) & 1, sz = (t.w >> 26) & 1;
  const int nc = rc.nc, nd = rc.nd, C = t.z;
  constexpr int ES = sizeof(T);
  const size_t dst = (size_t)t.x * nv * rc.ncell + lidx<ES>(nc, nd, i, j, k);
  const size_t pc = cell_index<ES>(nbr, rc, C, ci, cj, ck);
  size_t pm[3] = {pc, pc, pc}, pp[3] = {pc, pc, pc};
  pm[0] = cell_index<ES>(nbr, rc, C, ci - 1, cj, ck);
  pp[0] = cell_index<ES>(nbr, rc, C, ci + 1, cj, ck);
  pm[1] = cell_index<ES>(nbr, rc, C, ci, cj - 1, ck);
  pp[1] = cell_index<ES>(nbr, rc, C, ci, cj + 1, ck);
  if (nd == 3) {
    pm[2] = cell_index<ES>(nbr, rc, C, ci, cj, ck - 1);
    pp[2] = cell_index<ES>(nbr, rc, C, ci, cj, ck + 1);
  }
  auto at = [&](size_t packed, int v) {
    const size_t b = packed >> 32, c = packed & 0xffffffffull;
    return double(Q[(b * nv + v) * (size_t)rc.ncell + c]);
  };
  for (int v = 0; v < nv; v++) {
    const double qc = at(pc, v);
    double dq[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < nd; d++)
      dq[d] = limiter<double>(rc.prol_limiter, 0.5, qc - at(pm[d], v), at(pp[d], v) - qc);
    if (v == 0 || v == ipr) {
      if (qc > 0.0) {
        for (int it = 0; it < kMaxHalvings; it++) {
          double sa = fabs(dq[0]) + fabs(dq[1]);
          if (nd == 3) sa = sa + fabs(dq[2]);
          if (!(qc <= sa)) break;
          dq[0] *= 0.5; dq[1] *= 0.5; dq[2] *= 0.5;
        }
      } else {
        atomicAdd(&st->npos, 1ull);
        dq[0] = dq[1] = dq[2] = 0.0;
      }
    }
    dq[0] *= 0.5; dq[1] *= 0.5; dq[2] *= 0.5;
    double r;
    if (nd == 2) {
      const double dq1 = dq[0] + dq[1], dq2 = dq[0] - dq[1];
      const int s = sx + 2 * sy;
      r = s == 0 ? qc - dq1 : (s == 1 ? qc + dq2 : (s == 2 ? qc - dq2 : qc + dq1));
    } else {
      const double dq1 = (dq[0] + dq[1]) + dq[2], dq2 = (dq[0] - dq[1]) - dq[2];
      const double dq3 = (dq[0] - dq[1]) + dq[2], dq4 = (dq[0] + dq[1]) - dq[2];
      switch (sx + 2 * sy + 4 * sz) {
      case 0: r = qc - dq1; break;
      case 1: r = qc + dq2; break;
      case 2: r = qc - dq3; break;
      case 3: r = qc + dq4; break;
      case 4: r = qc - dq4; break;
      case 5: r = qc + dq3; break;
      case 6: r = qc - dq2; break;
      default: r = qc + dq1; break;
      }
    }
    Q[dst + (size_t)v * rc.ncell] = T(r);
  }
}

// Ghost blocks of the non-periodic domain sides (boundaries_specific,
// boundaries.F90:1089-1251, block_boundary_specific 5024-5637).  A virtual
// block stands next to a boundary block across one or more non-periodic
// sides; task = {virtual block, source block, side per axis (2 bits each:
// 0 none, 1 low, 2 high)}.  Its cells are the source block's, mirrored
// (reflective) or held at the edge cell (open, outflow) along every such
// axis, with the side types applied in x, y, z order: outflow clamps the
// normal velocity to point outwards, reflective negates the normal velocity
// and field.  One thread per cell.
template <int NV, typename T>
static __global__ void k_bnd_fill(const int4* __restrict__ task, int ntask, T* Q, RunConst rc,
                                  int ivx, int ibx) {
  const int nc = rc.nc, nd = rc.nd;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long long)ntask * rc.ncell) return;
  const int4 t = task[n / rc.ncell];
  const int c = (int)(n % rc.ncell);
  const int ci[3] = {c % nc, (c / nc) % nc, nd == 3 ? c / (nc * nc) : 0};
  int si[3], side[3], type[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    side[a] = (t.z >> (2 * a)) & 3;
    type[a] = side[a] ? rc.bnd[2 * a + side[a] - 1] : 0;
    if (!side[a]) si[a] = ci[a];
    else if (type[a] == AMUN_BND_REFLECTIVE) si[a] = nc - 1 - ci[a];
    else si[a] = side[a] == 1 ? 0 : nc - 1;
  }
  const size_t dst = (size_t)t.x * NV * rc.ncell + lidx<sizeof(T)>(nc, nd, ci[0], ci[1], ci[2]);
  const size_t src = (size_t)t.y * NV * rc.ncell + lidx<sizeof(T)>(nc, nd, si[0], si[1], si[2]);
  T q[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) q[v] = Q[src + (size_t)v * rc.ncell];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (!side[a]) continue;
    if (type[a] == AMUN_BND_OUTFLOW) {
      const T x = q[ivx + a];
      q[ivx + a] = side[a] == 1 ? (T(0) < x ? T(0) : x) : (T(0) > x ? T(0) : x);
    } else if (type[a] == AMUN_BND_REFLECTIVE) {
      q[ivx + a] = -q[ivx + a];
      if (ibx >= 0) q[ibx + a] = -q[ibx + a];
    }
  }
#pragma unroll
  for (int v = 0; v < NV; v++) Q[dst + (size_t)v * rc.ncell] = q[v];
}

// Flux through face c (0: the lower face of the pencil's first cell, nc:
// the upper face of its last) of pencil (c1, c2) of block b along DIR, as
// the sweep computes it (original component order).
template <class P, int REC, int RS, int ND, int DIR, typename T>
__device__ __forceinline__ void face_flux(const RunConst& rc, const T* __restrict__ Qp,
                                          const int* __restrict__ nbr, int b, int c, int c1,
                                          int c2, T cmax, T cmax2, DevState* st, double h,
                                          const T* mlps, const T* mlpg, T* f) {
  constexpr int NV = P::nv;
  using RC = Recon<P, REC, ND, DIR, T>;
  const EosConst<T> e{T(rc.gamma), T(rc.gammam1), T(rc.gammam1i), T(rc.csnd), T(rc.csnd2)};
  auto coords = [&](int cc, int& i, int& j, int& k) {
    if constexpr (DIR == 0) { i = cc; j = c1; k = c2; }
    else if constexpr (DIR == 1) { i = c1; j = cc; k = c2; }
    else { i = c1; j = c2; k = cc; }
  };
  auto loadq = [&](int cc, T* q) {
    int i, j, k;
    coords(cc, i, j, k);
    const size_t pk = cell_index<sizeof(T)>(nbr, rc, b, i, j, k);
#pragma unroll
    for (int v = 0; v < NV; v++) q[v] = Qp[var_index<NV>(rc, pk, v)];
  };
  RC rec;
  T qtmp[NV], flm[NV], frm[NV], flc[NV], frc[NV];
#pragma unroll
  for (int m = 0; m <= 2 * RC::R; m++) {
    loadq(c - 1 - RC::R + m, qtmp);
    rec.load(m, qtmp);
  }
  int i, j, k;
  const int pen = c1 + rc.nc * c2;
  coords(c - 1, i, j, k);
  rec.cell(rc, flm, frm, Qp, nbr, b, i, j, k, st, h);
  if (mlps)
    mlp_apply<P, ND, DIR, T>(rc, Qp, mlps, mlpg, nbr, b, c - 1, pen, i, j, k,
                             c - 1 == -1 && rc.ng == 2, flm, frm);
  rec.shift();
  loadq(c + RC::R, qtmp);
  rec.load(2 * RC::R, qtmp);
  coords(c, i, j, k);
  rec.cell(rc, flc, frc, Qp, nbr, b, i, j, k, st, h);
  if (mlps)
    mlp_apply<P, ND, DIR, T>(rc, Qp, mlps, mlpg, nbr, b, c, pen, i, j, k, false, flc, frc);
  T ql[NV], qr[NV], fr[NV];
#pragma unroll
  for (int p = 0; p < NV; p++) {
    ql[p] = flm[rot<P, DIR>(p)];
    qr[p] = frc[rot<P, DIR>(p)];
  }
  riemann<P, RS, T>(e, rc.sc, cmax, cmax2, ql, qr, fr);
#pragma unroll
  for (int p = 0; p < NV; p++) f[rot<P, DIR>(p)] = fr[p];
}

// Averaged fine fluxes of the coarse faces with finer neighbours, one
// thread per coarse face cell: x = correction slot, y = face cell of the
// coarse block (c1 + nc c2), z = fine block, w = its lower-corner fine face
// cell (f1 | f2 << 8) and face (0 or nc) << 16.  FB[slot][var][face cell].
template <class P, int REC, int RS, int ND, int DIR, typename T>
static __global__ void k_amr_fineflux(const int4* __restrict__ task, int ntask, const void* Q,
                                      const int* __restrict__ nbr, RunConst rc,
                                      const RunState* __restrict__ rs, int slot,
                                      double* __restrict__ FB, DevState* st,
                                      const int* __restrict__ blev, const void* mlps_,
                                      const void* mlpg_) {
  constexpr int NV = P::nv;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ntask) return;
  const int4 t = task[n];
  const int f1 = t.w & 0xff, f2 = (t.w >> 8) & 0xff, c = (t.w >> 16) & 0xff;
  const double cmaxd = cmax_of(rs, slot);
  const T cmax = T(cmaxd), cmax2 = T(cmaxd * cmaxd);
  const T* Qp = static_cast<const T*>(Q);
  const T* mlps = static_cast<const T*>(mlps_);
  const T* mlpg = static_cast<const T*>(mlpg_);
  T fa[NV], fb[NV];
  const int nf = (ND == 3) ? rc.nc * rc.nc : rc.nc;
  double* out = FB + (size_t)t.x * NV * nf + t.y;
  const double h = rs->h[DIR] / double(1 << blev[t.z]);   // the fine block's cell size
  face_flux<P, REC, RS, ND, DIR, T>(rc, Qp, nbr, t.z, c, f1, f2, cmax, cmax2, st, h, mlps, mlpg, fa);
  face_flux<P, REC, RS, ND, DIR, T>(rc, Qp, nbr, t.z, c, f1 + 1, f2, cmax, cmax2, st, h, mlps, mlpg, fb);
  if constexpr (ND == 2) {
#pragma unroll
    for (int v = 0; v < NV; v++) out[(size_t)v * nf] = (double(fa[v]) + double(fb[v])) / 2.0;
  } else {
    T fc[NV], fd[NV];
    face_flux<P, REC, RS, ND, DIR, T>(rc, Qp, nbr, t.z, c, f1, f2 + 1, cmax, cmax2, st, h, mlps, mlpg, fc);
    face_flux<P, REC, RS, ND, DIR, T>(rc, Qp, nbr, t.z, c, f1 + 1, f2 + 1, cmax, cmax2, st, h, mlps, mlpg, fd);
    // ((l,l) + (p,p)) + ((l,p) + (p,l)) over the two transverse axes
#pragma unroll
    for (int v = 0; v < NV; v++)
      out[(size_t)v * nf] = ((double(fa[v]) + double(fd[v])) + (double(fc[v]) + double(fb[v]))) / 4.0;
  }
}

}  // namespace amun_dev
