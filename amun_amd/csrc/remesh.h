// remesh.h -- the data movement of a mesh update on the device.
//
// update_mesh (mesh.F90:756-833) changes the leaf set by one level at a
// time: a block selected for derefinement is replaced by its parent, whose
// interior restrict_block (mesh.F90:1258-1454) fills with the mean of the
// 2^d covering child cells; a block selected for refinement is replaced by
// its 2^d children, which prolong_block (mesh.F90:1031-1257) fills from the
// parent's conserved state with limited slopes.  Every other leaf keeps its
// data.  Here the new block pool is written in one launch from the old one:
// one thread per new cell, all variables, cells in the pool's storage order
// so that the writes (and the copies' reads) are coalesced.
//
// prolong_block reads the parent's conserved variables one cell beyond its
// interior.  In the reference those ghost cells hold prim2cons of the
// primitive ghost values that boundary_variables filled
// (update_ghost_cells, boundaries.F90:6633-6736), not the neighbours'
// conserved values, so the kernel reads them the same way: the old mesh's
// primitive pool through its neighbour table (real blocks and the filled
// virtual blocks) and prim2cons in the reference's operation order.
#pragma once

namespace amun_dev {

enum { REMESH_COPY = 0, REMESH_PROLONG = 1, REMESH_RESTRICT = 2 };

// prim2cons_hd_adi / prim2cons_mhd_adi (equations.F90:1762, 2815) of one
// cell of an adiabatic HD (nv = 5) or MHD (nv = 9) primitive pool, variable
// by variable as the reference orders the operations; the isothermal
// systems (ien < 0, nv = 4 or 8: prim2cons_hd_iso / _mhd_iso, 1360, 2220)
// carry no energy
template <typename T>
__device__ __forceinline__ void ghost_cons(const T* __restrict__ Q, size_t b, size_t c,
                                           size_t ncell, int nv, int ien, double gammam1i,
                                           double* u) {
#pragma clang fp contract(off)
  double q[9];
  for (int v = 0; v < nv; v++) q[v] = double(Q[(b * nv + v) * ncell + c]);
  u[0] = q[0];
  u[1] = q[0] * q[1];
  u[2] = q[0] * q[2];
  u[3] = q[0] * q[3];
  if (ien < 0) {
    for (int v = 4; v < nv; v++) u[v] = q[v];
    return;
  }
  const double ek = 0.5 * ((u[1] * q[1] + u[2] * q[2]) + u[3] * q[3]);
  if (nv == 9) {
    for (int v = 4; v < 8; v++) u[v] = q[v];
    const double em = 0.5 * ((q[4] * q[4] + q[5] * q[5]) + q[6] * q[6]);
    u[8] = (q[8] * gammam1i + ek) + em;
  } else {
    u[4] = q[4] * gammam1i + ek;
  }
}

// position (i, j, k) of storage index s of a block's FP64 plane (the
// inverse of lidx<8>)
__device__ __forceinline__ void lidx8_inv(int nc, int nd, int s, int& i, int& j, int& k) {
  const int t = s >> 2;
  j = t % nc;
  if (nd == 2) {
    k = 0;
    i = ((t / nc) << 2) | (s & 3);
  } else {
    k = (t / nc) % nc;
    i = ((t / (nc * nc)) << 2) | (s & 3);
  }
}

// task[b] of new block b: x = kind, y = old source block (copy, prolong) or
// first of its 2^d entries in kids (restrict, child position px + 2 py +
// 4 pz), z = the block's position in its parent (px | py << 1 | pz << 2) for
// a prolongation.  rc is the old mesh's (nc, nd, ncell are the new one's too).
template <typename T>
static __global__ void __launch_bounds__(256)
k_remesh(const int4* __restrict__ task, const int* __restrict__ kids, int nblk,
         const double* __restrict__ U0o, const float* __restrict__ Co,
         const T* __restrict__ Qo, const int* __restrict__ nbro, RunConst rc,
         double* __restrict__ U0n, float* __restrict__ Cn, int nv, int ien,
         DevState* st) {
#pragma clang fp contract(off)
  const int nc = rc.nc, nd = rc.nd, ncell = rc.ncell;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)nblk * ncell) return;
  const int b = (int)(gid / ncell), s = (int)(gid % ncell);
  const int4 t = task[b];
  const size_t nb = (size_t)b * nv * ncell;
  if (t.x == REMESH_COPY) {
    const size_t ob = (size_t)t.y * nv * ncell;
    for (int v = 0; v < nv; v++) {
      U0n[nb + (size_t)v * ncell + s] = U0o[ob + (size_t)v * ncell + s];
      if (Cn) Cn[nb + (size_t)v * ncell + s] = Co[ob + (size_t)v * ncell + s];
    }
    return;
  }
  int i, j, k;
  lidx8_inv(nc, nd, s, i, j, k);
  if (Cn)
    for (int v = 0; v < nv; v++) Cn[nb + (size_t)v * ncell + s] = 0.0f;
  if (t.x == REMESH_RESTRICT) {
    const int half = nc >> 1;
    const int px = i >= half, py = j >= half, pz = (nd == 3) && k >= half;
    const int F = kids[t.y + px + 2 * py + 4 * pz];
    const int fi = 2 * (i - px * half), fj = 2 * (j - py * half);
    const int fk = nd == 3 ? 2 * (k - pz * half) : 0;
    const size_t fb = (size_t)F * nv * ncell;
    const int lll = lidx<8>(nc, nd, fi, fj, fk), pll = lidx<8>(nc, nd, fi + 1, fj, fk);
    const int lpl = lidx<8>(nc, nd, fi, fj + 1, fk), ppl = lidx<8>(nc, nd, fi + 1, fj + 1, fk);
    for (int v = 0; v < nv; v++) {
      const double* U = U0o + fb + (size_t)v * ncell;
      double r;
      if (nd == 2) {
        r = 2.50e-01 * ((U[lll] + U[ppl]) + (U[lpl] + U[pll]));
      } else {
        const int llp = lidx<8>(nc, nd, fi, fj, fk + 1);
        const int plp = lidx<8>(nc, nd, fi + 1, fj, fk + 1);
        const int lpp = lidx<8>(nc, nd, fi, fj + 1, fk + 1);
        const int ppp = lidx<8>(nc, nd, fi + 1, fj + 1, fk + 1);
        r = 1.25e-01 * (((U[lll] + U[ppp]) + (U[llp] + U[ppl])) +
                        ((U[lpp] + U[pll]) + (U[lpl] + U[plp])));
      }
      U0n[nb + (size_t)v * ncell + s] = r;
    }
    return;
  }
  // prolongation from old leaf P: parent cell and sub-cell position
  const int P = t.y;
  const int px = t.z & 1, py = (t.z >> 1) & 1, pz = (t.z >> 2) & 1;
  const int ci = (i + px * nc) >> 1, cj = (j + py * nc) >> 1;
  const int ck = nd == 3 ? (k + pz * nc) >> 1 : 0;
  const int sub = (i & 1) + 2 * (j & 1) + (nd == 3 ? 4 * (k & 1) : 0);
  const size_t pb = (size_t)P * nv * ncell;
  double um[3][9], up[3][9];
  for (int d = 0; d < nd; d++)
    for (int sg = -1; sg <= 1; sg += 2) {
      int q[3] = {ci, cj, ck};
      q[d] += sg;
      double* dst = sg < 0 ? um[d] : up[d];
      if (q[d] >= 0 && q[d] < nc) {
        const int c = lidx<8>(nc, nd, q[0], q[1], q[2]);
        for (int v = 0; v < nv; v++) dst[v] = U0o[pb + (size_t)v * ncell + c];
      } else {
        const size_t pk = cell_index<sizeof(T)>(nbro, rc, P, q[0], q[1], q[2]);
        ghost_cons<T>(Qo, pk >> 32, pk & 0xffffffffull, ncell, nv, ien, rc.gammam1i, dst);
      }
    }
  const int cc = lidx<8>(nc, nd, ci, cj, ck);
  for (int v = 0; v < nv; v++) {
    const double uc = U0o[pb + (size_t)v * ncell + cc];
    double du[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < nd; d++)
      du[d] = limiter<double>(rc.prol_limiter, 0.5, uc - um[d][v], up[d][v] - uc);
    if (v == 0 || v == ien) {
      if (uc > 0.0) {
        for (int it = 0; it < kMaxHalvings; it++) {
          double sa = fabs(du[0]) + fabs(du[1]);
          if (nd == 3) sa = sa + fabs(du[2]);
          if (!(uc <= sa)) break;
          du[0] *= 0.5; du[1] *= 0.5; du[2] *= 0.5;
        }
      } else {
        atomicAdd(&st->npos, 1ull);
        du[0] = du[1] = du[2] = 0.0;
      }
    }
    du[0] *= 0.5; du[1] *= 0.5; du[2] *= 0.5;
    double r;
    if (nd == 2) {
      const double du1 = du[0] + du[1], du2 = du[0] - du[1];
      r = sub == 0 ? uc - du1 : (sub == 1 ? uc + du2 : (sub == 2 ? uc - du2 : uc + du1));
    } else {
      const double du1 = (du[0] + du[1]) + du[2], du2 = (du[0] - du[1]) - du[2];
      const double du3 = (du[0] - du[1]) + du[2], du4 = (du[0] + du[1]) - du[2];
      switch (sub) {
      case 0: r = uc - du1; break;
      case 1: r = uc + du2; break;
      case 2: r = uc - du3; break;
      case 3: r = uc + du4; break;
      case 4: r = uc - du4; break;
      case 5: r = uc + du3; break;
      case 6: r = uc - du2; break;
      default: r = uc + du1; break;
      }
    }
    U0n[nb + (size_t)v * ncell + s] = r;
  }
}

}  // namespace amun_dev
