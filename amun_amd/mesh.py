"""Host-side multi-level block meshes for amun_dev_init_amr.

A mesh is an (nblk, 4) int array of leaf blocks: level (1 = base) and the
first interior cell (i0, j0, k0) in cells of that level, on the periodic base
grid of the parameters (xblocks x yblocks [x zblocks] blocks of ncells^d
cells).  ``refine`` builds one the way AMUN's generate_mesh does
(mesh.F90:431-740): level by level, leaves whose indicator asks for it are
split into 2^d children, and neighbours are refined further until every
leaf's face, edge and corner neighbours differ from it by one level at most
(the 2:1 balance the ghost-cell exchange relies on).  ``update_leaves`` is
update_mesh's selection (mesh.F90:756-833): from the per-block refinement
criterion (amun_dev_refinement) it returns the next leaf set, which
amun_dev_amr_remesh moves the device state onto.
"""
from __future__ import annotations

import numpy as np

from .config import Params
from .problems import prim_to_cons


def _nb(p: Params, level: int):
    f = 1 << (level - 1)
    return (p.xblocks * f, p.yblocks * f, (p.zblocks if p.ndims == 3 else 1) * f)


def refine(p: Params, wants, maxlev: int) -> np.ndarray:
    """wants(level, bx, by, bz) -> bool: refine the leaf at block position
    (bx, by, bz) of ``level``.  Returns the balanced leaf list as meta."""
    nd = p.ndims
    leaves = {(1, x, y, z) for x in range(p.xblocks) for y in range(p.yblocks)
              for z in range(p.zblocks if nd == 3 else 1)}

    def children(b):
        L, x, y, z = b
        return [(L + 1, 2 * x + i, 2 * y + j, 2 * z + k)
                for k in range(2 if nd == 3 else 1) for j in range(2) for i in range(2)]

    def covering(L, x, y, z):
        """Leaf covering block position (x, y, z) of level L, if any at <= L."""
        nbx, nby, nbz = _nb(p, L)
        x, y, z = x % nbx, y % nby, z % nbz
        for l in range(L, 0, -1):
            s = L - l
            b = (l, x >> s, y >> s, z >> s)
            if b in leaves:
                return b
        return None

    for L in range(1, maxlev):
        for b in sorted(b for b in leaves if b[0] == L and wants(*b)):
            leaves.remove(b)
            leaves.update(children(b))
        # 2:1 balance across faces, edges and corners
        changed = True
        while changed:
            changed = False
            for b in sorted(leaves, reverse=True):
                if b not in leaves:
                    continue
                Lb, x, y, z = b
                for dz in ((-1, 0, 1) if nd == 3 else (0,)):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            c = covering(Lb, x + dx, y + dy, z + dz)
                            if c is not None and c[0] < Lb - 1:
                                leaves.remove(c)
                                leaves.update(children(c))
                                changed = True
    n = p.ncells
    meta = np.array(sorted((L, x * n, y * n, z * n) for L, x, y, z in leaves), dtype=np.int32)
    return meta


def _touching(p: Params, leaves: dict, blk):
    """The leaves that share a face, edge or corner with leaf blk = (level,
    bx, by, bz) on the periodic base grid: same-level ones, the coarser one
    covering a neighbour position, or the finer ones of that position that
    touch blk (a 2:1 balanced mesh has no other).  These are the blocks the
    reference's neighbour pointers reach (faces, edges, corners;
    blocks.F90:3825-3851 iterate_over_neighbors)."""
    L, x, y, z = blk
    nd = p.ndims
    n = _nb(p, L)
    out = []
    for dz in ((-1, 0, 1) if nd == 3 else (0,)):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx == dy == dz == 0:
                    continue
                q = ((x + dx) % n[0], (y + dy) % n[1], (z + dz) % n[2])
                if (L,) + q in leaves:
                    out.append((L,) + q)
                    continue
                c = (L - 1, q[0] // 2, q[1] // 2, q[2] // 2)
                if L > 1 and c in leaves:
                    out.append(c)
                    continue
                # finer: the children of q on the side facing blk
                for k in range(1 << nd):
                    s = (k & 1, k >> 1 & 1, k >> 2 & 1)
                    if any(d != 0 and s[a] != (1 if d < 0 else 0)
                           for a, d in enumerate((dx, dy, dz))):
                        continue
                    f = (L + 1, 2 * q[0] + s[0], 2 * q[1] + s[1], 2 * q[2] + s[2])
                    if f in leaves:
                        out.append(f)
    return out


def update_leaves(p: Params, meta: np.ndarray, crit, minlev: int, maxlev: int) -> np.ndarray:
    """update_mesh's choice of the next leaf set (mesh.F90:756-833) from the
    refinement criterion crit[b] in {-1, 0, 1} of every leaf of meta:

    - the level limits (check_data_block_refinement, mesh.F90:1545-1550);
    - update_neighbor_refinement (1694-1748), finest level first, with the
      rules of reset_neighbors_refinement (blocks.F90:4300-4350): a leaf to
      be refined keeps its neighbours from derefining and refines its
      coarser neighbours, a kept leaf keeps its coarser neighbours, a leaf
      to be derefined is kept when a neighbour will be refined;
    - prepare_sibling_derefinement (1768-2055): a parent is derefined only
      when all its children are leaves that ask for it, else they are kept;
    - derefine_selected_blocks then refine_selected_blocks (2069-2257).

    The result of these rules does not depend on the order in which the
    leaves of one level are visited (a flag raised to 1 during a level's
    pass belongs to a coarser leaf, whose own pass then resets its
    neighbours), so the blocks are visited in meta order.  Returns the new
    meta: kept leaves in their old order, then the restored parents, then
    the children of refined leaves."""
    nd, nc = p.ndims, p.ncells
    order = [(int(m[0]), int(m[1]) // nc, int(m[2]) // nc, int(m[3]) // nc if nd == 3 else 0)
             for m in np.asarray(meta)]
    leaves = {b: i for i, b in enumerate(order)}
    flag = {}
    for b, c in zip(order, crit):
        r = int(c)
        if b[0] < minlev:
            r = 1
        if b[0] == minlev:
            r = max(0, r)
        if b[0] == maxlev:
            r = min(0, r)
        if b[0] > maxlev:
            r = -1
        flag[b] = r
    top = max(maxlev, max(b[0] for b in order))
    nbrs = {b: _touching(p, leaves, b) for b in order}
    for lev in range(top, 0, -1):
        for b in order:
            if b[0] != lev:
                continue
            for n in nbrs[b]:
                if flag[b] == 1:
                    flag[n] = max(0, flag[n])
                    if n[0] < b[0]:
                        flag[n] = 1
                elif flag[b] == 0:
                    if n[0] < b[0]:
                        flag[n] = max(0, flag[n])
                else:
                    if n[0] < b[0]:
                        flag[n] = max(0, flag[n])
                    if flag[n] == 1:
                        flag[b] = 0
    parents = []
    for lev in range(2, top + 1):
        for b in order:
            if b[0] != lev or flag[b] != -1:
                continue
            par = (lev - 1, b[1] // 2, b[2] // 2, b[3] // 2)
            kids = [(lev, 2 * par[1] + (k & 1), 2 * par[2] + (k >> 1 & 1),
                     2 * par[3] + (k >> 2 & 1)) for k in range(1 << nd)]
            if all(k in flag and flag[k] == -1 for k in kids):
                if par not in parents:
                    parents.append(par)
            else:
                for k in kids:
                    if k in flag:
                        flag[k] = max(0, flag[k])
    gone = set()
    for par in parents:
        gone.update((par[0] + 1, 2 * par[1] + (k & 1), 2 * par[2] + (k >> 1 & 1),
                     2 * par[3] + (k >> 2 & 1)) for k in range(1 << nd))
    kept = [b for b in order if b not in gone and flag[b] != 1]
    born = []
    for lev in range(1, top):
        for b in order:
            if b[0] == lev and flag[b] == 1 and b not in gone:
                born += [(lev + 1, 2 * b[1] + (k & 1), 2 * b[2] + (k >> 1 & 1),
                          2 * b[3] + (k >> 2 & 1)) for k in range(1 << nd)]
    out = kept + parents + born
    return np.array([(L, x * nc, y * nc, z * nc if nd == 3 else 0) for L, x, y, z in out],
                    dtype=np.int32)


def cell_centres(p: Params, meta: np.ndarray):
    """x, y, z cell centres of every block: arrays (nblk, [nc,] nc, nc)."""
    n = p.ncells
    lens = (p.xmax - p.xmin, p.ymax - p.ymin, p.zmax - p.zmin)
    lows = (p.xmin, p.ymin, p.zmin)
    base = (p.xblocks, p.yblocks, p.zblocks)
    out = []
    ar = np.arange(n) + 0.5
    for d in range(3):
        if d >= p.ndims:
            out.append(np.zeros((meta.shape[0],) + ((n,) * p.ndims)))
            continue
        h = np.array([lens[d] / (base[d] * n * (1 << (L - 1))) for L in meta[:, 0]])
        c = lows[d] + (meta[:, 1 + d][:, None] + ar[None, :]) * h[:, None]
        shape = [meta.shape[0]] + [1] * p.ndims
        shape[p.ndims - d] = n                     # (b, [k,] j, i) axes
        out.append(np.broadcast_to(c.reshape(shape), (meta.shape[0],) + (n,) * p.ndims))
    return out


def block_state(p: Params, meta: np.ndarray, prim) -> np.ndarray:
    """Conserved state of every block from prim(x, y, z) -> q[nv, ...] at the
    cell centres; returned as (nblk, [nc,] nc, nc, nv)."""
    x, y, z = cell_centres(p, meta)
    q = prim(x, y, z)
    u = prim_to_cons(p, q)
    return np.ascontiguousarray(np.moveaxis(u, 0, -1))


def blast_prim(p: Params, seed: int = 20261202, r0: float = 0.15, pin: float = 10.0,
               pout: float = 0.1, b0=(0.5, 0.3, 0.2), shift=None):
    """Blast with seeded density and velocity noise: p = pin inside r0 of
    the domain centre (moved by shift, if given), pout outside; uniform B =
    b0 for MHD.

    The state avoids the exact zeros at which the reference's HLLD switches
    branches discontinuously: all three field components are non-zero (at
    By = Bz = 0 it takes the zero-transverse-field branches,
    schemes.F90:4140-4189), and the gas moves (a gas at rest under uniform
    pressure and field keeps velocities of pure round-off, and the contact
    speed sm of the solver is then exactly zero or not depending on that
    round-off; at sm == 0 the reference returns the HLL flux, whose mass flux
    does not join the sm != 0 branches continuously)."""
    from .problems import uniform

    def prim(x, y, z):
        cx = 0.5 * (p.xmin + p.xmax)
        cy = 0.5 * (p.ymin + p.ymax)
        cz = 0.5 * (p.zmin + p.zmax) if p.ndims == 3 else 0.0
        if shift is not None:
            cx, cy, cz = cx + shift[0], cy + shift[1], cz + shift[2]
        r = np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2)
        q = np.zeros((p.nv,) + x.shape)
        xi = uniform(seed, 4 * x.size, stream=21).reshape((4,) + x.shape)
        q[0] = 1.0 + 0.01 * (2.0 * xi[0] - 1.0)
        q[1:4] = 0.01 * (2.0 * xi[1:4] - 1.0)
        if p.nv == 9:
            q[4], q[5], q[6] = b0
            q[8] = np.where(r < r0, pin, pout)
        else:
            q[4] = np.where(r < r0, pin, pout)
        return q
    return prim


def kh_prim(p: Params, seed: int = 20261202, width: float = 0.02):
    """MHD Kelvin-Helmholtz shear layer (vx = 0.5 tanh(y / width), seeded vy
    noise 1e-3, rho 1 above / 2 below, p = 2.5, Bx = 0.1)."""
    from .problems import uniform

    def prim(x, y, z):
        q = np.zeros((p.nv,) + x.shape)
        xi = uniform(seed, x.size, stream=22).reshape(x.shape)
        q[0] = np.where(y > 0.0, 1.0, 2.0)
        q[1] = 0.5 * np.tanh(y / width)
        q[2] = 1.0e-3 * (2.0 * xi - 1.0)
        if p.nv == 9:
            q[4] = 0.1
            q[8] = 2.5
        else:
            q[4] = 2.5
        return q
    return prim


def near_sphere(p: Params, r0: float, width: float):
    """Refinement indicator: the block touches the shell |r - r0| < width
    around the domain centre."""
    def wants(L, bx, by, bz):
        n = p.ncells
        f = 1 << (L - 1)
        lens = (p.xmax - p.xmin, p.ymax - p.ymin, p.zmax - p.zmin)
        lo = []
        hi = []
        for d, (b, base) in enumerate(zip((bx, by, bz), (p.xblocks, p.yblocks, p.zblocks))):
            if d >= p.ndims:
                lo.append(0.0)
                hi.append(0.0)
                continue
            h = lens[d] / (base * f)
            low = (p.xmin, p.ymin, p.zmin)[d]
            lo.append(low + b * h)
            hi.append(low + (b + 1) * h)
        c = (0.5 * (p.xmin + p.xmax), 0.5 * (p.ymin + p.ymax),
             0.5 * (p.zmin + p.zmax) if p.ndims == 3 else 0.0)
        near = np.sqrt(sum(max(lo[d] - c[d], 0.0, c[d] - hi[d]) ** 2 for d in range(3)))
        far = np.sqrt(sum(max(abs(lo[d] - c[d]), abs(hi[d] - c[d])) ** 2 for d in range(3)))
        return near < r0 + width and far > r0 - width
    return wants


def near_plane(p: Params, y0: float, width: float):
    """Refinement indicator: the block touches the slab |y - y0| < width."""
    def wants(L, bx, by, bz):
        h = (p.ymax - p.ymin) / (p.yblocks * (1 << (L - 1)))
        lo = p.ymin + by * h
        return lo - width < y0 < lo + h + width
    return wants


def c3_amr_params(precision: str = "fp64", nb: int = 8, r: float = 0.3):
    """The 3D block-AMR run of the batched/AMR configuration: three levels of
    16^3 blocks (+2 ghosts) on a periodic nb^3 base, refined towards a sphere
    of radius r (6224 leaves: 280 / 1272 / 4672 on levels 1 / 2 / 3 for
    nb = 8, r = 0.3), MHD, TVD-MC + HLLD + SSPRK(3,3), CFL 0.4.  Returns
    (Params, meta); the state is the blast of blast_prim(p, r0=0.2)."""
    p = Params(ndims=3, equation_system="mhd", reconstruction="tvd", limiter="mc",
               riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.4,
               ncells=16, nghosts=2, xblocks=nb, yblocks=nb, zblocks=nb, xmin=-0.5, xmax=0.5,
               ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5, precision=precision)
    meta = refine(p, near_sphere(p, r, 0.015), 3)
    return p, meta
