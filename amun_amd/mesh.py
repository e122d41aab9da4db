"""Host-side multi-level block meshes for amun_dev_init_amr.

A mesh is an (nblk, 4) int array of leaf blocks: level (1 = base) and the
first interior cell (i0, j0, k0) in cells of that level, on the periodic base
grid of the parameters (xblocks x yblocks [x zblocks] blocks of ncells^d
cells).  ``refine`` builds one the way AMUN's generate_mesh does
(mesh.F90:431-740): level by level, leaves whose indicator asks for it are
split into 2^d children, and neighbours are refined further until every
leaf's face, edge and corner neighbours differ from it by one level at most
(the 2:1 balance the ghost-cell exchange relies on).  The mesh stays on the
host; the device path takes it frozen.
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
               pout: float = 0.1, b0=(0.5, 0.3, 0.2)):
    """Blast with seeded density and velocity noise: p = pin inside r0 of
    the domain centre, pout outside; uniform B = b0 for MHD.

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
