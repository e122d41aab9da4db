"""GPU parity: the HIP path through the C-ABI against the C oracle (which is
itself pinned bit-exactly to the reference binaries, see test_oracle_*).

Tolerance (FP64 path): <= 1e-12 relative per step, where "relative" is the
maximum absolute difference of a variable divided by the maximum magnitude of
that variable over the grid (the GPU contracts a*b+c into FMAs, so results
are not bit-identical).  Multi-step runs are checked at a looser bound that
allows round-off growth through limiter switches.
"""
import numpy as np
import pytest

from amun_amd.config import Params, c1_params, c2_params
from amun_amd.problems import initial_state, prim_to_cons, cell_centres
from amun_amd.solver import Solver
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def relerr(a, b):
    """Per-variable max |a-b| over the max magnitude of that variable; vector
    components (momentum 1..3, magnetic field 4..6) share the magnitude of
    their vector, so a component that is zero in exact arithmetic is judged
    against its vector's scale."""
    ax = tuple(range(1, b.ndim))
    mag = np.abs(b).max(axis=ax)
    scale = mag.copy()
    scale[1:4] = mag[1:4].max()
    if b.shape[0] == 9:
        scale[4:7] = mag[4:7].max()
    scale = np.maximum(scale, 1e-300)
    return np.abs(a - b).max(axis=ax) / scale


def one_step(p, u0):
    print(f"[case] {p.ndims}D {p.equation_system} {p.reconstruction}/{p.limiter} "
          f"{p.riemann_solver} {p.time_advance} shape={p.shape}", flush=True)
    with Solver(p) as s:
        s.upload(u0)
        dt, cmax = s.advance()
        ug = s.download()
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(1)
    uo = o.get_state()
    return dt, cmax, ug, tr[0], uo


def test_c1_one_step():
    p = c1_params()
    u0 = initial_state(p)
    dt, cmax, ug, tr, uo = one_step(p, u0)
    assert abs(dt - tr[0]) <= 1e-13 * tr[0]
    assert abs(cmax - tr[1]) <= 1e-13 * tr[1]
    assert relerr(ug, uo).max() <= 1e-12


def test_c2_one_step_32():
    p = c2_params(32)
    u0 = initial_state(p)
    dt, cmax, ug, tr, uo = one_step(p, u0)
    assert abs(dt - tr[0]) <= 1e-13 * tr[0]
    assert relerr(ug, uo).max() <= 1e-12


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("eq,rs", [("hd", "hll"), ("hd", "hllc"), ("mhd", "hll"), ("mhd", "hlld")])
@pytest.mark.parametrize("rec,lim", [("tvd", "mc"), ("tvd", "mm"), ("weno5z", "mm")])
@pytest.mark.parametrize("ta", ["euler", "rk2", "rk3"])
def test_matrix_one_step(nd, eq, rs, rec, lim, ta):
    ng = 4 if rec == "weno5z" else 2
    p = Params(ndims=nd, equation_system=eq, riemann_solver=rs, reconstruction=rec,
               limiter=lim, time_advance=ta, ncells=8, nghosts=ng, xblocks=2,
               yblocks=3, zblocks=2, cfl=0.3, gamma=1.4 if eq == "hd" else 5.0 / 3.0,
               xmin=-0.5, xmax=0.5, ymin=-0.75, ymax=0.75, zmin=0.0, zmax=1.0)
    u0 = initial_state(p) if eq == "hd" else _mhd_state(p)
    dt, cmax, ug, tr, uo = one_step(p, u0)
    assert abs(dt - tr[0]) <= 1e-13 * tr[0]
    assert relerr(ug, uo).max() <= 1e-12


def _mhd_state(p):
    nz, ny, nx = p.shape
    z, y, x = cell_centres(p)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    q = np.zeros((9, nz, ny, nx))
    rng = np.random.default_rng(7)
    q[0] = 1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.05 * rng.random(X.shape)
    q[1] = 0.5 * np.sin(2 * np.pi * Y)
    q[2] = -0.5 * np.sin(2 * np.pi * X)
    q[3] = 0.1 * np.cos(2 * np.pi * (X + Z))
    q[4] = 0.3 + 0.1 * np.cos(2 * np.pi * Y)
    q[5] = 0.2 * np.sin(4 * np.pi * X)
    q[6] = 0.1 * np.cos(2 * np.pi * Z)
    q[7] = 0.01 * rng.standard_normal(X.shape)
    q[8] = 0.6 + 0.4 * (X * X + Y * Y < 0.05)
    return prim_to_cons(p, q)


def test_c1_multistep():
    """200 steps of the 128^2 blast: dt sequence and final state stay within
    round-off growth of the oracle."""
    p = c1_params()
    u0 = initial_state(p)
    with Solver(p) as s:
        s.upload(u0)
        dts = s.run(200)
        ug = s.download()
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(200)
    uo = o.get_state()
    assert np.abs(dts - tr[:, 0]).max() <= 1e-10 * tr[:, 0].max()
    assert relerr(ug, uo).max() <= 1e-8
