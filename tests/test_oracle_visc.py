"""The oracle's viscous and resistive source terms (update_sources,
sources.F90:340-598) and their time-step limit (new_time_step,
evolution.F90:478-483) against the reference's own runs
(tests/golden/make_golden.py visc), bit-exact; and against closed forms the
central differences reproduce exactly: for a velocity (field) that is
quadratic in the coordinates, the momentum (field) source of one Euler step
at a fixed dt is nu rho (lap v + grad div v / 3) (eta lap B)."""
import hashlib
import os
import sys

import numpy as np
import pytest

from amun_amd.config import Params
from amun_amd.problems import cell_centres, prim_to_cons
import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_golden import VISC_RUNS, visc_state  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


@pytest.mark.parametrize("name", [r[0] for r in VISC_RUNS])
def test_viscous_resistive_runs_bit_exact(name):
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    p = Params.from_text(str(g["params"]), ndims=int(g["ndims"]))
    assert p.viscosity > 0.0 or p.resistivity > 0.0
    u0 = visc_state(name, p)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


def _source_of_one_step(p, q, dt=1e-6):
    """(u(nu, eta) - u(0, 0)) / dt after one Euler step of size dt"""
    u0 = prim_to_cons(p, q)
    out = []
    for nu, eta in ((p.viscosity, p.resistivity), (0.0, 0.0)):
        pp = Params(**{k: v for k, v in p.__dict__.items() if k != "extra"})
        pp.viscosity, pp.resistivity = nu, eta
        o = ol.OracleRun(pp)
        o.set_state(u0, dtnext=dt)
        tr = o.run(1, dtnext=dt)
        assert tr[0, 0] == dt
        out.append(o.get_state())
    return (out[0] - out[1]) / dt


def test_viscous_source_closed_form():
    """vx = a y^2 + b x y, vy = c x^2, rho = 2, p uniform: the momentum
    source is nu rho (lap v + grad div v / 3) = nu rho (2 a, 2 c + b / 3),
    checked on the cells away from the periodic seam."""
    p = Params(ndims=2, equation_system="hd", reconstruction="tvd", limiter="zero",
               riemann_solver="hll", time_advance="euler", gamma=1.4, cfl=0.4, ncells=16,
               nghosts=2, xblocks=2, yblocks=2, viscosity=0.05)
    z, y, x = cell_centres(p)
    Y, X = np.meshgrid(y, x, indexing="ij")
    a, b, c = 0.3, 0.2, -0.4
    q = np.zeros((5, 1) + Y.shape)
    q[0] = 2.0
    q[1, 0] = a * Y ** 2 + b * X * Y
    q[2, 0] = c * X ** 2
    q[4] = 1.0
    s = _source_of_one_step(p, q)
    # nu rho (lap v + grad div v / 3) with div v = b y: x: nu rho 2 a,
    # y: nu rho (2 c + b / 3)
    nu, rho = p.viscosity, 2.0
    ex = nu * rho * (2 * a)
    ey = nu * rho * (2 * c + b / 3.0)
    inner = (slice(None), slice(3, -3), slice(3, -3))
    # the hyperbolic fluxes cancel between the two runs only up to the
    # change of the state; with dt = 1e-6 the difference is O(dt)
    np.testing.assert_allclose(s[1][inner], ex, rtol=0, atol=1e-5)
    np.testing.assert_allclose(s[2][inner], ey, rtol=0, atol=1e-5)


def test_resistive_source_closed_form():
    """Bx = a y^2, By = c x^2, Bz = d (x^2 + y^2): the field source is eta
    lap B = eta (2 a, 2 c, 4 d) on the interior cells."""
    p = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="zero",
               riemann_solver="hll", time_advance="euler", gamma=5.0 / 3.0, cfl=0.4,
               ncells=16, nghosts=2, xblocks=2, yblocks=2, resistivity=0.02)
    z, y, x = cell_centres(p)
    Y, X = np.meshgrid(y, x, indexing="ij")
    a, c, d = 0.5, -0.3, 0.2
    q = np.zeros((9, 1) + Y.shape)
    q[0] = 1.0
    q[4, 0] = a * Y ** 2
    q[5, 0] = c * X ** 2
    q[6, 0] = d * (X ** 2 + Y ** 2)
    q[8] = 1.0
    s = _source_of_one_step(p, q)
    eta = p.resistivity
    inner = (slice(None), slice(3, -3), slice(3, -3))
    np.testing.assert_allclose(s[4][inner], 2 * a * eta, rtol=0, atol=1e-5)
    np.testing.assert_allclose(s[5][inner], 2 * c * eta, rtol=0, atol=1e-5)
    np.testing.assert_allclose(s[6][inner], 4 * d * eta, rtol=0, atol=1e-5)


def test_viscous_time_step_limit():
    """dt = cfl dx_min / (cmax + 2 max(nu, eta) / dx_min) (evolution.F90:478-483)"""
    p = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
               riemann_solver="hlld", time_advance="rk2", gamma=5.0 / 3.0, cfl=0.4, ncells=8,
               nghosts=2, xblocks=2, yblocks=2, viscosity=0.01, resistivity=0.03)
    q = np.zeros((9, 1, 16, 16))
    q[0], q[8], q[4] = 1.0, 1.0, 0.5
    u0 = prim_to_cons(p, q)
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(1)
    dx = 1.0 / 16
    cmax = np.sqrt((5.0 / 3.0 * 1.0 + 0.25) / 1.0)      # fast speed at rest, |B| = 0.5
    assert tr[0, 0] == pytest.approx(0.4 * dx / (cmax + 2 * 0.03 / dx), rel=1e-14)
