"""The oracle's special-relativistic HD system (equations.F90:3478-4566,
schemes.F90:4433-4968) against the reference's own outputs
(tests/golden/make_golden.py srhd): equation-of-state vectors, the three
primitive solvers on physical, very fast and unphysical states, HLL and HLLC
through update_flux, 2D and 3D runs and the statistics file; bit-exact.
Beyond the reference: prim2cons -> cons2prim returns the state for every
solver, equal states give the exact flux, and the signal speeds stay below
the speed of light."""
import hashlib
import os
import sys

import numpy as np
import pytest

from amun_amd.config import Params
import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_golden import SRHD_RUNS, srhd_params, srhd_state, srhd_states  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def test_srhd_equation_of_state():
    g = load("unit_eqs_srhd")
    p = Params.from_text(str(g["params"]), ndims=2)
    u, q2, f, cm, cp = ol.oracle_eqs(p, g["q"], float(g["cmax"]))
    for a, b in ((u, g["u"]), (q2, g["q2"]), (f, g["f"]), (cm, g["cm"]), (cp, g["cp"])):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("ps", ["1dw", "2dwv", "2dwu"])
def test_srhd_primitive_solvers(ps):
    """Physical, W ~ 70 and unphysical states: the recovered primitives, and
    on failure the kept entries with p = 0, as the reference leaves them."""
    g = load("unit_eqs_srhd")
    p = Params.from_text(str(g["params"]), ndims=2)
    p.primitive_solver = ps
    q, _ = ol.oracle_cons2prim(p, g["ub"], g["q0"])
    np.testing.assert_array_equal(q, g["qb_" + ps])


@pytest.mark.parametrize("rs", ["hll", "hllc"])
def test_srhd_riemann_via_update_flux(rs):
    g = load(f"unit_flux_srhd_{rs}")
    p = Params.from_text(str(g["params"]), ndims=2)
    for b in range(g["q"].shape[0]):
        f = ol.oracle_update_flux(p, g["q"][b], float(g["cmax"]))
        np.testing.assert_array_equal(f, g["f"][b])


@pytest.mark.parametrize("name", [r[0] for r in SRHD_RUNS])
def test_srhd_runs_bit_exact(name):
    g = load(name)
    p = Params.from_text(str(g["params"]), ndims=int(g["ndims"]))
    u0 = srhd_state(name, p)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


def test_srhd_statistics_vs_reference_file():
    g = load("stats_srhd_ref")
    p = srhd_params(2, "tvd", "hllc", "1dw")
    p.xblocks, p.yblocks = 3, 2
    u0 = srhd_state("blast", p, seed=3)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    o.run(3)
    st, ref = np.asarray(o.statistics()), np.asarray(g["stats"])
    tol = 2e-8 * np.abs(ref) + 1e-300
    assert (np.abs(st - ref) <= tol).all(), (st, ref)


@pytest.mark.parametrize("ps", ["1dw", "2dwv", "2dwu"])
def test_srhd_round_trip(ps):
    """cons2prim(prim2cons(q)) = q up to the solver tolerance (1e-12)."""
    rng = np.random.default_rng(5)
    p = Params(ndims=2, equation_system="srhd", riemann_solver="hll", gamma=4.0 / 3.0,
               tolerance=1.0e-12, nr_maxit=50, primitive_solver=ps)
    q = srhd_states(rng, 300)
    q[:, 1:4] *= 0.95
    u, _, _, _, _ = ol.oracle_eqs(p, q, 1.0)
    q2, _ = ol.oracle_cons2prim(p, u, np.zeros_like(u))
    np.testing.assert_allclose(q2, q, rtol=1e-9, atol=1e-10 * np.abs(q).max())


@pytest.mark.parametrize("rs", ["hll", "hllc"])
def test_srhd_consistency_and_causality(rs):
    """Equal states give F(q) (rho W v, m vx + p, ..., m_x - D vx), and the
    wave speeds lie in (-1, 1)."""
    rng = np.random.default_rng(8)
    p = Params(ndims=2, equation_system="srhd", reconstruction="tvd", limiter="zero",
               riemann_solver=rs, gamma=4.0 / 3.0, ncells=8, nghosts=2)
    q = srhd_states(rng, 40)
    fo = ol.oracle_riemann(p, q, q.copy(), cmax=1.0)
    u, _, f, cm, cp = ol.oracle_eqs(p, q, 1.0)
    np.testing.assert_allclose(fo, f, rtol=1e-13, atol=1e-14 * np.abs(f).max())
    assert (cm > -1.0).all() and (cp < 1.0).all() and (cm <= cp).all()
