"""The oracle's gravity of the Rayleigh-Taylor problem (update_sources,
sources.F90:288-335; gravity.F90:141-150, 283-330) and the relativistic
systems' four-velocity reconstruction (state_variables = 4vec,
schemes.F90:4481-4517, 5040-5076) against the reference's own runs
(tests/golden/make_golden.py grav4v), bit-exact; and two checks that do not
rest on the reference: without the problem name the same runs differ (the
terms act), and a gas at rest in a uniform field gains momentum rho g dt per
Euler step."""
import hashlib
import os
import sys

import numpy as np
import pytest

from amun_amd.config import Params
from amun_amd.problems import prim_to_cons
import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_golden import GRAV_RUNS, V4_RUNS, grav_params, grav_state, v4_params, v4_state  # noqa: E402,E501


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def case(name):
    for n, nd, eq, rec, st in GRAV_RUNS:
        if n == name:
            p = grav_params(nd, eq, rec)
            return p, grav_state(p)
    for n, nd, eq, rec, rs, st in V4_RUNS:
        if n == name:
            p = v4_params(nd, eq, rec, rs)
            return p, v4_state(n, p)
    raise KeyError(name)


NAMES = [r[0] for r in GRAV_RUNS] + [r[0] for r in V4_RUNS]


@pytest.mark.parametrize("name", NAMES)
def test_grav4v_runs_bit_exact(name):
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    p, u0 = case(name)
    assert sha(u0) == str(g["sha_u0"])
    assert str(g["params"]) == p.to_params_in()
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


@pytest.mark.parametrize("name", ["grav_hd2d_tvd_rt_10", "v4_srhd2d_tvd_hllc_blast_20"])
def test_terms_act(name):
    """The same run without gravity (or with primitive states) ends
    elsewhere."""
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    p, u0 = case(name)
    if "problem" in p.extra:
        del p.extra["problem"]
    else:
        p.state_variables = "primitive"
    o = ol.OracleRun(p)
    o.set_state(u0)
    o.run(int(g["nsteps"]))
    assert sha(o.get_state()) != str(g["sha_u"])


def test_gravity_closed_form():
    """A uniform gas at rest: one Euler step of size dt adds rho g dt to
    the y momentum and nothing else (the fluxes of a uniform state cancel)."""
    p = Params(ndims=2, equation_system="hd", reconstruction="tvd", limiter="mc",
               riemann_solver="hllc", time_advance="euler", gamma=1.4, cfl=0.4, ncells=8,
               nghosts=2, xblocks=2, yblocks=2)
    p.extra.update(problem="rt", gacc="-0.7")
    q = np.zeros((5, 1, 16, 16))
    q[0], q[4] = 1.3, 1.0
    u0 = prim_to_cons(p, q)
    o = ol.OracleRun(p)
    o.set_state(u0, dtnext=1e-3)
    o.run(1, dtnext=1e-3)
    du = o.get_state() - u0
    np.testing.assert_allclose(du[2], 1.3 * -0.7 * 1e-3, rtol=1e-13)
    np.testing.assert_array_equal(du[[0, 1, 3, 4]], 0.0)
