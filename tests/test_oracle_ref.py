"""The C oracle against the live reference binaries on fresh random inputs
(skipped where oracle/_ref is not built, e.g. on hosts without the
reference sources)."""
import numpy as np
import pytest

from amun_amd.config import Params, c1_params, c2_params
from amun_amd import problems
import oracle_lib as ol

needs_ref = pytest.mark.skipif(not (ol.ref_available(2) and ol.ref_available(3)),
                               reason="reference binaries (oracle/_ref) not built")


@needs_ref
@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("ta", ["euler", "rk2", "rk3"])
@pytest.mark.parametrize("lim", ["mm", "mc", "sb", "vl", "va"])
def test_tvd_limiters_2d_mhd(seed, ta, lim):
    p = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter=lim,
               riemann_solver="hlld", time_advance=ta, gamma=5.0 / 3.0, cfl=0.4,
               ncells=8, nghosts=2, xblocks=3, yblocks=2)
    u0 = problems.orszag_tang_2d(p, seed=seed)
    ur, tr = ol.ref_run(p, u0, 3)
    o = ol.OracleRun(p)
    o.set_state(u0)
    to = o.run(3)
    np.testing.assert_array_equal(to, tr)
    np.testing.assert_array_equal(o.get_state(), ur)


@needs_ref
@pytest.mark.parametrize("rs", ["hll", "hlld"])
def test_weno_3d_mhd(rs):
    p = c2_params(32)
    p.riemann_solver = rs
    p.xblocks, p.yblocks, p.zblocks, p.ncells = 3, 2, 2, 8
    u0 = problems.turbulent_mhd_3d(p, seed=11)
    ur, tr = ol.ref_run(p, u0, 2)
    o = ol.OracleRun(p)
    o.set_state(u0)
    to = o.run(2)
    np.testing.assert_array_equal(to, tr)
    np.testing.assert_array_equal(o.get_state(), ur)


@needs_ref
def test_kh_hd_hllc_rk2():
    p = c1_params()
    p.xblocks = p.yblocks = 3
    u0 = problems.blast_hd_2d(p, seed=5)
    ur, tr = ol.ref_run(p, u0, 20)
    o = ol.OracleRun(p)
    o.set_state(u0)
    to = o.run(20)
    np.testing.assert_array_equal(to, tr)
    np.testing.assert_array_equal(o.get_state(), ur)
