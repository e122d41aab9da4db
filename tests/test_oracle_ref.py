"""The C oracle against the live reference binaries on fresh random inputs
(skipped where oracle/_ref is not built, e.g. on hosts without the
reference sources)."""
import numpy as np
import pytest

from amun_amd.config import Params, c1_params, c2_params, c4_params
from amun_amd import problems
import oracle_lib as ol

# each test skips on its own when the binary of its dimension is missing
# (oracle_lib._ref_exec); this mark only skips the module when neither is
needs_ref = pytest.mark.skipif(not (ol.ref_available(2) or ol.ref_available(3)),
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


@needs_ref
@pytest.mark.parametrize("ta,lim", [("rk2", "mm"), ("rk3", "mc"), ("euler", "vl")])
def test_srmhd_blast(ta, lim):
    p = c4_params(48)
    p.time_advance, p.limiter = ta, lim
    u0 = problems.srmhd_blast_2d(p, seed=3, noise=0.05)
    ur, tr = ol.ref_run(p, u0, 12)
    o = ol.OracleRun(p)
    o.set_state(u0)
    to = o.run(12)
    np.testing.assert_array_equal(to, tr)
    np.testing.assert_array_equal(o.get_state(), ur)


@needs_ref
def test_srmhd_eqs_and_roots_fresh_inputs():
    rng = np.random.default_rng(99)
    p = Params(ndims=2, equation_system="srmhd", riemann_solver="hll")
    q = np.zeros((500, 9))
    q[:, 0] = 10.0 ** rng.uniform(-2, 1, 500)
    d = rng.normal(size=(500, 3))
    q[:, 1:4] = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0, 0.999, (500, 1))
    q[:, 4:8] = rng.normal(size=(500, 4))
    q[:, 8] = 10.0 ** rng.uniform(-3, 1, 500)
    for a, b in zip(ol.oracle_eqs(p, q, 0.9), ol.ref_eqs(p, q, 0.9)):
        np.testing.assert_array_equal(a, b)
    co = rng.normal(size=(500, 5)) * 10.0 ** rng.integers(-4, 5, (500, 5))
    np.testing.assert_array_equal(ol.oracle_algebra(co), ol.ref_algebra(co))


@needs_ref
@pytest.mark.parametrize("sgp,kappa", [(9.0, 1.0), (3.0, 0.7), (1.5, 2.0)])
def test_gp_fresh_inputs(sgp, kappa):
    """GP weights for other kernel widths and MP limiter constants, on fresh
    pencils and a short 2D run."""
    rng = np.random.default_rng(int(sgp * 10))
    p = Params(ndims=2, equation_system="mhd", reconstruction="gp", limiter="mm",
               riemann_solver="hlld", time_advance="rk3", ncells=8, nghosts=3, xblocks=3,
               yblocks=2, cfl=0.3, sgp=sgp, kappa=kappa, kbeta=0.8)
    fp = np.cumsum(rng.normal(size=(20, 14)), axis=1)
    a, b = ol.oracle_reconstruct(p, fp), ol.ref_reconstruct(p, fp)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    u0 = problems.orszag_tang_2d(p, seed=4)
    ur, tr = ol.ref_run(p, u0, 3)
    o = ol.OracleRun(p)
    o.set_state(u0)
    np.testing.assert_array_equal(o.run(3), tr)
    np.testing.assert_array_equal(o.get_state(), ur)


@needs_ref
@pytest.mark.parametrize("ta,stages", [("ssprk(m,2)", 2), ("ssprk(m,2)", 3), ("ssprk(m,2)", 5),
                                       ("ssprk(m,2)", 12), ("ssprk(4,3)", 2), ("ssprk(5,3)", 2)])
@pytest.mark.parametrize("eq", ["mhd", "hd"])
def test_ssprk_integrators(ta, stages, eq):
    """SSPRK(m,2) (m clamped to 2..9), SSPRK(4,3) and SSPRK(5,3): stage
    combinations, CFL scaling and the psi decay at the scaled CFL."""
    if eq == "mhd":
        p = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
                   riemann_solver="hlld", time_advance=ta, gamma=5.0 / 3.0, cfl=0.3,
                   ncells=8, nghosts=2, xblocks=3, yblocks=2, stages=stages)
        u0 = problems.orszag_tang_2d(p, seed=4)
    else:
        p = Params(ndims=2, equation_system="hd", reconstruction="weno5z", limiter="mm",
                   riemann_solver="hllc", time_advance=ta, gamma=1.4, cfl=0.05,
                   ncells=8, nghosts=4, xblocks=2, yblocks=2, stages=stages)
        u0 = problems.blast_hd_2d(p)
    ur, tr = ol.ref_run(p, u0, 4)
    assert np.isfinite(ur).all()
    o = ol.OracleRun(p)
    o.set_state(u0)
    to = o.run(4)
    np.testing.assert_array_equal(to, tr)
    np.testing.assert_array_equal(o.get_state(), ur)


# ------------------------------------------------------- multi-level meshes
def _amr_params(nd, eq, rec, lim, ta, plim="mm", maxlev=3, base=2, problem="blast"):
    p = Params(ndims=nd, equation_system=eq, reconstruction=rec, limiter=lim,
               riemann_solver="hlld" if eq == "mhd" else "hllc", time_advance=ta,
               gamma=5.0 / 3.0 if eq == "mhd" else 1.4, cfl=0.4, ncells=8,
               nghosts=4 if rec == "weno5z" else 2, xblocks=base, yblocks=base, zblocks=base,
               xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5,
               prolongation_limiter=plim)
    if rec == "weno5z":
        p.ncells = 8
    p.extra.update({"problem": problem, "maxlev": maxlev, "radius": 0.15, "ratio": 100.0,
                    "refinement_variables": "dens pres", "crefmin": 0.1, "crefmax": 0.3})
    if eq == "mhd":
        p.extra["buni"] = 1.0
    return p


@needs_ref
@pytest.mark.parametrize("nd,eq,rec,lim,ta,plim,maxlev", [
    (2, "mhd", "tvd", "mc", "rk3", "mm", 3),
    (2, "hd", "tvd", "vl", "rk2", "mc", 3),
    (2, "hd", "weno5z", "mm", "rk3", "sb", 3),
    (2, "mhd", "tvd", "mm", "euler", "vl", 4),
    (2, "mhd", "tvd", "mc", "ssprk(4,3)", "zero", 3),
    (3, "mhd", "tvd", "mc", "rk3", "mm", 3),
    (3, "hd", "tvd", "mc", "rk2", "mc", 3),
    (2, "mhd", "limo3", "mm", "rk3", "mm", 3),
    (2, "hd", "weno3", "mm", "rk2", "mc", 3),
])
def test_amr_ghost_fill_and_run(nd, eq, rec, lim, ta, plim, maxlev):
    """The reference's own refined mesh (generate_mesh on a blast, frozen):
    every block's ghost layers after the initial boundary_variables, the dt
    trace and the block states after three steps, bit for bit."""
    p = _amr_params(nd, eq, rec, lim, ta, plim, maxlev)
    m0, u0, m1, u1, tr, q0 = ol.ref_amr(p, 3, qdump=True)
    assert len(np.unique(m0[:, 0])) >= 2, "the mesh should hold two levels at least"
    assert np.array_equal(m0, m1)
    o = ol.OracleAmr(p, m0)
    o.set_blocks(u0)
    uo, qo = o.get_blocks(with_q=True)
    np.testing.assert_array_equal(qo, q0)
    to = o.run(3)
    np.testing.assert_array_equal(to, tr)
    np.testing.assert_array_equal(o.get_blocks(), u1)


# ------------------------------------------------------ more 1D schemes
@needs_ref
@pytest.mark.parametrize("rec", ["weno5yc", "weno5ns", "mp5", "mp7", "weno3", "limo3", "ppm"])
def test_more_schemes_fresh_pencils(rec):
    rng = np.random.default_rng(17)
    fp = np.concatenate([rng.normal(size=(6, 24)), np.cumsum(rng.normal(size=(4, 24)), axis=1),
                         np.sin(np.linspace(0, 6, 24))[None, :] * np.ones((2, 1)),
                         np.where(np.arange(24) < 12, 1.0, 0.1)[None, :] * np.ones((2, 1))])
    for lim in ("mm", "mc"):
        p = Params(ndims=3, equation_system="mhd", reconstruction=rec, limiter=lim, ncells=16,
                   nghosts=4, cfl=0.3)
        for h in (1.0, 0.03125, 0.7):      # WENO3 and LimO3 weigh by the cell size
            fl, fr = ol.oracle_reconstruct(p, fp, h)
            rl, rr = ol.ref_reconstruct(p, fp, h)
            np.testing.assert_array_equal(fl, rl)
            np.testing.assert_array_equal(fr, rr)


@needs_ref
@pytest.mark.parametrize("rec", ["weno5yc", "weno5ns", "mp5", "mp7", "weno3", "limo3", "ppm"])
@pytest.mark.parametrize("nd,eq", [(2, "mhd"), (2, "hd"), (3, "mhd")])
def test_more_schemes_runs(rec, nd, eq):
    p = Params(ndims=nd, equation_system=eq, reconstruction=rec, limiter="mm",
               riemann_solver="hlld" if eq == "mhd" else "hllc", time_advance="rk3",
               gamma=5.0 / 3.0 if eq == "mhd" else 1.4, cfl=0.3, ncells=8,
               nghosts=2 if rec in ("weno3", "limo3") else 4,
               xblocks=3, yblocks=2, zblocks=2)
    if eq == "hd":
        u0 = problems.blast_hd_2d(p)
    elif nd == 2:
        u0 = problems.orszag_tang_2d(p, seed=5)
    else:
        u0 = problems.turbulent_mhd_3d(p, seed=5)
    ur, tr = ol.ref_run(p, u0, 3)
    o = ol.OracleRun(p)
    o.set_state(u0)
    to = o.run(3)
    np.testing.assert_array_equal(to, tr)
    np.testing.assert_array_equal(o.get_state(), ur)


# ------------------------------------------------------------ HD Roe
@needs_ref
@pytest.mark.parametrize("nd,rec,ta", [(2, "tvd", "rk2"), (2, "weno5z", "ssprk(4,3)"),
                                       (3, "tvd", "rk3")])
def test_roe_hd_runs(nd, rec, ta):
    p = Params(ndims=nd, equation_system="hd", reconstruction=rec, limiter="vl",
               riemann_solver="roe", time_advance=ta, gamma=1.4, cfl=0.3, ncells=8,
               nghosts=4 if rec == "weno5z" else 2, xblocks=3, yblocks=2, zblocks=2)
    u0 = problems.blast_hd_2d(p, seed=13)
    ur, tr = ol.ref_run(p, u0, 4)
    o = ol.OracleRun(p)
    o.set_state(u0)
    to = o.run(4)
    np.testing.assert_array_equal(to, tr)
    np.testing.assert_array_equal(o.get_state(), ur)


# ------------------------------------------------- non-periodic boundaries
def _bnd_params(nd, eq, bx, by, bz, rec="tvd", ng=2):
    return Params(ndims=nd, equation_system=eq, reconstruction=rec, limiter="mc",
                  riemann_solver="hlld" if eq == "mhd" else "hllc", time_advance="rk2",
                  gamma=5.0 / 3.0 if eq == "mhd" else 1.4, cfl=0.4, ncells=8, nghosts=ng,
                  xblocks=3, yblocks=2, zblocks=2, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5,
                  zmin=-0.5, zmax=0.5, xlbndry=bx, xubndry=bx, ylbndry=by, yubndry=by,
                  zlbndry=bz, zubndry=bz)


BND_CASES = [(2, "mhd", "reflective", "open", "periodic"), (2, "hd", "outflow", "reflective", "periodic"),
             (3, "mhd", "open", "periodic", "reflective"), (2, "mhd", "outflow", "outflow", "periodic"),
             (3, "hd", "reflective", "outflow", "open")]


def _bnd_state(p, eq, nd):
    if eq == "hd":
        return problems.blast_hd_2d(p, seed=3)
    return problems.orszag_tang_2d(p, seed=3) if nd == 2 else problems.turbulent_mhd_3d(p, seed=3)


@needs_ref
@pytest.mark.parametrize("nd,eq,bx,by,bz", BND_CASES)
def test_boundary_types_ghosts_and_runs(nd, eq, bx, by, bz):
    """boundaries_specific (open, outflow, reflective) on every side type and
    axis mix: every block's ghost layers, edges and corners included, after
    the initial boundary_variables, and 4 steps, bit for bit."""
    import os
    import tempfile
    p = _bnd_params(nd, eq, bx, by, bz)
    u0 = _bnd_state(p, eq, nd)
    im = p.ncells + 2 * p.nghosts
    km = im if nd == 3 else 1
    with tempfile.TemporaryDirectory() as d:
        u0.tofile(os.path.join(d, "in.bin"))
        ol._ref_exec(p, {"harness_mode": "ghosts", "harness_input": "in.bin",
                         "harness_output": "out", "harness_steps": 0}, d)
        raw = np.fromfile(os.path.join(d, "out_q.bin"))
    per = 3 + im * im * km * p.nv
    raw = raw.reshape(-1, per)
    o = ol.OracleRun(p)
    o.set_state(u0)
    qb = np.empty((raw.shape[0], km, im, im, p.nv))
    ol.lib().orc_get_blocks(o.h, ol._dp(np.empty(raw.shape[0] * p.ncells ** nd * p.nv)), ol._dp(qb))
    for row in raw:
        x0, y0, z0 = (int(round(v)) for v in row[:3])
        b = ((z0 // p.ncells if nd == 3 else 0) * p.yblocks + y0 // p.ncells) * p.xblocks + x0 // p.ncells
        np.testing.assert_array_equal(qb[b].ravel(), row[3:])
    ur, tr = ol.ref_run(p, u0, 4)
    to = o.run(4)
    np.testing.assert_array_equal(to, tr)
    np.testing.assert_array_equal(o.get_state(), ur)

