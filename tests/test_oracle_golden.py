"""The C oracle against the golden vectors written by the reference binaries
(tests/golden/make_golden.py).  On the x86-64 baseline ISA without FMA
contraction both evaluate the same operations in the same order, so the
comparisons are bit-exact (sha256 of the final state, exact dt traces)."""
import hashlib
import os

import numpy as np
import pytest

from amun_amd.config import Params
from amun_amd import problems
import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def params_of(g):
    return Params.from_text(str(g["params"]), ndims=int(g["ndims"]))


RUNS = {
    "c1_blast_hd2d_128_200": problems.blast_hd_2d,
    "c2_mhd3d_weno5z_hlld_32_4": problems.turbulent_mhd_3d,
    "orszag_tang_mhd2d_64_30": problems.orszag_tang_2d,
    "kh_mhd2d_64_30": problems.kelvin_helmholtz_2d,
}


@pytest.mark.parametrize("name", sorted(RUNS))
def test_run_bit_exact(name):
    g = load(name)
    p = params_of(g)
    u0 = RUNS[name](p)
    assert sha(u0) == str(g["sha_u0"]), "seeded initial state changed"
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    u = o.get_state()
    np.testing.assert_array_equal(tr, g["trace"])
    s = int(g["stride"])
    sub = u[:, ::s, ::s, ::s] if p.ndims == 3 else u[:, :, ::s, ::s]
    np.testing.assert_array_equal(sub, g["sub"])
    assert sha(u) == str(g["sha_u"])


def test_hd3d_weno_hllc_bit_exact():
    g = load("blast_hd3d_weno5z_hllc_4")
    p = params_of(g)
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import hd3_blast
    u0 = hd3_blast(p)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


def test_weno5z_pencils():
    g = load("unit_weno5z")
    p = Params(ndims=3, equation_system="mhd", reconstruction="weno5z", limiter="mm")
    fl, fr = ol.oracle_reconstruct(p, g["fp"])
    np.testing.assert_array_equal(fl, g["fl"])
    np.testing.assert_array_equal(fr, g["fr"])


@pytest.mark.parametrize("eq", ["hd", "mhd"])
def test_equation_of_state(eq):
    g = load(f"unit_eqs_{eq}")
    p = Params(ndims=3, equation_system=eq, gamma=1.4 if eq == "hd" else 5.0 / 3.0)
    u, q2, f, cm, cp = ol.oracle_eqs(p, g["q"], float(g["cmax"]))
    for a, b in ((u, g["u"]), (q2, g["q2"]), (f, g["f"]), (cm, g["cm"]), (cp, g["cp"])):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("eq,rs", [("hd", "hll"), ("hd", "hllc"), ("mhd", "hll"), ("mhd", "hlld"),
                                   ("mhd", "hllc"), ("hd", "roe"), ("srmhd", "hllc"),
                                   ("mhd", "roe")])
def test_riemann_via_update_flux(eq, rs):
    g = load(f"unit_flux_{eq}_{rs}")
    p = Params.from_text(str(g["params"]), ndims=2)
    for b in range(g["q"].shape[0]):
        f = ol.oracle_update_flux(p, g["q"][b], float(g["cmax"]))
        np.testing.assert_array_equal(f, g["f"][b])


@pytest.mark.parametrize("eq,rs", [("hd", "hll"), ("hd", "hllc"), ("hd", "roe"), ("mhd", "hll"),
                                   ("mhd", "hlld"), ("mhd", "hllc")])
def test_riemann_negative_pressure(eq, rs):
    """Cells of negative pressure: NaN wave speeds go through the
    reference's MIN/MAX order; the NaN fluxes land where the reference's do
    (assert_array_equal takes NaN as equal to NaN)."""
    g = load(f"unit_flux_{eq}_{rs}_negp")
    p = Params.from_text(str(g["params"]), ndims=2)
    for b in range(g["q"].shape[0]):
        f = ol.oracle_update_flux(p, g["q"][b], float(g["cmax"]))
        np.testing.assert_array_equal(f, g["f"][b])


# ----------------------------------------------------------------- SRMHD
def test_c4_srmhd_run_bit_exact():
    """Relativistic MHD blast (TVD minmod + HLL + SSPRK(2,2), 1D(W) primitive
    solver) on 64^2 for 100 steps."""
    g = load("c4_srmhd2d_64_100")
    p = params_of(g)
    assert p.equation_system == "srmhd" and p.tolerance == 1e-12 and p.nr_maxit == 50
    u0 = problems.srmhd_blast_2d(p, noise=0.01)
    assert sha(u0) == str(g["sha_u0"]), "initial state changed"
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    s = int(g["stride"])
    u = o.get_state()
    np.testing.assert_array_equal(u[:, :, ::s, ::s], g["sub"])
    assert sha(u) == str(g["sha_u"])
    assert o.c2p_events() == [0, 0, 0]


def test_algebra_root_finders():
    """quadratic, quadratic_normalized, cubic, cubic_normalized, quartic."""
    g = load("unit_algebra")
    r = ol.oracle_algebra(g["a"])
    np.testing.assert_array_equal(r, g["r"])


def test_srmhd_equation_of_state():
    g = load("unit_eqs_srmhd")
    p = Params.from_text(str(g["params"]), ndims=2)
    u, q2, f, cm, cp = ol.oracle_eqs(p, g["q"], float(g["cmax"]))
    for a, b in ((u, g["u"]), (q2, g["q2"]), (f, g["f"]), (cm, g["cm"]), (cp, g["cp"])):
        np.testing.assert_array_equal(a, b)
    # failing and hard conserved states: the reference writes only p = 0 for
    # a failed cell and keeps the rest of q
    qb, ev = ol.oracle_cons2prim(p, g["ub"], g["q0"])
    np.testing.assert_array_equal(qb, g["qb"])
    failed = np.all(qb[:, :8] == -7.0, axis=1) & (qb[:, 8] == 0.0)
    assert failed[:16].all() and ev[0] == failed.sum()


def test_srmhd_hll_via_update_flux():
    g = load("unit_flux_srmhd_hll")
    p = Params.from_text(str(g["params"]), ndims=2)
    for b in range(g["q"].shape[0]):
        f = ol.oracle_update_flux(p, g["q"][b], float(g["cmax"]))
        np.testing.assert_array_equal(f, g["f"][b])


# ------------------------------------------------------- Gaussian process
def test_gp_pencils():
    """reconstruct_gp with its quad-precision weights and mp_limiting."""
    g = load("unit_gp")
    p = Params.from_text(str(g["params"]), ndims=3)
    fl, fr = ol.oracle_reconstruct(p, g["fp"])
    np.testing.assert_array_equal(fl, g["fl"])
    np.testing.assert_array_equal(fr, g["fr"])


def test_gp_run_bit_exact():
    g = load("gp_mhd3d_hlld_3")
    p = params_of(g)
    assert p.reconstruction == "gp"
    u0 = problems.turbulent_mhd_3d(p, seed=12)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


# ------------------------------------------------------------ integrators
@pytest.mark.parametrize("name", ["ssprk_ssprkm2_ot2d_32_10", "ssprk_ssprk43_ot2d_32_10",
                                  "ssprk_ssprk53_ot2d_32_10"])
def test_ssprk_run_bit_exact(name):
    """SSPRK(4,2), SSPRK(4,3) and SSPRK(5,3) on Orszag-Tang (TVD-MC + HLLD)."""
    g = load(name)
    p = params_of(g)
    u0 = problems.orszag_tang_2d(p, seed=9)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    s = int(g["stride"])
    np.testing.assert_array_equal(o.get_state()[:, :, ::s, ::s], g["sub"])
    assert sha(o.get_state()) == str(g["sha_u"])


# ---------------------------------------------------- multi-level meshes
@pytest.mark.parametrize("name", ["amr_mhd2d_blast_10", "amr_mhd3d_blast_4"])
def test_amr_run_bit_exact(name):
    """Blast on the reference's own three-level mesh (frozen), this
    project's seeded state: dt trace and final state of the reference."""
    from amun_amd import mesh
    g = load(name)
    p = params_of(g)
    meta = g["meta"]
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    u = o.get_blocks()
    assert sha(u) == str(g["sha_u"])
    np.testing.assert_array_equal(u.reshape(u.shape[0], -1, u.shape[-1]).sum(axis=1),
                                  g["block_sums"])


# ------------------------------------------------ more 1D reconstructions
@pytest.mark.parametrize("rec", ["weno5yc", "weno5ns", "mp5", "mp7"])
def test_more_schemes_pencils(rec):
    g = load("unit_schemes")
    p = Params(ndims=3, equation_system="mhd", reconstruction=rec, limiter="mm", ncells=16,
               nghosts=4, cfl=0.3)
    fl, fr = ol.oracle_reconstruct(p, g["fp"])
    np.testing.assert_array_equal(fl, g["fl_" + rec])
    np.testing.assert_array_equal(fr, g["fr_" + rec])


@pytest.mark.parametrize("rec", ["weno5yc", "weno5ns", "mp5", "mp7"])
def test_more_schemes_runs_bit_exact(rec):
    g = load(f"{rec}_ot2d_32_8")
    p = params_of(g)
    u0 = problems.orszag_tang_2d(p, seed=6)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


@pytest.mark.parametrize("rec", ["weno3", "limo3", "ppm"])
def test_o3_schemes_pencils(rec):
    """reconstruct_weno3, _limo3 and _ppm on the reference's pencils at two
    cell sizes."""
    g = load("unit_schemes_o3")
    p = Params(ndims=3, equation_system="mhd", reconstruction=rec, limiter="mm", ncells=16,
               nghosts=4, cfl=0.3)
    for ih, h in enumerate(g["h"]):
        fl, fr = ol.oracle_reconstruct(p, g["fp"], float(h))
        np.testing.assert_array_equal(fl, g[f"fl_{rec}_{ih}"])
        np.testing.assert_array_equal(fr, g[f"fr_{rec}_{ih}"])


@pytest.mark.parametrize("name", ["weno3_ot2d_32_8", "limo3_ot2d_32_8", "ppm_otnoisy2d_32_8"])
def test_o3_schemes_runs_bit_exact(name):
    g = load(name)
    p = params_of(g)
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import ot_noisy
    u0 = ot_noisy(p) if "noisy" in name else problems.orszag_tang_2d(p, seed=6)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


def test_amr_limo3_run_bit_exact():
    """LimO3 on the reference's three-level blast mesh: its smoothness
    radius scales with each level's cell size."""
    from amun_amd import mesh
    g = load("amr_limo3_mhd2d_blast_6")
    p = params_of(g)
    assert p.reconstruction == "limo3"
    meta = g["meta"]
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_blocks()) == str(g["sha_u"])


def test_hllc_mhd_run_bit_exact():
    """riemann_mhd_adi_hllc in an Orszag-Tang run (TVD-MC, SSPRK(3,3))."""
    g = load("hllc_mhd_ot2d_32_10")
    p = params_of(g)
    u0 = problems.orszag_tang_2d(p, seed=8)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


@pytest.mark.parametrize("name", ["roe_hd2d_blast_20", "roe_hd3d_blast_4"])
def test_roe_hd_run_bit_exact(name):
    """riemann_hd_adi_roe in blast runs (2D TVD-MC + RK2, 3D WENO5-Z + RK3)."""
    g = load(name)
    p = params_of(g)
    assert p.riemann_solver == "roe"
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import hd3_blast
    u0 = problems.blast_hd_2d(p, seed=9) if p.ndims == 2 else hd3_blast(p)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


# ------------------------------------------------------ store_integrals
def _stats_close(a, ref, ndig_rel=2e-8):
    """Statistics against the reference's printed values (9 significant
    digits).  The mean of psi is a sum of round-off-sized terms of both
    signs: judged on the scale of psi's extremes."""
    a, ref = np.asarray(a), np.asarray(ref)
    tol = ndig_rel * np.abs(ref) + 1e-300
    tol[12] = 1e-12 * max(abs(ref[13]), abs(ref[14]), 1e-300)
    return np.abs(a - ref) <= tol


@pytest.mark.parametrize("name", ["hd2d", "mhd2d", "mhd3d", "srmhd2d"])
def test_statistics_vs_reference_files(name):
    """The oracle's statistics after three steps against the reference's
    statistics file of the same run."""
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import stat_case
    g = load("stats_ref")
    p, u0 = stat_case(name)
    assert sha(u0) == str(g["sha_u0_" + name])
    o = ol.OracleRun(p)
    o.set_state(u0)
    o.run(3)
    ok = _stats_close(o.statistics(), g["stats_" + name])
    assert ok.all(), (np.argwhere(~ok).ravel(), o.statistics(), g["stats_" + name])


# ------------------------------------------- 2D relativistic primitive solvers
@pytest.mark.parametrize("ps", ["2dwv", "2dwu"])
def test_srmhd_2d_primitive_solvers(ps):
    """nr_iterate_srmhd_adi_2dwv / _2dwu: the reference's cons2prim on random
    physical and on failing states (where the lower bracket becomes the
    solution), and an 8-step blast run, bit for bit."""
    g = load(f"srmhd_{ps}_blast2d_32_8")
    pe = Params.from_text(str(g["unit_params"]), ndims=2)
    assert pe.primitive_solver == ps
    qb, ev = ol.oracle_cons2prim(pe, g["ub"], g["q0"])
    np.testing.assert_array_equal(qb, g["qb"])
    assert ev[0] > 0          # the bracket fallback was taken
    p = params_of(g)
    u0 = problems.srmhd_blast_2d(p, noise=0.01)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


def test_srmhd_hllc_run_bit_exact():
    """riemann_srmhd_adi_hllc in an 8-step relativistic blast run."""
    g = load("srmhd_hllc_blast2d_32_8")
    p = params_of(g)
    assert p.riemann_solver == "hllc"
    u0 = problems.srmhd_blast_2d(p, noise=0.01, vnoise=0.01)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


def test_roe_mhd_run_bit_exact():
    """riemann_mhd_adi_roe in a 10-step Orszag-Tang run (noise on every field)."""
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import ot_noisy
    g = load("roe_mhd_otnoisy2d_32_10")
    p = params_of(g)
    assert p.riemann_solver == "roe"
    u0 = ot_noisy(p)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


# ------------------------------------------- multi-dimensional GP (MGP)
@pytest.mark.parametrize("name", ["mgp_otnoisy2d_32_4", "mgp3_otnoisy2d_32_4",
                                  "mgp_otnoisy3d_16_4", "mgp_noisy3d_16_4"])
def test_mgp_runs(name):
    """prepare_mgp (quad-precision covariance inverse over the ngp^nd
    stencil) and interfaces_mgp (GP face values, TVD fallback where one
    leaves its cells' range): 4-step Orszag-Tang runs (noise on every field),
    2D at ngp = 5 and 3, 3D at ngp = 5, and a fully three-dimensional noisy
    state in 3D, bit for bit with the reference."""
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import mgp_state
    r = load(name)
    p = params_of(r)
    u0 = mgp_state(name, p)
    assert sha(u0) == str(r["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(r["nsteps"]))
    np.testing.assert_array_equal(tr, r["trace"])
    assert sha(o.get_state()) == str(r["sha_u"])


# ------------------------------------------------------- compact WENO
@pytest.mark.parametrize("rec", ["crweno5z", "crweno5yc", "crweno5ns", "crmp5", "crmp5ld",
                                 "crmp7"])
def test_crweno_pencils_and_runs(rec):
    """reconstruct_crweno5z / _yc / _ns and reconstruct_crmp5 / _crmp5ld /
    _crmp7 with tridiag on the reference's pencils (ng = 2, 3, 4: the
    explicit end rows move with ng) and an 8-step Orszag-Tang run (noise on
    every field), bit for bit."""
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import ot_noisy
    g = load("unit_schemes_crmp" if rec.startswith("crmp") else "unit_schemes_cr")
    for ng in (2, 3, 4):
        p = Params(ndims=3, equation_system="mhd", reconstruction=rec, limiter="mm", ncells=16,
                   nghosts=ng, cfl=0.3)
        fl, fr = ol.oracle_reconstruct(p, g["fp"])
        np.testing.assert_array_equal(fl, g[f"fl_{rec}_{ng}"])
        np.testing.assert_array_equal(fr, g[f"fr_{rec}_{ng}"])
    r = load(f"{rec}_otnoisy2d_32_8")
    p = params_of(r)
    u0 = ot_noisy(p)
    assert sha(u0) == str(r["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(r["nsteps"]))
    np.testing.assert_array_equal(tr, r["trace"])
    assert sha(o.get_state()) == str(r["sha_u"])


# ----------------------------------------------- refinement criterion
@pytest.mark.parametrize("name", ["refine_mhd2d_dens_pres", "refine_mhd2d_pres_vort",
                                  "refine_mhd2d_velx_magy", "refine_mhd3d_all"])
def test_refinement_criterion(name):
    """check_refinement_criterion (second_derivative_error, vorticity and
    current-density magnitudes, refinement.F90:257-640) of every block of the
    reference's three-level blast meshes after a few steps, four sets of
    variables and thresholds: the oracle's run reproduces the reference's
    state (sha) and its criterion block for block."""
    from amun_amd import mesh
    from amun_amd.config import refine_cfg
    r = load(name)
    p = params_of(r)
    meta = r["meta"]
    o = ol.OracleAmr(p, meta)
    o.set_blocks(mesh.block_state(p, meta, mesh.blast_prim(p)))
    o.run(int(r["nsteps"]))
    assert sha(o.get_blocks()) == str(r["sha_u1"])
    crit, ind = o.refinement(refine_cfg(p))
    np.testing.assert_array_equal(crit, r["crit"])


# ----------------------------------------------- mesh update (remesh)
@pytest.mark.parametrize("name", ["remesh_mhd2d_l3", "remesh_mhd2d_l4", "remesh_mhd3d_l3"])
def test_amr_remesh_matches_update_mesh(name):
    """One update_mesh of the reference (mesh.F90:756-833) after a few steps
    on its generated blast mesh with the blast moved away: blocks kept,
    prolonged from a refined parent (prolong_block, 1031-1257) and restricted
    from derefined children (restrict_block, 1258-1454) all occur; the
    oracle's data movement onto the reference's new leaf set reproduces the
    reference's state after the update bit for bit (sha), and the host's
    selection (amun_amd.mesh.update_leaves) picks the reference's new leaf
    set from the reference's criterion."""
    from amun_amd import mesh
    r = load(name)
    p = params_of(r)
    meta, meta2 = r["meta"], r["meta2"]
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p, shift=tuple(r["shift"])))
    assert sha(u0) == str(r["sha_u0"])
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    o.run(int(r["nsteps"]))
    assert sha(o.get_blocks()) == str(r["sha_u1"])
    old = {tuple(m) for m in meta.tolist()}
    new = {tuple(m) for m in meta2.tolist()}
    nc = p.ncells
    refined = [m for m in new - old if (m[0] - 1, *((x // nc // 2) * nc for x in m[1:])) in old]
    assert refined and len(new & old) and len(new - old) > len(refined)   # all three kinds
    # the host's selection from the reference's criterion gives the same leaves
    sel = mesh.update_leaves(p, meta, r["crit"], 1, int(p.extra["maxlev"]))
    assert {tuple(m) for m in sel.tolist()} == new and len(sel) == len(meta2)
    n = o.remesh(meta2)
    u2 = n.get_blocks()
    np.testing.assert_array_equal(u2.reshape(u2.shape[0], -1, p.nv).sum(axis=1),
                                  r["block_sums2"])
    assert sha(u2) == str(r["sha_u2"])


# ----------------------------------------------- GLM source terms
@pytest.mark.parametrize("name", ["eglm_otnoisy2d_32_8", "heglm_otnoisy2d_32_8",
                                  "heglm_noisy3d_16_3"])
def test_glm_source_terms_golden(name):
    """The EGLM / HEGLM source terms of update_sources (sources.F90:479-539,
    with derivative_1st's central differences, operators.F90:181-308):
    Orszag-Tang with noise on the field in 2D and a fully three-dimensional
    noisy state in 3D, so that div B is far from zero; the oracle's run is
    the reference's bit for bit, and the same run without the terms is not."""
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import glm_state
    r = load(name)
    p = params_of(r)
    u0 = glm_state(name, p)
    assert sha(u0) == str(r["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(r["nsteps"]))
    np.testing.assert_array_equal(tr[:, 0], r["trace"][:, 0])
    assert sha(o.get_state()) == str(r["sha_u"])
    p.glm_source_terms = "none"
    o = ol.OracleRun(p)
    o.set_state(u0)
    o.run(int(r["nsteps"]))
    assert sha(o.get_state()) != str(r["sha_u"])


def test_glm_source_terms_amr_golden():
    """HEGLM on the reference's three-level blast mesh: per-level cell sizes
    in the derivatives; the oracle's run is the reference's bit for bit."""
    from amun_amd import mesh
    r = load("amr_heglm_mhd2d_blast_6")
    p = params_of(r)
    meta = r["meta"]
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
    assert sha(u0) == str(r["sha_u0"])
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    tr = o.run(int(r["nsteps"]))
    np.testing.assert_array_equal(tr[:, 0], r["trace"][:, 0])
    assert sha(o.get_blocks()) == str(r["sha_u"])


# ----------------------------------------------- multi-dimensional limiting
@pytest.mark.parametrize("name", ["mlp_tvd_otnoisy2d_32_8", "mlp_weno5z_otnoisy2d_32_8",
                                  "mlp_mgp_otnoisy2d_32_4", "mlp_tvd_noisy3d_16_3"])
def test_mlp_limiting_golden(name):
    """mlp_limiting (interpolations.F90:1171-1321) after the TVD, WENO5-Z and
    MGP face states, 2D and 3D: the oracle's run is the reference's bit for
    bit, and the same run without the process is not."""
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import mlp_state
    r = load(name)
    p = params_of(r)
    u0 = mlp_state(name, p)
    assert sha(u0) == str(r["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(r["nsteps"]))
    np.testing.assert_array_equal(tr[:, 0], r["trace"][:, 0])
    assert sha(o.get_state()) == str(r["sha_u"])
    p.mlp_limiting = "off"
    o = ol.OracleRun(p)
    o.set_state(u0)
    o.run(int(r["nsteps"]))
    assert sha(o.get_state()) != str(r["sha_u"])
