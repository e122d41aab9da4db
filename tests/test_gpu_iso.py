"""The isothermal HD and MHD systems on the GPU (equations.F90:399-830;
solvers HLL, HD Roe, MHD HLLD, HLLD-M and Roe of schemes.F90:656-2906)
against the reference's own outputs (tests/golden/*iso*, written by
make_golden.py iso) and against the C oracle, which matches those outputs
bit for bit (test_oracle_iso.py).

FP64 tolerances as for the adiabatic systems: <= 1e-12 of each vector's
magnitude for unit vectors and one step, <= 1e-10 after the multi-step runs
(round-off growth through limiter switches); the dt traces to 1e-12."""
import os
import sys

import numpy as np
import pytest

from amun_amd import mesh
from amun_amd.config import Params
from amun_amd.problems import prim_to_cons
from amun_amd.solver import AmrSolver, Solver, unit_eqs, unit_riemann
import oracle_lib as ol

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_golden import ISO_FLUX, ISO_RUNS, iso_state, iso_stat_case  # noqa: E402


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def close(a, b, tol):
    """|a - b| <= tol * (per-row max |b|), rows are state vectors."""
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300)
    return np.abs(a - b) <= tol * scale


def relerr(a, b, axis0=True):
    """Per-variable max |a - b| over the variable's largest magnitude; the
    momentum and field components share their vector's, psi the field's.
    Variables on axis 0 (global arrays) or on the last axis (block arrays)."""
    if not axis0:
        a, b = np.moveaxis(a, -1, 0), np.moveaxis(b, -1, 0)
    ax = tuple(range(1, b.ndim))
    mag = np.abs(b).max(axis=ax)
    scale = mag.copy()
    scale[1:4] = mag[1:4].max()
    if b.shape[0] == 8:
        scale[4:7] = mag[4:7].max()
        scale[7] = max(mag[7], mag[4:7].max())
    return np.abs(a - b).max(axis=ax) / np.maximum(scale, 1e-300)


@pytest.mark.parametrize("eq", ["hd", "mhd"])
def test_iso_equation_of_state_golden(eq):
    g = load(f"unit_eqs_{eq}_iso")
    p = Params.from_text(str(g["params"]), ndims=3)
    p.riemann_solver = "hll"      # the fixture's default name is an MHD solver's
    u, q2, f, cm, cp = unit_eqs(p, g["q"], float(g["cmax"]))
    for name, a, b in (("u", u, g["u"]), ("q2", q2, g["q2"]), ("f", f, g["f"])):
        ok = close(a, b, 1e-12)
        print(f"[case] {eq} iso {name}: max abs {np.abs(a - b).max():.3e}", flush=True)
        assert ok.all(), (name, np.argwhere(~ok)[:5])
    np.testing.assert_allclose(cm, g["cm"], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(cp, g["cp"], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("eq,rs", ISO_FLUX)
def test_iso_riemann_golden(eq, rs):
    """Every face of the zero-limiter update_flux fixtures is the Riemann
    problem of its two cell values: the device solver on those pairs
    against the reference's x fluxes, in FP64 and (FP32 round-off) mixed."""
    g = load(f"unit_flux_{eq}_iso_{rs}")
    p = Params.from_text(str(g["params"]), ndims=2)
    q, fref = g["q"], g["f"]
    lo, hi = p.nghosts - 1, p.nghosts + p.ncells
    ql = q[:, 0, lo:hi + 1, lo:hi, :].reshape(-1, p.nv)
    qr = q[:, 0, lo:hi + 1, lo + 1:hi + 1, :].reshape(-1, p.nv)
    fx = fref[:, 0, lo:hi + 1, lo:hi, :, 0].reshape(-1, p.nv)
    print(f"[case] iso {eq}/{rs}: {len(ql)} faces", flush=True)
    fg = unit_riemann(p, ql, qr, float(g["cmax"]))
    ok = close(fg, fx, 1e-12)
    print(f"  max abs {np.abs(fg - fx).max():.3e}")
    assert ok.all(), np.argwhere(~ok)[:5]
    p.precision = "mixed"
    fm = unit_riemann(p, ql, qr, float(g["cmax"]))
    assert close(fm, fx, 3e-5).mean() >= 0.99


@pytest.mark.parametrize("eq,rs", ISO_FLUX)
def test_iso_riemann_vs_oracle_random(eq, rs):
    """Random pairs with the degenerate cases (Bx = 0, no transverse field,
    weak Bx, fast flows, identical states) against the oracle."""
    rng = np.random.default_rng(41)
    p = Params(ndims=3, equation_system=eq, equation_of_state="iso", csnd=0.8,
               riemann_solver=rs)
    n = 2000

    def states():
        q = np.zeros((n, p.nv))
        q[:, 0] = rng.uniform(0.1, 3.0, n)
        q[:, 1:4] = rng.normal(scale=1.2, size=(n, 3))
        if p.nv == 8:
            q[:, 4:7] = rng.normal(size=(n, 3))
            q[:, 7] = rng.normal(scale=0.1, size=n)
            k = n // 8
            q[:k, 4] = 0.0
            q[k:2 * k, 5:7] = 0.0
            q[2 * k:3 * k, 4] *= 1e-6
            q[3 * k:4 * k, 1] *= 8.0
        return q
    ql, qr = states(), states()
    qr[:200] = ql[:200]
    fo = ol.oracle_riemann(p, ql, qr, cmax=1.7)
    fg = unit_riemann(p, ql, qr, cmax=1.7)
    ok = close(fg, fo, 1e-12)
    if eq == "mhd" and rs == "roe":
        # without transverse field alpha_s = sqrt((cfsq - c~^2) / ...) is the
        # square root of round-off (esystem_roe_mhd_iso, equations.F90:
        # 2640-2652), as in the adiabatic solver: those rows to sqrt(eps)
        rows = np.zeros(n, bool)
        rows[250:500] = True
        rows[:200] |= np.abs(ql[:200, 5:7]).sum(axis=1) == 0
        ok[rows] = close(fg[rows], fo[rows], 1e-7)
    print(f"[case] iso {eq}/{rs} random: max abs {np.abs(fg - fo).max():.3e}", flush=True)
    assert ok.all(), np.argwhere(~ok)[:5]


@pytest.mark.parametrize("name", [r[0] for r in ISO_RUNS])
def test_iso_golden_runs(name):
    """The reference's 2D and 3D runs (TVD, WENO5-Z, MP5; every solver):
    dt trace and a subsample of the final state."""
    g = load(name)
    p = Params.from_text(str(g["params"]), ndims=int(g["ndims"]))
    u0 = iso_state(name, p)
    print(f"[case] {name}", flush=True)
    with Solver(p) as s:
        s.upload(u0)
        dts = s.run(int(g["nsteps"]))
        u = s.download()
    dterr = np.abs(dts - g["trace"][:, 0]).max() / g["trace"][:, 0].max()
    st = int(g["stride"])
    sub = u[:, ::st, ::st, ::st] if p.ndims == 3 else u[:, :, ::st, ::st]
    err = relerr(sub, g["sub"]).max()
    print(f"  dt rel {dterr:.2e}, state max rel {err:.2e}")
    assert dterr <= 1e-12
    assert err <= 1e-10


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("eq,rs", [("hd", "roe"), ("mhd", "hlld"), ("mhd", "hlldm"),
                                   ("mhd", "roe")])
@pytest.mark.parametrize("rec", ["tvd", "weno5z", "ppm", "crweno5z", "mgp"])
def test_iso_one_step_vs_oracle(nd, eq, rs, rec):
    p = Params(ndims=nd, equation_system=eq, equation_of_state="iso", csnd=0.9,
               riemann_solver=rs, reconstruction=rec, limiter="mc" if rec == "tvd" else "mm",
               time_advance="rk3", ncells=8, nghosts=2 if rec == "tvd" else 4,
               xblocks=2, yblocks=2, zblocks=2, cfl=0.3, xmin=-0.5, xmax=0.5, ymin=-0.5,
               ymax=0.5, zmin=-0.5, zmax=0.5)
    u0 = iso_state("iso_noisy", p, seed=9)
    print(f"[case] iso {nd}D {eq}/{rs} {rec}", flush=True)
    with Solver(p) as s:
        s.upload(u0)
        dt, _ = s.advance()
        ug = s.download()
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(1)
    uo = o.get_state()
    err = relerr(ug, uo).max()
    print(f"  dt {dt:.15e} vs {tr[0, 0]:.15e}; max rel {err:.2e}")
    assert abs(dt - tr[0, 0]) <= 1e-13 * tr[0, 0]
    assert err <= 1e-12


@pytest.mark.parametrize("name", ["hd2d_iso", "mhd2d_iso"])
def test_iso_statistics_vs_reference_files(name):
    """store_integrals' statistics after three steps against the reference's
    file: no pressure entries, the Mach number |v| / csnd."""
    g = load("stats_iso_ref")
    p, u0 = iso_stat_case(name)
    with Solver(p) as s:
        s.upload(u0)
        s.run(3)
        st = np.asarray(s.statistics())
    ref = np.asarray(g["stats_" + name])
    tol = 2e-8 * np.abs(ref) + 1e-300
    tol[12] = 1e-12 * max(abs(ref[13]), abs(ref[14]), 1e-300)
    print(f"[case] stats {name}: {st[:6]}", flush=True)
    assert (np.abs(st - ref) <= tol).all(), (st, ref)
    assert st[3] == st[4] == st[5] == 0.0


def test_iso_reflective_sides_one_step_vs_oracle():
    """Reflective and outflow sides with the isothermal MHD system."""
    p = Params(ndims=2, equation_system="mhd", equation_of_state="iso", csnd=0.9,
               riemann_solver="hlld", reconstruction="tvd", limiter="mc", time_advance="rk2",
               ncells=8, nghosts=2, xblocks=3, yblocks=2, cfl=0.4, xmin=-0.5, xmax=0.5,
               ymin=-0.5, ymax=0.5, xlbndry="reflective", xubndry="reflective",
               ylbndry="outflow", yubndry="outflow")
    u0 = iso_state("iso_noisy", p, seed=4)
    print("[case] iso reflective/outflow", flush=True)
    with Solver(p) as s:
        s.upload(u0)
        s.advance()
        ug = s.download()
    o = ol.OracleRun(p)
    o.set_state(u0)
    o.run(1)
    err = relerr(ug, o.get_state()).max()
    print(f"  max rel {err:.2e}")
    assert err <= 1e-12


def _iso_amr_case(nd, eq, rs, precision="fp64"):
    p = Params(ndims=nd, equation_system=eq, equation_of_state="iso", csnd=0.8,
               reconstruction="tvd", limiter="mc", riemann_solver=rs, time_advance="rk3",
               cfl=0.4, ncells=8, nghosts=2, xblocks=2, yblocks=2, zblocks=2, xmin=-0.5,
               xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5,
               prolongation_limiter="mm", precision=precision)
    meta = mesh.refine(p, mesh.near_sphere(p, 0.15, 0.05), 3)

    def prim(x, y, z):
        r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
        q = np.zeros((p.nv,) + x.shape)
        q[0] = np.where(r < 0.15, 4.0, 1.0) * (1.0 + 0.05 * np.sin(7.0 * x + 3.0 * y))
        q[1] = 0.1 * np.sin(5.0 * y) + 0.05
        q[2] = 0.1 * np.cos(4.0 * x) - 0.03
        q[3] = 0.02
        if p.nv == 8:
            q[4], q[5], q[6] = 0.5, 0.3 + 0.1 * np.sin(3.0 * x), 0.2
        return q
    return p, meta, mesh.block_state(p, meta, prim)


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("eq,rs", [("hd", "roe"), ("mhd", "hlld")])
def test_iso_amr_one_step_vs_oracle(nd, eq, rs):
    """A three-level mesh: prolongations without a pressure to keep positive,
    restrictions and the fine-face flux correction."""
    p, meta, u0 = _iso_amr_case(nd, eq, rs)
    print(f"[case] iso AMR {nd}D {eq}/{rs} blocks={len(meta)}", flush=True)
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    tr = o.run(1)
    uo = o.get_blocks()
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        dt, _ = s.advance()
        ug = s.download()
    err = relerr(ug.reshape(-1, p.nv), uo.reshape(-1, p.nv), axis0=False).max()
    print(f"  dt {dt:.15e} vs {tr[0, 0]:.15e}; max rel {err:.2e}")
    assert abs(dt - tr[0, 0]) <= 1e-13 * tr[0, 0]
    assert err <= 1e-12


def test_iso_mixed_vs_fp64():
    """The mixed-precision path on the 3D isothermal MHD WENO5-Z run: FP32
    face states and fluxes, FP64 conserved sums."""
    name = "iso_mhd3d_weno5z_hlld_noisy_4"
    g = load(name)
    p = Params.from_text(str(g["params"]), ndims=3)
    u0 = iso_state(name, p)
    with Solver(p) as s:
        s.upload(u0)
        s.run(4)
        u64 = s.download()
    p.precision = "mixed"
    print("[case] iso mixed vs fp64", flush=True)
    with Solver(p) as s:
        s.upload(u0)
        s.run(4)
        u32 = s.download()
    err = relerr(u32, u64).max()
    print(f"  mixed vs fp64 max rel {err:.2e}")
    assert err <= 1e-5


def test_iso_uniform_state_is_preserved():
    """A uniform isothermal MHD state with a uniform field stays uniform."""
    p = Params(ndims=3, equation_system="mhd", equation_of_state="iso", csnd=1.3,
               riemann_solver="hlld", reconstruction="weno5z", limiter="mm", ncells=8,
               nghosts=4, xblocks=2, yblocks=2, zblocks=2)
    q = np.zeros((8,) + p.shape)
    q[0], q[1], q[2], q[3] = 1.7, 0.3, -0.2, 0.1
    q[4], q[5], q[6] = 0.4, -0.6, 0.25
    u0 = prim_to_cons(p, q)
    with Solver(p) as s:
        s.upload(u0)
        s.run(3)
        u = s.download()
    assert relerr(u, u0).max() <= 1e-14


@pytest.mark.parametrize("nd", [2, 3])
def test_iso_amr_remesh_vs_oracle(nd):
    """A mesh update of an isothermal MHD run (random refine / keep /
    derefine flags through update_mesh's selection): the prolongations take
    the ghost cells as prim2cons of the isothermal primitives and keep only
    the density positive; within round-off of the oracle's update."""
    p, meta, u0 = _iso_amr_case(nd, "mhd", "hlld")
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    o.run(2)
    uo = o.get_blocks()
    crit = np.random.default_rng(3 + nd).integers(-1, 2, len(meta))
    meta2 = mesh.update_leaves(p, meta, crit, 1, 3)
    assert len(meta2) != len(meta)
    ref = o.remesh(meta2).get_blocks()
    print(f"[case] iso remesh {nd}D blocks {len(meta)} -> {len(meta2)}", flush=True)
    with AmrSolver(p, meta) as s:
        s.upload(uo)
        s.remesh(meta2)
        ud = s.download()
    err = relerr(ud.reshape(-1, p.nv), ref.reshape(-1, p.nv), axis0=False).max()
    print(f"  max rel {err:.2e}")
    assert err <= 1e-13
