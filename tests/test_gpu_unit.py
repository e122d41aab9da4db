"""Unit-level parity of the device routines (amun_dev_unit_riemann,
amun_dev_unit_eqs, amun_dev_unit_reconstruct) on vectors of states and
pencils: against the golden vectors written
by the reference binaries and against the oracle on random and degenerate
states.  FP64: <= 1e-12 of the magnitude of each vector (FMA contraction and
refined reciprocals are the only differences); mixed: FP32 round-off."""
import os

import numpy as np
import pytest

from amun_amd.config import Params
from amun_amd.solver import unit_eqs, unit_reconstruct, unit_riemann
import oracle_lib as ol

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def close(a, b, tol):
    """|a - b| <= tol * (per-row max |b|), rows are state vectors."""
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300)
    return np.abs(a - b) <= tol * scale


def _states(nv, n, rng, degenerate=True):
    q = np.zeros((n, nv))
    q[:, 0] = rng.uniform(0.1, 3.0, n)
    q[:, 1:4] = rng.normal(scale=1.2, size=(n, 3))
    q[:, -1] = rng.uniform(0.05, 3.0, n)
    if nv == 9:
        q[:, 4:7] = rng.normal(size=(n, 3))
        q[:, 7] = rng.normal(scale=0.1, size=n)
        if degenerate:
            k = n // 8
            q[:k, 4] = 0.0                       # Bx = 0
            q[k:2 * k, 5:7] = 0.0                # no transverse field
            q[2 * k:3 * k, 4] = 1e-6 * q[2 * k:3 * k, 4]   # weak Bx
            q[3 * k:4 * k, 1] *= 8.0             # fast flows
    return q


CASES = [("hd", "hll"), ("hd", "hllc"), ("mhd", "hll"), ("mhd", "hlld"), ("mhd", "hllc"),
         ("srmhd", "hll"), ("hd", "roe"), ("srmhd", "hllc"), ("mhd", "roe")]


@pytest.mark.parametrize("eq,rs", CASES)
def test_riemann_vs_oracle(eq, rs):
    rng = np.random.default_rng(31)
    p = Params(ndims=3, equation_system=eq, riemann_solver=rs,
               gamma=1.4 if eq == "hd" else 5.0 / 3.0, tolerance=1e-12, nr_maxit=50)
    if eq == "srmhd":
        from test_oracle_invariants import _sr_states
        ql, qr = _sr_states(rng, 2000), _sr_states(rng, 2000)
    else:
        ql, qr = _states(p.nv, 2000, rng), _states(p.nv, 2000, rng)
        qr[:200] = ql[:200]                      # identical states
        qr[200:400] = ql[200:400] * (1 + 1e-9)   # nearly identical
    fo = ol.oracle_riemann(p, ql, qr, cmax=1.7)
    fg = unit_riemann(p, ql, qr, cmax=1.7)
    ok = close(fg, fo, 1e-12)
    if eq == "mhd" and rs == "roe":
        # without transverse field esystem_roe_mhd_adi's cfsq equals a~^2 or
        # c_A^2 up to round-off, and alpha_s = sqrt((cfsq - a~^2) / ...) is the
        # square root of that round-off (equations.F90:3276-3289): those rows
        # (250:500, also the identical and near-identical pairs among the
        # first 400) agree to sqrt(eps)
        rows = np.zeros(len(fo), bool)
        rows[250:500] = True
        rows[:400] |= (np.abs(ql[:400, 5:7]).sum(axis=1) == 0)
        ok[rows] = close(fg[rows], fo[rows], 1e-7)
    print(f"[case] {eq}/{rs}: max rel {np.abs(fg - fo).max():.3e}")
    assert ok.all(), np.argwhere(~ok)[:5]
    pm = Params(**{k: v for k, v in p.__dict__.items() if k != "extra"})
    pm.precision = "mixed"
    fm = unit_riemann(pm, ql, qr, cmax=1.7)
    assert close(fm, fo, 3e-5).mean() >= 0.999


@pytest.mark.parametrize("eq", ["hd", "mhd", "srmhd"])
def test_equation_of_state_golden(eq):
    """prim2cons, fluxspeed and cons2prim against the reference's own outputs."""
    g = np.load(os.path.join(GOLD, f"unit_eqs_{eq}.npz"), allow_pickle=False)
    if eq == "srmhd":
        p = Params.from_text(str(g["params"]), ndims=2)
    else:
        p = Params(ndims=3, equation_system=eq, gamma=1.4 if eq == "hd" else 5.0 / 3.0,
                   riemann_solver="hllc" if eq == "hd" else "hlld")
    u, q2, f, cm, cp = unit_eqs(p, g["q"], float(g["cmax"]))
    for name, a, b in (("u", u, g["u"]), ("q2", q2, g["q2"]), ("f", f, g["f"])):
        tol = 1e-12 if name != "q2" or eq != "srmhd" else 1e-11
        ok = close(a, b, tol)
        print(f"[case] {eq} {name}: max abs {np.abs(a - b).max():.3e}")
        assert ok.all(), (name, np.argwhere(~ok)[:5])
    np.testing.assert_allclose(cm, g["cm"], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(cp, g["cp"], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("eq,rs", [("hd", "hllc"), ("mhd", "hlld"), ("mhd", "hllc"),
                                   ("srmhd", "hll"), ("hd", "roe"), ("srmhd", "hllc"), ("mhd", "roe")])
def test_riemann_golden_via_update_flux_states(eq, rs):
    """The golden update_flux fixtures use the zero limiter, so every face is
    the Riemann problem of its two cell values: the device solver on those
    pairs against the reference's fluxes (x faces)."""
    g = np.load(os.path.join(GOLD, f"unit_flux_{eq}_{rs}.npz"), allow_pickle=False)
    p = Params.from_text(str(g["params"]), ndims=2)
    q, fref = g["q"], g["f"]               # (nb, 1, jm, im, nv), (..., nv, nd)
    # faces whose both cells are reconstructed: cells ibl..ieu (one ghost
    # layer each side) of rows jbl..jeu
    lo, hi = p.nghosts - 1, p.nghosts + p.ncells
    ql = q[:, 0, lo:hi + 1, lo:hi, :].reshape(-1, p.nv)
    qr = q[:, 0, lo:hi + 1, lo + 1:hi + 1, :].reshape(-1, p.nv)
    fx = fref[:, 0, lo:hi + 1, lo:hi, :, 0].reshape(-1, p.nv)
    fg = unit_riemann(p, ql, qr, float(g["cmax"]))
    ok = close(fg, fx, 1e-12)
    print(f"[case] golden {eq}/{rs}: max abs {np.abs(fg - fx).max():.3e}")
    assert ok.all(), np.argwhere(~ok)[:5]


@pytest.mark.parametrize("eq,rs", [("hd", "hll"), ("hd", "hllc"), ("hd", "roe"), ("mhd", "hll"),
                                   ("mhd", "hlld"), ("mhd", "hllc")])
def test_riemann_negative_pressure_golden(eq, rs):
    """Face pairs with negative pressure on either side: the device solvers
    take the reference's MIN/MAX order with NaN wave speeds, so the NaN
    fluxes sit where the reference's do and the rest agree to 1e-12."""
    g = np.load(os.path.join(GOLD, f"unit_flux_{eq}_{rs}_negp.npz"), allow_pickle=False)
    p = Params.from_text(str(g["params"]), ndims=2)
    q, fref = g["q"], g["f"]
    lo, hi = p.nghosts - 1, p.nghosts + p.ncells
    ql = q[:, 0, lo:hi + 1, lo:hi, :].reshape(-1, p.nv)
    qr = q[:, 0, lo:hi + 1, lo + 1:hi + 1, :].reshape(-1, p.nv)
    fx = fref[:, 0, lo:hi + 1, lo:hi, :, 0].reshape(-1, p.nv)
    fg = unit_riemann(p, ql, qr, float(g["cmax"]))
    nan_ref, nan_dev = ~np.isfinite(fx), ~np.isfinite(fg)
    print(f"[case] {eq}/{rs}: non-finite ref {nan_ref.sum()} device {nan_dev.sum()}")
    np.testing.assert_array_equal(nan_dev, nan_ref)
    ok = close(np.where(nan_ref, 0.0, fg), np.where(nan_ref, 0.0, fx), 1e-12)
    assert ok.all(), np.argwhere(~ok)[:5]


def _pencil_cases():
    """(name, params, fp, h, fl, fr) of the reference's golden pencils."""
    out = []
    g = np.load(os.path.join(GOLD, "unit_weno5z.npz"), allow_pickle=False)
    out.append(("weno5z", Params(ndims=3, equation_system="mhd", reconstruction="weno5z",
                                 limiter="mm", ncells=16, nghosts=4, cfl=0.3),
                g["fp"], 1.0, g["fl"], g["fr"]))
    g = np.load(os.path.join(GOLD, "unit_gp.npz"), allow_pickle=False)
    out.append(("gp", Params.from_text(str(g["params"]), ndims=3), g["fp"], 1.0, g["fl"], g["fr"]))
    g = np.load(os.path.join(GOLD, "unit_schemes.npz"), allow_pickle=False)
    for rec in ("weno5yc", "weno5ns", "mp5", "mp7"):
        out.append((rec, Params(ndims=3, equation_system="mhd", reconstruction=rec, limiter="mm",
                                ncells=16, nghosts=4, cfl=0.3),
                    g["fp"], 1.0, g["fl_" + rec], g["fr_" + rec]))
    g = np.load(os.path.join(GOLD, "unit_schemes_o3.npz"), allow_pickle=False)
    for rec in ("weno3", "limo3", "ppm"):
        for ih, h in enumerate(g["h"]):
            out.append((f"{rec}_h{ih}", Params(ndims=3, equation_system="mhd", reconstruction=rec,
                                               limiter="mm", ncells=16, nghosts=4, cfl=0.3),
                        g["fp"], float(h), g[f"fl_{rec}_{ih}"], g[f"fr_{rec}_{ih}"]))
    return out


@pytest.mark.parametrize("case", _pencil_cases(), ids=lambda c: c[0])
def test_reconstruction_pencils_golden(case):
    """The device reconstruction of every 1D scheme on the reference's own
    pencils (interior cells, four from either end): FP64 to 1e-12 of the
    pencil's range; FP32 to 1e-5."""
    name, p, fp, h, rl, rr = case
    fl, fr = unit_reconstruct(p, fp, h)
    sl = slice(4, fp.shape[1] - 4)
    scale = np.abs(fp).max(axis=1, keepdims=True)
    el = np.abs(fl[:, sl] - rl[:, sl]) / scale
    er = np.abs(fr[:, 3:fp.shape[1] - 5] - rr[:, 3:fp.shape[1] - 5]) / scale
    print(f"[case] {name}: max rel fl {el.max():.2e} fr {er.max():.2e}")
    # the GP weights come from the quad-precision solve of prepare_gp on the
    # host (rounded once to FP64), the reference rounds its own way
    tol = 1e-11 if name == "gp" else 1e-12
    assert el.max() <= tol, np.argwhere(el > tol)[:5]
    assert er.max() <= tol, np.argwhere(er > tol)[:5]
    pm = Params(**{k: v for k, v in p.__dict__.items() if k != "extra"})
    pm.precision = "mixed"
    ml, mr = unit_reconstruct(pm, fp, h)
    # FP32 round-off can move a value across a limiter switch (the PPM and
    # LimO3 branches, the MP bounds): judged by the share of close values
    em = np.abs(ml[:, sl] - rl[:, sl]) / scale
    print(f"  mixed: share within 1e-5 {np.mean(em <= 1e-5):.4f}")
    assert np.mean(em <= 1e-5) >= 0.97


@pytest.mark.parametrize("ps", ["2dwv", "2dwu"])
def test_srmhd_2d_primitive_solvers_golden(ps):
    """The device's 2D(W,v^2) and 2D(W,u^2) recoveries (FP64) on the
    reference's own vectors: random physical states and states without a
    bracket (the lower bracket becomes the solution)."""
    g = np.load(os.path.join(GOLD, f"srmhd_{ps}_blast2d_32_8.npz"), allow_pickle=False)
    p = Params.from_text(str(g["unit_params"]), ndims=2)
    ub, qb = g["ub"], g["qb"]
    from amun_amd.solver import unit_cons2prim
    q2, ev = unit_cons2prim(p, ub, g["q0"], events=True)
    qo, evo = ol.oracle_cons2prim(p, ub, g["q0"])
    print(f"[case] {ps}: max abs {np.abs(q2 - qb).max():.3e}, events {ev} (oracle {list(evo)})")
    # 400 random physical states: converged Newton solutions
    ok = close(q2[:400], qb[:400], 1e-12)
    assert ok.all(), np.argwhere(~ok)[:5]
    # the hard states: many take the lower bracket (roots of the positivity
    # polynomials) as W, and p = W - E + ... cancels there
    ok = close(q2[400:], qb[400:], 1e-9)
    assert ok.all(), np.argwhere(~ok)[:5]
    assert ev == list(evo)
