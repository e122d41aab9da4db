"""The oracle's isothermal HD and MHD systems (equations.F90:399-830 and the
solvers of schemes.F90:656-2906) against the reference's own outputs
(tests/golden/make_golden.py iso): equation-of-state vectors, every solver
through update_flux, 2D and 3D runs and the statistics file.  Bit-exact, as
the other oracle comparisons; plus two checks that do not rest on the
reference: the isothermal Roe flux equals the exact flux for a pair of equal
states, and the HLLD-M flux reduces to HLL when Bx = 0 and the contact
speed vanishes."""
import hashlib
import os
import sys

import numpy as np
import pytest

from amun_amd.config import Params
import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_golden import ISO_FLUX, ISO_RUNS, iso_state, iso_stat_case  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("eq", ["hd", "mhd"])
def test_iso_equation_of_state(eq):
    g = load(f"unit_eqs_{eq}_iso")
    p = Params.from_text(str(g["params"]), ndims=3)
    assert p.equation_of_state == "iso" and p.nv == (4 if eq == "hd" else 8)
    u, q2, f, cm, cp = ol.oracle_eqs(p, g["q"], float(g["cmax"]))
    for a, b in ((u, g["u"]), (q2, g["q2"]), (f, g["f"]), (cm, g["cm"]), (cp, g["cp"])):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("eq,rs", ISO_FLUX)
def test_iso_riemann_via_update_flux(eq, rs):
    g = load(f"unit_flux_{eq}_iso_{rs}")
    p = Params.from_text(str(g["params"]), ndims=2)
    for b in range(g["q"].shape[0]):
        f = ol.oracle_update_flux(p, g["q"][b], float(g["cmax"]))
        np.testing.assert_array_equal(f, g["f"][b])


@pytest.mark.parametrize("name", [r[0] for r in ISO_RUNS])
def test_iso_runs_bit_exact(name):
    g = load(name)
    p = Params.from_text(str(g["params"]), ndims=int(g["ndims"]))
    u0 = iso_state(name, p)
    assert sha(u0) == str(g["sha_u0"])
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(int(g["nsteps"]))
    np.testing.assert_array_equal(tr, g["trace"])
    assert sha(o.get_state()) == str(g["sha_u"])


@pytest.mark.parametrize("name", ["hd2d_iso", "mhd2d_iso"])
def test_iso_statistics_vs_reference_files(name):
    g = load("stats_iso_ref")
    p, u0 = iso_stat_case(name)
    assert sha(u0) == str(g["sha_u0_" + name])
    o = ol.OracleRun(p)
    o.set_state(u0)
    o.run(3)
    st, ref = np.asarray(o.statistics()), np.asarray(g["stats_" + name])
    # 9 significant digits in the reference's file; psi's mean on the scale
    # of its extremes
    tol = 2e-8 * np.abs(ref) + 1e-300
    tol[12] = 1e-12 * max(abs(ref[13]), abs(ref[14]), 1e-300)
    assert (np.abs(st - ref) <= tol).all(), (st, ref)
    assert st[3] == st[4] == st[5] == 0.0      # no pressure


def _flux_x(p, ql, qr, cmax=2.3):
    """the x flux of the face between two piecewise-constant states"""
    nv = p.nv
    q = np.empty((1, 12, 12, nv))
    q[:, :, :6] = ql
    q[:, :, 6:] = qr
    f = ol.oracle_update_flux(p, q, cmax)
    return f[0, 6, 5, :, 0]


@pytest.mark.parametrize("eq,rs", ISO_FLUX)
def test_iso_consistency_equal_states(eq, rs):
    """Every solver returns the exact flux F(q) for equal states:
    rho v, rho v vx + csnd^2 rho - Bx^2 + |B|^2 / 2, ... (closed form)."""
    c = 0.7
    p = Params(ndims=2, equation_system=eq, equation_of_state="iso", csnd=c,
               reconstruction="tvd", limiter="zero", riemann_solver=rs, ncells=8, nghosts=2)
    rng = np.random.default_rng(3)
    for _ in range(5):
        q = np.zeros(p.nv)
        q[0] = rng.uniform(0.3, 2.0)
        q[1:4] = rng.normal(scale=0.3, size=3)
        if p.nv == 8:
            q[4:7] = rng.normal(size=3)
        f = _flux_x(p, q, q)
        rho, v = q[0], q[1:4]
        ex = np.zeros(p.nv)
        ex[0] = rho * v[0]
        ex[1:4] = rho * v * v[0]
        ex[1] += c * c * rho
        if p.nv == 8:
            b = q[4:7]
            ex[1:4] -= b[0] * b
            ex[1] += 0.5 * (b @ b)
            ex[5:7] = v[0] * b[1:] - b[0] * v[1:]
            ex[4] = q[7]
            ex[7] = 2.3 * 2.3 * b[0]
        np.testing.assert_allclose(f, ex, rtol=1e-13, atol=1e-13)


def test_iso_hlldm_reduces_to_hll_without_bx():
    """HLLD-M with Bx = 0 and psi = 0 at a face whose HLL contact speed is
    exactly zero (mirror-symmetric states) takes the HLL branch."""
    base = dict(ndims=2, equation_system="mhd", equation_of_state="iso", csnd=0.7,
                reconstruction="tvd", limiter="zero", ncells=8, nghosts=2)
    pm = Params(riemann_solver="hlldm", **base)
    ph = Params(riemann_solver="hll", **base)
    ql = np.array([1.3, 0.4, 0.2, -0.1, 0.0, 0.5, -0.3, 0.0])
    qr = ql.copy()
    qr[1] = -ql[1]
    fm, fh = _flux_x(pm, ql, qr), _flux_x(ph, ql, qr)
    np.testing.assert_array_equal(fm, fh)
