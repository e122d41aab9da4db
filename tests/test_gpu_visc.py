"""The viscous and resistive source terms (update_sources,
sources.F90:340-598) and their time-step limit (evolution.F90:478-483) on
the GPU: the reference's own runs (tests/golden/visc_*, resist_*,
viscres_*) and one step against the C oracle, which reproduces those runs
bit for bit (test_oracle_visc.py), on single- and multi-level meshes.
FP64: <= 1e-12 relative per step, <= 1e-10 after the runs."""
import os
import sys

import numpy as np
import pytest

from amun_amd import mesh
from amun_amd.config import Params
from amun_amd.solver import AmrSolver, AmunError, Solver
import oracle_lib as ol

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_golden import VISC_RUNS, visc_params, visc_state  # noqa: E402


def relerr(a, b):
    """per-variable max |a - b| over the variable's (vector's) magnitude,
    variables on axis 0"""
    ax = tuple(range(1, b.ndim))
    mag = np.abs(b).max(axis=ax)
    scale = mag.copy()
    scale[1:4] = mag[1:4].max()
    if b.shape[0] >= 8:
        scale[4:7] = mag[4:7].max()
        scale[7] = max(mag[7], mag[4:7].max())
    return np.abs(a - b).max(axis=ax) / np.maximum(scale, 1e-300)


@pytest.mark.parametrize("name", [r[0] for r in VISC_RUNS])
def test_viscous_golden_runs(name):
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    p = Params.from_text(str(g["params"]), ndims=int(g["ndims"]))
    u0 = visc_state(name, p)
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
@pytest.mark.parametrize("eq", ["hd", "mhd", "mhdiso"])
@pytest.mark.parametrize("rec", ["tvd", "weno5z"])
def test_viscous_one_step_vs_oracle(nd, eq, rec):
    p = visc_params(nd, eq, rec, 2e-3, 3e-3 if eq != "hd" else 0.0)
    p.ncells = 8
    name = "noisy" if nd == 3 else ("blast" if eq == "hd" else "x_ot_")
    u0 = visc_state(name, p)
    print(f"[case] visc {nd}D {eq} {rec}", flush=True)
    with Solver(p) as s:
        s.upload(u0)
        dt, _ = s.advance()
        ug = s.download()
    o = ol.OracleRun(p)
    o.set_state(u0)
    tr = o.run(1)
    err = relerr(ug, o.get_state()).max()
    print(f"  dt {dt:.15e} vs {tr[0, 0]:.15e}; max rel {err:.2e}")
    assert abs(dt - tr[0, 0]) <= 1e-13 * tr[0, 0]
    assert err <= 1e-12


@pytest.mark.parametrize("nd", [2, 3])
def test_viscous_amr_one_step_vs_oracle(nd):
    """A three-level mesh: the stencil of the stress reaches two cells into
    the filled virtual blocks; each level's cell size."""
    p = Params(ndims=nd, equation_system="mhd", reconstruction="tvd", limiter="mc",
               riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.4, ncells=8,
               nghosts=2, xblocks=2, yblocks=2, zblocks=2, xmin=-0.5, xmax=0.5, ymin=-0.5,
               ymax=0.5, zmin=-0.5, zmax=0.5, prolongation_limiter="mm", viscosity=2e-3,
               resistivity=3e-3)
    meta = mesh.refine(p, mesh.near_sphere(p, 0.15, 0.05), 3)
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
    print(f"[case] visc AMR {nd}D blocks={len(meta)}", flush=True)
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    tr = o.run(1)
    uo = o.get_blocks()
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        dt, _ = s.advance()
        ug = s.download()
    err = relerr(np.moveaxis(ug.reshape(-1, p.nv), -1, 0),
                 np.moveaxis(uo.reshape(-1, p.nv), -1, 0)).max()
    print(f"  dt {dt:.15e} vs {tr[0, 0]:.15e}; max rel {err:.2e}")
    assert abs(dt - tr[0, 0]) <= 1e-13 * tr[0, 0]
    assert err <= 1e-12


def test_viscous_mixed_vs_fp64():
    name = "viscres_mhd3d_weno5z_noisy_4"
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    p = Params.from_text(str(g["params"]), ndims=3)
    u0 = visc_state(name, p)
    with Solver(p) as s:
        s.upload(u0)
        s.run(4)
        u64 = s.download()
    p.precision = "mixed"
    print("[case] visc mixed vs fp64", flush=True)
    with Solver(p) as s:
        s.upload(u0)
        s.run(4)
        u32 = s.download()
    err = relerr(u32, u64).max()
    print(f"  mixed vs fp64 max rel {err:.2e}")
    assert err <= 1e-5


def test_gradients_refuse_viscous_terms():
    p = visc_params(2, "mhd", "tvd", 1e-3, 0.0)
    u0 = visc_state("x_ot_", p)
    with Solver(p) as s:
        s.upload(u0)
        with pytest.raises(AmunError) as e:
            s.tape_forward(1)
        assert e.value.code == 600
