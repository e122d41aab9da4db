"""Multi-level (AMR) meshes on the GPU (amun_dev_init_amr) against the C
oracle, which reproduces the reference's own refined runs bit for bit
(test_oracle_golden.py, test_oracle_ref.py): the ghost layers of every block
after the fill (restriction, prolongation), one step of several
configurations (<= 1e-12 relative), the reference's multi-step runs from
the golden fixtures, conservation through the flux correction, mixed
precision, and HIP-graph replay."""
import os

import numpy as np
import pytest

from amun_amd import mesh
from amun_amd.config import Params
from amun_amd.solver import AmrSolver
import oracle_lib as ol

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _case(nd, eq="mhd", rec="tvd", ta="rk3", lim="mc", plim="mm", precision="fp64",
          problem="blast"):
    p = Params(ndims=nd, equation_system=eq, reconstruction=rec, limiter=lim,
               riemann_solver="hlld" if eq == "mhd" else "hllc", time_advance=ta,
               gamma=5.0 / 3.0 if eq == "mhd" else 1.4, cfl=0.4, ncells=8,
               nghosts=4 if rec in ("weno5z", "ppm") else 2, xblocks=2, yblocks=2, zblocks=2,
               xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5,
               prolongation_limiter=plim, precision=precision)
    if problem == "blast":
        meta = mesh.refine(p, mesh.near_sphere(p, 0.15, 0.05), 3)
        u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
    else:
        meta = mesh.refine(p, mesh.near_plane(p, 0.0, 0.05), 3)
        u0 = mesh.block_state(p, meta, mesh.kh_prim(p))
    return p, meta, u0


def relerr(a, b):
    """Per-variable max |a - b| over the variable's largest magnitude (vector
    components share their vector's)."""
    ax = tuple(range(b.ndim - 1))
    mag = np.abs(b).max(axis=ax)
    scale = mag.copy()
    scale[1:4] = mag[1:4].max()
    if b.shape[-1] == 9:
        scale[4:7] = mag[4:7].max()
        scale[7] = max(mag[7], mag[4:7].max())
    return np.abs(a - b).max(axis=ax) / np.maximum(scale, 1e-300)


@pytest.mark.parametrize("nd", [2, 3])
def test_ghost_fill_matches_oracle(nd):
    """q of every block with its ghost layers after the fill: same-level
    copies exact, restrictions and prolongations to round-off."""
    p, meta, u0 = _case(nd)
    print(f"[case] ghosts {nd}D blocks={len(meta)}", flush=True)
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    _, qo = o.get_blocks(with_q=True)
    with AmrSolver(p, meta) as s:
        info = s.info()
        assert info["virtual_blocks"] > 0 and info["prolonged"] > 0 and info["restricted"] > 0
        s.upload(u0)
        _, qg = s.download(with_q=True)
    qo = qo.reshape(qg.shape)
    err = relerr(qg.reshape(-1, p.nv), qo.reshape(-1, p.nv))
    print(f"  {info}: ghost q max rel {err.max():.2e}")
    assert err.max() <= 1e-14


CASES = [(2, "mhd", "tvd", "rk3", "mc", "mm"), (2, "hd", "weno5z", "rk2", "mm", "mc"),
         (2, "mhd", "tvd", "ssprk(4,3)", "vl", "sb"), (2, "hd", "tvd", "euler", "sb", "vl"),
         (3, "mhd", "tvd", "rk3", "mc", "mm"), (3, "hd", "weno5z", "rk3", "mm", "mm"),
         (2, "mhd", "limo3", "rk3", "mm", "mm"), (2, "hd", "weno3", "rk2", "mm", "mc"),
         (3, "mhd", "limo3", "rk2", "mm", "vl")]


@pytest.mark.parametrize("case", CASES)
def test_one_step_vs_oracle(case):
    nd, eq, rec, ta, lim, plim = case
    p, meta, u0 = _case(nd, eq, rec, ta, lim, plim)
    print(f"[case] {case} blocks={len(meta)}", flush=True)
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    tr = o.run(1)
    uo = o.get_blocks()
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        dt, cmax = s.advance()
        ug = s.download()
    err = relerr(ug.reshape(-1, p.nv), uo.reshape(-1, p.nv))
    print(f"  dt {dt:.15e} vs {tr[0, 0]:.15e}; max rel {err.max():.2e}")
    assert abs(dt - tr[0, 0]) <= 1e-13 * tr[0, 0]
    assert err.max() <= 1e-12


@pytest.mark.parametrize("name", ["amr_mhd2d_blast_10", "amr_mhd3d_blast_4",
                                  "amr_limo3_mhd2d_blast_6"])
def test_reference_runs(name):
    """The reference's own three-level blast runs (golden fixtures)."""
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    p = Params.from_text(str(g["params"]), ndims=int(g["ndims"]))
    meta = g["meta"]
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
    print(f"[case] {name}", flush=True)
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        dts = s.run(int(g["nsteps"]))
        u = s.download()
    assert np.abs(dts - g["trace"][:, 0]).max() <= 1e-11 * g["trace"][:, 0].max()
    sums = u.reshape(u.shape[0], -1, u.shape[-1]).sum(axis=1)
    ref = g["block_sums"]
    err = np.abs(sums - ref).max(axis=0) / np.maximum(np.abs(ref).max(axis=0), 1e-300)
    print(f"  block sums max rel {err.max():.2e}")
    assert np.delete(err, 7).max() <= 1e-10
    if "u1" in g:
        e2 = relerr(u.reshape(-1, p.nv), g["u1"].reshape(-1, p.nv))
        print(f"  state max rel {e2.max():.2e}")
        assert e2.max() <= 1e-9


@pytest.mark.parametrize("nd", [2, 3])
def test_conservation(nd):
    """Mass, momentum, energy and B summed over the leaves with their cell
    volumes: constant to round-off over 20 steps (the flux correction makes
    the coarse side of every level boundary use the fine fluxes)."""
    p, meta, u0 = _case(nd)
    vol = np.array([np.prod(p.dx()) / float(2 ** ((L - 1) * nd)) for L in meta[:, 0]])

    def totals(u):
        return (u.reshape(u.shape[0], -1, u.shape[-1]).sum(axis=1) * vol[:, None]).sum(axis=0)
    print(f"[case] conservation {nd}D", flush=True)
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        s.run(20)
        u = s.download()
        ints = s.integrals()
    d = np.abs(totals(u) - totals(u0))[[0, 1, 2, 3, 4, 5, 6, 8]] / np.abs(u0).max()
    print(f"  drift {d.max():.2e}")
    assert d.max() <= 1e-13
    # the device integrals weigh the levels the same way
    assert ints["mass"] == pytest.approx(totals(u)[0], rel=1e-12)


@pytest.mark.parametrize("nd", [2, 3])
def test_statistics_vs_oracle(nd):
    """amun_dev_statistics on a multi-level mesh: level-weighted means,
    minima and maxima of the leaves against the oracle after 3 steps."""
    p, meta, u0 = _case(nd)
    print(f"[case] statistics {nd}D blocks={len(meta)}", flush=True)
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    o.run(3)
    so = o.statistics()
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        s.run(3)
        st = s.statistics()
    # each (mean, min, max) triple on the scale of its largest magnitude
    scale = np.repeat(np.abs(so).reshape(7, 3).max(axis=1), 3) + 1e-300
    assert (np.abs(st - so) <= 1e-12 * scale).all(), (st, so)


def test_mixed_precision_vs_fp64():
    """FP32 reconstruction, fluxes and ghost fill with the FP64 state: 3D
    blast on three levels, 10 steps, density L1 against the FP64 run."""
    res = []
    for prec in ("fp64", "mixed"):
        p, meta, u0 = _case(3, precision=prec)
        print(f"[case] mixed {prec}", flush=True)
        with AmrSolver(p, meta) as s:
            s.upload(u0)
            s.run(10)
            res.append(s.download())
    l1 = np.abs(res[1][..., 0] - res[0][..., 0]).mean() / np.abs(res[0][..., 0]).mean()
    print(f"  density L1 mixed vs fp64 {l1:.2e}")
    assert l1 <= 1e-6


def test_graph_replay_bitwise():
    """Steps replayed from captured graphs equal the launched ones."""
    p, meta, u0 = _case(2, problem="kh")
    out = []
    for graph in (False, True):
        print(f"[case] graph={graph}", flush=True)
        with AmrSolver(p, meta) as s:
            s.set_graph(graph)
            s.upload(u0)
            s.run(6)
            out.append(s.download())
    np.testing.assert_array_equal(out[0], out[1])


# ----------------------------------------------- the batched/AMR configuration
def c3_amr_case(precision="fp64", nb=8, r=0.3):
    """The 3D block-AMR run of the batched/AMR configuration: three levels of
    16^3 blocks (+2 ghosts) on a periodic nb^3 base, refined towards a sphere
    of radius r (6224 leaves: 280 / 1272 / 4672 on levels 1 / 2 / 3 for
    nb = 8, r = 0.3), MHD blast with noise, TVD-MC + HLLD + SSPRK(3,3),
    CFL 0.4."""
    p = Params(ndims=3, equation_system="mhd", reconstruction="tvd", limiter="mc",
               riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.4,
               ncells=16, nghosts=2, xblocks=nb, yblocks=nb, zblocks=nb, xmin=-0.5, xmax=0.5,
               ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5, precision=precision)
    meta = mesh.refine(p, mesh.near_sphere(p, r, 0.015), 3)
    return p, meta


def _volumes(p, meta):
    return np.array([np.prod(p.dx()) / float(2 ** ((L - 1) * p.ndims)) for L in meta[:, 0]])


def test_c3_amr_full_size():
    """6224 leaves of 16^3 on three levels (25.5 M cells): FP64 and mixed
    runs of 3 steps.  Conservation of mass, momentum, energy and B over the
    leaves to round-off, the graph-replayed step bit for bit with the
    launched one, and the mixed run's density within 1e-6 (L1) of FP64."""
    p, meta = c3_amr_case()
    assert len(meta) == 6224 and np.bincount(meta[:, 0]).tolist() == [0, 280, 1272, 4672]
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p, r0=0.2))
    vol = _volumes(p, meta)

    def totals(u):
        return (u.reshape(u.shape[0], -1, u.shape[-1]).sum(axis=1) * vol[:, None]).sum(axis=0)
    res = {}
    for prec, graph in (("fp64", False), ("fp64", True), ("mixed", True)):
        p.precision = prec
        print(f"[case] c3 amr {prec} graph={graph} leaves={len(meta)}", flush=True)
        with AmrSolver(p, meta) as s:
            s.set_graph(graph)
            s.upload(u0)
            dts = s.run(3)
            res[(prec, graph)] = (s.download(), dts)
    u, dts = res[("fp64", False)]
    d = np.abs(totals(u) - totals(u0))[[0, 1, 2, 3, 4, 5, 6, 8]] / np.abs(u0).max()
    print(f"  dt {dts}, conservation drift {d.max():.2e}")
    assert d.max() <= 1e-12
    np.testing.assert_array_equal(res[("fp64", True)][0], u)
    um = res[("mixed", True)][0]
    l1 = np.abs(um[..., 0] - u[..., 0]).mean() / np.abs(u[..., 0]).mean()
    print(f"  density L1 mixed vs fp64 {l1:.2e}")
    assert l1 <= 1e-6


def test_c3_amr_uniform_state_preserved():
    """A uniform moving state on the 6224-leaf mesh stays uniform through the
    restrictions, limited prolongations and flux corrections."""
    p, meta = c3_amr_case()
    q = np.array([1.3, 0.2, -0.1, 0.05, 0.5, 0.3, 0.2, 0.0, 0.8])
    u0 = mesh.block_state(p, meta, lambda x, y, z: q.reshape(-1, *([1] * x.ndim)) *
                          np.ones((1,) + x.shape))
    print(f"[case] c3 amr uniform leaves={len(meta)}", flush=True)
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        s.run(2)
        u = s.download()
    err = np.abs(u - u0).max(axis=tuple(range(u.ndim - 1))) / np.abs(u0).max()
    print(f"  max change {err.max():.2e}")
    assert err.max() <= 1e-14


def test_c3_amr_blocks_one_step_vs_oracle():
    """The same 16^3 blocks on three levels on a 4^3 base (176 leaves:
    56 / 56 / 64, refined towards a sphere of radius 0.1), one step against
    the oracle."""
    p, meta = c3_amr_case(nb=4, r=0.1)
    assert np.bincount(meta[:, 0]).tolist() == [0, 56, 56, 64]
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p, r0=0.1))
    print(f"[case] c3 amr small leaves={len(meta)}", flush=True)
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    tr = o.run(1)
    uo = o.get_blocks()
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        dt, cmax = s.advance()
        ug = s.download()
    err = relerr(ug.reshape(-1, p.nv), uo.reshape(-1, p.nv))
    print(f"  dt {dt:.15e} vs {tr[0, 0]:.15e}; max rel {err.max():.2e}")
    assert abs(dt - tr[0, 0]) <= 1e-13 * tr[0, 0]
    assert err.max() <= 1e-12
