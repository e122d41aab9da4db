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


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("extra", [{"mlp_limiting": "on"}, {"glm_source_terms": "heglm"}])
def test_one_step_vs_oracle_mlp_glm(nd, extra):
    """mlp_limiting and the HEGLM source terms on a three-level mesh: the
    process reads the virtual blocks (and the fine-face fluxes of the flux
    correction apply it too); the source terms take each level's cell size."""
    p, meta, u0 = _case(nd)
    for k, v in extra.items():
        setattr(p, k, v)
    print(f"[case] {nd}D {extra} blocks={len(meta)}", flush=True)
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    tr = o.run(1)
    uo = o.get_blocks()
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        dt, cmax = s.advance()
        ug = s.download()
    err = relerr(ug.reshape(-1, p.nv), uo.reshape(-1, p.nv))
    print(f"  max rel {err.max():.2e}")
    assert abs(dt - tr[0, 0]) <= 1e-13 * tr[0, 0]
    assert err.max() <= 1e-12


@pytest.mark.parametrize("name", ["amr_mhd2d_blast_10", "amr_mhd3d_blast_4",
                                  "amr_limo3_mhd2d_blast_6", "amr_heglm_mhd2d_blast_6"])
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
    """The 3D block-AMR run of the batched/AMR configuration (6224 leaves of
    16^3 on three levels for nb = 8, r = 0.3): mesh.c3_amr_params."""
    return mesh.c3_amr_params(precision, nb, r)


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


# ----------------------------------------------------------- refinement
REFINE = ["refine_mhd2d_dens_pres", "refine_mhd2d_pres_vort", "refine_mhd2d_velx_magy",
          "refine_mhd3d_all"]


@pytest.mark.parametrize("name", REFINE)
def test_refinement_criterion_vs_reference(name):
    """amun_dev_refinement on the reference's three-level blast meshes: on the
    oracle's final state (bit for bit the reference's) the device criterion
    equals the reference's block for block and its indicators the oracle's
    bit for bit; after the same steps run on the device (a state within
    round-off of the reference's) the criterion is the reference's again and
    the indicators agree to float round-off."""
    from amun_amd.config import refine_cfg
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    p = Params.from_text(str(g["params"]), ndims=int(g["ndims"]))
    meta = g["meta"]
    nsteps = int(g["nsteps"])
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    o.run(nsteps)
    rc = refine_cfg(p)
    co, io = o.refinement(rc)
    np.testing.assert_array_equal(co, g["crit"])
    print(f"[case] refinement {name} blocks={len(meta)}", flush=True)
    with AmrSolver(p, meta) as s:
        s.upload(o.get_blocks())
        cg, ig = s.refinement(rc)
        np.testing.assert_array_equal(cg, g["crit"])
        np.testing.assert_array_equal(ig, io)
        s.upload(u0)
        s.run(nsteps)
        cr, ir = s.refinement(rc)
    np.testing.assert_array_equal(cr, g["crit"])
    np.testing.assert_allclose(ir, io, rtol=1e-5, atol=1e-7)


def test_refinement_level_limits():
    """minlev / maxlev as update_mesh applies them (mesh.F90:1547-1550): a
    block below minlev refines, at minlev it does not derefine, at maxlev it
    does not refine, above maxlev it derefines."""
    g = np.load(os.path.join(GOLD, "refine_mhd2d_dens_pres.npz"), allow_pickle=False)
    p = Params.from_text(str(g["params"]), ndims=2)
    meta = g["meta"]
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        c0, _ = s.refinement()
        c1, _ = s.refinement(minlev=2, maxlev=2)
    lev = meta[:, 0]
    assert (c1[lev < 2] == 1).all()
    np.testing.assert_array_equal(c1[lev == 2], np.minimum(0, np.maximum(0, c0[lev == 2])))
    assert (c1[lev > 2] == -1).all()


# ----------------------------------------------------------- mesh update
REMESH = ["remesh_mhd2d_l3", "remesh_mhd2d_l4", "remesh_mhd3d_l3"]


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def _edge_mask(p, meta, meta2):
    """Per new block and cell: True where the block is a child of an old
    leaf and the parent cell's slope stencil reaches outside the parent (it
    then reads prim2cons of the primitive ghost values)."""
    nc, nd = p.ncells, p.ndims
    old = {tuple(m) for m in meta.tolist()}
    shp = (nc, nc) if nd == 2 else (nc, nc, nc)
    mask = np.zeros((len(meta2),) + shp, dtype=bool)
    idx = np.indices(shp)                       # (k,) j, i order of the blocks' arrays
    for b, (L, x, y, z) in enumerate(meta2.tolist()):
        if (L, x, y, z) in old:
            continue
        par = (L - 1, (x // nc // 2) * nc, (y // nc // 2) * nc, (z // nc // 2) * nc if nd == 3 else 0)
        if par not in old:
            continue
        off = [(x // nc) & 1, (y // nc) & 1, (z // nc) & 1]
        m = np.zeros(shp, dtype=bool)
        for a in range(nd):                   # axis a of (x, y, z) is array axis nd - 1 - a
            c = (idx[nd - 1 - a] + off[a] * nc) // 2
            m |= (c == 0) | (c == nc - 1)
        mask[b] = m
    return mask


@pytest.mark.parametrize("name", REMESH)
def test_remesh_reference_update(name):
    """amun_dev_amr_remesh onto the leaf set of one reference update_mesh
    (copies, prolong_block and restrict_block all occur).  From the oracle's
    state before the update (bit for bit the reference's): kept and restricted
    blocks, and prolonged cells whose parent stencil stays inside the parent,
    are the reference's bit for bit; prolonged cells whose parent stencil
    reads a ghost cell (prim2cons of the device's primitive ghosts, whose
    cons2prim rounds differently) within 1e-14.  From the device's own run,
    the host selection (mesh.update_leaves) from the device criterion gives
    the reference's leaves, and the run continues on the new mesh as the
    oracle's does."""
    from amun_amd.config import refine_cfg
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    p = Params.from_text(str(g["params"]), ndims=int(g["ndims"]))
    meta, meta2 = g["meta"], g["meta2"]
    nsteps = int(g["nsteps"])
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p, shift=tuple(g["shift"])))
    o = ol.OracleAmr(p, meta)
    o.set_blocks(u0)
    o.run(nsteps)
    uo = o.get_blocks()
    assert _sha(uo) == str(g["sha_u1"])
    ref = o.remesh(meta2).get_blocks()
    assert _sha(ref) == str(g["sha_u2"])
    edge = _edge_mask(p, meta, meta2)
    print(f"[case] remesh {name} blocks {len(meta)} -> {len(meta2)}", flush=True)
    with AmrSolver(p, meta) as s:
        s.upload(uo)
        s.remesh(meta2)
        ud = s.download()
    same = ud == ref
    print(f"  from the reference's state: {int((~same).any(axis=-1).sum())} of "
          f"{same[..., 0].size} cells differ, {int(edge.sum())} cells read ghosts")
    assert same[~edge].all()
    e = relerr(ud[edge], ref[edge]) if edge.any() else np.zeros(1)
    print(f"  ghost-reading cells max rel {e.max():.2e}")
    assert e.max() <= 1e-14
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        s.run(nsteps)
        crit, _ = s.refinement(refine_cfg(p))
        sel = mesh.update_leaves(p, meta, crit, 1, int(p.extra["maxlev"]))
        assert {tuple(m) for m in sel.tolist()} == {tuple(m) for m in meta2.tolist()}
        ud = s.download()
        t0, k0 = s.time, s.step
        s.remesh(meta2)
        assert s.time == t0 and s.step == k0 == nsteps
        un = s.download()
        on = ol.OracleAmr(p, meta)
        on.set_blocks(ud)
        on = on.remesh(meta2)
        e = relerr(un.reshape(-1, p.nv), on.get_blocks().reshape(-1, p.nv))
        print(f"  from the device's run: vs the oracle's update of that state max rel {e.max():.2e}")
        assert e.max() <= 1e-14
        dts = s.run(2)
        on.run(2)
        e2 = relerr(s.download().reshape(-1, p.nv), on.get_blocks().reshape(-1, p.nv))
        print(f"  two more steps: dt {dts}, vs oracle max rel {e2.max():.2e}")
        assert np.delete(e2, 7).max() <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_remesh_identity_continues_bitwise(precision):
    """A mesh update onto the same leaves (every block copied, in a new
    order) changes nothing: the run continues bit for bit as the one that
    was not updated (state, dt, time; the FP32 compensation of the mixed
    mode is carried with its blocks)."""
    p, meta, u0 = _case(2, precision=precision)
    perm = np.random.default_rng(4).permutation(len(meta))
    with AmrSolver(p, meta) as a, AmrSolver(p, meta) as b:
        a.upload(u0)
        b.upload(u0)
        a.run(3)
        b.run(3)
        b.remesh(meta[perm])
        da, db = a.run(3), b.run(3)
        np.testing.assert_array_equal(db, da)
        np.testing.assert_array_equal(b.download(), a.download()[perm])
        assert a.time == b.time and a.step == b.step


def test_remesh_rejects_level_jumps_and_keeps_state():
    """A new leaf set that moves a block by two levels, or is not 2:1
    balanced, is refused (702 / 700) and the context runs on unchanged."""
    from amun_amd.solver import AmunError
    p, meta, u0 = _case(2)
    lv = meta[:, 0]
    b3 = int(np.flatnonzero(lv == 3)[0])
    L, x, y, z = (int(v) for v in meta[b3])
    bad = meta.copy()
    bad[b3] = (1, (x // 4) // p.ncells * p.ncells, (y // 4) // p.ncells * p.ncells, 0)
    with AmrSolver(p, meta) as a, AmrSolver(p, meta) as b:
        a.upload(u0)
        b.upload(u0)
        with pytest.raises(AmunError) as ei:
            b.remesh(bad)
        assert ei.value.code in (500, 700, 702)
        np.testing.assert_array_equal(b.run(2), a.run(2))
        np.testing.assert_array_equal(b.download(), a.download())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c3_amr_adapts_on_device(precision):
    """The C3 AMR mesh (6224 leaves of 16^3) adapting without a state
    transfer through the host: two steps, the device criterion, the host
    selection, the device mesh update, two more steps.  The volume integrals
    of mass, momentum, energy and B are conserved through the update to
    round-off, refining every leaf of a region and derefining it again
    returns the state to round-off, and the update's duration is printed."""
    import time
    from amun_amd.config import refine_cfg
    p, meta = c3_amr_case(precision)
    p.extra.update({"refinement_variables": "dens pres", "crefmin": 0.2, "crefmax": 0.5,
                    "epsref": 0.5})
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p, r0=0.2, shift=(0.05, 0.03, 0.0)))
    print(f"[case] c3 amr adapt {precision} leaves={len(meta)}", flush=True)

    def totals(u, m):
        vol = _volumes(p, m)
        return (u.reshape(u.shape[0], -1, u.shape[-1]).sum(axis=1) * vol[:, None]).sum(axis=0)
    with AmrSolver(p, meta) as s:
        s.upload(u0)
        s.run(2)
        crit, _ = s.refinement(refine_cfg(p))
        new = mesh.update_leaves(p, meta, crit, 1, 3)
        ub = s.download()
        s.sync()
        t = time.perf_counter()
        s.remesh(new)
        s.sync()
        ms = (time.perf_counter() - t) * 1e3
        ua = s.download()
        lv0, lv1 = np.bincount(meta[:, 0]), np.bincount(new[:, 0])
        d = np.abs(totals(ua, new) - totals(ub, meta))[[0, 1, 2, 3, 4, 5, 6, 8]]
        print(f"  leaves per level {lv0[1:].tolist()} -> {lv1[1:].tolist()}, update "
              f"{ms:.1f} ms (host tables + one launch), conservation {d.max() / np.abs(ub).max():.2e}")
        assert (lv1[1:4] != lv0[1:4]).any()
        assert d.max() <= 1e-12 * np.abs(ub).max()
        dts = s.run(2)
        assert np.isfinite(s.download()).all() and (dts > 0).all()
        # refine eight level-1 leaves (their neighbours are on levels 1 and 2,
        # so the children keep the 2:1 balance) and derefine them again
        lev1 = [b for b in range(len(new)) if new[b, 0] == 1][:8]
        kids = []
        nc = p.ncells
        for b in lev1:
            _, x, y, z = (int(v) for v in new[b])
            kids += [(2, 2 * x + (q & 1) * nc, 2 * y + (q >> 1 & 1) * nc, 2 * z + (q >> 2 & 1) * nc)
                     for q in range(8)]
        fine = np.array([tuple(m) for i, m in enumerate(new.tolist()) if i not in lev1] + kids,
                        dtype=np.int32)
        uc = s.download()
        s.remesh(fine)
        s.remesh(new)
        ur = s.download()
        e = relerr(ur.reshape(-1, p.nv), uc.reshape(-1, p.nv))
        print(f"  refine + derefine of {len(lev1)} leaves: max rel {e.max():.2e}")
        assert e.max() <= 1e-14
