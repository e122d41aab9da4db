#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ from the reference binaries
(oracle/_ref, built from the unmodified reference sources by
oracle/build_ref.sh).  Run in the container that holds the reference:

    python tests/golden/make_golden.py        # every fixture
    python tests/golden/make_golden.py sr     # the relativistic ones only
    python tests/golden/make_golden.py gp     # the Gaussian-process ones only

Every fixture stores the configuration (as AMUN parameter text), the seed,
the reference outputs needed by the tests (sha256 of the exact bytes, the
(dt, cmax) trace, a strided subsample of the final state, or full small
arrays for the unit-level vectors).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from amun_amd.config import Params, c1_params, c2_params, c4_params  # noqa: E402
from amun_amd import problems  # noqa: E402
import oracle_lib as ol  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def run_cases():
    """Full runs: (name, params, initial-state builder, steps, stride)."""
    p_c1 = c1_params()
    p_c2 = c2_params(32)
    p_ot = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
                  riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.4,
                  ncells=16, nghosts=2, xblocks=4, yblocks=4)
    p_kh = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
                  riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.4,
                  ncells=16, nghosts=2, xblocks=4, yblocks=4, xmin=-0.5, xmax=0.5,
                  ymin=-0.5, ymax=0.5)
    p_hd3 = Params(ndims=3, equation_system="hd", reconstruction="weno5z", limiter="mm",
                   riemann_solver="hllc", time_advance="rk2", gamma=1.4, cfl=0.4,
                   ncells=8, nghosts=4, xblocks=3, yblocks=2, zblocks=2, xmin=-0.5,
                   xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5)
    return [
        ("c1_blast_hd2d_128_200", p_c1, problems.blast_hd_2d, 200, 4),
        ("c2_mhd3d_weno5z_hlld_32_4", p_c2, problems.turbulent_mhd_3d, 4, 4),
        ("orszag_tang_mhd2d_64_30", p_ot, problems.orszag_tang_2d, 30, 4),
        ("kh_mhd2d_64_30", p_kh, problems.kelvin_helmholtz_2d, 30, 4),
        ("blast_hd3d_weno5z_hllc_4", p_hd3, None, 4, 2),
    ]


def hd3_blast(p):
    nz, ny, nx = p.shape
    z, y, x = problems.cell_centres(p)
    r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2)
    q = np.zeros((5, nz, ny, nx))
    q[0] = 1.0 + 0.01 * (2 * problems.uniform(7, nz * ny * nx).reshape(nz, ny, nx) - 1)
    q[4] = np.where(r < 0.2, 10.0, 0.1)
    return problems.prim_to_cons(p, q)


def sr_states(rng, n):
    """Random special-relativistic primitive states (n, 9): |v| < 0.99,
    densities and pressures over decades, field strengths from weak to
    dominant; the first rows have v = 0, then Bx = 0, then both, then B = 0."""
    q = np.zeros((n, 9))
    q[:, 0] = 10.0 ** rng.uniform(-2, 1, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    q[:, 1:4] = d * rng.uniform(0.0, 0.99, n)[:, None]
    q[:, 4:7] = rng.normal(size=(n, 3)) * 10.0 ** rng.uniform(-2, 0.5, (n, 1))
    q[:, 7] = rng.normal(scale=0.1, size=n)
    q[:, 8] = 10.0 ** rng.uniform(-3, 1, n)
    k = n // 16
    q[:k, 1:4] = 0.0
    q[k:2 * k, 4] = 0.0
    q[2 * k:3 * k, 1:4] = 0.0
    q[2 * k:3 * k, 4] = 0.0
    q[3 * k:3 * k + 4, 4:7] = 0.0
    return q


def algebra_sets(rng, n):
    """Coefficient sets a(5) (increasing powers) for the root finders:
    random magnitudes over decades, zero constant / leading terms, and
    polynomials with known multiple roots."""
    a = rng.normal(size=(n, 5)) * 10.0 ** rng.integers(-3, 4, (n, 5))
    a[10:40, 0] = 0.0
    a[40:70, 4] = 0.0
    a[70:90, 3:] = 0.0
    a[90:100, 2:] = 0.0
    known = [(1, 1, 2, -3), (2, 2, 2, 2), (0.5, -0.5, 1, -1), (1, 2, 3, 4), (-1, -1, -1, 5)]
    for i, r in enumerate(known):
        a[i] = np.poly(r)[::-1]
    a[5, :4] = np.poly([2, 2, 2])[::-1]
    a[5, 4] = 0.0
    a[6, :3] = np.poly([3, 3])[::-1]
    a[6, 3:] = 0.0
    return a


def main_sr():
    """Fixtures of the special-relativistic MHD path (own generator stream,
    so the other fixtures are unaffected)."""
    rng = np.random.default_rng(20261203)
    a = algebra_sets(rng, 600)
    np.savez_compressed(os.path.join(HERE, "unit_algebra.npz"), a=a, r=ol.ref_algebra(a))

    pe = Params(ndims=2, equation_system="srmhd", riemann_solver="hll",
                tolerance=1.0e-12, nr_maxit=50)
    q = sr_states(rng, 256)
    u, q2, f, cm, cp = ol.ref_eqs(pe, q, cmax=0.8)
    # conserved states for which the primitive solver fails (energy below
    # the rest mass) or copes (large momenta); the output starts at -7
    ub = u[:64].copy()
    ub[:16, 8] = -2.0 * ub[:16, 0]
    ub[16:32, 1:4] *= 50.0
    q0 = np.full_like(ub, -7.0)
    qb, _ = ol.ref_cons2prim(pe, ub, q0)
    np.savez_compressed(os.path.join(HERE, "unit_eqs_srmhd.npz"), q=q, u=u, q2=q2, f=f,
                        cm=cm, cp=cp, cmax=0.8, ub=ub, q0=q0, qb=qb,
                        params=np.array(pe.to_params_in()))

    pf = Params(ndims=2, equation_system="srmhd", reconstruction="tvd", limiter="zero",
                riemann_solver="hll", gamma=5.0 / 3.0, ncells=8, nghosts=2,
                tolerance=1.0e-12, nr_maxit=50)
    nb = 3
    qb = sr_states(rng, nb * 144).reshape(nb, 1, 12, 12, 9)
    f = ol.ref_update_flux(pf, qb, cmax=1.0)
    np.savez_compressed(os.path.join(HERE, "unit_flux_srmhd_hll.npz"), q=qb, f=f,
                        cmax=1.0, params=np.array(pf.to_params_in()))

    p4 = c4_params(64)
    u0 = problems.srmhd_blast_2d(p4, noise=0.01)
    nsteps, stride = 100, 4
    u, tr = ol.ref_run(p4, u0, nsteps)
    name = "c4_srmhd2d_64_100"
    np.savez_compressed(os.path.join(HERE, name + ".npz"),
                        params=np.array(p4.to_params_in()), ndims=2, nsteps=nsteps,
                        stride=stride, trace=tr, sub=u[:, :, ::stride, ::stride],
                        sha_u0=sha(u0), sha_u=sha(u))
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)
    print(name, p4.shape, "dt", tr[-1, 0])


def main_gp():
    """Fixtures of the Gaussian-process reconstruction (own generator)."""
    rng = np.random.default_rng(20261204)
    pg = Params(ndims=3, equation_system="mhd", reconstruction="gp", limiter="mm",
                ncells=16, nghosts=4, cfl=0.3)
    fp = np.concatenate([rng.normal(size=(8, 24)),
                         np.cumsum(rng.normal(size=(4, 24)), axis=1),
                         np.sin(np.linspace(0, 6, 24))[None, :] * np.ones((2, 1)),
                         np.where(np.arange(24) < 12, 1.0, 0.1)[None, :] * np.ones((2, 1)),
                         np.linspace(0, 3, 24)[None, :] ** 2])
    fl, fr = ol.ref_reconstruct(pg, fp)
    np.savez_compressed(os.path.join(HERE, "unit_gp.npz"), fp=fp, fl=fl, fr=fr,
                        params=np.array(pg.to_params_in()))
    p3 = c2_params(32)
    p3.reconstruction = "gp"
    p3.xblocks, p3.yblocks, p3.zblocks, p3.ncells = 3, 2, 2, 8
    u0 = problems.turbulent_mhd_3d(p3, seed=12)
    nsteps, stride = 3, 2
    u, tr = ol.ref_run(p3, u0, nsteps)
    name = "gp_mhd3d_hlld_3"
    np.savez_compressed(os.path.join(HERE, name + ".npz"),
                        params=np.array(p3.to_params_in()), ndims=3, nsteps=nsteps,
                        stride=stride, trace=tr, sub=u[:, ::stride, ::stride, ::stride],
                        sha_u0=sha(u0), sha_u=sha(u))
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)
    print(name, p3.shape, "dt", tr[-1, 0])


SSPRK_RUNS = [("ssprk(m,2)", 4), ("ssprk(4,3)", 2), ("ssprk(5,3)", 2)]


def ssprk_params(ta, stages):
    return Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
                  riemann_solver="hlld", time_advance=ta, gamma=5.0 / 3.0, cfl=0.3,
                  ncells=16, nghosts=2, xblocks=2, yblocks=2, stages=stages)


def main_ssprk():
    """Orszag-Tang runs of the three SSPRK families (own generator)."""
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for ta, stages in SSPRK_RUNS:
        p = ssprk_params(ta, stages)
        u0 = problems.orszag_tang_2d(p, seed=9)
        nsteps, stride = 10, 3
        u, tr = ol.ref_run(p, u0, nsteps)
        name = "ssprk_" + ta.replace("(", "").replace(")", "").replace(",", "") + "_ot2d_32_10"
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=2, nsteps=nsteps,
                            stride=stride, trace=tr, sub=u[:, :, ::stride, ::stride],
                            sha_u0=sha(u0), sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, p.shape, "dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


def amr_params(nd):
    """Blast on a three-level mesh that the reference's generate_mesh refines
    by its second-derivative indicator (crefmin/crefmax on density and
    pressure), frozen for the run."""
    p = Params(ndims=nd, equation_system="mhd", reconstruction="tvd", limiter="mc",
               riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.4,
               ncells=8, nghosts=2, xblocks=2, yblocks=2, zblocks=2,
               xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5)
    p.extra.update({"problem": "blast", "maxlev": 3, "radius": 0.15, "ratio": 100.0,
                    "buni": 1.0, "refinement_variables": "dens pres", "crefmin": 0.1,
                    "crefmax": 0.3})
    return p


def main_amr():
    """Multi-level runs of the reference: mesh from its generate_mesh, the
    state replaced by this project's seeded blast (amun_amd.mesh), steps on
    the frozen mesh with restriction, prolongation and flux correction."""
    from amun_amd import mesh
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for nd, nsteps in ((2, 10), (3, 4)):
        p = amr_params(nd)
        meta, _, _, _, _ = ol.ref_amr(p, 1)
        u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
        m0, u0r, m1, u1, tr = ol.ref_amr(p, nsteps, u_in=u0)
        assert np.array_equal(m0, meta) and np.array_equal(u0r, u0)
        name = f"amr_mhd{nd}d_blast_{nsteps}"
        sums = u1.reshape(u1.shape[0], -1, u1.shape[-1]).sum(axis=1)
        extra = {"u1": u1} if nd == 2 else {}
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            meta=meta, trace=tr, block_sums=sums, sha_u0=sha(u0),
                            sha_u=sha(u1), **extra)
        index[name] = {"steps": nsteps, "sha_u": sha(u1), "dt_last": float(tr[-1, 0]),
                       "blocks": int(meta.shape[0])}
        print(name, meta.shape[0], "blocks, dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


REFINE_CASES = (
    ("refine_mhd2d_dens_pres", 2, 10, {"refinement_variables": "dens pres", "crefmin": 0.05,
                                       "crefmax": 0.3, "epsref": 0.5}),
    ("refine_mhd2d_pres_vort", 2, 10, {"refinement_variables": "pres vort", "crefmin": 0.05,
                                       "crefmax": 0.3, "epsref": 0.5, "vortmin": 0.01,
                                       "vortmax": 0.03}),
    ("refine_mhd2d_velx_magy", 2, 10, {"refinement_variables": "dens velx magy vort jabs",
                                       "crefmin": 0.05, "crefmax": 0.3, "epsref": 0.5,
                                       "vortmin": 0.005, "vortmax": 0.02, "jabsmin": 0.003,
                                       "jabsmax": 0.01}),
    ("refine_mhd3d_all", 3, 4, {"refinement_variables": "dens pres vort jabs", "crefmin": 0.05,
                                "crefmax": 0.3, "epsref": 0.5, "vortmin": 0.02, "vortmax": 0.1,
                                "jabsmin": 0.02, "jabsmax": 0.1}),
)


def main_refine():
    """check_refinement_criterion (refinement.F90:257-462) of every block of
    the reference's three-level blast meshes after a few steps, with several
    sets of refinement variables and thresholds (own generator)."""
    from amun_amd import mesh
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for name, nd, nsteps, rp in REFINE_CASES:
        p = amr_params(nd)
        p.extra.update(rp)
        meta, _, _, _, _ = ol.ref_amr(p, 1)
        u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
        m0, _, m1, u1, tr, crit = ol.ref_amr(p, nsteps, u_in=u0, criterion=True)
        assert np.array_equal(m1, meta)
        # the state after the steps is the oracle's (bit for bit, pinned by
        # sha_u1); the fixture keeps the mesh and the criterion
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            meta=meta, crit=crit, sha_u1=sha(u1))
        index[name] = {"steps": nsteps, "blocks": int(meta.shape[0]),
                       "crit": [int((crit == c).sum()) for c in (-1, 0, 1)]}
        print(name, meta.shape[0], "blocks, criterion -1/0/1:", index[name]["crit"])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


# (name, ndims, maxlev, steps, blast centre shift, refinement parameters):
# the reference's mesh is generated for the blast at the domain centre, the
# state is replaced by the seeded blast moved by the shift, and after the
# steps one update_mesh refines around the new blast and derefines around
# the old position
REMESH_CASES = (
    ("remesh_mhd2d_l3", 2, 3, 2, (0.25, 0.15, 0.0), {"crefmin": 0.05, "crefmax": 0.3,
                                                     "epsref": 0.5}),
    ("remesh_mhd2d_l4", 2, 4, 2, (0.3, 0.2, 0.0), {"crefmin": 0.1, "crefmax": 0.4,
                                                   "epsref": 0.5}),
    ("remesh_mhd3d_l3", 3, 3, 1, (0.25, 0.15, 0.1), {"crefmin": 0.05, "crefmax": 0.3,
                                                     "epsref": 0.5}),
)


def remesh_params(nd, maxlev, rp):
    p = amr_params(nd)
    p.extra["maxlev"] = maxlev
    p.extra.update(rp)
    return p


def main_remesh():
    """One update_mesh of the reference (mesh.F90:756-833: the criterion,
    the neighbour and sibling rules, derefine_selected_blocks with
    restrict_block and refine_selected_blocks with prolong_block) after a
    few steps on its generated mesh, the blast moved away (own generator)."""
    from amun_amd import mesh
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for name, nd, maxlev, nsteps, shift, rp in REMESH_CASES:
        p = remesh_params(nd, maxlev, rp)
        meta, _, _, _, _ = ol.ref_amr(p, 0)
        u0 = mesh.block_state(p, meta, mesh.blast_prim(p, shift=shift))
        _, _, m1, u1, tr, crit, (m2, u2) = ol.ref_amr(p, nsteps, u_in=u0, criterion=True,
                                                      update=True)
        assert np.array_equal(m1, meta)
        sums = u2.reshape(u2.shape[0], -1, u2.shape[-1]).sum(axis=1)
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            shift=np.array(shift), meta=meta, crit=crit, meta2=m2,
                            block_sums2=sums, sha_u0=sha(u0), sha_u1=sha(u1), sha_u2=sha(u2))
        s0 = {tuple(r) for r in meta.tolist()}
        kept = sum(tuple(r) in s0 for r in m2.tolist())
        index[name] = {"steps": nsteps, "blocks": int(meta.shape[0]),
                       "blocks_after": int(m2.shape[0]), "kept": int(kept),
                       "crit": [int((crit == c).sum()) for c in (-1, 0, 1)]}
        print(name, index[name])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


# (name, ndims, glm source terms, reconstruction, steps): Orszag-Tang with
# noise on the field (2D) and the fully three-dimensional noisy state (3D),
# so that div B is far from zero
GLM_CASES = (("eglm_otnoisy2d_32_8", 2, "eglm", "tvd", 8),
             ("heglm_otnoisy2d_32_8", 2, "heglm", "tvd", 8),
             ("heglm_noisy3d_16_3", 3, "heglm", "weno5z", 3))


def glm_params(nd, glm, rec):
    p = Params(ndims=nd, equation_system="mhd", reconstruction=rec, limiter="mc",
               riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.3, ncells=8,
               nghosts=4 if rec == "weno5z" else 2, xblocks=4 if nd == 2 else 2,
               yblocks=4 if nd == 2 else 2, zblocks=2)
    p.glm_source_terms = glm
    return p


def glm_state(name, p):
    return noisy3d(p) if "noisy3d" in name else ot_noisy(p)


def main_glm():
    """The EGLM / HEGLM source terms (sources.F90:479-539): short runs of
    the reference (own generator), and one on its three-level blast mesh."""
    from amun_amd import mesh
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for name, nd, glm, rec, nsteps in GLM_CASES:
        p = glm_params(nd, glm, rec)
        u0 = glm_state(name, p)
        u, tr = ol.ref_run(p, u0, nsteps)
        st = 3 if nd == 2 else 2
        sub = u[:, :, ::st, ::st] if nd == 2 else u[:, ::st, ::st, ::st]
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            stride=st, trace=tr, sub=sub, sha_u0=sha(u0), sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    p = amr_params(2)
    p.glm_source_terms = "heglm"
    meta, _, _, _, _ = ol.ref_amr(p, 1)
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
    nsteps = 6
    m0, u0r, m1, u1, tr = ol.ref_amr(p, nsteps, u_in=u0)
    name = f"amr_heglm_mhd2d_blast_{nsteps}"
    sums = u1.reshape(u1.shape[0], -1, u1.shape[-1]).sum(axis=1)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), params=np.array(p.to_params_in()),
                        ndims=2, nsteps=nsteps, meta=meta, trace=tr, block_sums=sums,
                        sha_u0=sha(u0), sha_u=sha(u1))
    index[name] = {"steps": nsteps, "sha_u": sha(u1), "dt_last": float(tr[-1, 0]),
                   "blocks": int(meta.shape[0])}
    print(name, meta.shape[0], "blocks, dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


# (name, ndims, reconstruction, ghosts, steps): runs with mlp_limiting on
MLP_CASES = (("mlp_tvd_otnoisy2d_32_8", 2, "tvd", 2, 8),
             ("mlp_weno5z_otnoisy2d_32_8", 2, "weno5z", 4, 8),
             ("mlp_mgp_otnoisy2d_32_4", 2, "mgp", 4, 4),
             ("mlp_tvd_noisy3d_16_3", 3, "tvd", 2, 3))


def mlp_params(nd, rec, ng):
    p = Params(ndims=nd, equation_system="mhd", reconstruction=rec, limiter="mc",
               riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.3, ncells=8,
               nghosts=ng, xblocks=4 if nd == 2 else 2, yblocks=4 if nd == 2 else 2,
               zblocks=2)
    p.mlp_limiting = "on"
    return p


def mlp_state(name, p):
    return noisy3d(p) if "noisy3d" in name else ot_noisy(p)


def main_mlp():
    """The multi-dimensional limiting process (mlp_limiting,
    interpolations.F90:1171-1321) after TVD, WENO5-Z and MGP reconstructions:
    short runs of the reference (own generator)."""
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for name, nd, rec, ng, nsteps in MLP_CASES:
        p = mlp_params(nd, rec, ng)
        u0 = mlp_state(name, p)
        u, tr = ol.ref_run(p, u0, nsteps)
        st = 3 if nd == 2 else 2
        sub = u[:, :, ::st, ::st] if nd == 2 else u[:, ::st, ::st, ::st]
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            stride=st, trace=tr, sub=sub, sha_u0=sha(u0), sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


def main_hllc_mhd():
    """riemann_mhd_adi_hllc through update_flux (zero limiter), random states
    with weak, zero and strong Bx, and an Orszag-Tang run (own generator)."""
    rng = np.random.default_rng(20261206)
    pf = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="zero",
                riemann_solver="hllc", gamma=5.0 / 3.0, ncells=8, nghosts=2)
    qb = np.zeros((4, 1, 12, 12, 9))
    qb[..., 0] = rng.uniform(0.2, 2.0, qb.shape[:-1])
    qb[..., 1:4] = rng.normal(scale=1.5, size=qb.shape[:-1] + (3,))
    qb[..., 8] = rng.uniform(0.1, 2.0, qb.shape[:-1])
    qb[..., 4:7] = rng.normal(size=qb.shape[:-1] + (3,))
    qb[..., 7] = rng.normal(scale=0.1, size=qb.shape[:-1])
    qb[1, ..., 4] = 1e-6 * rng.normal(size=qb.shape[1:-1])
    qb[2, ..., 4] = 0.0
    qb[2, ..., 7] = 0.0                 # Bx* = 0 exactly: the Bx = 0 branch
    qb[3, ..., 1] = qb[3, ..., 1] * 5.0
    f = ol.ref_update_flux(pf, qb, cmax=2.3)
    np.savez_compressed(os.path.join(HERE, "unit_flux_mhd_hllc.npz"), q=qb, f=f, cmax=2.3,
                        params=np.array(pf.to_params_in()))
    p = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
               riemann_solver="hllc", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.4,
               ncells=16, nghosts=2, xblocks=2, yblocks=2)
    u0 = problems.orszag_tang_2d(p, seed=8)
    nsteps = 10
    u, tr = ol.ref_run(p, u0, nsteps)
    name = f"hllc_mhd_ot2d_32_{nsteps}"
    np.savez_compressed(os.path.join(HERE, name + ".npz"), params=np.array(p.to_params_in()),
                        ndims=2, nsteps=nsteps, stride=3, trace=tr, sub=u[:, :, ::3, ::3],
                        sha_u0=sha(u0), sha_u=sha(u))
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)
    print(name, "dt", tr[-1, 0])


def main_roe_hd():
    """riemann_hd_adi_roe through update_flux (zero limiter): random states,
    supersonic pairs on both sides (the pure left and right fluxes), Roe
    velocities of both signs and pairs at rest; and a blast run (own
    generator)."""
    rng = np.random.default_rng(20261207)
    pf = Params(ndims=2, equation_system="hd", reconstruction="tvd", limiter="zero",
                riemann_solver="roe", gamma=1.4, ncells=8, nghosts=2)
    qb = np.zeros((4, 1, 12, 12, 5))
    qb[..., 0] = rng.uniform(0.2, 2.0, qb.shape[:-1])
    qb[..., 1:4] = rng.normal(scale=1.5, size=qb.shape[:-1] + (3,))
    qb[..., 4] = rng.uniform(0.1, 2.0, qb.shape[:-1])
    qb[1, ..., 1] = qb[1, ..., 1] * 6.0           # supersonic both ways
    qb[2, ..., 1:4] = 0.0                         # gas at rest
    qb[3, ..., 1] = 1e-3 * rng.normal(size=qb.shape[1:-1])   # Roe velocity ~ 0
    f = ol.ref_update_flux(pf, qb, cmax=2.3)
    np.savez_compressed(os.path.join(HERE, "unit_flux_hd_roe.npz"), q=qb, f=f, cmax=2.3,
                        params=np.array(pf.to_params_in()))
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for nd, rec, ta, nsteps in ((2, "tvd", "rk2", 20), (3, "weno5z", "rk3", 4)):
        p = Params(ndims=nd, equation_system="hd", reconstruction=rec,
                   limiter="mc" if rec == "tvd" else "mm", riemann_solver="roe",
                   time_advance=ta, gamma=1.4, cfl=0.4 if nd == 2 else 0.3,
                   ncells=16 if nd == 2 else 8, nghosts=2 if rec == "tvd" else 4,
                   xblocks=2, yblocks=2, zblocks=2)
        u0 = problems.blast_hd_2d(p, seed=9) if nd == 2 else hd3_blast(p)
        u, tr = ol.ref_run(p, u0, nsteps)
        name = f"roe_hd{nd}d_blast_{nsteps}"
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            stride=3, trace=tr, sha_u0=sha(u0), sha_u=sha(u),
                            sub=u[:, ::3, ::3, ::3] if nd == 3 else u[:, :, ::3, ::3])
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


SCHEMES = ("weno5yc", "weno5ns", "mp5", "mp7")


def main_schemes():
    """WENO5-YC, WENO5-NS, MP5 and MP7: pencils and short Orszag-Tang runs
    of the reference (own generator)."""
    rng = np.random.default_rng(20261205)
    fp = np.concatenate([rng.normal(size=(6, 24)), np.cumsum(rng.normal(size=(4, 24)), axis=1),
                         np.sin(np.linspace(0, 6, 24))[None, :] * np.ones((2, 1)),
                         np.where(np.arange(24) < 12, 1.0, 0.1)[None, :] * np.ones((2, 1))])
    out = {"fp": fp}
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for rec in SCHEMES:
        pu = Params(ndims=3, equation_system="mhd", reconstruction=rec, limiter="mm",
                    ncells=16, nghosts=4, cfl=0.3)
        out["fl_" + rec], out["fr_" + rec] = ol.ref_reconstruct(pu, fp)
        p = Params(ndims=2, equation_system="mhd", reconstruction=rec, limiter="mm",
                   riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.3,
                   ncells=8, nghosts=4, xblocks=4, yblocks=4)
        u0 = problems.orszag_tang_2d(p, seed=6)
        nsteps = 8
        u, tr = ol.ref_run(p, u0, nsteps)
        name = f"{rec}_ot2d_32_{nsteps}"
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=2, nsteps=nsteps,
                            stride=3, trace=tr, sub=u[:, :, ::3, ::3], sha_u0=sha(u0),
                            sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    np.savez_compressed(os.path.join(HERE, "unit_schemes.npz"), **out)
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


O3_SCHEMES = ("weno3", "limo3", "ppm")


def main_negp():
    """Riemann solvers through update_flux (zero limiter) with a few cells of
    negative pressure: the reference's NaN wave speeds, its MIN/MAX order
    (a NaN first operand yields the second, a NaN second operand
    propagates) and the NaN fluxes that follow (own generator)."""
    rng = np.random.default_rng(20261209)
    for eq, rs in (("hd", "hll"), ("hd", "hllc"), ("hd", "roe"), ("mhd", "hll"),
                   ("mhd", "hlld"), ("mhd", "hllc")):
        nv = 5 if eq == "hd" else 9
        pf = Params(ndims=2, equation_system=eq, reconstruction="tvd", limiter="zero",
                    riemann_solver=rs, gamma=1.4 if eq == "hd" else 5.0 / 3.0, ncells=8,
                    nghosts=2)
        qb = np.zeros((2, 1, 12, 12, nv))
        qb[..., 0] = rng.uniform(0.5, 2.0, qb.shape[:-1])
        qb[..., 1:4] = rng.normal(scale=0.5, size=qb.shape[:-1] + (3,))
        qb[..., -1] = rng.uniform(0.1, 2.0, qb.shape[:-1])
        if nv == 9:
            qb[..., 4:7] = 0.3 * rng.normal(size=qb.shape[:-1] + (3,))
        neg = rng.random(qb.shape[:-1]) < 0.08
        qb[..., -1] = np.where(neg, -rng.uniform(0.01, 0.3, qb.shape[:-1]), qb[..., -1])
        f = ol.ref_update_flux(pf, qb, cmax=2.3)
        np.savez_compressed(os.path.join(HERE, f"unit_flux_{eq}_{rs}_negp.npz"), q=qb, f=f,
                            cmax=2.3, params=np.array(pf.to_params_in()))
        print(eq, rs, "non-finite fluxes", int((~np.isfinite(f)).sum()))


def ot_noisy(p, seed=6):
    """Orszag-Tang with 1 % seeded noise on density, velocity, field and
    pressure.  PPM's extremum test classifies round-off noise on a flat
    profile as extrema, so its result on a uniform pressure depends on the
    last bit of the primitive recovery; the noise leaves no flat profile."""
    nz, ny, nx = p.shape
    z, y, x = problems.cell_centres(p)
    Y, X = np.meshgrid(y, x, indexing="ij")
    xi = [problems.uniform(seed, nz * ny * nx, stream=s).reshape(nz, ny, nx) - 0.5
          for s in range(20, 25)]
    q = np.zeros((p.nv, nz, ny, nx))
    b0 = 1.0 / np.sqrt(4.0 * np.pi)
    q[0] = 25.0 / (36.0 * np.pi) * (1.0 + 0.02 * xi[0])
    q[1] = -np.sin(2.0 * np.pi * Y) + 0.02 * xi[1]
    q[2] = np.sin(2.0 * np.pi * X) + 0.02 * xi[2]
    q[4] = -b0 * np.sin(2.0 * np.pi * Y) + 0.02 * b0 * xi[3]
    q[5] = b0 * np.sin(4.0 * np.pi * X)
    q[8] = 5.0 / (12.0 * np.pi) * (1.0 + 0.02 * xi[4])
    return problems.prim_to_cons(p, q)


def main_o3():
    """WENO3, LimO3 and PPM: pencils at two cell sizes (WENO3 and LimO3 weigh
    by it), short Orszag-Tang runs, and a LimO3 run on the reference's
    three-level blast mesh (per-level cell sizes) (own generator)."""
    from amun_amd import mesh
    rng = np.random.default_rng(20261208)
    fp = np.concatenate([rng.normal(size=(6, 24)), np.cumsum(rng.normal(size=(4, 24)), axis=1),
                         np.sin(np.linspace(0, 6, 24))[None, :] * np.ones((2, 1)),
                         np.where(np.arange(24) < 12, 1.0, 0.1)[None, :] * np.ones((2, 1)),
                         1e-3 * np.cumsum(rng.normal(size=(2, 24)), axis=1)])
    out = {"fp": fp, "h": np.array([1.0, 0.0078125])}
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for rec in O3_SCHEMES:
        pu = Params(ndims=3, equation_system="mhd", reconstruction=rec, limiter="mm",
                    ncells=16, nghosts=4, cfl=0.3)
        for ih, h in enumerate(out["h"]):
            out[f"fl_{rec}_{ih}"], out[f"fr_{rec}_{ih}"] = ol.ref_reconstruct(pu, fp, float(h))
        p = Params(ndims=2, equation_system="mhd", reconstruction=rec, limiter="mm",
                   riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.3,
                   ncells=8, nghosts=4 if rec == "ppm" else 2, xblocks=4, yblocks=4)
        u0 = ot_noisy(p) if rec == "ppm" else problems.orszag_tang_2d(p, seed=6)
        nsteps = 8
        u, tr = ol.ref_run(p, u0, nsteps)
        name = f"{rec}_ot{'noisy' if rec == 'ppm' else ''}2d_32_{nsteps}"
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=2, nsteps=nsteps,
                            stride=3, trace=tr, sub=u[:, :, ::3, ::3], sha_u0=sha(u0),
                            sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    np.savez_compressed(os.path.join(HERE, "unit_schemes_o3.npz"), **out)
    p = amr_params(2)
    p.reconstruction = "limo3"
    meta, _, _, _, _ = ol.ref_amr(p, 1)
    u0 = mesh.block_state(p, meta, mesh.blast_prim(p))
    nsteps = 6
    m0, u0r, m1, u1, tr = ol.ref_amr(p, nsteps, u_in=u0)
    assert np.array_equal(m0, meta) and np.array_equal(u0r, u0)
    name = f"amr_limo3_mhd2d_blast_{nsteps}"
    sums = u1.reshape(u1.shape[0], -1, u1.shape[-1]).sum(axis=1)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), params=np.array(p.to_params_in()),
                        ndims=2, nsteps=nsteps, meta=meta, trace=tr, block_sums=sums,
                        sha_u0=sha(u0), sha_u=sha(u1), u1=u1)
    index[name] = {"steps": nsteps, "sha_u": sha(u1), "dt_last": float(tr[-1, 0]),
                   "blocks": int(meta.shape[0])}
    print(name, meta.shape[0], "blocks, dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


def main_c2p2d():
    """The 2D(W,v^2) and 2D(W,u^2) primitive solvers of the relativistic
    system: the reference's cons2prim on random physical states and on the
    failing and hard states of unit_eqs_srmhd, and short blast runs (own
    generator)."""
    rng = np.random.default_rng(20261210)
    g = np.load(os.path.join(HERE, "unit_eqs_srmhd.npz"), allow_pickle=False)
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for ps in ("2dwv", "2dwu"):
        pe = Params(ndims=2, equation_system="srmhd", riemann_solver="hll", gamma=5.0 / 3.0,
                    tolerance=1.0e-12, nr_maxit=50, primitive_solver=ps)
        q = sr_states(rng, 400)
        u, _, _, _, _ = ol.ref_eqs(pe, q, cmax=0.8)
        ub = np.concatenate([u, g["ub"]])
        q0 = np.full_like(ub, -7.0)
        qb, _ = ol.ref_cons2prim(pe, ub, q0)
        p = Params(ndims=2, equation_system="srmhd", reconstruction="tvd", limiter="mm",
                   riemann_solver="hll", time_advance="rk2", gamma=5.0 / 3.0, cfl=0.25,
                   ncells=8, nghosts=2, xblocks=4, yblocks=4, xmin=-0.5, xmax=0.5, ymin=-0.5,
                   ymax=0.5, tolerance=1.0e-12, nr_maxit=50, primitive_solver=ps)
        u0 = problems.srmhd_blast_2d(p, noise=0.01)
        nsteps = 8
        uf, tr = ol.ref_run(p, u0, nsteps)
        name = f"srmhd_{ps}_blast2d_32_{nsteps}"
        np.savez_compressed(os.path.join(HERE, name + ".npz"), params=np.array(p.to_params_in()),
                            ndims=2, nsteps=nsteps, stride=3, trace=tr, sub=uf[:, :, ::3, ::3],
                            sha_u0=sha(u0), sha_u=sha(uf), ub=ub, q0=q0, qb=qb,
                            unit_params=np.array(pe.to_params_in()))
        index[name] = {"steps": nsteps, "sha_u": sha(uf), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0], "failed cells", int((qb[:, 8] == 0.0).sum()))
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


def main_sr_hllc():
    """riemann_srmhd_adi_hllc: update_flux with the zero limiter on random
    relativistic states (v = 0, Bx = 0 and B = 0 rows included), and a blast
    run (own generator)."""
    rng = np.random.default_rng(20261211)
    pf = Params(ndims=2, equation_system="srmhd", reconstruction="tvd", limiter="zero",
                riemann_solver="hllc", gamma=5.0 / 3.0, ncells=8, nghosts=2,
                tolerance=1.0e-12, nr_maxit=50)
    qb = sr_states(rng, 4 * 144).reshape(4, 1, 12, 12, 9)
    f = ol.ref_update_flux(pf, qb, cmax=0.9)
    np.savez_compressed(os.path.join(HERE, "unit_flux_srmhd_hllc.npz"), q=qb, f=f, cmax=0.9,
                        params=np.array(pf.to_params_in()))
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    p = Params(ndims=2, equation_system="srmhd", reconstruction="tvd", limiter="mm",
               riemann_solver="hllc", time_advance="rk2", gamma=5.0 / 3.0, cfl=0.25, ncells=8,
               nghosts=2, xblocks=4, yblocks=4, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5,
               tolerance=1.0e-12, nr_maxit=50)
    # velocities and transverse fields off zero: at a gas at rest the
    # contact speed is round-off and its sign picks the star state or, at an
    # exact zero, the HLL flux (schemes.F90:5560-5600)
    u0 = problems.srmhd_blast_2d(p, noise=0.01, vnoise=0.01)
    nsteps = 8
    u, tr = ol.ref_run(p, u0, nsteps)
    name = f"srmhd_hllc_blast2d_32_{nsteps}"
    np.savez_compressed(os.path.join(HERE, name + ".npz"), params=np.array(p.to_params_in()),
                        ndims=2, nsteps=nsteps, stride=3, trace=tr, sub=u[:, :, ::3, ::3],
                        sha_u0=sha(u0), sha_u=sha(u))
    index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)
    print(name, "dt", tr[-1, 0])


def main_roe_mhd():
    """riemann_mhd_adi_roe through update_flux (zero limiter): random states
    with zero transverse field, Bx = 0 and fast flows; and an Orszag-Tang run
    with noise on every field (own generator)."""
    rng = np.random.default_rng(20261212)
    pf = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="zero",
                riemann_solver="roe", gamma=5.0 / 3.0, ncells=8, nghosts=2)
    qb = np.zeros((4, 1, 12, 12, 9))
    qb[..., 0] = rng.uniform(0.2, 2.0, qb.shape[:-1])
    qb[..., 1:4] = rng.normal(scale=1.0, size=qb.shape[:-1] + (3,))
    qb[..., 8] = rng.uniform(0.1, 2.0, qb.shape[:-1])
    qb[..., 4:7] = rng.normal(size=qb.shape[:-1] + (3,))
    qb[..., 7] = rng.normal(scale=0.1, size=qb.shape[:-1])
    qb[1, ..., 5:7] = 0.0                 # no transverse field
    qb[2, ..., 4] = 0.0                   # Bx = 0
    qb[3, ..., 1] = qb[3, ..., 1] * 5.0   # fast flows
    f = ol.ref_update_flux(pf, qb, cmax=2.3)
    np.savez_compressed(os.path.join(HERE, "unit_flux_mhd_roe.npz"), q=qb, f=f, cmax=2.3,
                        params=np.array(pf.to_params_in()))
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    p = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
               riemann_solver="roe", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.4, ncells=16,
               nghosts=2, xblocks=2, yblocks=2)
    u0 = ot_noisy(p)
    nsteps = 10
    u, tr = ol.ref_run(p, u0, nsteps)
    name = f"roe_mhd_otnoisy2d_32_{nsteps}"
    np.savez_compressed(os.path.join(HERE, name + ".npz"), params=np.array(p.to_params_in()),
                        ndims=2, nsteps=nsteps, stride=3, trace=tr, sub=u[:, :, ::3, ::3],
                        sha_u0=sha(u0), sha_u=sha(u))
    index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)
    print(name, "dt", tr[-1, 0])


CR_SCHEMES = ("crweno5z", "crweno5yc", "crweno5ns")
CRMP_SCHEMES = ("crmp5", "crmp5ld", "crmp7")


def main_crweno(schemes=CR_SCHEMES, outname="unit_schemes_cr"):
    """The compact WENO (or compact MP) schemes: the reference's pencils at
    ng = 2, 3, 4 and Orszag-Tang runs (noise on every field) (own
    generator)."""
    rng = np.random.default_rng(20261213)
    fp = np.concatenate([rng.normal(size=(6, 24)), np.cumsum(rng.normal(size=(4, 24)), axis=1),
                         np.sin(np.linspace(0, 6, 24))[None, :] * np.ones((2, 1)),
                         np.where(np.arange(24) < 12, 1.0, 0.1)[None, :] * np.ones((2, 1))])
    out = {"fp": fp}
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for rec in schemes:
        for ng in (2, 3, 4):
            pu = Params(ndims=3, equation_system="mhd", reconstruction=rec, limiter="mm",
                        ncells=16, nghosts=ng, cfl=0.3)
            out[f"fl_{rec}_{ng}"], out[f"fr_{rec}_{ng}"] = ol.ref_reconstruct(pu, fp)
        p = Params(ndims=2, equation_system="mhd", reconstruction=rec, limiter="mm",
                   riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.3,
                   ncells=8, nghosts=4, xblocks=4, yblocks=4)
        u0 = ot_noisy(p)
        nsteps = 8
        u, tr = ol.ref_run(p, u0, nsteps)
        name = f"{rec}_otnoisy2d_32_{nsteps}"
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=2, nsteps=nsteps,
                            stride=3, trace=tr, sub=u[:, :, ::3, ::3], sha_u0=sha(u0),
                            sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    np.savez_compressed(os.path.join(HERE, outname + ".npz"), **out)
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


MGP_CASES = (("mgp_otnoisy2d_32_4", 2, 5, 4), ("mgp3_otnoisy2d_32_4", 2, 3, 2),
             ("mgp_otnoisy3d_16_4", 3, 5, 3), ("mgp_noisy3d_16_4", 3, 5, 3))


def noisy3d(p, seed=6):
    """A fully three-dimensional MHD state: smooth O(0.1-1) fields varying
    along x, y and z in every primitive variable, psi included, plus 2 %
    seeded noise on each.  The extruded Orszag-Tang state leaves vz, Bz and
    psi at round-off size in 3D, where interfaces_mgp's monotonicity products
    sit at its eps threshold and the TVD fallback flips with the last bit of
    the face sums (an FMA-contracted build of the same code takes other
    branches); here every product is far from it."""
    nz, ny, nx = p.shape
    z, y, x = problems.cell_centres(p)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    xi = [problems.uniform(seed, nz * ny * nx, stream=s).reshape(nz, ny, nx) - 0.5
          for s in range(30, 39)]
    tp = 2.0 * np.pi
    a = 0.02
    q = np.zeros((p.nv, nz, ny, nx))
    q[0] = 1.0 + 0.2 * np.sin(tp * (X + Y + Z)) + a * xi[0]
    q[1] = 0.5 * np.sin(tp * Y) + 0.3 * np.cos(tp * Z) + a * xi[1]
    q[2] = 0.5 * np.sin(tp * Z) + 0.3 * np.cos(tp * X) + a * xi[2]
    q[3] = 0.5 * np.sin(tp * X) + 0.3 * np.cos(tp * Y) + a * xi[3]
    q[4] = 0.3 * np.sin(tp * (Y + Z)) + 0.2 + a * xi[4]
    q[5] = 0.3 * np.sin(tp * (Z + X)) - 0.1 + a * xi[5]
    q[6] = 0.3 * np.sin(tp * (X + Y)) + 0.15 + a * xi[6]
    q[7] = a * xi[7]
    q[8] = 1.0 + 0.1 * np.cos(tp * (X - Y + Z)) + a * xi[8]
    return problems.prim_to_cons(p, q)


def mgp_state(name, p):
    return noisy3d(p) if name.startswith("mgp_noisy3d") else ot_noisy(p)


def mgp_params(nd, ngp, ng):
    p = Params(ndims=nd, equation_system="mhd", reconstruction="mgp", limiter="mm",
               riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.3, ncells=8,
               nghosts=ng, xblocks=4 if nd == 2 else 2, yblocks=4 if nd == 2 else 2,
               zblocks=2)
    p.ngp = ngp
    return p


def main_mgp():
    """Multi-dimensional GP: Orszag-Tang runs (noise on every field), 2D at
    ngp = 5 and 3, 3D at ngp = 5 (own generator)."""
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for name, nd, ngp, ng in MGP_CASES:
        p = mgp_params(nd, ngp, ng)
        u0 = mgp_state(name, p)
        nsteps = 4
        u, tr = ol.ref_run(p, u0, nsteps)
        sub = u[:, ::3, ::3, ::3] if nd == 3 else u[:, :, ::3, ::3]
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            stride=3, trace=tr, sub=sub, sha_u0=sha(u0), sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


STAT_CASES = ("hd2d", "mhd2d", "mhd3d", "srmhd2d")


def stat_case(name):
    """Parameters and initial state of the statistics fixtures."""
    if name == "hd2d":
        p = Params(ndims=2, equation_system="hd", reconstruction="tvd", limiter="mc",
                   riemann_solver="hllc", time_advance="rk2", gamma=1.4, cfl=0.4, ncells=16,
                   nghosts=2, xblocks=3, yblocks=3, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5)
        return p, problems.blast_hd_2d(p, seed=5)
    if name == "mhd2d":
        p = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
                   riemann_solver="hlld", time_advance="rk3", gamma=5.0 / 3.0, cfl=0.4,
                   ncells=8, nghosts=2, xblocks=3, yblocks=2)
        return p, problems.orszag_tang_2d(p, seed=3)
    if name == "mhd3d":
        p = c2_params(32)
        p.xblocks, p.yblocks, p.zblocks, p.ncells = 3, 2, 2, 8
        return p, problems.turbulent_mhd_3d(p, seed=11)
    p = Params(ndims=2, equation_system="srmhd", reconstruction="tvd", limiter="mm",
               riemann_solver="hll", time_advance="rk2", gamma=5.0 / 3.0, cfl=0.25, ncells=8,
               nghosts=2, xblocks=4, yblocks=4, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5,
               tolerance=1.0e-12, nr_maxit=50)
    return p, problems.srmhd_blast_2d(p, noise=0.01)


def main_stats():
    """The reference's integrals and statistics files (store_integrals, 9
    significant digits) after three steps of four configurations (own
    generator)."""
    out = {}
    for name in STAT_CASES:
        p, u0 = stat_case(name)
        ints, st = ol.ref_statistics(p, u0, 3)
        out["ints_" + name], out["stats_" + name] = ints, st
        out["params_" + name] = np.array(p.to_params_in())
        out["sha_u0_" + name] = np.array(sha(u0))
        print(name, st[:3])
    np.savez_compressed(os.path.join(HERE, "stats_ref.npz"), **out)


ISO_FLUX = (("hd", "hll"), ("hd", "roe"), ("mhd", "hll"), ("mhd", "hlld"), ("mhd", "hlldm"),
            ("mhd", "roe"))
# (name, ndims, equation system, reconstruction, Riemann solver, steps)
ISO_RUNS = (("iso_hd2d_tvd_roe_blast_20", 2, "hd", "tvd", "roe", 20),
            ("iso_hd2d_tvd_hll_blast_20", 2, "hd", "tvd", "hll", 20),
            ("iso_hd3d_weno5z_hll_noisy_4", 3, "hd", "weno5z", "hll", 4),
            ("iso_hd3d_mp5_roe_noisy_4", 3, "hd", "mp5", "roe", 4),
            ("iso_mhd2d_tvd_hlld_ot_20", 2, "mhd", "tvd", "hlld", 20),
            ("iso_mhd2d_tvd_hlldm_ot_20", 2, "mhd", "tvd", "hlldm", 20),
            ("iso_mhd2d_tvd_roe_ot_20", 2, "mhd", "tvd", "roe", 20),
            ("iso_mhd2d_tvd_hll_ot_20", 2, "mhd", "tvd", "hll", 20),
            ("iso_mhd3d_weno5z_hlld_noisy_4", 3, "mhd", "weno5z", "hlld", 4),
            ("iso_mhd3d_weno5z_roe_noisy_4", 3, "mhd", "weno5z", "roe", 4))
ISO_CSND = 0.7


def iso_params(nd, eq, rec, rs):
    return Params(ndims=nd, equation_system=eq, equation_of_state="iso", csnd=ISO_CSND,
                  reconstruction=rec, limiter="mc" if rec == "tvd" else "mm",
                  riemann_solver=rs, time_advance="rk2" if eq == "hd" else "rk3",
                  cfl=0.4 if nd == 2 else 0.3, ncells=16 if nd == 2 else 8,
                  nghosts=2 if rec == "tvd" else 4, xblocks=2, yblocks=2, zblocks=2,
                  xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5)


def iso_state(name, p, seed=6):
    """Seeded isothermal states: a 10:1 density blast at rest (HD, r < 0.1),
    the Orszag-Tang field with 2 % noise (2D MHD), or the smooth fully
    three-dimensional fields of noisy3d without pressure (3D)."""
    nz, ny, nx = p.shape
    z, y, x = problems.cell_centres(p)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    xi = [problems.uniform(seed, nz * ny * nx, stream=s).reshape(nz, ny, nx) - 0.5
          for s in range(40, 48)]
    q = np.zeros((p.nv, nz, ny, nx))
    tp = 2.0 * np.pi
    if "blast" in name:
        r = np.sqrt(X ** 2 + Y ** 2)
        q[0] = np.where(r < 0.1, 10.0, 1.0) * (1.0 + 0.02 * xi[0])
        q[1] = 0.01 * xi[1]
        q[2] = 0.01 * xi[2]
    elif "_ot_" in name:
        b0 = 1.0 / np.sqrt(4.0 * np.pi)
        q[0] = 1.0 + 0.02 * xi[0]
        q[1] = -np.sin(tp * Y) + 0.02 * xi[1]
        q[2] = np.sin(tp * X) + 0.02 * xi[2]
        q[4] = -b0 * np.sin(tp * Y) + 0.02 * b0 * xi[3]
        q[5] = b0 * np.sin(2.0 * tp * X)
    else:
        a = 0.02
        q[0] = 1.0 + 0.2 * np.sin(tp * (X + Y + Z)) + a * xi[0]
        q[1] = 0.5 * np.sin(tp * Y) + 0.3 * np.cos(tp * Z) + a * xi[1]
        q[2] = 0.5 * np.sin(tp * Z) + 0.3 * np.cos(tp * X) + a * xi[2]
        q[3] = 0.5 * np.sin(tp * X) + 0.3 * np.cos(tp * Y) + a * xi[3]
        if p.nv == 8:
            q[4] = 0.3 * np.sin(tp * (Y + Z)) + 0.2 + a * xi[4]
            q[5] = 0.3 * np.sin(tp * (Z + X)) - 0.1 + a * xi[5]
            q[6] = 0.3 * np.sin(tp * (X + Y)) + 0.15 + a * xi[6]
            q[7] = a * xi[7]
    return problems.prim_to_cons(p, q)


def iso_stat_case(name):
    nd, eq = (2, "hd") if name == "hd2d_iso" else (2, "mhd")
    p = iso_params(nd, eq, "tvd", "roe" if eq == "hd" else "hlld")
    p.xblocks, p.yblocks = 3, 2
    return p, iso_state("iso_blast" if eq == "hd" else "iso_ot_", p, seed=3)


def main_iso():
    """The isothermal HD and MHD systems (equations.F90:399-830): equation of
    state vectors, every Riemann solver through update_flux (zero limiter,
    random states with weak, zero and strong Bx and fast flows), runs of TVD,
    WENO5-Z and MP5 in 2D and 3D, and the integrals and statistics files
    (own generator)."""
    rng = np.random.default_rng(20261211)
    for eq, nv in (("hd", 4), ("mhd", 8)):
        pe = Params(ndims=3, equation_system=eq, equation_of_state="iso", csnd=ISO_CSND)
        q = np.zeros((64, nv))
        q[:, 0] = rng.uniform(0.1, 3.0, 64)
        q[:, 1:4] = rng.normal(size=(64, 3))
        if nv == 8:
            q[:, 4:8] = rng.normal(size=(64, 4))
            q[:5, 4] = 0.0
        u, q2, f, cm, cp = ol.ref_eqs(pe, q, cmax=1.7)
        np.savez_compressed(os.path.join(HERE, f"unit_eqs_{eq}_iso.npz"), q=q, u=u, q2=q2,
                            f=f, cm=cm, cp=cp, cmax=1.7, params=np.array(pe.to_params_in()))
    for eq, rs in ISO_FLUX:
        pf = Params(ndims=2, equation_system=eq, equation_of_state="iso", csnd=ISO_CSND,
                    reconstruction="tvd", limiter="zero", riemann_solver=rs, ncells=8,
                    nghosts=2)
        nv = pf.nv
        qb = np.zeros((4, 1, 12, 12, nv))
        qb[..., 0] = rng.uniform(0.2, 2.0, qb.shape[:-1])
        qb[..., 1:4] = rng.normal(scale=1.0, size=qb.shape[:-1] + (3,))
        if nv == 8:
            qb[..., 4:7] = rng.normal(size=qb.shape[:-1] + (3,))
            qb[..., 7] = rng.normal(scale=0.1, size=qb.shape[:-1])
            qb[1, ..., 4] = 1e-6 * rng.normal(size=qb.shape[1:-1])   # weak Bx
            qb[2, ..., 4] = 0.0                                       # Bx* = 0
            qb[2, ..., 7] = 0.0
            qb[3, ..., 1] = qb[3, ..., 1] * 5.0                       # fast flows
        else:
            qb[1, ..., 1] = qb[1, ..., 1] * 6.0                       # supersonic
            qb[2, ..., 1:4] = 0.0                                     # at rest
            qb[3, ..., 1] = 1e-3 * rng.normal(size=qb.shape[1:-1])   # Roe v ~ 0
        f = ol.ref_update_flux(pf, qb, cmax=2.3)
        np.savez_compressed(os.path.join(HERE, f"unit_flux_{eq}_iso_{rs}.npz"), q=qb, f=f,
                            cmax=2.3, params=np.array(pf.to_params_in()))
        print("unit_flux", eq, rs)
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for name, nd, eq, rec, rs, nsteps in ISO_RUNS:
        p = iso_params(nd, eq, rec, rs)
        u0 = iso_state(name, p)
        u, tr = ol.ref_run(p, u0, nsteps)
        sub = u[:, ::2, ::2, ::2] if nd == 3 else u[:, :, ::3, ::3]
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            stride=2 if nd == 3 else 3, trace=tr, sub=sub, sha_u0=sha(u0),
                            sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)
    out = {}
    for name in ("hd2d_iso", "mhd2d_iso"):
        p, u0 = iso_stat_case(name)
        ints, st = ol.ref_statistics(p, u0, 3)
        out["ints_" + name], out["stats_" + name] = ints, st
        out["params_" + name] = np.array(p.to_params_in())
        out["sha_u0_" + name] = np.array(sha(u0))
        print(name, st[:3])
    np.savez_compressed(os.path.join(HERE, "stats_iso_ref.npz"), **out)


# (name, ndims, reconstruction, Riemann solver, primitive solver, steps)
SRHD_RUNS = (("srhd_hd2d_tvd_hllc_1dw_blast_20", 2, "tvd", "hllc", "1dw", 20),
             ("srhd_hd2d_tvd_hll_2dwv_blast_20", 2, "tvd", "hll", "2dwv", 20),
             ("srhd_hd2d_mp5_hllc_2dwu_blast_10", 2, "mp5", "hllc", "2dwu", 10),
             ("srhd_hd3d_weno5z_hllc_1dw_noisy_4", 3, "weno5z", "hllc", "1dw", 4),
             ("srhd_hd3d_tvd_hll_2dwu_noisy_4", 3, "tvd", "hll", "2dwu", 4))


def srhd_states(rng, n):
    """Random special-relativistic HD primitive states (n, 5): |v| < 0.99,
    densities and pressures over decades; the first rows at rest."""
    q = np.zeros((n, 5))
    q[:, 0] = 10.0 ** rng.uniform(-2, 1, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    q[:, 1:4] = d * rng.uniform(0.0, 0.99, n)[:, None]
    q[:, 4] = 10.0 ** rng.uniform(-3, 1, n)
    q[: n // 16, 1:4] = 0.0
    return q


def srhd_params(nd, rec, rs, c2p):
    return Params(ndims=nd, equation_system="srhd", reconstruction=rec,
                  limiter="mc" if rec == "tvd" else "mm", riemann_solver=rs,
                  time_advance="rk2" if nd == 2 else "rk3", gamma=4.0 / 3.0,
                  cfl=0.4 if nd == 2 else 0.3, ncells=16 if nd == 2 else 8,
                  nghosts=2 if rec == "tvd" else 4, xblocks=2, yblocks=2, zblocks=2,
                  xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5,
                  tolerance=1.0e-12, nr_maxit=50, primitive_solver=c2p)


def srhd_state(name, p, seed=6):
    """Relativistic HD states: a cylindrical blast (rho = 1, p = 1 inside
    r < 0.1, 0.01 outside, 1 % seeded density noise and velocities of 1e-3)
    or smooth fully three-dimensional flows with |v| up to about 0.6."""
    nz, ny, nx = p.shape
    z, y, x = problems.cell_centres(p)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    xi = [problems.uniform(seed, nz * ny * nx, stream=s).reshape(nz, ny, nx) - 0.5
          for s in range(50, 55)]
    q = np.zeros((5, nz, ny, nx))
    tp = 2.0 * np.pi
    if "blast" in name:
        r = np.sqrt(X ** 2 + Y ** 2)
        q[0] = 1.0 + 0.02 * xi[0]
        q[1] = 2e-3 * xi[1]
        q[2] = 2e-3 * xi[2]
        q[4] = np.where(r < 0.1, 1.0, 0.01)
    else:
        a = 0.02
        q[0] = 1.0 + 0.2 * np.sin(tp * (X + Y + Z)) + a * xi[0]
        q[1] = 0.3 * np.sin(tp * Y) + 0.2 * np.cos(tp * Z) + a * xi[1]
        q[2] = 0.3 * np.sin(tp * Z) + 0.2 * np.cos(tp * X) + a * xi[2]
        q[3] = 0.3 * np.sin(tp * X) + 0.2 * np.cos(tp * Y) + a * xi[3]
        q[4] = 1.0 + 0.1 * np.cos(tp * (X - Y + Z)) + a * xi[4]
    return problems.prim_to_cons(p, q)


def main_srhd():
    """The special-relativistic HD system (equations.F90:3478-4566,
    schemes.F90:4433-4968): equation-of-state vectors with the three
    primitive solvers (random, at-rest and failing states), HLL and HLLC
    through update_flux, runs of TVD, MP5 and WENO5-Z in 2D and 3D, and the
    statistics file (own generator)."""
    rng = np.random.default_rng(20261212)
    out = {}
    pe = Params(ndims=2, equation_system="srhd", riemann_solver="hll", gamma=4.0 / 3.0,
                tolerance=1.0e-12, nr_maxit=50)
    q = srhd_states(rng, 400)
    u, q2, f, cm, cp = ol.ref_eqs(pe, q, cmax=0.8)
    out.update(q=q, u=u, q2=q2, f=f, cm=cm, cp=cp, cmax=0.8,
               params=np.array(pe.to_params_in()))
    # conserved states for the primitive solvers: the physical ones, very
    # fast flows (W up to ~70), and unphysical ones (E below sqrt(m^2 + D^2))
    fast = srhd_states(rng, 32)
    d = rng.normal(size=(32, 3))
    fast[:, 1:4] = d / np.linalg.norm(d, axis=1)[:, None] * 0.9999
    ub = np.concatenate([u, ol.ref_eqs(pe, fast, cmax=0.8)[0]])
    ub[:16, 4] = -0.5 * ub[:16, 0]          # E + D below sqrt(m^2 + D^2)
    ub[16:32, 1:4] *= 1.0 + 1e-3
    for ps in ("1dw", "2dwv", "2dwu"):
        pp = Params(ndims=2, equation_system="srhd", riemann_solver="hll", gamma=4.0 / 3.0,
                    tolerance=1.0e-12, nr_maxit=50, primitive_solver=ps)
        q0 = np.full_like(ub, -7.0)
        qb, _ = ol.ref_cons2prim(pp, ub, q0)
        out["ub"], out["q0"], out["qb_" + ps] = ub, q0, qb
    np.savez_compressed(os.path.join(HERE, "unit_eqs_srhd.npz"), **out)
    for rs in ("hll", "hllc"):
        pf = Params(ndims=2, equation_system="srhd", reconstruction="tvd", limiter="zero",
                    riemann_solver=rs, gamma=4.0 / 3.0, ncells=8, nghosts=2,
                    tolerance=1.0e-12, nr_maxit=50)
        qb = srhd_states(rng, 3 * 144).reshape(3, 1, 12, 12, 5)
        f = ol.ref_update_flux(pf, qb, cmax=1.0)
        np.savez_compressed(os.path.join(HERE, f"unit_flux_srhd_{rs}.npz"), q=qb, f=f,
                            cmax=1.0, params=np.array(pf.to_params_in()))
        print("unit_flux srhd", rs)
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for name, nd, rec, rs, c2p, nsteps in SRHD_RUNS:
        p = srhd_params(nd, rec, rs, c2p)
        u0 = srhd_state(name, p)
        u, tr = ol.ref_run(p, u0, nsteps)
        sub = u[:, ::2, ::2, ::2] if nd == 3 else u[:, :, ::3, ::3]
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            stride=2 if nd == 3 else 3, trace=tr, sub=sub, sha_u0=sha(u0),
                            sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)
    p = srhd_params(2, "tvd", "hllc", "1dw")
    p.xblocks, p.yblocks = 3, 2
    u0 = srhd_state("blast", p, seed=3)
    ints, st = ol.ref_statistics(p, u0, 3)
    np.savez_compressed(os.path.join(HERE, "stats_srhd_ref.npz"), ints=ints, stats=st,
                        params=np.array(p.to_params_in()), sha_u0=np.array(sha(u0)))
    print("stats srhd", st[:6])


# (name, ndims, equation system, reconstruction, viscosity, resistivity, steps)
VISC_RUNS = (("visc_hd2d_tvd_blast_10", 2, "hd", "tvd", 2e-3, 0.0, 10),
             ("visc_mhd2d_tvd_ot_10", 2, "mhd", "tvd", 1e-3, 0.0, 10),
             ("resist_mhd2d_tvd_ot_10", 2, "mhd", "tvd", 0.0, 2e-3, 10),
             ("viscres_mhd3d_weno5z_noisy_4", 3, "mhd", "weno5z", 1e-3, 1e-3, 4),
             ("viscres_mhd2d_iso_ot_10", 2, "mhdiso", "tvd", 1e-3, 2e-3, 10),
             ("visc_hd3d_weno5z_noisy_4", 3, "hd", "weno5z", 3e-3, 0.0, 4))


def visc_params(nd, eq, rec, nu, eta):
    iso = eq.endswith("iso")
    eqs = eq[:-3] if iso else eq
    p = Params(ndims=nd, equation_system=eqs, reconstruction=rec,
               limiter="mc" if rec == "tvd" else "mm",
               riemann_solver="hllc" if eqs == "hd" else "hlld", time_advance="rk3",
               gamma=1.4 if eqs == "hd" else 5.0 / 3.0, cfl=0.4 if nd == 2 else 0.3,
               ncells=16 if nd == 2 else 8, nghosts=2 if rec == "tvd" else 4, xblocks=2,
               yblocks=2, zblocks=2, viscosity=nu, resistivity=eta)
    if iso:
        p.equation_of_state, p.csnd = "iso", 0.8
    return p


def visc_state(name, p):
    if "blast" in name:
        return problems.blast_hd_2d(p, seed=9)
    if p.equation_of_state == "iso":
        return iso_state("iso" + name[name.index("_ot_"):] if "_ot_" in name else name, p)
    if "_ot_" in name:
        return ot_noisy(p)
    if p.equation_system == "hd":
        return srhd_state("noisy", p)        # smooth 3D fields, |v| < 0.6
    return noisy3d(p)


def main_visc():
    """The viscous and resistive source terms of update_sources
    (sources.F90:340-598) and their time-step limit (evolution.F90:478-483):
    2D and 3D HD and MHD runs, adiabatic and isothermal (own generator)."""
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for name, nd, eq, rec, nu, eta, nsteps in VISC_RUNS:
        p = visc_params(nd, eq, rec, nu, eta)
        u0 = visc_state(name, p)
        u, tr = ol.ref_run(p, u0, nsteps)
        sub = u[:, ::2, ::2, ::2] if nd == 3 else u[:, :, ::3, ::3]
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            stride=2 if nd == 3 else 3, trace=tr, sub=sub, sha_u0=sha(u0),
                            sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


# (name, ndims, equation system, reconstruction, steps): gravity of the
# Rayleigh-Taylor problem
GRAV_RUNS = (("grav_hd2d_tvd_rt_10", 2, "hd", "tvd", 10),
             ("grav_mhd2d_tvd_rt_10", 2, "mhd", "tvd", 10),
             ("grav_hdiso2d_tvd_rt_10", 2, "hdiso", "tvd", 10),
             ("grav_hd3d_weno5z_rt_4", 3, "hd", "weno5z", 4))
# (name, ndims, equation system, reconstruction, Riemann solver, steps):
# the four-velocity reconstruction of the relativistic systems
V4_RUNS = (("v4_srhd2d_tvd_hllc_blast_20", 2, "srhd", "tvd", "hllc", 20),
           ("v4_srhd3d_weno5z_hll_noisy_4", 3, "srhd", "weno5z", "hll", 4),
           ("v4_srmhd2d_tvd_hllc_blast_10", 2, "srmhd", "tvd", "hllc", 10))


def grav_params(nd, eq, rec):
    iso = eq.endswith("iso")
    eqs = eq[:-3] if iso else eq
    p = Params(ndims=nd, equation_system=eqs, reconstruction=rec,
               limiter="mc" if rec == "tvd" else "mm",
               riemann_solver="hllc" if eqs == "hd" else "hlld", time_advance="rk2",
               gamma=1.4 if eqs == "hd" else 5.0 / 3.0, cfl=0.4 if nd == 2 else 0.3,
               ncells=16 if nd == 2 else 8, nghosts=2 if rec == "tvd" else 4, xblocks=2,
               yblocks=2, zblocks=2, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5,
               zmax=0.5, ylbndry="reflective", yubndry="reflective")
    if iso:
        p.equation_of_state, p.csnd = "iso", 1.0
    p.extra["problem"] = "rt"
    p.extra["gacc"] = "-0.5"
    return p


def grav_state(p, seed=6):
    """Heavy fluid (rho 2) above light (rho 1) with the interface displaced
    by 0.02 cos(2 pi x) (times cos(2 pi z) in 3D), 2 % density noise,
    velocities of 1e-2, p = 2.5 - 0.5 rho y where there is a pressure, and
    B = (0.05, 0.03, 0.02) where there is a field.  The gas moves and every
    field component is non-zero: at v = 0 with By = Bz = 0 the HLLD contact
    speed is exactly zero or not depending on round-off, where the
    reference's branches do not join continuously."""
    nz, ny, nx = p.shape
    z, y, x = problems.cell_centres(p)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    xi = [problems.uniform(seed, nz * ny * nx, stream=s).reshape(nz, ny, nx) - 0.5
          for s in range(60, 64)]
    q = np.zeros((p.nv, nz, ny, nx))
    yi = 0.02 * np.cos(2.0 * np.pi * X) * (np.cos(2.0 * np.pi * Z) if p.ndims == 3 else 1.0)
    q[0] = np.where(Y > yi, 2.0, 1.0) * (1.0 + 0.02 * xi[0])
    q[1:4] = 0.02 * np.stack(xi[1:4])
    if p.nv in (5, 9):
        q[-1] = 2.5 - 0.5 * q[0] * Y
    if p.nv >= 8:
        q[4], q[5], q[6] = 0.05, 0.03, 0.02
    return problems.prim_to_cons(p, q)


def v4_params(nd, eq, rec, rs):
    p = srhd_params(nd, rec, rs, "1dw")
    p.equation_system = eq
    if eq == "srmhd":
        p.gamma = 5.0 / 3.0
    p.state_variables = "4vec"
    return p


def v4_state(name, p):
    if eq_of(p) == "srmhd":
        return problems.srmhd_blast_2d(p, noise=0.01, vnoise=0.05)
    return srhd_state(name, p)


def eq_of(p):
    return p.equation_system.lower()


def main_grav4v():
    """Gravity of the Rayleigh-Taylor problem (sources.F90:288-335,
    gravity.F90:141-150, 283-330) and the relativistic systems'
    four-velocity reconstruction (state_variables = 4vec,
    schemes.F90:4481-4517, 5040-5076): short runs (own generator)."""
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    runs = [(n, grav_params(nd, eq, rec), None, st) for n, nd, eq, rec, st in GRAV_RUNS]
    runs += [(n, v4_params(nd, eq, rec, rs), n, st) for n, nd, eq, rec, rs, st in V4_RUNS]
    for name, p, v4, nsteps in runs:
        u0 = grav_state(p) if v4 is None else v4_state(name, p)
        u, tr = ol.ref_run(p, u0, nsteps)
        nd = p.ndims
        sub = u[:, ::2, ::2, ::2] if nd == 3 else u[:, :, ::3, ::3]
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=nd, nsteps=nsteps,
                            stride=2 if nd == 3 else 3, trace=tr, sub=sub, sha_u0=sha(u0),
                            sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, "dt", tr[-1, 0])
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)


def main():
    if not (ol.ref_available(2) and ol.ref_available(3)):
        sys.exit("reference binaries missing: run oracle/build_ref.sh first")
    if sys.argv[1:] == ["sr"]:
        main_sr()
        return
    if sys.argv[1:] == ["gp"]:
        main_gp()
        return
    if sys.argv[1:] == ["ssprk"]:
        main_ssprk()
        return
    if sys.argv[1:] == ["refine"]:
        main_refine()
        sys.exit(0)
    if sys.argv[1:] == ["grav4v"]:
        main_grav4v()
        return
    if sys.argv[1:] == ["visc"]:
        main_visc()
        return
    if sys.argv[1:] == ["srhd"]:
        main_srhd()
        return
    if sys.argv[1:] == ["iso"]:
        main_iso()
        return
    if sys.argv[1:] == ["mlp"]:
        main_mlp()
        return
    if sys.argv[1:] == ["glm"]:
        main_glm()
        return
    if sys.argv[1:] == ["remesh"]:
        main_remesh()
        return
    if sys.argv[1:] == ["amr"]:
        main_amr()
        return
    if sys.argv[1:] == ["schemes"]:
        main_schemes()
        return
    if sys.argv[1:] == ["hllc_mhd"]:
        main_hllc_mhd()
        return
    if sys.argv[1:] == ["roe_hd"]:
        main_roe_hd()
        return
    if sys.argv[1:] == ["o3"]:
        main_o3()
        return
    if sys.argv[1:] == ["negp"]:
        main_negp()
        return
    if sys.argv[1:] == ["stats"]:
        main_stats()
        return
    if sys.argv[1:] == ["c2p2d"]:
        main_c2p2d()
        return
    if sys.argv[1:] == ["sr_hllc"]:
        main_sr_hllc()
        return
    if sys.argv[1:] == ["roe_mhd"]:
        main_roe_mhd()
        return
    if sys.argv[1:] == ["crweno"]:
        main_crweno()
    if sys.argv[1:] == ["mgp"]:
        main_mgp()
        return
    if sys.argv[1:] == ["crmp"]:
        main_crweno(CRMP_SCHEMES, "unit_schemes_crmp")
        return
    index = {}
    for name, p, builder, nsteps, stride in run_cases():
        u0 = builder(p) if builder else hd3_blast(p)
        u, tr = ol.ref_run(p, u0, nsteps)
        sub = u[:, ::stride, ::stride, ::stride] if p.ndims == 3 else u[:, :, ::stride, ::stride]
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            params=np.array(p.to_params_in()), ndims=p.ndims,
                            nsteps=nsteps, stride=stride, trace=tr, sub=sub,
                            sha_u0=sha(u0), sha_u=sha(u))
        index[name] = {"steps": nsteps, "sha_u": sha(u), "dt_last": float(tr[-1, 0])}
        print(name, p.shape, "dt", tr[-1, 0])

    rng = np.random.default_rng(20261202)
    # WENO5-Z pencils (reconstruct_weno5z), including steps, extrema and
    # smooth data
    pw = c2_params(32)
    fp = np.concatenate([rng.normal(size=(8, 24)),
                         np.cumsum(rng.normal(size=(4, 24)), axis=1),
                         np.sin(np.linspace(0, 6, 24))[None, :] * np.ones((2, 1)),
                         np.where(np.arange(24) < 12, 1.0, 0.1)[None, :] * np.ones((2, 1))])
    fl, fr = ol.ref_reconstruct(pw, fp)
    np.savez_compressed(os.path.join(HERE, "unit_weno5z.npz"), fp=fp, fl=fl, fr=fr)

    # equation of state vectors, HD and MHD
    for eq, nv in (("hd", 5), ("mhd", 9)):
        pe = Params(ndims=3, equation_system=eq, gamma=1.4 if eq == "hd" else 5.0 / 3.0)
        q = np.zeros((64, nv))
        q[:, 0] = rng.uniform(0.1, 3.0, 64)
        q[:, 1:4] = rng.normal(size=(64, 3))
        if nv == 9:
            q[:, 4:8] = rng.normal(size=(64, 4))
            q[:5, 4] = 0.0
        q[:, -1] = rng.uniform(0.05, 3.0, 64)
        u, q2, f, cm, cp = ol.ref_eqs(pe, q, cmax=1.7)
        np.savez_compressed(os.path.join(HERE, f"unit_eqs_{eq}.npz"), q=q, u=u, q2=q2, f=f,
                            cm=cm, cp=cp, cmax=1.7)

    # Riemann solvers through update_flux with the zero TVD limiter (piecewise
    # constant states, so every face solves the Riemann problem of its two
    # cell values): 2D blocks with random states, including weak Bx, sm ~ 0
    # and strongly supersonic pairs
    for eq, rs in (("hd", "hll"), ("hd", "hllc"), ("mhd", "hll"), ("mhd", "hlld")):
        pf = Params(ndims=2, equation_system=eq, reconstruction="tvd", limiter="zero",
                    riemann_solver=rs, gamma=1.4 if eq == "hd" else 5.0 / 3.0,
                    ncells=8, nghosts=2)
        nv = pf.nv
        nb = 4
        qb = np.zeros((nb, 1, 12, 12, nv))
        qb[..., 0] = rng.uniform(0.2, 2.0, qb.shape[:-1])
        qb[..., 1:4] = rng.normal(scale=1.5, size=qb.shape[:-1] + (3,))
        qb[..., -1] = rng.uniform(0.1, 2.0, qb.shape[:-1])
        if nv == 9:
            qb[..., 4:7] = rng.normal(size=qb.shape[:-1] + (3,))
            qb[..., 7] = rng.normal(scale=0.1, size=qb.shape[:-1])
            qb[1, ..., 4] = 1e-6 * rng.normal(size=qb.shape[1:-1])   # weak Bx
            qb[2, ..., 4] = 0.0                                       # Bx = 0
            qb[3, ..., 1] = qb[3, ..., 1] * 5.0                       # fast flows
        f = ol.ref_update_flux(pf, qb, cmax=2.3)
        np.savez_compressed(os.path.join(HERE, f"unit_flux_{eq}_{rs}.npz"), q=qb, f=f,
                            cmax=2.3, params=np.array(pf.to_params_in()))
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1)
    main_sr()
    main_gp()
    main_ssprk()
    main_amr()
    main_schemes()
    main_hllc_mhd()
    main_roe_hd()
    main_o3()
    main_negp()
    main_stats()
    main_c2p2d()
    main_sr_hllc()
    main_roe_mhd()
    main_crweno()
    main_crweno(CRMP_SCHEMES, "unit_schemes_crmp")
    main_mgp()


if __name__ == "__main__":
    main()
