"""ctypes access to the C oracle (oracle/liboracle.so) and the reference
harness binaries (oracle/_ref/amun_ref_{2,3}d.x).  Test infrastructure: the
oracle is only ever the checker."""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

from amun_amd.config import Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.environ.get("AMUN_ORACLE_SO", os.path.join(ROOT, "oracle", "liboracle.so"))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

_P = ctypes.POINTER(ctypes.c_double)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_P)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_state.argtypes = [ctypes.c_void_p, _P, ctypes.c_double]
        L.orc_get_state.argtypes = [ctypes.c_void_p, _P, _P]
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _P]
        L.orc_statistics.argtypes = [ctypes.c_void_p, _P]
        L.orc_positivity_events.argtypes = [ctypes.c_void_p]
        L.orc_positivity_events.restype = ctypes.c_long
        L.orc_mgp_fallbacks.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_ulonglong)]
        L.orc_refinement.argtypes = [ctypes.c_void_p, _P, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)]
        L.orc_update_flux.argtypes = [ctypes.c_void_p, _P, _P, ctypes.c_double, _P]
        L.orc_reconstruct.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, _P, _P, _P]
        L.orc_eqs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _P, _P, _P, _P, _P, _P]
        L.orc_riemann.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _P, _P, _P]
        L.orc_c2p_events.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        L.orc_algebra.argtypes = [ctypes.c_int, _P, _P]
        L.orc_gp_weights.argtypes = [ctypes.c_void_p, _P]
        L.orc_cons2prim.argtypes = [ctypes.c_void_p, ctypes.c_int, _P, _P,
                                    ctypes.POINTER(ctypes.c_long)]
        L.orc_create_amr.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_void_p)]
        L.orc_set_blocks.argtypes = [ctypes.c_void_p, _P, ctypes.c_double]
        L.orc_get_blocks.argtypes = [ctypes.c_void_p, _P, _P]
        L.orc_amr_skip_correction.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_amr_remesh.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_void_p)]
        _lib = L
    return _lib


class OracleRun:
    """A block-structured CPU run of the oracle."""

    def __init__(self, p: Params):
        self.p = p
        self.cfg = p.to_cfg()
        h = ctypes.c_void_p()
        r = lib().orc_create(ctypes.byref(self.cfg), ctypes.byref(h))
        if r:
            raise RuntimeError(f"orc_create failed with code {r}")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_state(self, u: np.ndarray, dtnext: float = 2.0e30):
        u = np.ascontiguousarray(u, dtype=np.float64)
        lib().orc_set_state(self.h, _dp(u), dtnext)

    def get_state(self, with_q: bool = False):
        nz, ny, nx = self.p.shape
        u = np.empty((self.p.nv, nz, ny, nx))
        q = np.empty_like(u) if with_q else None
        lib().orc_get_state(self.h, _dp(u), _dp(q) if with_q else None)
        return (u, q) if with_q else u

    def run(self, nsteps: int, dtnext: float = 2.0e30) -> np.ndarray:
        tr = np.zeros((nsteps, 2))
        lib().orc_run(self.h, nsteps, dtnext, _dp(tr))
        return tr

    def statistics(self) -> np.ndarray:
        """store_integrals' statistics: (mean, min, max) of density,
        pressure, |v|, |B|, psi, sonic and Alfvenic Mach number."""
        out = np.empty(21)
        lib().orc_statistics(self.h, _dp(out))
        return out

    def positivity_events(self) -> int:
        return lib().orc_positivity_events(self.h)

    def mgp_fallbacks(self):
        """(count, position hash) of the cells where interfaces_mgp took the
        TVD slopes so far."""
        out = (ctypes.c_ulonglong * 2)()
        lib().orc_mgp_fallbacks(self.h, out)
        return int(out[0]), int(out[1])

    def refinement(self, rc):
        """check_refinement_criterion of every block (refinement.F90:257)
        with the parameters of rc (config.AmunRefineCfg): (criterion per
        block, indicators (block, 3): largest second-derivative error,
        vorticity, current density)."""
        nb = getattr(self, "meta", None)
        nb = nb.shape[0] if nb is not None else self.p.nblocks
        par = np.array([rc.crefmin, rc.crefmax, rc.epsref, rc.vortmin, rc.vortmax, rc.jabsmin,
                        rc.jabsmax])
        flags = (ctypes.c_int * 3)(rc.varmask, rc.vort, rc.jabs)
        crit = np.zeros(nb, dtype=np.int32)
        ind = np.zeros((nb, 3), dtype=np.float32)
        lib().orc_refinement(self.h, _dp(par), flags,
                             crit.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             ind.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return crit, ind

    def c2p_events(self):
        ev = (ctypes.c_long * 3)()
        lib().orc_c2p_events(self.h, ev)
        return list(ev)


class OracleAmr(OracleRun):
    """The oracle on a multi-level mesh: meta (nblk, 4) = level and the
    first interior cell (i0, j0, k0) of every block in cells of its level."""

    def __init__(self, p: Params, meta: np.ndarray):
        self.p = p
        self.cfg = p.to_cfg()
        self.meta = np.ascontiguousarray(meta, dtype=np.int32)
        h = ctypes.c_void_p()
        r = lib().orc_create_amr(ctypes.byref(self.cfg), self.meta.shape[0],
                                 self.meta.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                 ctypes.byref(h))
        if r:
            raise RuntimeError(f"orc_create_amr failed with code {r}")
        self.h = h

    def set_blocks(self, u: np.ndarray, dtnext: float = 2.0e30):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = lib().orc_set_blocks(self.h, _dp(u), dtnext)
        if r:
            raise RuntimeError(f"orc_set_blocks failed with code {r}")

    def get_blocks(self, with_q: bool = False):
        nb, nc, ng = self.meta.shape[0], self.p.ncells, self.p.nghosts
        shp = (nb, nc, nc, self.p.nv) if self.p.ndims == 2 else (nb, nc, nc, nc, self.p.nv)
        u = np.empty(shp)
        im = nc + 2 * ng
        qs = (nb, im, im, self.p.nv) if self.p.ndims == 2 else (nb, im, im, im, self.p.nv)
        q = np.empty(qs) if with_q else None
        r = lib().orc_get_blocks(self.h, _dp(u), _dp(q) if with_q else None)
        if r:
            raise RuntimeError(f"orc_get_blocks failed with code {r}")
        return (u, q) if with_q else u

    def remesh(self, meta: np.ndarray) -> "OracleAmr":
        """A new oracle grid on the leaf set meta with the state moved as
        update_mesh moves it (copies, prolong_block, restrict_block); time,
        step and dt carry over."""
        new = OracleAmr.__new__(OracleAmr)
        new.p = self.p
        new.cfg = self.cfg
        new.meta = np.ascontiguousarray(meta, dtype=np.int32)
        h = ctypes.c_void_p()
        r = lib().orc_amr_remesh(ctypes.byref(self.cfg), self.h, new.meta.shape[0],
                                 new.meta.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                 ctypes.byref(h))
        if r:
            raise RuntimeError(f"orc_amr_remesh failed with code {r}")
        new.h = h
        return new


def amr_meta(m: np.ndarray, nc: int) -> np.ndarray:
    """Harness records (level, i0, j0, k0) -> the oracle's block meta."""
    return np.ascontiguousarray(m, dtype=np.int32)


def oracle_riemann(p: Params, ql: np.ndarray, qr: np.ndarray, cmax: float) -> np.ndarray:
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    f = np.empty_like(ql)
    cfg = p.to_cfg()
    r = lib().orc_riemann(ctypes.byref(cfg), ql.shape[0], cmax, _dp(ql.copy()), _dp(qr.copy()), _dp(f))
    assert r == 0, r
    return f


def oracle_update_flux(p: Params, q: np.ndarray, cmax: float) -> np.ndarray:
    """q: (km, jm, im, nv) C order == Fortran q(nv,im,jm,km); returns f with
    C shape (km, jm, im, nv, ndims) == Fortran f(NDIMS,nv,im,jm,km)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    f = np.zeros(q.shape + (p.ndims,))
    dx = np.array(p.dx())
    cfg = p.to_cfg()
    r = lib().orc_update_flux(ctypes.byref(cfg), _dp(q), _dp(dx), cmax, _dp(f))
    assert r == 0, r
    return f


def oracle_reconstruct(p: Params, fp: np.ndarray, h: float = 1.0):
    fp = np.ascontiguousarray(fp, dtype=np.float64)
    npen, n = fp.shape
    fl = np.empty_like(fp)
    fr = np.empty_like(fp)
    cfg = p.to_cfg()
    lib().orc_reconstruct(ctypes.byref(cfg), n, npen, h, _dp(fp), _dp(fl), _dp(fr))
    return fl, fr


def oracle_eqs(p: Params, q: np.ndarray, cmax: float = 1.0):
    q = np.ascontiguousarray(q, dtype=np.float64)
    n = q.shape[0]
    u = np.empty_like(q); q2 = np.empty_like(q); f = np.empty_like(q)
    cm = np.empty(n); cp = np.empty(n)
    cfg = p.to_cfg()
    lib().orc_eqs(ctypes.byref(cfg), n, cmax, _dp(q), _dp(u), _dp(q2), _dp(f), _dp(cm), _dp(cp))
    return u, q2, f, cm, cp


def oracle_algebra(a: np.ndarray) -> np.ndarray:
    """Root finders of algebra.F90 on coefficient sets a (n, 5), increasing
    powers; returns (n, 19) as laid out by orc_algebra."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros((a.shape[0], 19))
    lib().orc_algebra(a.shape[0], _dp(a), _dp(out))
    return out


def oracle_gp_weights(p: Params) -> np.ndarray:
    w = np.zeros(p.ngp)
    cfg = p.to_cfg()
    r = lib().orc_gp_weights(ctypes.byref(cfg), _dp(w))
    assert r == 0, r
    return w


def oracle_cons2prim(p: Params, u: np.ndarray, q0: np.ndarray):
    """cons2prim of the equation system on u (n, nv); q starts as q0."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    q = np.array(q0, dtype=np.float64, order="C", copy=True)
    ev = (ctypes.c_long * 3)()
    cfg = p.to_cfg()
    r = lib().orc_cons2prim(ctypes.byref(cfg), u.shape[0], _dp(u), _dp(q), ev)
    assert r == 0, r
    return q, list(ev)


# ------------------------------------------------------------------ reference
def ref_available(ndims: int) -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"amun_ref_{ndims}d.x"))


def _ref_exec(p: Params, extra: dict, workdir: str):
    exe = os.path.join(REF_DIR, f"amun_ref_{p.ndims}d.x")
    if not os.path.exists(exe):
        # a test that needs the binary of this dimension skips (the other
        # dimension's pins still run)
        import pytest
        pytest.skip(f"reference binary {os.path.basename(exe)} (oracle/_ref) not built")
    txt = p.to_params_in() + "".join(f"{k} = {v}\n" for k, v in extra.items())
    with open(os.path.join(workdir, "params.in"), "w") as fh:
        fh.write(txt)
    r = subprocess.run([exe, "-i", "params.in"], cwd=workdir, capture_output=True,
                       text=True, timeout=3600)
    if r.returncode != 0:
        raise RuntimeError(f"reference harness failed: {r.stdout}\n{r.stderr}")
    return r


def ref_run(p: Params, u0: np.ndarray, nsteps: int):
    """Run the reference binary on the global conserved state u0; returns the
    final state and the (dt, cmax) trace."""
    with tempfile.TemporaryDirectory() as d:
        np.ascontiguousarray(u0, dtype=np.float64).tofile(os.path.join(d, "in.bin"))
        _ref_exec(p, {"harness_mode": "run", "harness_input": "in.bin",
                      "harness_output": "out", "harness_steps": nsteps}, d)
        u = np.fromfile(os.path.join(d, "out_u.bin")).reshape(u0.shape)
        tr = np.fromfile(os.path.join(d, "out_dt.bin")).reshape(nsteps, 2)
    return u, tr


def ref_statistics(p: Params, u0: np.ndarray, nsteps: int):
    """Run the reference binary nsteps from u0 and let its store_integrals
    write the integrals and statistics files of the final state; returns
    (integrals[8], statistics[21]) as printed (9 significant digits)."""
    with tempfile.TemporaryDirectory() as d:
        np.ascontiguousarray(u0, dtype=np.float64).tofile(os.path.join(d, "in.bin"))
        _ref_exec(p, {"harness_mode": "run", "harness_input": "in.bin",
                      "harness_output": "out", "harness_steps": nsteps,
                      "harness_integrals": 1, "integrals_interval": 1}, d)

        def last_row(name):
            rows = [ln.split() for ln in open(os.path.join(d, name)) if not ln.startswith("#")]
            return np.array([float(x) for x in rows[-1]])
        ints = last_row("integrals_01.dat")
        stats = last_row("statistics_01.dat")
    return ints[3:], stats[2:]


def ref_update_flux(p: Params, qblocks: np.ndarray, cmax: float) -> np.ndarray:
    """qblocks: (nb, km, jm, im, nv); returns (nb, km, jm, im, nv, ndims)."""
    nb = qblocks.shape[0]
    dx = np.array(p.dx())
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "in.bin"), "wb") as fh:
            for b in range(nb):
                np.ascontiguousarray(qblocks[b], dtype=np.float64).tofile(fh)
                dx.tofile(fh)
        _ref_exec(p, {"harness_mode": "flux", "harness_input": "in.bin",
                      "harness_output": "out", "harness_count": nb,
                      "harness_cmax": repr(float(cmax))}, d)
        f = np.fromfile(os.path.join(d, "out_f.bin")).reshape(qblocks.shape + (p.ndims,))
    return f


def ref_reconstruct(p: Params, fp: np.ndarray, h: float = 1.0):
    npen, n = fp.shape
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "in.bin"), "wb") as fh:
            np.ascontiguousarray(fp, dtype=np.float64).tofile(fh)
            np.array([h]).tofile(fh)
        _ref_exec(p, {"harness_mode": "recon", "harness_input": "in.bin",
                      "harness_output": "out", "harness_count": npen,
                      "harness_n": n}, d)
        fl = np.fromfile(os.path.join(d, "out_fl.bin")).reshape(npen, n)
        fr = np.fromfile(os.path.join(d, "out_fr.bin")).reshape(npen, n)
    return fl, fr


def ref_eqs(p: Params, q: np.ndarray, cmax: float = 1.0):
    n = q.shape[0]
    with tempfile.TemporaryDirectory() as d:
        np.ascontiguousarray(q, dtype=np.float64).tofile(os.path.join(d, "in.bin"))
        _ref_exec(p, {"harness_mode": "eqs", "harness_input": "in.bin",
                      "harness_output": "out", "harness_n": n,
                      "harness_cmax": repr(float(cmax))}, d)
        a = np.fromfile(os.path.join(d, "out_eqs.bin"))
    nv = p.nv
    u = a[: nv * n].reshape(n, nv)
    q2 = a[nv * n: 2 * nv * n].reshape(n, nv)
    f = a[2 * nv * n: 3 * nv * n].reshape(n, nv)
    cm = a[3 * nv * n: 3 * nv * n + n]
    cp = a[3 * nv * n + n:]
    return u, q2, f, cm, cp


def ref_algebra(a: np.ndarray, ndims: int = 2) -> np.ndarray:
    """The reference's root finders (harness mode "algebra")."""
    n = a.shape[0]
    p = Params(ndims=ndims, equation_system="hd", riemann_solver="hllc")
    with tempfile.TemporaryDirectory() as d:
        np.ascontiguousarray(a, dtype=np.float64).tofile(os.path.join(d, "in.bin"))
        _ref_exec(p, {"harness_mode": "algebra", "harness_input": "in.bin",
                      "harness_output": "out", "harness_n": n}, d)
        r = np.fromfile(os.path.join(d, "out_alg.bin")).reshape(n, 19)
    return r


def ref_cons2prim(p: Params, u: np.ndarray, q0: np.ndarray):
    """The reference's cons2prim on u (n, nv) with q starting as q0; also
    returns the stderr text (the solver's messages)."""
    n = u.shape[0]
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "in.bin"), "wb") as fh:
            np.ascontiguousarray(u, dtype=np.float64).tofile(fh)
            np.ascontiguousarray(q0, dtype=np.float64).tofile(fh)
        r = _ref_exec(p, {"harness_mode": "c2p", "harness_input": "in.bin",
                          "harness_output": "out", "harness_n": n}, d)
        q = np.fromfile(os.path.join(d, "out_c2p.bin")).reshape(n, p.nv)
    return q, r.stderr


def _read_blocks(path, ndims, nv, nc):
    """Per-block records of the harness's AMR dumps: (level, i0, j0, k0) and
    the interior u as (nc[,nc],nc, nv) arrays in (k, j, i, var) order."""
    ncell = nc ** ndims
    rec = 4 + nv * ncell
    raw = np.fromfile(path)
    assert raw.size % rec == 0
    raw = raw.reshape(-1, rec)
    meta = raw[:, :4].astype(np.int64)
    shape = (-1, nc, nc, nv) if ndims == 2 else (-1, nc, nc, nc, nv)
    u = raw[:, 4:].reshape(shape)
    return meta, u


def ref_amr(p: Params, nsteps: int, qdump: bool = False, u_in=None, criterion: bool = False,
            update: bool = False):
    """The reference's own refined mesh (generate_mesh on the problem named
    in p.extra["problem"], up to p.extra["maxlev"]) and nsteps steps on it,
    frozen.  u_in (per-block interiors in the order of the mesh the same
    parameters generate) replaces the problem's state.  Returns (meta0, u0,
    meta1, u1, trace[, q0 with ghosts][, check_refinement_criterion of every
    block after the steps][, (meta2, u2) after one update_mesh that follows
    the steps])."""
    nv = {"hd": 5, "mhd": 9}[p.equation_system]
    with tempfile.TemporaryDirectory() as d:
        extra = {"harness_mode": "amr", "harness_output": "out", "harness_steps": nsteps}
        if u_in is not None:
            np.ascontiguousarray(u_in, dtype=np.float64).tofile(os.path.join(d, "in.bin"))
            extra.update({"harness_amr_input": 1, "harness_input": "in.bin"})
        if criterion:
            extra["harness_criterion"] = 1
        if update:
            extra["harness_update"] = 1
        _ref_exec(p, extra, d)
        m0, u0 = _read_blocks(os.path.join(d, "out_mesh0.bin"), p.ndims, nv, p.ncells)
        m1, u1 = _read_blocks(os.path.join(d, "out_mesh1.bin"), p.ndims, nv, p.ncells)
        tr = np.fromfile(os.path.join(d, "out_dt.bin")).reshape(nsteps, 2)
        out = [m0, u0, m1, u1, tr]
        if qdump:
            im = p.ncells + 2 * p.nghosts
            shp = (-1, im, im, nv) if p.ndims == 2 else (-1, im, im, im, nv)
            out.append(np.fromfile(os.path.join(d, "out_q0.bin")).reshape(shp))
        if criterion:
            out.append(np.fromfile(os.path.join(d, "out_crit.bin")).astype(np.int64))
        if update:
            out.append(_read_blocks(os.path.join(d, "out_mesh2.bin"), p.ndims, nv, p.ncells))
    return tuple(out)
