"""ctypes binding of the C-ABI (include/amun_dev.h) of libamun_hip.so.

``Solver`` mirrors the reference's step-level interface: ``new_time_step``
(evolution.F90:379), ``advance`` (evolution.F90:292), the ``time``/``step``
counters (evolution.F90:75), ``store_integrals`` (integrals.F90:254) and the
host transfers of block data.  Every call goes to the HIP library; there is
no CPU fallback -- a missing library raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .config import AmunCfg, AmunRefineCfg, Params

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamun_hip.so")
_P = ctypes.POINTER(ctypes.c_double)

EXPORTS = (
    "amun_dev_default_cfg", "amun_dev_init", "amun_dev_finalize",
    "amun_dev_upload_blocks", "amun_dev_upload_global",
    "amun_dev_download_blocks", "amun_dev_download_global",
    "amun_dev_new_time_step", "amun_dev_step", "amun_dev_run",
    "amun_dev_integrals", "amun_dev_get_time", "amun_dev_nblocks",
    "amun_dev_block_coords", "amun_dev_last_error", "amun_dev_sync",
    "amun_dev_positivity_events", "amun_dev_stream", "amun_dev_timed_run",
    "amun_dev_c2p_events", "amun_dev_init_batch", "amun_dev_nruns",
    "amun_dev_run_blocks", "amun_dev_upload_run", "amun_dev_download_run",
    "amun_dev_run_state", "amun_dev_integrals_run", "amun_dev_set_graph",
    "amun_dev_autotune", "amun_dev_set_launch", "amun_dev_unit_riemann",
    "amun_dev_unit_reconstruct", "amun_dev_unit_cons2prim",
    "amun_dev_unit_eqs", "amun_dev_tape_forward", "amun_dev_tape_steps",
    "amun_dev_tape_backward", "amun_dev_tape_tangent", "amun_dev_run_fixed",
    "amun_dev_init_amr", "amun_dev_amr_info", "amun_dev_statistics",
    "amun_dev_statistics_run", "amun_dev_set_stage_fusion",
)


class AmunGrid(ctypes.Structure):
    """ctypes mirror of ``amun_grid`` (include/amun_dev.h)."""
    _fields_ = [
        ("xblocks", ctypes.c_int32), ("yblocks", ctypes.c_int32),
        ("zblocks", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
        ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
    ]

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libamun_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: the HIP library must be built "
            "(__graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.amun_dev_default_cfg.argtypes = [ctypes.POINTER(AmunCfg)]
    L.amun_dev_default_cfg.restype = None
    L.amun_dev_init.argtypes = [ctypes.POINTER(AmunCfg), ctypes.POINTER(vp)]
    L.amun_dev_finalize.argtypes = [vp]
    L.amun_dev_finalize.restype = None
    L.amun_dev_upload_blocks.argtypes = [vp, _P]
    L.amun_dev_upload_global.argtypes = [vp, _P]
    L.amun_dev_download_blocks.argtypes = [vp, _P, _P]
    L.amun_dev_download_global.argtypes = [vp, _P, _P]
    L.amun_dev_new_time_step.argtypes = [vp, ctypes.c_double, _P, _P]
    L.amun_dev_step.argtypes = [vp, ctypes.c_double, _P, _P]
    L.amun_dev_run.argtypes = [vp, ctypes.c_int32, ctypes.c_double, _P]
    L.amun_dev_integrals.argtypes = [vp, _P]
    L.amun_dev_statistics.argtypes = [vp, _P]
    L.amun_dev_statistics_run.argtypes = [vp, ctypes.c_int32, _P]
    L.amun_dev_get_time.argtypes = [vp, _P, ctypes.POINTER(ctypes.c_int64)]
    L.amun_dev_nblocks.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
    L.amun_dev_block_coords.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
    L.amun_dev_timed_run.argtypes = [vp, ctypes.c_int32, ctypes.c_double, _P, _P,
                                     ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.amun_dev_stream.argtypes = [vp]
    L.amun_dev_c2p_events.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.amun_dev_init_batch.argtypes = [ctypes.POINTER(AmunCfg), ctypes.c_int32,
                                      ctypes.POINTER(AmunGrid), ctypes.POINTER(vp)]
    L.amun_dev_nruns.argtypes = [vp, i32p]
    L.amun_dev_run_blocks.argtypes = [vp, ctypes.c_int32, i32p, i32p]
    L.amun_dev_upload_run.argtypes = [vp, ctypes.c_int32, _P]
    L.amun_dev_download_run.argtypes = [vp, ctypes.c_int32, _P, _P]
    L.amun_dev_run_state.argtypes = [vp, _P, _P, ctypes.POINTER(ctypes.c_int64)]
    L.amun_dev_integrals_run.argtypes = [vp, ctypes.c_int32, _P]
    L.amun_dev_set_graph.argtypes = [vp, ctypes.c_int32]
    L.amun_dev_autotune.argtypes = [vp, ctypes.c_int32, i32p]
    L.amun_dev_set_stage_fusion.argtypes = [vp, ctypes.c_int32]
    L.amun_dev_set_launch.argtypes = [vp, i32p]
    L.amun_dev_tape_forward.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double]
    L.amun_dev_tape_steps.argtypes = [vp, _P, _P]
    L.amun_dev_tape_backward.argtypes = [vp, _P, _P, _P]
    L.amun_dev_tape_tangent.argtypes = [vp, _P, ctypes.c_double, _P]
    L.amun_dev_run_fixed.argtypes = [vp, ctypes.c_int32, _P, _P]
    L.amun_dev_unit_riemann.argtypes = [ctypes.POINTER(AmunCfg), ctypes.c_int32,
                                        ctypes.c_double, _P, _P, _P]
    L.amun_dev_unit_eqs.argtypes = [ctypes.POINTER(AmunCfg), ctypes.c_int32, ctypes.c_double,
                                    _P, _P, _P, _P, _P, _P]
    L.amun_dev_unit_reconstruct.argtypes = [ctypes.POINTER(AmunCfg), ctypes.c_int32,
                                            ctypes.c_int32, ctypes.c_double, _P, _P, _P]
    L.amun_dev_unit_cons2prim.argtypes = [ctypes.POINTER(AmunCfg), ctypes.c_int32, _P, _P,
                                          ctypes.POINTER(ctypes.c_int64)]
    L.amun_dev_refinement.argtypes = [vp, ctypes.POINTER(AmunRefineCfg),
                                      ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_float)]
    L.amun_dev_cons2prim_noalloc.argtypes = [ctypes.POINTER(AmunCfg), ctypes.c_int32,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]
    L.amun_dev_init_amr.argtypes = [ctypes.POINTER(AmunCfg), ctypes.c_int32,
                                    ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(vp)]
    L.amun_dev_amr_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
    L.amun_dev_amr_remesh.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.amun_dev_stream.restype = ctypes.c_void_p
    L.amun_dev_last_error.argtypes = []
    L.amun_dev_last_error.restype = ctypes.c_char_p
    for name in ("amun_dev_sync", "amun_dev_positivity_events", "amun_dev_tridiag_events"):
        getattr(L, name).argtypes = [vp] + ([ctypes.POINTER(ctypes.c_uint64)]
                                            if name.endswith("events") else [])
    _lib = L
    return L


class AmunError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"amun_dev error {code}: {msg}")
        self.code = code


def _dp(a):
    return None if a is None else a.ctypes.data_as(_P)


class Solver:
    """One device context: the per-block update of one run on ``cuda:0``."""

    nruns = 1

    def __init__(self, params: Params):
        self.lib = load_library()
        self.params = params
        self.cfg = params.to_cfg()
        h = ctypes.c_void_p()
        self._check(self.lib.amun_dev_init(ctypes.byref(self.cfg), ctypes.byref(h)))
        self.h = h

    def _check(self, rc: int):
        if rc != 0:
            raise AmunError(rc, self.lib.amun_dev_last_error().decode(errors="replace"))

    def close(self):
        if getattr(self, "h", None):
            self.lib.amun_dev_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ transfers
    def upload(self, u: np.ndarray):
        """Global conserved state U[nv][nz][ny][nx]."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        nz, ny, nx = self.params.shape
        if u.shape != (self.params.nv, nz, ny, nx):
            raise ValueError(f"state shape {u.shape} != {(self.params.nv, nz, ny, nx)}")
        self._check(self.lib.amun_dev_upload_global(self.h, _dp(u)))

    def download(self, with_q: bool = False):
        nz, ny, nx = self.params.shape
        u = np.empty((self.params.nv, nz, ny, nx))
        q = np.empty_like(u) if with_q else None
        self._check(self.lib.amun_dev_download_global(self.h, _dp(u), _dp(q)))
        return (u, q) if with_q else u

    def _block_shape(self):
        p = self.params
        im = p.ncells + 2 * p.nghosts
        km = im if p.ndims == 3 else 1
        return (p.nblocks, km, im, im, p.nv)

    def upload_blocks(self, u: np.ndarray):
        """Host blocks u0(nv,im,jm,km) (Fortran order) one after another."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != self._block_shape():
            raise ValueError(f"block array shape {u.shape} != {self._block_shape()}")
        self._check(self.lib.amun_dev_upload_blocks(self.h, _dp(u)))

    def download_blocks(self, with_q: bool = False):
        u = np.empty(self._block_shape())
        q = np.empty_like(u) if with_q else None
        self._check(self.lib.amun_dev_download_blocks(self.h, _dp(u), _dp(q)))
        return (u, q) if with_q else u

    def block_coords(self) -> np.ndarray:
        n = ctypes.c_int32()
        self._check(self.lib.amun_dev_nblocks(self.h, ctypes.byref(n)))
        c = np.empty((n.value, 3), dtype=np.int32)
        self._check(self.lib.amun_dev_block_coords(
            self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return c

    # ------------------------------------------------------------ stepping
    def new_time_step(self, dtnext: float = 2.0e30):
        dt, cm = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.amun_dev_new_time_step(self.h, dtnext, ctypes.byref(dt),
                                                    ctypes.byref(cm)))
        return dt.value, cm.value

    def advance(self, dtnext: float = 2.0e30):
        """One step; returns (dt, cmax) used by it."""
        dt, cm = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.amun_dev_step(self.h, dtnext, ctypes.byref(dt), ctypes.byref(cm)))
        return dt.value, cm.value

    def run(self, nsteps: int, dtnext: float = 2.0e30) -> np.ndarray:
        """nsteps steps queued back to back; returns the dt of each step
        (shape (nsteps,) for one run, (nsteps, nruns) for a batch)."""
        dts = np.zeros((max(nsteps, 1), self.nruns))
        self._check(self.lib.amun_dev_run(self.h, nsteps, dtnext, _dp(dts)))
        return dts[:nsteps, 0] if self.nruns == 1 else dts[:nsteps]

    def set_graph(self, enable: bool = True):
        """Capture steps into HIP graphs and replay them."""
        self._check(self.lib.amun_dev_set_graph(self.h, int(enable)))

    def set_stage_fusion(self, enable: bool = True):
        """One launch per stage: the x, y (and z) sweeps of a stage in one
        kernel whose workgroups hold whole blocks (bitwise the same
        results)."""
        self._check(self.lib.amun_dev_set_stage_fusion(self.h, int(enable)))

    def autotune(self, nsteps: int = 3):
        """Time every launch shape on the live state (which is restored) and
        keep the fastest per sweep direction; returns [(wg, xcd_order)...]."""
        ch = np.zeros(6, dtype=np.int32)
        self._check(self.lib.amun_dev_autotune(
            self.h, nsteps, ch.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return [(int(ch[2 * d]), int(ch[2 * d + 1])) for d in range(self.params.ndims)]

    # ----------------------------------------------------------- gradients
    def tape_forward(self, nsteps: int, interval: int = 20, dtnext: float = 2.0e30):
        """nsteps steps with checkpoints every interval steps (for
        tape_backward / tape_tangent)."""
        self._taped = nsteps
        self._check(self.lib.amun_dev_tape_forward(self.h, nsteps, interval, dtnext))

    def tape_steps(self):
        """(dt, cmax) of every taped step."""
        dt, cm = np.zeros(self._taped), np.zeros(self._taped)
        self._check(self.lib.amun_dev_tape_steps(self.h, _dp(dt), _dp(cm)))
        return dt, cm

    def tape_backward(self, lam: np.ndarray):
        """dJ/dU0 and dJ/dgamma from lam = dJ/dU of the final state."""
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        lam0 = np.empty_like(lam)
        dg = ctypes.c_double()
        self._check(self.lib.amun_dev_tape_backward(self.h, _dp(lam), _dp(lam0), ctypes.byref(dg)))
        return lam0, dg.value

    def tape_tangent(self, du0: np.ndarray, dgamma: float = 0.0) -> np.ndarray:
        """dU of the final state for the initial perturbation (du0, dgamma)."""
        du0 = np.ascontiguousarray(du0, dtype=np.float64)
        du = np.empty_like(du0)
        self._check(self.lib.amun_dev_tape_tangent(self.h, _dp(du0), float(dgamma), _dp(du)))
        return du

    def run_fixed(self, dt: np.ndarray, cmax: np.ndarray):
        """Ordinary steps with dt and cmax imposed per step."""
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        cmax = np.ascontiguousarray(cmax, dtype=np.float64)
        self._check(self.lib.amun_dev_run_fixed(self.h, len(dt), _dp(dt), _dp(cmax)))

    def set_launch(self, choice):
        ch = np.zeros(6, dtype=np.int32)
        for d, (wg, z) in enumerate(choice):
            ch[2 * d], ch[2 * d + 1] = wg, z
        self._check(self.lib.amun_dev_set_launch(
            self.h, ch.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))

    def launch(self, nsteps: int, dtnext: float = 2.0e30):
        """Queue nsteps without reading anything back (timing loops)."""
        self._check(self.lib.amun_dev_run(self.h, nsteps, dtnext, None))

    def timed_run(self, nsteps: int, dtnext: float = 2.0e30, per_kernel: bool = False):
        """Queue nsteps with HIP events on the solver stream; returns the
        device time in ms and (per_kernel) the average ms of each kernel of a
        step in launch order."""
        ms = ctypes.c_double()
        nk = ctypes.c_int32()
        kms = np.zeros(64)
        self._check(self.lib.amun_dev_timed_run(
            self.h, nsteps, dtnext, ctypes.byref(ms), _dp(kms) if per_kernel else None,
            64 if per_kernel else 0, ctypes.byref(nk)))
        return ms.value, (kms[: nk.value].copy() if per_kernel else None)

    def sync(self):
        self._check(self.lib.amun_dev_sync(self.h))

    @property
    def time(self) -> float:
        t, s = ctypes.c_double(), ctypes.c_int64()
        self._check(self.lib.amun_dev_get_time(self.h, ctypes.byref(t), ctypes.byref(s)))
        return t.value

    @property
    def step(self) -> int:
        t, s = ctypes.c_double(), ctypes.c_int64()
        self._check(self.lib.amun_dev_get_time(self.h, ctypes.byref(t), ctypes.byref(s)))
        return s.value

    def integrals(self) -> dict:
        out = np.zeros(8)
        self._check(self.lib.amun_dev_integrals(self.h, _dp(out)))
        keys = ("mass", "momx", "momy", "momz", "ener", "ekin", "emag", "eint")
        return dict(zip(keys, out.tolist()))

    def statistics(self) -> np.ndarray:
        """store_integrals' statistics: (mean, min, max) of density, pressure,
        |v|, |B|, psi, sonic and Alfvenic Mach number, shape (21,)."""
        out = np.zeros(21)
        self._check(self.lib.amun_dev_statistics(self.h, _dp(out)))
        return out

    def positivity_events(self) -> int:
        n = ctypes.c_uint64()
        self._check(self.lib.amun_dev_positivity_events(self.h, ctypes.byref(n)))
        return n.value

    def refinement(self, rc=None, minlev: int = 0, maxlev: int = 0):
        """check_refinement_criterion of every block of the current state
        (refinement.F90:257): (criterion per block in {-1, 0, 1}, indicators
        (block, 3): largest second-derivative error, vorticity, current
        density).  rc: an AmunRefineCfg, by default from the run parameters
        (config.refine_cfg)."""
        from .config import refine_cfg
        if rc is None:
            rc = refine_cfg(self.params, minlev, maxlev)
        n = ctypes.c_int32()
        self._check(self.lib.amun_dev_nblocks(self.h, ctypes.byref(n)))
        crit = np.zeros(n.value, dtype=np.int32)
        ind = np.zeros((n.value, 3), dtype=np.float32)
        self._check(self.lib.amun_dev_refinement(
            self.h, ctypes.byref(rc), crit.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            ind.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return crit, ind

    def tridiag_events(self) -> int:
        """Compact-scheme pencils whose tridiagonal solve met a zero pivot."""
        n = ctypes.c_uint64()
        self._check(self.lib.amun_dev_tridiag_events(self.h, ctypes.byref(n)))
        return n.value

    def c2p_events(self) -> list:
        """[bracketing failed, p <= 0, not converged] of the relativistic
        primitive solver so far."""
        n = (ctypes.c_uint64 * 3)()
        self._check(self.lib.amun_dev_c2p_events(self.h, n))
        return list(n)


class AmrSolver(Solver):
    """A multi-level block mesh (amun_dev_init_amr), frozen: ``meta`` holds
    (level, i0, j0, k0) per leaf block, the first interior cell in cells of
    the block's level.  State moves as per-block interiors in meta order,
    u[b][k][j][i][v] (the Fortran u(nv, ib:ie, jb:je, kb:ke) of every block,
    one after another)."""

    def __init__(self, params: Params, meta: np.ndarray):
        self.lib = load_library()
        self.params = params
        self.cfg = params.to_cfg()
        self.meta = np.ascontiguousarray(meta, dtype=np.int32)
        h = ctypes.c_void_p()
        self._check(self.lib.amun_dev_init_amr(
            ctypes.byref(self.cfg), self.meta.shape[0],
            self.meta.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(h)))
        self.h = h

    def remesh(self, meta: np.ndarray):
        """Move the state onto the leaf set meta on the device (update_mesh's
        copies, prolong_block and restrict_block); the solver then runs on
        the new mesh from the same time and step."""
        m = np.ascontiguousarray(meta, dtype=np.int32)
        self._check(self.lib.amun_dev_amr_remesh(
            self.h, m.shape[0], m.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        self.meta = m

    def info(self) -> dict:
        a = np.zeros(6, dtype=np.int32)
        self._check(self.lib.amun_dev_amr_info(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        keys = ("blocks", "virtual_blocks", "levels", "restricted", "prolonged", "corrected_faces")
        return dict(zip(keys, (int(x) for x in a)))

    def _block_shape(self):
        p = self.params
        im = p.ncells + 2 * p.nghosts
        km = im if p.ndims == 3 else 1
        return (self.meta.shape[0], km, im, im, p.nv)

    def _interior(self, a):
        p = self.params
        g = p.nghosts
        if p.ndims == 2:
            return a[:, 0, g:g + p.ncells, g:g + p.ncells, :]
        return a[:, g:g + p.ncells, g:g + p.ncells, g:g + p.ncells, :]

    def upload(self, u: np.ndarray):
        full = np.zeros(self._block_shape())
        inner = self._interior(full)
        inner[...] = np.asarray(u, dtype=np.float64).reshape(inner.shape)
        self.upload_blocks(full)

    def download(self, with_q: bool = False):
        """Interiors of u (and q with its ghost layers, as after
        boundary_variables, if with_q)."""
        if with_q:
            u, q = self.download_blocks(with_q=True)
        else:
            u, q = self.download_blocks(), None
        ui = np.ascontiguousarray(self._interior(u))
        if self.params.ndims == 2:
            ui = ui.reshape(ui.shape[0], self.params.ncells, self.params.ncells, -1)
            if with_q:
                q = q[:, 0]
        return (ui, q) if with_q else ui


class Batch(Solver):
    """Many independent uniform grids in one device context
    (amun_dev_init_batch): one launch per sweep over all their blocks, each
    grid with its own cell size, time step and time.  ``grids`` holds one
    Params per run; only their block counts and bounds may differ from
    ``params``."""

    def __init__(self, params: Params, grids):
        self.lib = load_library()
        self.params = params
        self.grids = list(grids)
        self.nruns = len(self.grids)
        self.cfg = params.to_cfg()
        arr = (AmunGrid * self.nruns)()
        for g, p in zip(arr, self.grids):
            g.xblocks, g.yblocks = p.xblocks, p.yblocks
            g.zblocks = p.zblocks if p.ndims == 3 else 1
            g.xmin, g.xmax, g.ymin, g.ymax = p.xmin, p.xmax, p.ymin, p.ymax
            g.zmin, g.zmax = (p.zmin, p.zmax) if p.ndims == 3 else (0.0, 1.0)
        h = ctypes.c_void_p()
        self._check(self.lib.amun_dev_init_batch(ctypes.byref(self.cfg), self.nruns, arr,
                                                 ctypes.byref(h)))
        self.h = h

    def upload(self, r: int, u: np.ndarray):
        """Global conserved state of run r, U[nv][nz][ny][nx]."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        nz, ny, nx = self.grids[r].shape
        if u.shape != (self.params.nv, nz, ny, nx):
            raise ValueError(f"state shape {u.shape} != {(self.params.nv, nz, ny, nx)}")
        self._check(self.lib.amun_dev_upload_run(self.h, r, _dp(u)))

    def download(self, r: int, with_q: bool = False):
        nz, ny, nx = self.grids[r].shape
        u = np.empty((self.params.nv, nz, ny, nx))
        q = np.empty_like(u) if with_q else None
        self._check(self.lib.amun_dev_download_run(self.h, r, _dp(u), _dp(q)))
        return (u, q) if with_q else u

    def state(self):
        """(time, dt, step) arrays over the runs."""
        t, d = np.zeros(self.nruns), np.zeros(self.nruns)
        s = np.zeros(self.nruns, dtype=np.int64)
        self._check(self.lib.amun_dev_run_state(
            self.h, _dp(t), _dp(d), s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return t, d, s

    def integrals_run(self, r: int) -> dict:
        out = np.zeros(8)
        self._check(self.lib.amun_dev_integrals_run(self.h, r, _dp(out)))
        keys = ("mass", "momx", "momy", "momz", "ener", "ekin", "emag", "eint")
        return dict(zip(keys, out.tolist()))

    def statistics_run(self, r: int) -> np.ndarray:
        out = np.zeros(21)
        self._check(self.lib.amun_dev_statistics_run(self.h, r, _dp(out)))
        return out

    def run_blocks(self, r: int):
        f, n = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.lib.amun_dev_run_blocks(self.h, r, ctypes.byref(f), ctypes.byref(n)))
        return f.value, n.value


def _unit_check(lib, rc):
    if rc != 0:
        raise AmunError(rc, lib.amun_dev_last_error().decode(errors="replace"))


def unit_riemann(params: Params, ql: np.ndarray, qr: np.ndarray, cmax: float) -> np.ndarray:
    """Device Riemann solver of the configuration on state pairs (n, nv)."""
    lib = load_library()
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    f = np.empty_like(ql)
    cfg = params.to_cfg()
    _unit_check(lib, lib.amun_dev_unit_riemann(ctypes.byref(cfg), ql.shape[0], cmax, _dp(ql),
                                               _dp(qr), _dp(f)))
    return f


def unit_reconstruct(params: Params, fp: np.ndarray, h: float = 1.0):
    """Device 1D reconstruction of the configuration on pencils fp (npen, n):
    returns fl, fr (NaN where the window leaves the pencil)."""
    lib = load_library()
    fp = np.ascontiguousarray(fp, dtype=np.float64)
    fl = np.full_like(fp, np.nan)
    fr = np.full_like(fp, np.nan)
    cfg = params.to_cfg()
    _unit_check(lib, lib.amun_dev_unit_reconstruct(ctypes.byref(cfg), fp.shape[1], fp.shape[0],
                                                   h, _dp(fp), _dp(fl), _dp(fr)))
    return fl, fr


def unit_cons2prim(params: Params, u: np.ndarray, q0: np.ndarray, events: bool = False):
    """Device FP64 cons2prim on conserved vectors u (n, nv); q0 holds the
    previous primitives.  Returns q (and the event counts)."""
    lib = load_library()
    u = np.ascontiguousarray(u, dtype=np.float64)
    q = np.array(q0, dtype=np.float64, copy=True, order="C")
    ev = (ctypes.c_int64 * 3)()
    cfg = params.to_cfg()
    _unit_check(lib, lib.amun_dev_unit_cons2prim(ctypes.byref(cfg), u.shape[0], _dp(u), _dp(q),
                                                 ev))
    return (q, list(ev)) if events else q


def unit_eqs(params: Params, q: np.ndarray, cmax: float = 1.0, q2_init=None):
    """Device prim2cons / fluxspeed / cons2prim on states q (n, nv):
    returns u, q2, f, cm, cp."""
    lib = load_library()
    q = np.ascontiguousarray(q, dtype=np.float64)
    n = q.shape[0]
    u = np.empty_like(q)
    q2 = np.zeros_like(q) if q2_init is None else np.array(q2_init, dtype=np.float64, copy=True)
    f = np.empty_like(q)
    cm, cp = np.empty(n), np.empty(n)
    cfg = params.to_cfg()
    _unit_check(lib, lib.amun_dev_unit_eqs(ctypes.byref(cfg), n, cmax, _dp(q), _dp(u), _dp(q2),
                                           _dp(f), _dp(cm), _dp(cp)))
    return u, q2, f, cm, cp
