"""Run configuration of the AMUN per-block update, keyed by the reference's
parameter names.

``Params`` holds the keys that select the hot path (``equation_system``,
``reconstruction``, ``limiter``, ``riemann_solver``, ``time_advance``, ``cfl``,
``gamma``, ``ncells``, ``nghosts``, ``xblocks``... exactly as they appear in an
AMUN ``params.in``, parsed by parameters.F90:99-686) and converts them to the
``amun_cfg`` C struct of ``include/amun_dev.h``.  Unknown names raise the same
integer error codes as the reference initialisers (equations.F90:399-830,
interpolations.F90:382).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field, fields

# constants mirrored from include/amun_dev.h
EQSYS = {"hd": 1, "hydrodynamic": 1, "mhd": 2, "magnetohydrodynamic": 2,
         "srhd": 3, "srmhd": 4}
EOS = {"adi": 1, "adiabatic": 1, "iso": 2, "isothermal": 2}
RECON = {"tvd": 1, "weno3": 2, "limo3": 3, "ppm": 4, "crweno5z": 8, "crweno5-z": 8,
         "crweno5yc": 9, "crweno5-yc": 9, "crweno5ns": 10, "crweno5-ns": 10, "crmp5": 13, "mgp": 21,
         "crmp5l": 14, "crmp5ld": 14, "crmp7": 15, "weno5z": 5, "weno5-z": 5, "weno5yc": 6, "weno5-yc": 6, "weno5ns": 7,
         "weno5-ns": 7, "mp5": 11, "mp7": 12, "gp": 20}
LIMITER = {"zero": 0, "mm": 1, "minmod": 1, "mc": 2, "monotonized_central": 2,
           "sb": 3, "superbee": 3, "vl": 4, "vanleer": 4, "va": 5,
           "vanalbada": 5}
RIEMANN = {"hll": 1, "hllc": 2, "hlld": 3, "roe": 4, "hlldm": 5, "hlld-m": 5}
INTEGRATOR = {"euler": 0, "rk2": 1, "rk3": 2, "ssprk(m,2)": 3, "rk3.4": 4, "ssprk(4,3)": 4,
              "rk3.5": 5, "ssprk(5,3)": 5}
PRECISION = {"fp64": 0, "mixed": 1}
# equations.F90:747-789
# boundaries.F90:5024-5637 (hydrostatic and user types are not built)
BNDTYPE = {"periodic": 0, "open": 1, "outflow": 2, "out": 2, "reflective": 3, "reflecting": 3,
           "reflect": 3}
PRIMSOLVER = {"1dw": 0, "1d(w)": 0, "2dwv": 1, "2d(w,v)": 1, "2dwu": 2, "2d(w,u)": 2}


class AmunCfg(ctypes.Structure):
    """ctypes mirror of ``amun_cfg`` (include/amun_dev.h)."""
    _fields_ = [
        ("ndims", ctypes.c_int32), ("eqsys", ctypes.c_int32),
        ("eos", ctypes.c_int32), ("nv", ctypes.c_int32),
        ("ncells", ctypes.c_int32), ("nghosts", ctypes.c_int32),
        ("xblocks", ctypes.c_int32), ("yblocks", ctypes.c_int32),
        ("zblocks", ctypes.c_int32),
        ("reconstruction", ctypes.c_int32), ("limiter", ctypes.c_int32),
        ("riemann_solver", ctypes.c_int32), ("time_advance", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("fix_positivity", ctypes.c_int32), ("clip_extrema", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
        ("eps", ctypes.c_double), ("alpha", ctypes.c_double),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
        ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
        ("tolerance", ctypes.c_double), ("pmin", ctypes.c_double),
        ("nr_maxit", ctypes.c_int32), ("nr_extra", ctypes.c_int32),
        ("sgp", ctypes.c_double), ("kappa", ctypes.c_double), ("kbeta", ctypes.c_double),
        ("ngp", ctypes.c_int32), ("stages", ctypes.c_int32),
        ("prol_limiter", ctypes.c_int32), ("primitive_solver", ctypes.c_int32),
        ("limo3_rad", ctypes.c_double), ("ppm_const", ctypes.c_double),
        ("bnd_type", ctypes.c_int32 * 6),
        ("mlp_limiting", ctypes.c_int32), ("fix_unphysical_cells", ctypes.c_int32),
        ("glm_source_terms", ctypes.c_int32), ("reserved_cfg", ctypes.c_int32),
        ("viscosity", ctypes.c_double), ("resistivity", ctypes.c_double),
        ("csnd", ctypes.c_double),
        ("states_4vec", ctypes.c_int32), ("reserved_cfg2", ctypes.c_int32),
        ("gravity", ctypes.c_int32), ("reserved_cfg3", ctypes.c_int32),
        ("gacc", ctypes.c_double),
    ]


class AmunRefineCfg(ctypes.Structure):
    """ctypes mirror of ``amun_refine_cfg`` (include/amun_dev.h)."""
    _fields_ = [
        ("crefmin", ctypes.c_double), ("crefmax", ctypes.c_double),
        ("epsref", ctypes.c_double), ("vortmin", ctypes.c_double),
        ("vortmax", ctypes.c_double), ("jabsmin", ctypes.c_double),
        ("jabsmax", ctypes.c_double), ("varmask", ctypes.c_int32),
        ("vort", ctypes.c_int32), ("jabs", ctypes.c_int32),
        ("minlev", ctypes.c_int32), ("maxlev", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


# primitive variable names of equations.F90:410-414, 513-521
PVARS = {1: ["dens", "velx", "vely", "velz", "pres"],
         3: ["dens", "velx", "vely", "velz", "pres"],
         2: ["dens", "velx", "vely", "velz", "magx", "magy", "magz", "bpsi", "pres"],
         4: ["dens", "velx", "vely", "velz", "magx", "magy", "magz", "bpsi", "pres"]}


def refine_cfg(p: "Params", minlev: int = 0, maxlev: int = 0) -> AmunRefineCfg:
    """The refinement criterion's parameters from the run parameters
    (p.extra), with the reference's defaults and its selection of variables
    by name: variable p counts when its name occurs in
    "refinement_variables" (refinement.F90:136-178, index(variables,
    pvars(p)) > 0); "vort" and (magnetic systems) "jabs" switch on the
    vorticity and current-density criteria."""
    x = p.extra

    def num(key, default):
        return float(str(x.get(key, default)).replace("d", "e").replace("D", "e"))
    names = str(x.get("refinement_variables", "dens pres"))
    eq = EQSYS[p.equation_system.lower()]
    c = AmunRefineCfg()
    c.crefmin = num("crefmin", 0.2)
    c.crefmax = num("crefmax", 0.8)
    c.epsref = num("epsref", 0.01)
    c.vortmin = num("vortmin", 1.0e-3)
    c.vortmax = num("vortmax", 0.1)
    c.jabsmin = num("jabsmin", 1.0e-3)
    c.jabsmax = num("jabsmax", 0.1)
    pv = PVARS[eq]
    if EOS.get(p.equation_of_state.lower(), 1) == 2:
        pv = pv[:-1]      # the isothermal systems have no pressure (equations.F90:410-414)
    c.varmask = sum(1 << i for i, n in enumerate(pv) if n in names)
    c.vort = int("vort" in names)
    c.jabs = int("jabs" in names and eq != 1)
    c.minlev, c.maxlev = int(minlev), int(maxlev)
    return c


class ConfigError(ValueError):
    """Invalid parameter, carrying the reference's integer error code."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _onoff(v) -> int:
    if isinstance(v, bool):
        return int(v)
    return int(str(v).strip().lower() in ("on", "t", "y", "true", "yes", "1"))


@dataclass
class Params:
    """The hot-path subset of an AMUN parameter file (defaults = reference
    defaults: equations.F90:130-200, interpolations.F90:80-118, 220-238,
    evolution.F90:65-80, coordinates.F90:42-80)."""
    ndims: int = 3
    equation_system: str = "hd"
    equation_of_state: str = "adi"
    csnd: float = 1.0
    state_variables: str = "primitive"
    reconstruction: str = "tvd"
    limiter: str = "mm"
    riemann_solver: str = "hlld"
    time_advance: str = "rk2"
    precision: str = "fp64"
    fix_positivity: str = "off"
    clip_extrema: str = "off"
    gamma: float = 5.0 / 3.0
    cfl: float = 0.5
    eps: float = 2.220446049250313e-16
    alpha: float = 2.0
    ncells: int = 8
    nghosts: int = 2
    xblocks: int = 1
    yblocks: int = 1
    zblocks: int = 1
    xmin: float = 0.0
    xmax: float = 1.0
    ymin: float = 0.0
    ymax: float = 1.0
    zmin: float = 0.0
    zmax: float = 1.0
    tolerance: float = 1.0e-10
    pmin: float = 1.0e-16
    nr_maxit: int = 100
    nr_extra: int = 2
    sgp: float = 9.0
    kappa: float = 1.0
    kbeta: float = 1.0
    ngp: int = 5
    stages: int = 2
    prolongation_limiter: str = "mm"
    primitive_solver: str = "1Dw"
    xlbndry: str = "periodic"
    xubndry: str = "periodic"
    ylbndry: str = "periodic"
    yubndry: str = "periodic"
    zlbndry: str = "periodic"
    zubndry: str = "periodic"
    limo3_rad: float = 0.5
    ppm_const: float = 1.25
    mlp_limiting: str = "off"
    fix_unphysical_cells: str = "off"
    glm_source_terms: str = "none"
    viscosity: float = 0.0
    resistivity: float = 0.0
    extra: dict = field(default_factory=dict)

    # ---------------------------------------------------------------- derived
    @property
    def nv(self) -> int:
        iso = EOS.get(self.equation_of_state.lower(), 1) == 2
        return (5 if EQSYS[self.equation_system.lower()] in (1, 3) else 9) - (1 if iso else 0)

    @property
    def shape(self):
        """Global grid shape (nz, ny, nx) of the base level."""
        nx = self.xblocks * self.ncells
        ny = self.yblocks * self.ncells
        nz = self.zblocks * self.ncells if self.ndims == 3 else 1
        return (nz, ny, nx)

    @property
    def nblocks(self) -> int:
        return self.xblocks * self.yblocks * (self.zblocks if self.ndims == 3 else 1)

    @property
    def ncells_total(self) -> int:
        nz, ny, nx = self.shape
        return nz * ny * nx

    def to_cfg(self) -> AmunCfg:
        def pick(table, key, name, code):
            k = str(key).strip().lower()
            if k not in table:
                raise ConfigError(code, f"unsupported {name}: {key}")
            return table[k]

        c = AmunCfg()
        c.ndims = int(self.ndims)
        c.eqsys = pick(EQSYS, self.equation_system, "equation_system", 100)
        c.eos = pick(EOS, self.equation_of_state, "equation_of_state", 110)
        c.nv = self.nv
        c.ncells = int(self.ncells)
        c.nghosts = int(self.nghosts)
        c.xblocks, c.yblocks, c.zblocks = int(self.xblocks), int(self.yblocks), int(self.zblocks)
        c.reconstruction = pick(RECON, self.reconstruction, "reconstruction", 200)
        # unknown limiter names select limiter_zero (interpolations.F90:394-397)
        c.limiter = LIMITER.get(str(self.limiter).strip().lower(), 0)
        c.riemann_solver = pick(RIEMANN, self.riemann_solver, "riemann_solver", 300)
        c.time_advance = pick(INTEGRATOR, self.time_advance, "time_advance", 400)
        c.precision = pick(PRECISION, self.precision, "precision", 500)
        c.fix_positivity = _onoff(self.fix_positivity)
        c.clip_extrema = _onoff(self.clip_extrema)
        c.gamma, c.cfl, c.eps, c.alpha = self.gamma, self.cfl, self.eps, self.alpha
        c.xmin, c.xmax, c.ymin, c.ymax = self.xmin, self.xmax, self.ymin, self.ymax
        c.zmin, c.zmax = self.zmin, self.zmax
        c.tolerance, c.pmin = self.tolerance, self.pmin
        c.nr_maxit, c.nr_extra = int(self.nr_maxit), int(self.nr_extra)
        c.sgp, c.kappa, c.kbeta, c.ngp = self.sgp, self.kappa, self.kbeta, int(self.ngp)
        c.stages = int(self.stages)
        c.limo3_rad, c.ppm_const = self.limo3_rad, self.ppm_const
        c.primitive_solver = pick(PRIMSOLVER, self.primitive_solver, "primitive_solver", 110)
        for n, side in enumerate((self.xlbndry, self.xubndry, self.ylbndry, self.yubndry,
                                  self.zlbndry, self.zubndry)):
            c.bnd_type[n] = pick(BNDTYPE, side, "boundary type", 600)
        # interpolations.F90:412-428: mm, mc, sb, vl, anything else the zero slope
        c.mlp_limiting = _onoff(self.mlp_limiting)
        c.fix_unphysical_cells = _onoff(self.fix_unphysical_cells)
        # sources.F90:128-141: eglm, heglm, anything else none
        c.glm_source_terms = {"eglm": 1, "heglm": 2}.get(
            str(self.glm_source_terms).strip().lower(), 0)
        c.viscosity, c.resistivity = float(self.viscosity), float(self.resistivity)
        c.csnd = float(self.csnd)
        c.states_4vec = int(self.state_variables.lower() in ("4vec", "4-vector"))
        # gravity.F90:141-150: the Rayleigh-Taylor problem turns gravity on
        c.gravity = int(str(self.extra.get("problem", "")).strip().lower()
                        in ("rt", "rayleightaylor", "rayleigh-taylor"))
        c.gacc = float(str(self.extra.get("gacc", "-1.0")).replace("d", "e").replace("D", "e"))
        c.prol_limiter = {"mm": 1, "minmod": 1, "mc": 2, "monotonized_central": 2,
                          "sb": 3, "superbee": 3, "vl": 4, "vanleer": 4}.get(
                              self.prolongation_limiter.lower(), 0)
        return c

    def dx(self):
        nz, ny, nx = self.shape
        d = [(self.xmax - self.xmin) / nx, (self.ymax - self.ymin) / ny]
        if self.ndims == 3:
            d.append((self.zmax - self.zmin) / nz)
        return d

    def to_params_in(self) -> str:
        """Text of an AMUN params.in with these settings (periodic, one level)."""
        eq = {1: "hd", 2: "mhd", 3: "srhd", 4: "srmhd"}[EQSYS[self.equation_system.lower()]]
        lines = [
            f"equation_system = {eq}", f"equation_of_state = {self.equation_of_state}",
            f"csnd = {self.csnd!r}", f"state_variables = {self.state_variables}",
            f"reconstruction = {self.reconstruction}", f"limiter = {self.limiter}",
            f"riemann_solver = {self.riemann_solver}",
            f"time_advance = {self.time_advance}",
            f"fix_positivity = {self.fix_positivity}",
            f"clip_extrema = {self.clip_extrema}",
            f"gamma = {self.gamma!r}", f"cfl = {self.cfl!r}", f"eps = {self.eps!r}",
            f"alpha = {self.alpha!r}",
            f"ncells = {self.ncells}", f"nghosts = {self.nghosts}",
            f"xblocks = {self.xblocks}", f"yblocks = {self.yblocks}",
            f"zblocks = {self.zblocks}",
            f"xmin = {self.xmin!r}", f"xmax = {self.xmax!r}",
            f"ymin = {self.ymin!r}", f"ymax = {self.ymax!r}",
            f"zmin = {self.zmin!r}", f"zmax = {self.zmax!r}",
            f"tolerance = {self.tolerance!r}", f"pmin = {self.pmin!r}",
            f"nr_maxit = {self.nr_maxit}", f"nr_extra = {self.nr_extra}",
            f"sgp = {self.sgp!r}", f"kappa = {self.kappa!r}", f"kbeta = {self.kbeta!r}",
            f"ngp = {self.ngp}", f"stages = {self.stages}",
            f"prolongation_limiter = {self.prolongation_limiter}",
            f"limo3_rad = {self.limo3_rad!r}", f"ppm_const = {self.ppm_const!r}",
            f"primitive_solver = {self.primitive_solver}",
            f"mlp_limiting = {self.mlp_limiting}",
            f"fix_unphysical_cells = {self.fix_unphysical_cells}",
            f"glm_source_terms = {self.glm_source_terms}",
            f"viscosity = {self.viscosity!r}", f"resistivity = {self.resistivity!r}",
            "minlev = 1", "maxlev = 1",
            f"xlbndry = {self.xlbndry}", f"xubndry = {self.xubndry}",
            f"ylbndry = {self.ylbndry}", f"yubndry = {self.yubndry}",
            f"zlbndry = {self.zlbndry}", f"zubndry = {self.zubndry}",
            "problem = blast", "snapshot_interval = 0", "restart_interval = 0",
            "integrals_interval = 1000000", "tmax = 1.0e+30", "nmax = 1000000000",
        ]
        # keys given in extra (e.g. maxlev, problem) replace the defaults
        lines = [ln for ln in lines if ln.split(" = ")[0] not in self.extra]
        for k, v in self.extra.items():
            lines.append(f"{k} = {v}")
        return "\n".join(lines) + "\n"

    @classmethod
    def from_text(cls, text: str, ndims: int = 3) -> "Params":
        """Parse ``key = value`` lines the way parameters.F90 does: ``#``
        comments, optional quotes, last occurrence wins, unknown keys kept in
        ``extra``."""
        p = cls(ndims=ndims)
        names = {f.name for f in fields(cls)} - {"extra", "ndims"}
        for raw in text.splitlines():
            line = raw.split("#", 1)[0].strip()
            if "=" not in line:
                continue
            k, v = (s.strip() for s in line.split("=", 1))
            v = v.strip("'\"")
            if k in names:
                cur = getattr(p, k)
                if isinstance(cur, bool) or isinstance(cur, str):
                    setattr(p, k, v)
                elif isinstance(cur, int):
                    setattr(p, k, int(v))
                else:
                    setattr(p, k, float(v.replace("d", "e").replace("D", "e")))
            else:
                p.extra[k] = v
        return p


def c1_params() -> Params:
    """2D adiabatic hydro blast, 128^2 as 8x8 blocks of 16^2 (+2 ghosts),
    TVD-MC + HLLC + SSPRK(2,2), CFL 0.4, periodic, gamma 1.4."""
    return Params(ndims=2, equation_system="hd", reconstruction="tvd",
                  limiter="mc", riemann_solver="hllc", time_advance="rk2",
                  gamma=1.4, cfl=0.4, ncells=16, nghosts=2, xblocks=8,
                  yblocks=8, zblocks=1, xmin=-0.5, xmax=0.5, ymin=-0.5,
                  ymax=0.5)


def c2_params(n: int = 256, precision: str = "fp64") -> Params:
    """3D ideal MHD with GLM cleaning, n^3 as (n/16)^3 blocks of 16^3
    (+4 ghosts), WENO5-Z + HLLD + SSPRK(3,3), CFL 0.3, periodic, gamma 5/3.
    The TVD limiter of the WENO5-Z edge points keeps the reference default
    (minmod)."""
    nb = n // 16
    return Params(ndims=3, equation_system="mhd", reconstruction="weno5z",
                  limiter="mm", riemann_solver="hlld", time_advance="rk3",
                  precision=precision, gamma=5.0 / 3.0, cfl=0.3, ncells=16,
                  nghosts=4, xblocks=nb, yblocks=nb, zblocks=nb, xmin=-0.5,
                  xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5)


def c4_params(n: int = 1024, precision: str = "fp64") -> Params:
    """2D special-relativistic MHD blast, n^2 as (n/16)^2 blocks of 16^2
    (+2 ghosts), TVD minmod + HLL + SSPRK(2,2), CFL 0.25, periodic, primitive
    solver 1D(W) with tolerance 1e-12 and at most 50 Newton iterations.  The
    adiabatic index is the reference default 5/3 (equations.F90:167)."""
    nb = n // 16
    return Params(ndims=2, equation_system="srmhd", reconstruction="tvd",
                  limiter="mm", riemann_solver="hll", time_advance="rk2",
                  precision=precision, gamma=5.0 / 3.0, cfl=0.25, ncells=16,
                  nghosts=2, xblocks=nb, yblocks=nb, zblocks=1, xmin=-0.5,
                  xmax=0.5, ymin=-0.5, ymax=0.5, tolerance=1.0e-12,
                  nr_maxit=50)


C3_SIZES = (64, 128, 256, 512)


def c3_batch_params(nruns: int = 64, seed: int = 20261202, precision: str = "fp64"):
    """64 independent 2D MHD Kelvin-Helmholtz runs whose grid sizes are drawn
    by seed from {64^2, 128^2, 256^2, 512^2} (SplitMix64 stream 11), each on
    [-0.5, 0.5]^2 in blocks of 16^2 (+2 ghosts), TVD-MC + HLLD + SSPRK(3,3),
    CFL 0.4.  Returns (common Params, [Params of every run])."""
    from .problems import uniform
    xi = uniform(seed, nruns, stream=11)
    common = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
                    riemann_solver="hlld", time_advance="rk3", precision=precision,
                    gamma=5.0 / 3.0, cfl=0.4, ncells=16, nghosts=2, xmin=-0.5,
                    xmax=0.5, ymin=-0.5, ymax=0.5)
    runs = []
    for x in xi:
        n = C3_SIZES[min(int(4.0 * x), 3)]
        p = Params(**{f.name: getattr(common, f.name) for f in fields(Params) if f.name != "extra"})
        p.xblocks = p.yblocks = n // 16
        runs.append(p)
    return common, runs


def psi_decay(p: Params) -> float:
    return math.exp(-p.alpha * p.cfl)
