"""The C-ABI boundary without a GPU: the library loads, exports every entry
point declared in include/amun_dev.h, and rejects bad configurations with
the reference's error codes before touching the device."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

from amun_amd.config import Params, AmunCfg, ConfigError, c1_params, c2_params
from amun_amd import solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "amun_dev.h")).read()
    return sorted(set(re.findall(r"\b(amun_dev_\w+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    lib = solver.load_library()
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(solver.EXPORTS) <= set(syms)


def test_default_cfg_matches_reference_defaults():
    lib = solver.load_library()
    c = AmunCfg()
    lib.amun_dev_default_cfg(ctypes.byref(c))
    assert c.gamma == pytest.approx(5.0 / 3.0)       # equations.F90:167
    assert c.eps == 2.220446049250313e-16            # interpolations.F90:82
    assert c.alpha == 2.0                            # evolution.F90:68
    assert c.limiter == 1                            # "mm" default
    assert c.ncells == 8 and c.nghosts == 2          # coordinates.F90:42-45


def _init_code(p):
    lib = solver.load_library()
    cfg = p.to_cfg()
    h = ctypes.c_void_p()
    rc = lib.amun_dev_init(ctypes.byref(cfg), ctypes.byref(h))
    if rc == 0:
        lib.amun_dev_finalize(h)
    return rc


@pytest.mark.parametrize("field,value,code", [
    ("eqsys", 7, 100), ("eos", 3, 110), ("reconstruction", 16, 200),
    ("riemann_solver", 3, 300), ("time_advance", 9, 400), ("nghosts", 1, 600),
    ("ncells", 2, 500),
])
def test_config_errors(field, value, code):
    p = c1_params()
    cfg = p.to_cfg()
    setattr(cfg, field, value)
    lib = solver.load_library()
    h = ctypes.c_void_p()
    assert lib.amun_dev_init(ctypes.byref(cfg), ctypes.byref(h)) == code
    assert lib.amun_dev_last_error()


@pytest.mark.parametrize("key,value", [("fix_unphysical_cells", "on")])
def test_unsupported_terms_are_refused(key, value):
    """Parameters of terms outside the device path (the unphysical-cell fix)
    are carried into amun_cfg and refused with 600 instead of being
    ignored."""
    p = Params.from_text(f"reconstruction = tvd\nriemann_solver = hllc\n{key} = {value}\n",
                         ndims=2)
    assert key not in p.extra
    cfg = p.to_cfg()
    lib = solver.load_library()
    h = ctypes.c_void_p()
    assert lib.amun_dev_init(ctypes.byref(cfg), ctypes.byref(h)) == 600
    assert key.split("_")[0].encode() in lib.amun_dev_last_error()


@pytest.mark.parametrize("value,code", [("eglm", 1), ("heglm", 2), ("none", 0)])
def test_glm_source_terms_carried(value, code):
    """glm_source_terms reaches amun_cfg (eglm 1, heglm 2); values outside
    0..2 are refused with 500."""
    p = Params.from_text(f"equation_system = mhd\nriemann_solver = hlld\n"
                         f"glm_source_terms = {value}\n", ndims=2)
    cfg = p.to_cfg()
    assert cfg.glm_source_terms == code
    cfg.glm_source_terms = 3
    lib = solver.load_library()
    h = ctypes.c_void_p()
    assert lib.amun_dev_init(ctypes.byref(cfg), ctypes.byref(h)) == 500


@pytest.mark.parametrize("text,code", [
    ("equation_system = srmhd\nequation_of_state = iso\n", 110),
    ("equation_system = mhd\nequation_of_state = iso\ncsnd = 0.0\n", 500),
    ("equation_system = mhd\nequation_of_state = adi\nriemann_solver = hlldm\n", 300),
    ("equation_system = hd\nequation_of_state = iso\nriemann_solver = hlld\n", 300),
])
def test_isothermal_configuration_errors(text, code):
    """The isothermal systems are HD and MHD (equations.F90:399-830), with a
    positive sound speed; HLLD-M belongs to isothermal MHD
    (schemes.F90:219-221)."""
    p = Params.from_text(text, ndims=2)
    cfg = p.to_cfg()
    lib = solver.load_library()
    h = ctypes.c_void_p()
    assert lib.amun_dev_init(ctypes.byref(cfg), ctypes.byref(h)) == code
    assert lib.amun_dev_last_error()


def test_isothermal_configuration_carried():
    """equation_of_state and csnd reach amun_cfg; four and eight variables."""
    p = Params.from_text("equation_system = mhd\nequation_of_state = isothermal\n"
                         "csnd = 0.35\nriemann_solver = hlld-m\n", ndims=2)
    cfg = p.to_cfg()
    assert cfg.eos == 2 and cfg.csnd == 0.35 and cfg.riemann_solver == 5
    assert p.nv == 8
    assert Params(equation_system="hd", equation_of_state="iso").nv == 4
    lib = solver.load_library()
    d = AmunCfg()
    lib.amun_dev_default_cfg(ctypes.byref(d))
    assert d.csnd == 1.0                             # equations.F90:176


def test_four_vector_states_and_gravity_carried():
    """state_variables = 4vec (the relativistic systems' four-velocity
    reconstruction, schemes.F90:4481-4517) and the Rayleigh-Taylor gravity
    (problem = rt, gacc; gravity.F90:141-150, 283-330) reach amun_cfg."""
    p = Params.from_text("equation_system = srhd\nriemann_solver = hllc\n"
                         "state_variables = 4vec\n", ndims=2)
    assert "state_variables" not in p.extra
    cfg = p.to_cfg()
    assert cfg.states_4vec == 1 and cfg.gravity == 0
    p = Params.from_text("problem = rt\ngacc = -0.25\n", ndims=2)
    cfg = p.to_cfg()
    assert cfg.gravity == 1 and cfg.gacc == -0.25
    assert Params.from_text("problem = blast\n", ndims=2).to_cfg().gravity == 0


@pytest.mark.parametrize("key", ["viscosity", "resistivity"])
def test_viscous_coefficients_carried(key):
    """viscosity and resistivity reach amun_cfg; negative values are refused
    with 500 (zero switches the terms off, as sources.F90:340, 543)."""
    p = Params.from_text(f"equation_system = mhd\nriemann_solver = hlld\n{key} = 2.5e-3\n",
                         ndims=2)
    cfg = p.to_cfg()
    assert getattr(cfg, key) == 2.5e-3
    setattr(cfg, key, -1.0)
    lib = solver.load_library()
    h = ctypes.c_void_p()
    assert lib.amun_dev_init(ctypes.byref(cfg), ctypes.byref(h)) == 500


def test_weno_needs_four_ghosts():
    p = c2_params(32)
    p.nghosts = 2
    assert _init_code(p) == 600


def test_valid_config_reaches_the_device_query():
    rc = _init_code(c1_params())
    assert rc == 0 or rc >= 1000


def test_params_names_and_errors():
    text = """
    # AMUN parameter file
    equation_system = "mhd"
    reconstruction  = weno5z
    limiter         = mc
    riemann_solver  = hlld
    time_advance    = rk3
    cfl             = 0.3d0
    ncells = 16
    nghosts = 4
    limiter = mm
    problem = blast
    """
    p = Params.from_text(text)
    assert p.limiter == "mm" and p.cfl == 0.3 and p.extra["problem"] == "blast"
    assert p.to_cfg().riemann_solver == 3
    p.riemann_solver = "roe2"
    with pytest.raises(ConfigError) as e:
        p.to_cfg()
    assert e.value.code == 300


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "4"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_mlp_limiting_ghost_layers():
    """With two ghost layers the multi-dimensional limiting is built for the
    TVD and MGP reconstructions only (the reference's process reads the face
    state of each block's outermost cell); others need nghosts >= 3."""
    lib = solver.load_library()
    for rec, ng, code in (("weno3", 2, 600), ("limo3", 2, 600), ("weno3", 3, 0),
                          ("tvd", 2, 0)):
        p = Params(ndims=2, equation_system="mhd", riemann_solver="hlld", reconstruction=rec,
                   nghosts=ng, mlp_limiting="on")
        cfg = p.to_cfg()
        h = ctypes.c_void_p()
        rc = lib.amun_dev_init(ctypes.byref(cfg), ctypes.byref(h))
        assert rc == code or (code == 0 and rc >= 1000), (rec, ng, rc)
