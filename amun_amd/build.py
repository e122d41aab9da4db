"""Build steps of the native parts (used by __graft_entry__.build()).

- amun_amd/libamun_hip.so : the HIP kernels + C-ABI, hipcc for gfx950
- amun_amd/host/amun_host.x : the Fortran 2003 host driver (amdflang)
- oracle/liboracle.so      : the C oracle (test infrastructure, gcc)
- oracle/_ref/*            : the reference binaries, only where the
                             reference sources exist (this container)
"""
from __future__ import annotations

import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "amun_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
ORACLE = os.path.join(ROOT, "oracle")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
FC = shutil.which("amdflang") or os.path.join(ROCM, "bin", "amdflang")

# -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 VALU ops;
# --offload-compress: the device code objects are stored zstd-compressed in
# the fat binary (about a third of the size; the HIP runtime unpacks them
# when the library is loaded)
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "--offload-compress", "-Wno-unused-value", "-Wno-unused-result"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def inst_entries():
    """(name, rec, eqsys, precision) of every kernel-table unit, read from
    csrc/inst_list.h."""
    import re
    text = open(os.path.join(CSRC, "inst_list.h")).read()
    return re.findall(r"X\((pick_\w+), (\w+), (\w+), (\w+)\)", text)


def hip_jobs():
    """(object, command) of every HIP compilation: the host unit of the C-ABI
    and one kernel-table unit per entry of inst_list.h."""
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    obj = os.path.join(bdir, "amun_dev.o")
    jobs = [(obj, [HIPCC, *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, "amun_dev.hip")])]
    for name, rec, eq, prec in inst_entries():
        obj = os.path.join(bdir, name + ".o")
        jobs.append((obj, [HIPCC, *HIP_FLAGS, f"-DAMUN_INST_ENTRY={name}",
                           f"-DAMUN_INST_REC={rec}", f"-DAMUN_INST_EQ={eq}",
                           f"-DAMUN_INST_PREC={prec}", "-c", "-o", obj,
                           os.path.join(CSRC, "inst_unit.hip")]))
    # host-only quad-precision weights (libquadmath): g++, not hipcc
    qobj = os.path.join(bdir, "gp_weights.o")
    jobs.append((qobj, ["g++", "-O2", "-ffp-contract=off", "-fPIC", "-c", "-o", qobj,
                        os.path.join(CSRC, "gp_weights.cpp")]))
    return jobs


def link_hip(objs):
    out = os.path.join(PKG, "libamun_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp",
          *objs, "-lquadmath"])
    os.replace(out + ".tmp", out)
    return out


def build_hip(jobs_max=None):
    """Compile the host unit and the kernel-table units in parallel (largest
    units first), then link libamun_hip.so."""
    from concurrent.futures import ThreadPoolExecutor
    jobs = hip_jobs()
    # the MHD and SRMHD tables hold the most kernels: start them first
    order = sorted(jobs, key=lambda j: (("_srmhd_" not in j[0]), ("_mhd_" not in j[0])))
    workers = jobs_max or max(1, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for f in [ex.submit(_run, cmd) for _, cmd in order]:
            f.result()
    return link_hip([obj for obj, _ in jobs])


def build_oracle():
    out = os.path.join(ORACLE, "liboracle.so")
    _run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", out,
          os.path.join(ORACLE, "amun_oracle.c"), "-lquadmath", "-lm"])
    return out


def build_host():
    """Fortran host driver linked against libamun_hip.so."""
    src = [os.path.join(HOST, "amun_dev_mod.F90"), os.path.join(HOST, "amun_host.F90")]
    if not all(os.path.exists(s) for s in src):
        return None
    out = os.path.join(HOST, "amun_host.x")
    mod = os.path.join(HOST, "_mod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for s in src:
        o = os.path.join(mod, os.path.basename(s) + ".o")
        _run([FC, "-cpp", "-O2", "-module-dir", mod, "-c", s, "-o", o])
        objs.append(o)
    _run([FC, "-O2", "-o", out, *objs, "-L" + PKG, "-lamun_hip",
          "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + PKG])
    return out


def build_reference():
    """Reference binaries for golden vectors; skipped where the reference
    sources are absent."""
    ref = os.environ.get("AMUN_REFERENCE_ROOT", "/root/reference")
    if not os.path.exists(os.path.join(ref, "src", "evolution.F90")):
        return None
    _run(["sh", os.path.join(ORACLE, "build_ref.sh"), ref])
    out = os.path.join(ORACLE, "_ref")
    missing = [x for x in ("amun_ref_2d.x", "amun_ref_3d.x")
               if not os.path.exists(os.path.join(out, x))]
    if missing:
        raise RuntimeError(f"reference build finished without {', '.join(missing)}")
    return out


def build_all():
    built = {"hip": build_hip(), "oracle": build_oracle(), "host": build_host(),
             "reference": build_reference()}
    return built
