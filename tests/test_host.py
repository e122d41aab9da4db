"""The Fortran 2003 host driver (amun_amd/host): parameter parsing, seeded
problem set-up bit-identical to the Python side, configuration errors
(CPU), and a full run through the C-ABI against the Python binding (GPU)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from amun_amd.config import Params, c1_params, c4_params
from amun_amd import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "amun_amd", "host", "amun_host.x")

needs_host = pytest.mark.skipif(not os.path.exists(HOST), reason="host driver not built")


def run_host(p: Params, extra: dict, workdir: str, timeout=600):
    txt = p.to_params_in() + f"ndims = {p.ndims}\nprecision = {p.precision}\n"
    txt += "".join(f"{k} = {v}\n" for k, v in extra.items())
    with open(os.path.join(workdir, "params.in"), "w") as fh:
        fh.write(txt)
    return subprocess.run([HOST, "-i", "params.in"], cwd=workdir, capture_output=True,
                          text=True, timeout=timeout)


@needs_host
@pytest.mark.parametrize("problem", ["blast", "orszag_tang"])
def test_host_initial_state_matches_python(problem):
    if problem == "blast":
        p = c1_params()
        ref = problems.blast_hd_2d(p)
    else:
        p = Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
                   riemann_solver="hlld", time_advance="rk3", ncells=16, nghosts=2,
                   xblocks=4, yblocks=4)
        ref = problems.orszag_tang_2d(p)
    with tempfile.TemporaryDirectory() as d:
        r = run_host(p, {"problem": problem, "dump_initial": "ic.bin"}, d)
        assert r.returncode == 0, r.stderr
        ug = np.fromfile(os.path.join(d, "ic.bin")).reshape(ref.shape)
    # the SplitMix64 noise is bit-identical; sin/cos/sqrt of the two
    # runtimes may differ in the last ulp
    np.testing.assert_allclose(ug, ref, rtol=4e-16, atol=1e-300)
    if problem == "blast":
        np.testing.assert_array_equal(ug, ref)


@needs_host
def test_host_srblast_initial_state_matches_python():
    p = c4_params(64)
    ref = problems.srmhd_blast_2d(p)
    with tempfile.TemporaryDirectory() as d:
        r = run_host(p, {"problem": "srblast", "dump_initial": "ic.bin"}, d)
        assert r.returncode == 0, r.stderr
        ug = np.fromfile(os.path.join(d, "ic.bin")).reshape(ref.shape)
    np.testing.assert_array_equal(ug, ref)


@needs_host
def test_host_rejects_bad_configuration():
    p = c1_params()
    p.riemann_solver = "hlld"          # HLLD needs MHD
    with tempfile.TemporaryDirectory() as d:
        r = run_host(p, {"problem": "blast"}, d)
    assert r.returncode != 0
    assert "300" in r.stderr


@needs_host
@pytest.mark.gpu
def test_host_run_matches_python_binding():
    """C1 at 64^2 for 20 steps through the Fortran host (blocks in the
    reference layout, amun_dev_upload_blocks/step/download_blocks) equals the
    same run through the Python binding bit for bit."""
    from amun_amd.solver import Solver
    p = c1_params()
    p.xblocks = p.yblocks = 4
    u0 = problems.blast_hd_2d(p)
    with tempfile.TemporaryDirectory() as d:
        r = run_host(p, {"problem": "blast", "nmax": 20, "output": "out.bin"}, d)
        assert r.returncode == 0, r.stderr
        print(r.stdout)
        uh = np.fromfile(os.path.join(d, "out.bin")).reshape(u0.shape)
    with Solver(p) as s:
        s.upload(u0)
        for _ in range(20):
            s.advance()
        up = s.download()
    np.testing.assert_array_equal(uh, up)


@needs_host
@pytest.mark.gpu
def test_host_srmhd_run_matches_python_binding():
    """The relativistic blast at 64^2 for 20 steps through the Fortran host
    equals the same run through the Python binding bit for bit, and the host
    reports no primitive-recovery failures."""
    from amun_amd.solver import Solver
    p = c4_params(64)
    u0 = problems.srmhd_blast_2d(p)
    with tempfile.TemporaryDirectory() as d:
        r = run_host(p, {"problem": "srblast", "nmax": 20, "output": "out.bin"}, d)
        assert r.returncode == 0, r.stderr
        print(r.stdout)
        uh = np.fromfile(os.path.join(d, "out.bin")).reshape(u0.shape)
    assert "not converged = 0 0 0" in r.stdout
    with Solver(p) as s:
        s.upload(u0)
        for _ in range(20):
            s.advance()
        up = s.download()
    np.testing.assert_array_equal(uh, up)
