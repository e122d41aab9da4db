"""The allocation-free cons2prim entry point (amun_dev_cons2prim_noalloc)
against amun_dev_unit_cons2prim, through the C-ABI: relativistic states with
failing recoveries (the event counters live in the library's static device
array), two calls in a row, the second on fewer states, bit for bit; and an
HD call on the same counters afterwards."""
import ctypes
import os

import numpy as np
import pytest

from amun_amd.config import Params
from amun_amd import solver

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _noalloc(lib, p, u, q0, torch):
    """One call on device copies of u and q0 on torch's current stream; the
    result and the event counts, read back after the stream is done."""
    du = torch.from_numpy(np.ascontiguousarray(u)).to("cuda")
    dq = torch.from_numpy(np.ascontiguousarray(q0)).to("cuda")
    dev = torch.full((3,), -1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    cfg = p.to_cfg()
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.amun_dev_cons2prim_noalloc(ctypes.byref(cfg), u.shape[0], du.data_ptr(),
                                        dq.data_ptr(), dev.data_ptr(), stream)
    assert rc == 0, lib.amun_dev_last_error()
    torch.cuda.synchronize()
    return dq.cpu().numpy(), dev.cpu().numpy().tolist()


@pytest.mark.gpu
def test_cons2prim_noalloc_matches_unit_cons2prim():
    import torch
    assert torch.cuda.is_available()
    lib = solver.load_library()
    g = np.load(os.path.join(GOLD, "srmhd_2dwv_blast2d_32_8.npz"), allow_pickle=False)
    p = Params.from_text(str(g["unit_params"]), ndims=2)
    ub, q0 = g["ub"], g["q0"]
    # all 464 states (31 failed recoveries), then the last 50 (17 of them):
    # the counters start from zero on every call
    n = ub.shape[0]
    evs = []
    for sl in (slice(0, n), slice(n - 50, n)):
        print(f"[case] srmhd noalloc cons2prim states {sl.start}..{sl.stop}", flush=True)
        q_ref, ev_ref = solver.unit_cons2prim(p, ub[sl], q0[sl], events=True)
        q_na, ev_na = _noalloc(lib, p, ub[sl], q0[sl], torch)
        np.testing.assert_array_equal(q_na, q_ref)
        assert ev_na == ev_ref
        evs.append(ev_na)
    assert sum(evs[0]) > sum(evs[1]) > 0
    ph = Params(ndims=3, equation_system="hd", reconstruction="tvd", riemann_solver="hllc")
    rng = np.random.default_rng(3)
    qh = np.column_stack([rng.uniform(0.5, 2, 200), rng.normal(size=(200, 3)),
                          rng.uniform(0.1, 1, 200)])
    from amun_amd.problems import prim_to_cons
    uh = np.ascontiguousarray(prim_to_cons(ph, qh.T).T)
    print("[case] hd noalloc cons2prim n=200", flush=True)
    q_ref = solver.unit_cons2prim(ph, uh, np.zeros_like(qh))
    q_na, ev_na = _noalloc(lib, ph, uh, np.zeros_like(qh), torch)
    np.testing.assert_array_equal(q_na, q_ref)
    assert ev_na == [0, 0, 0]


def test_cons2prim_noalloc_rejects_bad_arguments():
    lib = solver.load_library()
    p = Params(ndims=2, equation_system="hd", reconstruction="tvd", riemann_solver="hllc")
    cfg = p.to_cfg()
    assert lib.amun_dev_cons2prim_noalloc(ctypes.byref(cfg), 0, 1, 1, None, None) == 500
    assert lib.amun_dev_cons2prim_noalloc(ctypes.byref(cfg), 4, None, None, None, None) == 500
    cfg.eqsys = 7
    assert lib.amun_dev_cons2prim_noalloc(ctypes.byref(cfg), 4, 1, 1, None, None) == 100
