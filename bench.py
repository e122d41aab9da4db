#!/usr/bin/env python3
"""Benchmark of the AMUN per-block finite-volume update on one MI355X.

Workload: 3D ideal MHD with GLM cleaning, 256^3 cells as 4096 blocks of 16^3,
WENO5-Z + HLLD + SSPRK(3,3), CFL 0.3, periodic (the 256^3 MHD configuration),
seeded turbulent initial state (synthetic).  One "step" is one full SSPRK(3,3)
step of the whole grid, including the CFL reduction, on the device.

Prints one JSON line: value = zone-updates/s (interior cells x steps / wall
time of the K timed steps, bracketed by device synchronisation), the roofline
of the step kernels against the MI355X HBM peak, and a CPU baseline timed on
the host cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
FP64_VEC_TFLOPS = 78.6
FP32_VEC_TFLOPS = 157.3


def bytes_per_zone_update(p) -> int:
    """Algorithmic bytes: state-array passes per step x nv x 8 B (SSPRK(3,3):
    8 passes, SSPRK(2,2): 5, Euler: 2), perfect on-chip stencil reuse."""
    passes = {"rk3": 8, "rk2": 5, "euler": 2}[p.time_advance]
    return passes * p.nv * 8


def ensure_built():
    lib = os.path.join(ROOT, "amun_amd", "libamun_hip.so")
    if not os.path.exists(lib):
        import __graft_entry__
        __graft_entry__.build()


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_oracle(p, nsteps):
    """Seconds of nsteps steps of the C oracle (one host thread) from the
    seeded initial state of p."""
    from amun_amd.problems import initial_state
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol

    o = ol.OracleRun(p)
    o.set_state(initial_state(p))
    t0 = time.perf_counter()
    o.run(nsteps)
    return time.perf_counter() - t0


def cpu_baseline(p_full):
    """The C oracle port of the reference update (one host thread) on a 64^3
    sample of the same 3D MHD configuration, 2 steps, and on the whole 2D HD
    128^2 x 200-step configuration beside it."""
    from amun_amd.config import c1_params, c2_params

    p = c2_params(64, precision="fp64")
    nsteps = 2
    secs = _time_oracle(p, nsteps)
    p1 = c1_params()
    n1 = 200
    secs1 = _time_oracle(p1, n1)
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return {"value": p.ncells_total * nsteps / secs, "unit": "zone-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"64^3 (64 blocks of 16^3), {nsteps} SSPRK(3,3) steps, FP64, "
                      f"{secs:.2f} s on one host thread",
            "cpu_model": _cpu_model(), "nproc": nproc,
            "hd2d_128": {"value": p1.ncells_total * n1 / secs1, "unit": "zone-updates/s",
                         "cores": 1,
                         "sample": f"2D HD TVD-MC+HLLC SSPRK(2,2) 128^2 (64 blocks of 16^2), "
                                   f"{n1} steps, FP64, {secs1:.2f} s on one host thread"}}


def c3_legs(steps: int, warmup: int, launch: str):
    """The batched/AMR configuration beside the headline: the 64-grid KH
    batch (one block pool, per-run dt) and the 3-level 3D AMR mesh of 6224
    leaves of 16^3, FP64 and mixed, and the 2D HD 128^2 grid (launch-bound),
    steps replayed from captured graphs; zone-updates/s over all interior
    cells.  Each is timed with one launch per sweep direction and with one
    launch per stage; `value` and `ms_per_step` follow `launch`."""
    from amun_amd.config import c1_params, c3_batch_params
    from amun_amd import mesh
    from amun_amd.problems import initial_state, kelvin_helmholtz_2d
    from amun_amd.solver import AmrSolver, Batch, Solver

    def timed(make, upload, fused, nsteps):
        with make() as s:
            upload(s)
            if fused:
                s.set_stage_fusion(True)
            s.set_graph(True)
            s.run(warmup)
            s.sync()
            t0 = time.perf_counter()
            s.run(nsteps)
            s.sync()
            return time.perf_counter() - t0, s.info() if hasattr(s, "info") else {}

    def leg(make, upload, cells, nsteps=steps, **extra):
        ws, info = timed(make, upload, False, nsteps)
        wf, _ = timed(make, upload, True, nsteps)
        w = wf if launch == "fused" else ws
        return {"value": cells * nsteps / w, "unit": "zone-updates/s",
                "ms_per_step": w / nsteps * 1e3, "ms_per_step_split": ws / nsteps * 1e3,
                "ms_per_step_fused": wf / nsteps * 1e3, "cells": cells, **extra}, info

    def upload_batch(runs):
        def up(b):
            for r, p in enumerate(runs):
                b.upload(r, kelvin_helmholtz_2d(p, seed=r))
        return up

    out = {}
    for prec in ("fp64", "mixed"):
        common, runs = c3_batch_params(precision=prec)
        cells = sum(p.ncells_total for p in runs)
        print(f"[bench] c3 batch {prec}: {len(runs)} runs, {cells} cells", file=sys.stderr,
              flush=True)
        out[f"batch_{prec}"], _ = leg(lambda: Batch(common, runs), upload_batch(runs), cells,
                                      runs=len(runs))
        p, meta = mesh.c3_amr_params(prec)
        leaves = len(meta)
        cells = leaves * p.ncells ** 3
        print(f"[bench] c3 amr {prec}: {leaves} leaves, {cells} cells", file=sys.stderr,
              flush=True)
        u0 = mesh.block_state(p, meta, mesh.blast_prim(p, r0=0.2))
        out[f"amr_{prec}"], info = leg(lambda: AmrSolver(p, meta), lambda s: s.upload(u0),
                                       cells, leaves=leaves)
        out[f"amr_{prec}"]["virtual_blocks"] = info["virtual_blocks"]
    p1 = c1_params()
    u1 = initial_state(p1)
    print("[bench] c1 2D HD 128^2 fp64", file=sys.stderr, flush=True)
    out["c1_fp64"], _ = leg(lambda: Solver(p1), lambda s: s.upload(u1), p1.ncells_total,
                            nsteps=20 * steps)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision leg reported beside the FP64 value")
    ap.add_argument("--launch", default="fused", choices=["fused", "split"],
                    help="one launch per stage (x, y, z sweeps in one kernel, "
                         "amun_dev_set_stage_fusion) or one launch per sweep direction; "
                         "the other is reported beside the value")
    ap.add_argument("--no-other-launch", action="store_true",
                    help="skip the FP64 leg of the other launch mode")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the batched/AMR legs (64 KH grids, 6224-leaf AMR mesh)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run three 1 GiB device-to-device copies (the calibration "
                         "of the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.copy_probe:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print("copied 3 x 1 GiB", flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (the multi-GPU decomposition is a design note only)",
              file=sys.stderr)
        sys.exit(2)

    ensure_built()
    import numpy as np
    from amun_amd.config import c2_params
    from amun_amd.problems import initial_state
    from amun_amd.solver import Solver

    try:
        import torch
        sync_all = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync_all = lambda: None  # noqa: E731

    p = c2_params(args.size, precision=args.precision)
    cells = p.ncells_total
    print(f"[bench] building {args.size}^3 initial state", file=sys.stderr, flush=True)
    u0 = initial_state(p)

    def measure(params, launch):
        fused = launch == "fused"
        s = Solver(params)
        if fused:
            s.set_stage_fusion(True)
        s.upload(u0)
        tag = params.precision + (" fused" if fused else "")
        print(f"[bench] {tag}: warmup {args.warmup} steps", file=sys.stderr, flush=True)
        s.run(args.warmup)
        s.sync()
        sync_all()
        t0 = time.perf_counter()
        dev_ms, _ = s.timed_run(args.steps)
        s.sync()
        sync_all()
        wall = time.perf_counter() - t0
        print(f"[bench] {tag}: {args.steps} steps: wall {wall:.3f} s, "
              f"device {dev_ms:.1f} ms", file=sys.stderr, flush=True)
        # per-kernel averages over a few more steps (same stream, HIP events)
        _, kms = s.timed_run(3, per_kernel=True)
        s.close()
        return wall, dev_ms, kms

    wall, dev_ms, kms = measure(p, args.launch)
    # the same step in the other launch mode (bitwise the same results)
    other_name = "split" if args.launch == "fused" else "fused"
    other = None
    if not args.no_other_launch:
        wo, do, ko = measure(p, other_name)
        other = {"value": cells * args.steps / wo, "ms_per_step": wo / args.steps * 1e3,
                 "device_ms_per_step": do / args.steps,
                 f"speedup_of_{args.launch}": wo / wall,
                 "kernel_ms_per_step": [round(float(x), 4) for x in ko]}
    mixed = None
    if args.precision == "fp64" and not args.no_mixed:
        pm = c2_params(args.size, precision="mixed")
        wm, dm, km = measure(pm, args.launch)
        mixed = {"value": cells * args.steps / wm, "ms_per_step": wm / args.steps * 1e3,
                 "speedup_vs_fp64": wall / wm,
                 "roofline_frac": bytes_per_zone_update(pm) * cells / (dm / args.steps * 1e-3)
                 / 1e9 / HBM_PEAK_GBS,
                 "kernel_ms_per_step": [round(float(x), 4) for x in km]}
    del u0

    bpz = bytes_per_zone_update(p)
    step_ms = dev_ms / args.steps
    achieved = bpz * cells / (step_ms * 1e-3) / 1e9
    traffic = None
    valu = None
    tf = os.path.join(ROOT, "profiles",
                      f"traffic_c2_{args.size}_{args.precision}_{args.launch}.json")
    if os.path.exists(tf):
        prof = json.load(open(tf))
        traffic = prof.get("hbm_bytes_per_step")
        if args.precision == "fp64" and prof.get("fp64_flop_per_step"):
            ach = prof["fp64_flop_per_step"] / (step_ms * 1e-3) / 1e12
            valu = {"achieved": ach, "peak": FP64_VEC_TFLOPS, "unit": "TFLOP/s",
                    "frac": ach / FP64_VEC_TFLOPS, "dtype": "f64",
                    "flop_per_zone_update": prof["fp64_flop_per_step"] / cells}
    value = cells * args.steps / wall
    # the bound of the step: the larger of its HBM time (algorithmic bytes at
    # the HBM peak) and its VALU time (counted FP64 flops at the FP64 vector
    # peak); both fractions are reported
    hbm = {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS}
    roofline = {"bound": "hbm", **hbm}
    if valu is not None and valu["achieved"] / valu["peak"] > hbm["frac"]:
        roofline = {"bound": "valu", "achieved": valu["achieved"], "peak": valu["peak"],
                    "unit": valu["unit"], "frac": valu["frac"], "hbm": hbm}
    roofline.update({"traffic": traffic,
                     "kernel": ("one SSPRK(3,3) step = 3 k_stage launches + k_end_step"
                                if args.launch == "fused" else
                                "one SSPRK(3,3) step = 9 k_sweep launches + k_end_step"),
                     "bytes_per_zone_update": bpz,
                     "device_ms_per_step": step_ms,
                     "kernel_ms_per_step": [round(float(x), 4) for x in kms]})
    out = {
        "metric": "zone-updates/s (3D MHD WENO5+HLLD 256^3) FP64 vs mixed, % HBM roofline",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-flux/f64-state",
        "data": "synthetic (seeded random-phase Fourier v and B, rho=1, p=0.6)",
        "config": {"workload": f"3D MHD-GLM WENO5-Z+HLLD SSPRK(3,3) {args.size}^3 "
                               f"as {p.nblocks} blocks of 16^3, CFL 0.3, periodic",
                   "precision": args.precision, "launch": args.launch, "cells": cells},
        "roofline": roofline,
    }
    if valu is not None:
        out["valu"] = valu
    if other is not None:
        out[other_name] = other
    if mixed is not None:
        out["mixed"] = mixed
    if not args.no_c3:
        out["c3"] = c3_legs(args.steps, args.warmup, args.launch)
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(p)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
