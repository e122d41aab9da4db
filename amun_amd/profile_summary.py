"""Summaries of rocprofv3 runs of bench.py (the text kept under profiles/).

    python -m amun_amd.profile_summary --precision fp64 --size 256 \\
        --trace DIR --fetch DIR --write DIR --valu DIR \\
        --copy-fetch DIR --copy-write DIR --out profiles/NAME.txt --json profiles/NAME.json

Each DIR is the -d output directory of one rocprofv3 run (its CSV files are
found below it):
  --trace       rocprofv3 --kernel-trace --stats       (average durations)
  --fetch       rocprofv3 --pmc FETCH_SIZE             (KiB read from HBM)
  --write       rocprofv3 --pmc WRITE_SIZE             (KiB written to HBM)
  --valu        rocprofv3 --pmc <SQ VALU counters>     (flops and VALU activity)
  --copy-*      the same two counters over bench.py --copy-probe (three 1 GiB
                device copies), which calibrate FETCH_SIZE / WRITE_SIZE to bytes.
The text gives per-kernel durations, HBM bytes per launch, FP64/FP32 flops
per launch, VALU activity and register use, and the step totals against the
MI355X peaks; the JSON holds the step totals bench.py reads for `traffic`.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
from collections import defaultdict

HBM_PEAK_GBS = 8000.0
FP64_TFLOPS = 78.6
FP32_TFLOPS = 157.3
GIB = float(1 << 30)


def _rows(d, suffix):
    out = []
    for f in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
        with open(f, newline="") as fh:
            out += list(csv.DictReader(fh))
    return out


def kernel_stats(d):
    """{kernel name: (calls, average us, min us, max us)}"""
    st = {}
    for r in _rows(d, "kernel_stats.csv"):
        st[r["Name"]] = (int(r["Calls"]), float(r["AverageNs"]) / 1e3,
                         float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3)
    return st


def counters(d):
    """{kernel name: {counter: [value per dispatch]}} and {kernel name:
    (VGPR, AGPR, SGPR, scratch bytes, LDS bytes)}"""
    per = defaultdict(lambda: defaultdict(lambda: defaultdict(float)))
    res = {}
    for r in _rows(d, "counter_collection.csv"):
        k = r["Kernel_Name"]
        disp = r.get("Dispatch_Id", r.get("Correlation_Id", ""))
        per[k][r["Counter_Name"]][disp] += float(r["Counter_Value"])
        res[k] = (r.get("VGPR_Count", "?"), r.get("Accum_VGPR_Count", "?"),
                  r.get("SGPR_Count", "?"), r.get("Scratch_Size", "?"),
                  r.get("LDS_Block_Size", "?"))
    return ({k: {c: list(v.values()) for c, v in cs.items()} for k, cs in per.items()}, res)


def _mean(v):
    return sum(v) / len(v) if v else 0.0


def calibration(d, counter):
    """GiB copied per GiB the counter reports for the largest copy dispatch
    (FETCH_SIZE and WRITE_SIZE count KiB)."""
    cs, _ = counters(d)
    best = 0.0
    for k, c in cs.items():
        if counter in c and ("copy" in k.lower() or "Copy" in k):
            best = max(best, max(c[counter]))
    if best <= 0.0:
        raise SystemExit(f"no copy dispatch with {counter} under {d}")
    return GIB / (best * 1024.0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", required=True)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--valu", required=True)
    ap.add_argument("--copy-fetch", required=True)
    ap.add_argument("--copy-write", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--json", required=True)
    ap.add_argument("--commands", default="", help="text of the rocprofv3 commands (header)")
    a = ap.parse_args()

    cells = a.size ** 3
    st = kernel_stats(a.trace)
    fcal = calibration(a.copy_fetch, "FETCH_SIZE")
    wcal = calibration(a.copy_write, "WRITE_SIZE")
    fc, res = counters(a.fetch)
    wc, _ = counters(a.write)
    vc, vres = counters(a.valu)
    res.update(vres)
    sweeps = sorted(k for k in st if "k_sweep" in k or "k_stage" in k)
    lines = [f"# rocprofv3 summary: 3D MHD-GLM WENO5-Z+HLLD SSPRK(3,3), {a.size}^3 "
             f"({cells // 4096} blocks of 16^3), precision {a.precision}"]
    lines += ["# " + c for c in a.commands.split("\n") if c.strip()]
    lines.append(f"# HBM calibration: FETCH_SIZE x {fcal:.3f}, WRITE_SIZE x {wcal:.3f} (bytes "
                 "copied per byte counted, 1 GiB device copies of bench.py --copy-probe)")
    lines.append("")
    lines.append(f"{'kernel':66s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s}")
    for k, (n, av, mn, mx) in sorted(st.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        lines.append(f"{k[:66]:66s} {n:6d} {av:9.1f} {mn:9.1f} {mx:9.1f}")
    lines.append("")
    lines.append("HBM traffic per launch (calibrated), GB; the register, scratch and LDS "
                 "fields of rocprofv3's dispatch records (allocation granules and static LDS; "
                 "the compiler's resource report gives the registers):")
    rd = wr = 0.0
    for k in sweeps:
        r = _mean(fc.get(k, {}).get("FETCH_SIZE", [])) * 1024.0 * fcal / 1e9
        w = _mean(wc.get(k, {}).get("WRITE_SIZE", [])) * 1024.0 * wcal / 1e9
        rd += r
        wr += w
        v, ag, sg, sc, lds = res.get(k, ("?",) * 5)
        lines.append(f"  {k[:62]:62s} read {r:6.3f}  write {w:6.3f}   "
                     f"VGPR {v} AGPR {ag} SGPR {sg} scratch {sc} LDS {lds}")
    lines.append("")
    lines.append("VALU per launch (SQ counters summed over the GPU; the FLOPS counters "
                 "count per wave, x 64 lanes):")
    f64 = f32 = 0.0
    for k in sweeps:
        c = vc.get(k, {})
        a64 = _mean(c.get("SQ_INSTS_VALU_FLOPS_FP64", [])) * 64 / 1e9
        a32 = _mean(c.get("SQ_INSTS_VALU_FLOPS_FP32", [])) * 64 / 1e9
        act = _mean(c.get("SQ_ACTIVE_INST_VALU", []))
        cyc = _mean(c.get("SQ_WAVE_CYCLES", []))
        wait = _mean(c.get("SQ_WAIT_INST_ANY", []))
        f64 += a64
        f32 += a32
        lines.append(f"  {k[:62]:62s} fp64 {a64:7.2f} GFLOP  fp32 {a32:7.2f} GFLOP  "
                     f"VALU-active/wave-cycles {act / cyc if cyc else 0:5.2f}  "
                     f"wait-any/wave-cycles {wait / cyc if cyc else 0:5.2f}")
    # a step is three stages of one launch per direction (k_sweep) or one
    # launch each (k_stage)
    nst = 3
    step_us = sum(st[k][1] for k in sweeps) * nst
    step_bytes = (rd + wr) * 1e9 * nst
    bpz = 8 * 9 * 8
    lines.append("")
    lines.append(f"step ({3 * len(sweeps)} launches): {step_us:.1f} us; HBM traffic "
                 f"{step_bytes / 1e9:.2f} GB = {step_bytes / cells:.0f} B/zone-update "
                 f"(algorithmic {bpz} B) -> {step_bytes / step_us / 1e3:.0f} GB/s measured; "
                 f"algorithmic {bpz * cells / step_us / 1e3:.0f} GB/s = "
                 f"{bpz * cells / step_us / 1e3 / HBM_PEAK_GBS:.1%} of 8 TB/s")
    t64 = f64 * nst / step_us * 1e3          # GFLOP per us -> TFLOP/s
    t32 = f32 * nst / step_us * 1e3
    lines.append(f"FP64 {f64 * 1e9 * nst / cells:.0f} flop/zone-update -> {t64:.2f} TFLOP/s = "
                 f"{t64 / FP64_TFLOPS:.1%} of {FP64_TFLOPS}; FP32 "
                 f"{f32 * 1e9 * nst / cells:.0f} flop/zone-update -> {t32:.2f} TFLOP/s = "
                 f"{t32 / FP32_TFLOPS:.1%} of {FP32_TFLOPS}")
    with open(a.out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    with open(a.json, "w") as fh:
        json.dump({"hbm_bytes_per_step": step_bytes,
                   "bytes_per_zone_update_measured": step_bytes / cells,
                   "step_us_kernel_trace": step_us,
                   "fp64_flop_per_step": f64 * 1e9 * nst,
                   "fp32_flop_per_step": f32 * 1e9 * nst,
                   "fetch_calibration": fcal, "write_calibration": wcal}, fh, indent=1)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
