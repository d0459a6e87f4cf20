"""Condense rocprofv3 output into the short summaries kept under profiles/.

    python profile_summary.py stats <kernel_stats.csv> [--title T] [--command C]
    python profile_summary.py traffic --fetch F.csv --write W.csv \
        --copy-fetch CF.csv --copy-write CW.csv --copy-bytes B [--command C]
    python profile_summary.py mfma --pmc P.csv --stats S.csv [--elements k=n,...]

`stats` prints a table of the per-kernel summary (`--kernel-trace --stats`).
`traffic` turns the FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) of two
PMC runs into HBM bytes per launch, each scaled by the bytes-per-counter-unit
factor that the same counter shows for a plain device copy of known size, and
prints the JSON that bench.py reads (profiles/traffic.json).  `mfma` turns the
MFMA / VALU counters of a PMC run and the durations of a kernel-trace run into
MFMA TFLOP/s against the dense peak, the MFMA-busy fraction of the SIMDs, and
VALU instructions per element.
"""
from __future__ import annotations

import argparse
import csv
import json
import sys
from collections import defaultdict


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _get(row, *names):
    for n in names:
        if n in row and row[n] != "":
            return row[n]
    raise KeyError(names)


def stats(path, title="", command=""):
    rows = _rows(path)
    out = []
    if title:
        out.append(f"# {title}")
    if command:
        out.append(f"command: {command}")
    out.append(f"{'kernel':58s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s} {'pct':>6s}")
    for r in rows:
        name = _get(r, "Name", "KernelName", "Kernel_Name")
        calls = int(float(_get(r, "Calls", "Count")))
        tot = float(_get(r, "TotalDurationNs", "Total_Duration_ns"))
        avg = float(_get(r, "AverageNs", "Average_ns"))
        pct = float(_get(r, "Percentage", "Percent"))
        out.append(f"{name[:58]:58s} {calls:6d} {avg / 1e3:10.2f} {tot / 1e6:10.3f} {pct:6.2f}")
    return "\n".join(out)


def _per_kernel(path, counter):
    acc = defaultdict(list)
    for r in _rows(path):
        if _get(r, "Counter_Name", "Counter") != counter:
            continue
        acc[_get(r, "Kernel_Name", "KernelName", "Name")].append(float(_get(r, "Counter_Value", "Value")))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def _copy_factor(path, counter, copy_bytes):
    per = _per_kernel(path, counter)
    cand = {k: v for k, v in per.items() if "copy" in k.lower() or "elementwise" in k.lower()}
    if not cand:
        raise SystemExit(f"no copy kernel found in {path}: {list(per)}")
    # the copy of `copy_bytes` each way: bytes per counter unit
    name, val = max(cand.items(), key=lambda kv: kv[1])
    return copy_bytes / val, name, val


def traffic(a):
    ff, fname, fval = _copy_factor(a.copy_fetch, "FETCH_SIZE", a.copy_bytes)
    wf, wname, wval = _copy_factor(a.copy_write, "WRITE_SIZE", a.copy_bytes)
    fetch = _per_kernel(a.fetch, "FETCH_SIZE")
    write = _per_kernel(a.write, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        if not k.startswith("mcfft_"):
            continue
        fb = fetch.get(k, 0.0) * ff
        wb = write.get(k, 0.0) * wf
        kernels[k] = {"fetch_bytes": fb, "write_bytes": wb, "hbm_bytes_per_launch": fb + wb,
                      "fetch_counter": fetch.get(k), "write_counter": write.get(k)}
    return {"method": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (separate runs), each scaled by bytes per "
                      "counter unit of a plain device copy of known size measured the same way",
            "copy_bytes_each_way": a.copy_bytes,
            "fetch_factor_bytes_per_unit": ff, "write_factor_bytes_per_unit": wf,
            "copy_kernel": fname, "copy_fetch_counter": fval, "copy_write_counter": wval,
            "command": a.command, "kernels": kernels}


def mfma(a):
    """MFMA work and activity per kernel from one PMC run with
    SQ_INSTS_VALU_MFMA_MOPS_F16, SQ_INSTS_VALU_MFMA_MOPS_BF16,
    SQ_VALU_MFMA_BUSY_CYCLES, SQ_INSTS_VALU and GRBM_GUI_ACTIVE, plus the kernel
    durations of a separate --kernel-trace --stats run (counter collection slows
    the kernels it measures)."""
    rows = _rows(a.pmc)
    acc = defaultdict(lambda: defaultdict(list))
    for r in rows:
        k = _get(r, "Kernel_Name", "KernelName", "Name")
        if not k.startswith("mcfft_"):
            continue
        acc[k][_get(r, "Counter_Name", "Counter")].append(float(_get(r, "Counter_Value", "Value")))
    dur = {}
    for r in _rows(a.stats):
        dur[_get(r, "Name", "KernelName", "Kernel_Name")] = float(_get(r, "AverageNs", "Average_ns")) * 1e-9
    mean = lambda v: sum(v) / len(v) if v else 0.0
    out = []
    if a.title:
        out.append(f"# {a.title}")
    if a.command:
        out.append(f"command: {a.command}")
    out.append(f"peak dense FP16/BF16 MFMA: {a.peak_tflops:.0f} TFLOP/s; SIMDs: {a.simds}")
    out.append(f"{'kernel':30s} {'launches':>8s} {'MFMA GFLOP/launch':>17s} {'avg us':>8s} {'MFMA TFLOP/s':>12s} "
               f"{'of peak':>8s} {'MFMA busy':>9s} {'VALU inst/elem':>14s}")
    for k in sorted(acc):
        c = acc[k]
        mops = mean(c.get("SQ_INSTS_VALU_MFMA_MOPS_F16", [])) + mean(c.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", []))
        flops = mops * 512.0
        busy = mean(c.get("SQ_VALU_MFMA_BUSY_CYCLES", []))
        gui = mean(c.get("GRBM_GUI_ACTIVE", []))
        valu = mean(c.get("SQ_INSTS_VALU", []))
        n = max(len(v) for v in c.values())
        t = dur.get(k)
        tf = flops / t / 1e12 if t else float("nan")
        # the CSV sums GRBM_GUI_ACTIVE over the XCDs (one graphics block each);
        # the busy fraction divides by one XCD's active cycles, as rocprofv3's
        # MfmaUtil (reduce(GRBM_GUI_ACTIVE, max)) does
        util = 100.0 * busy / (gui / a.xcds * a.simds) if gui else float("nan")
        elems = a.elements.get(k) if a.elements else None
        vpe = valu * 64.0 / elems if elems else float("nan")
        out.append(f"{k:30s} {n:8d} {flops / 1e9:17.2f} {t * 1e6 if t else float('nan'):8.2f} {tf:12.1f} "
                   f"{100.0 * tf / a.peak_tflops:7.2f}% {util:8.2f}% {vpe:14.1f}")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--title", default="")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-bytes", type=float, default=float(256 << 20))
    t.add_argument("--command", default="")
    m = sub.add_parser("mfma")
    m.add_argument("--pmc", required=True, help="counter_collection.csv of the MFMA/VALU counter run")
    m.add_argument("--stats", required=True, help="kernel_stats.csv of a --kernel-trace --stats run")
    m.add_argument("--elements", default="", help="kernel=elements,... (complex elements per launch)")
    m.add_argument("--peak-tflops", type=float, default=2500.0)
    m.add_argument("--simds", type=int, default=1024)
    m.add_argument("--xcds", type=int, default=8)
    m.add_argument("--title", default="")
    m.add_argument("--command", default="")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats(a.csv, a.title, a.command))
    elif a.cmd == "mfma":
        a.elements = {kv.split("=")[0]: float(kv.split("=")[1]) for kv in a.elements.split(",") if "=" in kv}
        print(mfma(a))
    else:
        json.dump(traffic(a), sys.stdout, indent=1)
        print()


if __name__ == "__main__":
    main()
