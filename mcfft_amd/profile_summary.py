"""Condense rocprofv3 output into the short summaries kept under profiles/.

    python -m mcfft_amd.profile_summary stats <kernel_stats.csv> [--title T] [--command C]
    python -m mcfft_amd.profile_summary traffic --fetch F.csv --write W.csv \
        --copy-fetch CF.csv --copy-write CW.csv --copy-bytes B [--command C]

`stats` prints a table of the per-kernel summary (`--kernel-trace --stats`).
`traffic` turns the FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) of two
PMC runs into HBM bytes per launch, each scaled by the bytes-per-counter-unit
factor that the same counter shows for a plain device copy of known size, and
prints the JSON that bench.py reads (profiles/traffic.json).
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
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats(a.csv, a.title, a.command))
    else:
        json.dump(traffic(a), sys.stdout, indent=1)
        print()


if __name__ == "__main__":
    main()
