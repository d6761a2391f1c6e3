"""HBM traffic per kernel launch from two rocprofv3 counter runs (FETCH_SIZE, then WRITE_SIZE).

    python -m fwi_amd.pmc FETCH_DIR WRITE_DIR --workload NAME --command "..." \
        --bench-log FETCH_RUN_OUTPUT --bench-log WRITE_RUN_OUTPUT > traffic.json

Each directory is the -d output of one `rocprofv3 --pmc <counter> --output-format csv` run of the same program,
which first copies 1 GiB device to device (`bench.py --calibrate-copy`): the first `__amd_rocclr_copyBuffer`
dispatch gives the scale from each counter's unit to bytes (FETCH_SIZE counts L2 fills, some of which the MALL
serves).  Per kernel: the median over its last 60 dispatches, scaled, in bytes per launch.  The output carries
a hash of the library's kernel sources (`source_hash`), so a reader can tell whether the counts belong to the
kernels it runs.
"""
from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import os
import statistics
import sys
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
CALIBRATION_KIB = 1048576.0  # the 1 GiB calibration copy, in KiB read and written


def source_hash() -> str:
    """sha256 (first 16 hex digits) of the HIP/C++ sources the library is built from."""
    h = hashlib.sha256()
    for p in sorted(glob.glob(os.path.join(HERE, "csrc", "*"))):
        if p.endswith((".hip", ".h", ".cpp")):
            h.update(os.path.basename(p).encode())
            with open(p, "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def short_name(kernel: str) -> str:
    """'void fwi::k_fwd_main<0, float, true>(fwi::FwdArgs)' -> 'k_fwd_main<0, float, true>'."""
    k = kernel.strip().replace("(anonymous namespace)::", "")
    if k.startswith("void "):
        k = k[5:]
    depth = 0
    for i, ch in enumerate(k):  # cut the argument list: the first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            k = k[:i]
            break
    return k.replace("fwi::", "")


def _load(d: str):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no *counter_collection.csv under {d}")
    per = defaultdict(list)
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                per[r["Kernel_Name"]].append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    for v in per.values():
        v.sort()
    return per


def summarize(fetch_dir: str, write_dir: str) -> dict:
    f, w = _load(fetch_dir), _load(write_dir)
    cal = [k for k in f if "copyBuffer" in k]
    if not cal or not any("copyBuffer" in k for k in w):
        raise ValueError("no calibration copy (run the program with --calibrate-copy)")
    sf = CALIBRATION_KIB / f[cal[0]][0][1]
    sw = CALIBRATION_KIB / [w[k] for k in w if "copyBuffer" in k][0][0][1]
    out = {}
    for k in f:
        if not k.lstrip().startswith(("void fwi::", "fwi::")) or k not in w:
            continue
        fv = statistics.median(v for _, v in f[k][-60:])
        wv = statistics.median(v for _, v in w[k][-60:])
        out[short_name(k)] = {"read_bytes": round(fv * sf * 1024.0), "write_bytes": round(wv * sw * 1024.0),
                              "dispatches": len(f[k])}
    return {"fetch_scale": round(sf, 4), "write_scale": round(sw, 4), "kernels": out}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("fetch_dir")
    ap.add_argument("write_dir")
    ap.add_argument("--workload", default="")
    ap.add_argument("--command", default="")
    ap.add_argument("--bench-log", action="append", default=[],
                    help="output of a profiled bench.py run (once per counter run): its JSON line gives the work-item "
                         "tiling, which every run must share")
    a = ap.parse_args(argv)
    tilings = []
    for f in a.bench_log:
        with open(f) as fh:
            lines = [ln for ln in fh if ln.startswith('{"metric"')]
        if not lines:
            sys.exit(f"{f}: no bench JSON line")
        c = json.loads(lines[-1])["config"]
        tilings.append({k: c[k] for k in ("rows_per_item", "item_order", "adjoint_tiling", "pair_rows", "adj_pair_rows")
                        if k in c})
    if any(t != tilings[0] for t in tilings):
        sys.exit(f"the counter runs used different tilings: {tilings}")
    res = summarize(a.fetch_dir, a.write_dir)
    res = {"source_hash": source_hash(), "workload": a.workload, "command": a.command,
           "tiling": tilings[0] if tilings else None, **res}
    json.dump(res, sys.stdout, indent=1)
    print()


if __name__ == "__main__":
    main()
