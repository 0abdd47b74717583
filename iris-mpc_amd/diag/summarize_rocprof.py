"""Summaries of rocprofv3 CSV output for the share GEMM roofline.

    python3 summarize_rocprof.py stats <kernel_stats.csv> <command...>
        kernel-trace --stats table (top kernels, average durations).
    python3 summarize_rocprof.py traffic <fetch.csv> <write.csv> <copy_fetch.csv>
        <copy_write.csv> <kernel-substring> <out.json> <command...>
        HBM bytes per dispatch of one kernel: FETCH_SIZE and WRITE_SIZE (KiB per
        dispatch) of that kernel, each scaled by the factor that makes the same
        counter read exactly 2 GiB for the plain 2 GiB copy of roofline_probe.
"""
import csv
import json
import sys
from collections import defaultdict

COPY_BYTES = 2 << 30


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def stats(path, command):
    rows = _rows(path)
    print(f"# rocprofv3 --kernel-trace --stats summary\n# command: {command}\n")
    print(f"{'calls':>6} {'avg ms':>10} {'total ms':>11} {'share %':>8}  kernel")
    for r in sorted(rows, key=lambda r: -float(r.get("TotalDurationNs", 0)))[:12]:
        avg = float(r["AverageNs"]) / 1e6
        tot = float(r["TotalDurationNs"]) / 1e6
        print(f"{int(r['Calls']):>6} {avg:>10.3f} {tot:>11.3f} {float(r['Percentage']):>8.2f}  {r['Name'][:110]}")


def _per_dispatch(path, counter, match):
    out = defaultdict(float)
    names = {}
    for r in _rows(path):
        if r["Counter_Name"] != counter or match not in r["Kernel_Name"]:
            continue
        out[int(r["Dispatch_Id"])] += float(r["Counter_Value"])
        names[int(r["Dispatch_Id"])] = r["Kernel_Name"]
    return [out[k] for k in sorted(out)], names


def traffic(fetch, write, cfetch, cwrite, match, out_json, command):
    f, names = _per_dispatch(fetch, "FETCH_SIZE", match)
    w, _ = _per_dispatch(write, "WRITE_SIZE", match)
    cf, _ = _per_dispatch(cfetch, "FETCH_SIZE", "copy16")
    cw, _ = _per_dispatch(cwrite, "WRITE_SIZE", "copy16")
    # the probe's second copy16 dispatch is the timed one; both move 2 GiB each way
    fs = COPY_BYTES / (cf[-1] * 1024.0)
    ws = COPY_BYTES / (cw[-1] * 1024.0)
    per = [fs * a * 1024.0 + ws * b * 1024.0 for a, b in zip(f, w)]
    res = {
        "kernel": sorted(set(names.values())),
        "copy_calibration": {"bytes_each_way": COPY_BYTES, "FETCH_SIZE_KiB": cf[-1], "WRITE_SIZE_KiB": cw[-1],
                             "fetch_scale": fs, "write_scale": ws},
        "dispatches": [{"FETCH_SIZE_KiB": a, "WRITE_SIZE_KiB": b, "hbm_bytes": p} for a, b, p in zip(f, w, per)],
        "traffic_bytes_per_launch": sum(per) / len(per),
        "command": command,
    }
    with open(out_json, "w") as fo:
        json.dump(res, fo, indent=1)
    print(f"# HBM traffic per dispatch ({match}); counters scaled by the 2 GiB copy calibration")
    print(f"# command: {command}")
    print(f"# copy16 2 GiB: FETCH_SIZE {cf[-1]:.0f} KiB (scale {fs:.4f}), WRITE_SIZE {cw[-1]:.0f} KiB (scale {ws:.4f})")
    for d in res["dispatches"]:
        print(f"FETCH {d['FETCH_SIZE_KiB']:.0f} KiB  WRITE {d['WRITE_SIZE_KiB']:.0f} KiB  -> {d['hbm_bytes'] / 1e9:.3f} GB")
    print(f"average per launch: {res['traffic_bytes_per_launch'] / 1e9:.3f} GB")


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(sys.argv[2], " ".join(sys.argv[3:]))
    elif sys.argv[1] == "traffic":
        traffic(*sys.argv[2:8], " ".join(sys.argv[8:]))
    else:
        sys.exit("usage: summarize_rocprof.py stats|traffic ...")
