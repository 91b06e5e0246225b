"""Summaries of rocprofv3 runs of bench.py for profiles/ (reads the rocpd databases).

    python profile_summary.py trace <run_results.db>              -> kernel table (markdown)
    python profile_summary.py trace-last <run_results.db> <n>     -> the same over the last n
        dispatches of each kernel (after an autotuner's trial launches), with grid and LDS
    python profile_summary.py traffic <fetch.db> <write.db> <out.json>
        FETCH_SIZE / WRITE_SIZE per dispatch of each sweep kernel, scaled by the ratio of the
        true bytes of the 1 GiB calibration copy (bench.py --copy-calibration) to what the
        same counter reads for it; writes {kernel: bytes per launch} for bench.py and prints
        a markdown table.
"""
import json
import sqlite3
import sys
from collections import defaultdict

COPY_BYTES = 128 * 1024 * 1024 * 8  # the calibration copy reads and writes 1 GiB


def short(name):
    return name.replace("void bfvmb::(anonymous namespace)::", "").replace("bfvmb::(anonymous namespace)::", "").split("(")[0]


def trace(db):
    c = sqlite3.connect(db)
    print("| kernel | calls | total (us) | average (us) | % |")
    print("|---|---|---|---|---|")
    for r in c.execute("select name,total_calls,total_duration,average,percentage from top_kernels"):
        print(f"| `{short(r[0])}` | {r[1]} | {r[2]:.1f} | {r[3]:.1f} | {r[4]:.1f} |")


def trace_last(db, n):
    c = sqlite3.connect(db)
    rows = defaultdict(list)
    for name, start, dur, grid, lds in c.execute(
            "select name, start, duration, grid_x, lds_size from kernels order by start"):
        rows[short(name)].append((dur / 1e3, grid, lds))
    print(f"| kernel | calls (last) | average (us) | min (us) | max (us) | grid (threads) | LDS (bytes) |")
    print("|---|---|---|---|---|---|---|")
    for k, v in sorted(rows.items(), key=lambda kv: -sum(d for d, _, _ in kv[1][-n:])):
        t = v[-n:]
        d = [x[0] for x in t]
        shapes = sorted({(g, l) for _, g, l in t})
        g = ", ".join(str(x[0]) for x in shapes)
        l = ", ".join(str(x[1]) for x in shapes)
        print(f"| `{k}` | {len(t)} | {sum(d) / len(d):.1f} | {min(d):.1f} | {max(d):.1f} | {g} | {l} |")


def counter(db, name):
    c = sqlite3.connect(db)
    per = defaultdict(list)
    for kname, cn, v in c.execute("select name, counter_name, counter_value from pmc_events"):
        if cn == name:
            per[kname].append(v * 1024.0)  # KiB -> bytes
    return per


def traffic(fetch_db, write_db, out):
    f, w = counter(fetch_db, "FETCH_SIZE"), counter(write_db, "WRITE_SIZE")
    cal = {}
    for label, d in (("fetch", f), ("write", w)):
        vals = [max(v) for k, v in d.items() if "copyBuffer" in k]  # torch's device copy
        big = max(vals) if vals else None
        cal[label] = COPY_BYTES / big if big else 1.0
    res, rows = {}, []
    # k_sweep_x<T, PC, REGC, TL, VF, AG[, NT]>: AG = the step's agent update in the same launch
    phase = {"k_sweep_x": "sweep_x", "k_sweep_yz<double, 1": "sweep_y", "k_sweep_yz<double, 2": "sweep_z"}

    def template_args(name):
        return name[name.index("<") + 1:name.rindex(">")].split(", ") if "<" in name else []

    for k in f:
        s = short(k)
        if "k_sweep" not in s and "k_agents" not in s:
            continue
        fb = sum(f[k]) / len(f[k]) * cal["fetch"]
        wb = sum(w.get(k, [0.0])) / max(1, len(w.get(k, [0.0]))) * cal["write"]
        rows.append((s, fb, wb))
        for pre, ph in phase.items():
            if s.startswith(pre) and "double" in s:
                args = template_args(s)
                if ph == "sweep_x" and len(args) >= 6 and args[5] == "true":
                    ph = "sweep_x_with_agents"
                res[ph] = fb + wb
    # provenance: the sweep source these counters describe (bench.py reports the traffic only
    # while biofvm-b_amd/csrc/lod.hip still has this hash) and the commit checked out
    import hashlib
    import os
    import subprocess
    root = os.path.dirname(os.path.abspath(__file__))
    sha = hashlib.sha256(open(os.path.join(root, "biofvm-b_amd", "csrc", "lod.hip"), "rb").read()).hexdigest()
    try:
        commit = subprocess.run(["git", "-C", root, "rev-parse", "--short=12", "HEAD"], capture_output=True,
                                text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        commit = None
    json.dump({"bytes_per_launch": res, "calibration": cal, "lod_hip_sha256": sha, "commit": commit},
              open(out, "w"), indent=1)
    print(f"calibration (true bytes / counter bytes of the 1 GiB copy): fetch {cal['fetch']:.3f}, write {cal['write']:.3f}")
    print("| kernel | fetched (GB/launch) | written (GB/launch) | total (GB) |")
    print("|---|---|---|---|")
    for s, fb, wb in sorted(rows):
        print(f"| `{s}` | {fb / 1e9:.3f} | {wb / 1e9:.3f} | {(fb + wb) / 1e9:.3f} |")


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        trace(sys.argv[2])
    elif sys.argv[1] == "trace-last":
        trace_last(sys.argv[2], int(sys.argv[3]))
    else:
        traffic(sys.argv[2], sys.argv[3], sys.argv[4])
