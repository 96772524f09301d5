"""Summarise rocprofv3 CSV output into the short text profiles kept in profiles/.

Inputs are the directories rocprofv3 wrote with ``--output-format csv``:
  --stats      a ``--kernel-trace --stats`` run (``*_kernel_stats.csv``)
  --fetch      a ``--pmc FETCH_SIZE`` run of the same command
  --write      a ``--pmc WRITE_SIZE`` run of the same command
  --cal-fetch / --cal-write   the same two counters over ``bench.py
               --calibrate-copy MIB`` (a plain elementwise copy of known size)
  --pmc        (repeatable) any other ``--pmc`` run of the same command; every
               counter it holds for the kernel is listed per launch, per wave
               (when SQ_WAVES is among them) and per work unit (--units)

The counters are scaled by what they read for the copy kernel: factor =
bytes the copy moved / bytes the counter reported.  The summary ends with
machine-readable lines (``key = value``) that bench.py reads to fill
``roofline.traffic`` when the kernel's average duration still agrees with the
live measurement.

    python -m gnomon_amd.profile_summary --kernel k_score_mfma --workload ... \\
        --stats D --fetch D --write D --cal-fetch D --cal-write D \\
        --cal-mib 2048 --alg-bytes N --command "..." -o profiles/NAME.md
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import statistics


def _one(d, suffix):
    hits = glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return hits[0]


def kernel_stats(d, kernel):
    rows = list(csv.DictReader(open(_one(d, "kernel_stats.csv"))))
    out = []
    for r in rows:
        out.append((r["Name"], int(r["Calls"]), float(r["AverageNs"]), float(r["Percentage"])))
    mine = [o for o in out if kernel in o[0]]
    return out, mine


def counter_per_dispatch(d, kernel):
    vals = [float(r["Counter_Value"]) for r in csv.DictReader(open(_one(d, "counter_collection.csv")))
            if kernel in r["Kernel_Name"]]
    if not vals:
        raise ValueError(f"kernel {kernel!r} not found in {d}")
    return statistics.mean(vals), len(vals)


def counters_per_dispatch(d, kernel):
    """{counter: mean per dispatch} for every counter of the kernel in a --pmc run."""
    per = {}
    for r in csv.DictReader(open(_one(d, "counter_collection.csv"))):
        if kernel in r["Kernel_Name"]:
            key = (r["Counter_Name"], r["Dispatch_Id"])
            per[key] = per.get(key, 0.0) + float(r["Counter_Value"])   # summed over dimensions
    out = {}
    for (c, _), v in per.items():
        out.setdefault(c, []).append(v)
    return {c: statistics.mean(v) for c, v in out.items()}


def calibration(d, cal_bytes):
    # the largest dispatches of the copy are the copy itself (torch splits a
    # tensor over 2^31 elements into equal launches)
    rows = [r for r in csv.DictReader(open(_one(d, "counter_collection.csv")))
            if "bitwise" in r["Kernel_Name"]]
    if not rows:
        raise ValueError(f"no copy kernel in {d}")
    grid = [int(r["Grid_Size"]) for r in rows]
    per_launch_bytes = cal_bytes * grid[0] / sum(grid[:len(grid) // 3 or 1])  # three copies were run
    kib = statistics.mean(float(r["Counter_Value"]) for r in rows)
    return per_launch_bytes / (kib * 1024.0), kib, per_launch_bytes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kernel", required=True)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--precision", default="")
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--cal-fetch")
    ap.add_argument("--cal-write")
    ap.add_argument("--cal-mib", type=int, default=2048)
    ap.add_argument("--alg-bytes", type=float, required=True)
    ap.add_argument("--pmc", action="append", default=[])
    ap.add_argument("--units", type=float, default=0.0, help="work units per launch, for per-unit counts")
    ap.add_argument("--unit-name", default="unit")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--mfma-ops", type=float, default=0.0,
                    help="executed matrix operations per launch (prints achieved TFLOP/s against --mfma-peak)")
    ap.add_argument("--mfma-peak", type=float, default=2500.0, help="dense MFMA peak, TFLOP/s (or TOP/s)")
    ap.add_argument("--simds", type=int, default=1024, help="SIMDs of the device (256 CUs x 4)")
    ap.add_argument("--xcds", type=int, default=8, help="GRBM counters are summed over this many XCDs")
    ap.add_argument("-o", "--output", required=True)
    a = ap.parse_args()

    allk, mine = kernel_stats(a.stats, a.kernel)
    if not mine:
        raise SystemExit(f"kernel {a.kernel!r} not in the stats")
    name, calls, avg_ns, pct = max(mine, key=lambda m: m[1] * m[2])
    avg_ms = avg_ns / 1e6
    lines = [f"# {a.kernel} -- {a.workload}" + (f" ({a.precision})" if a.precision else ""), ""]
    lines += ["Commands (run from the repository root on one MI355X; rocprofv3 writes CSV):", "", "```"]
    lines += a.command + ["```", ""]
    lines += ["## Kernel trace (rocprofv3 --kernel-trace --stats)", "",
              "| kernel | calls | avg ms | % of GPU time |", "|---|---|---|---|"]
    for n, c, ns, p in sorted(allk, key=lambda x: -x[3])[:8]:
        short = n.split("(")[0].replace("void ", "")
        lines.append(f"| `{short}` | {c} | {ns / 1e6:.4f} | {p:.2f} |")
    lines += ["", f"Dominant kernel `{name.split('(')[0].replace('void ', '')}`: average {avg_ms:.4f} ms over "
              f"{calls} launches.", ""]
    traffic = None
    if a.fetch and a.write and a.cal_fetch and a.cal_write:
        cal_bytes = a.cal_mib * (1 << 20)
        ff, fk, fb = calibration(a.cal_fetch, cal_bytes)
        wf, wk, wb = calibration(a.cal_write, cal_bytes)
        fetch_kib, nf = counter_per_dispatch(a.fetch, a.kernel)
        write_kib, nw = counter_per_dispatch(a.write, a.kernel)
        fetch_b, write_b = fetch_kib * 1024 * ff, write_kib * 1024 * wf
        traffic = fetch_b + write_b
        lines += ["## HBM traffic (rocprofv3 --pmc, one counter per run)", "",
                  "| counter | copy kernel: bytes per launch | copy kernel: counter KiB | scale factor |",
                  "|---|---|---|---|",
                  f"| FETCH_SIZE | {fb:.0f} | {fk:.1f} | {ff:.4f} |",
                  f"| WRITE_SIZE | {wb:.0f} | {wk:.1f} | {wf:.4f} |", "",
                  "| counter | kernel KiB per launch (mean) | launches | scaled bytes |", "|---|---|---|---|",
                  f"| FETCH_SIZE | {fetch_kib:.1f} | {nf} | {fetch_b:.4e} |",
                  f"| WRITE_SIZE | {write_kib:.1f} | {nw} | {write_b:.4e} |", "",
                  f"Traffic per launch {traffic:.4e} B against {a.alg_bytes:.4e} algorithmic bytes "
                  f"(ratio {traffic / a.alg_bytes:.3f}).", ""]
    if a.pmc:
        cnt = {}
        for d in a.pmc:
            cnt.update(counters_per_dispatch(d, a.kernel))
        waves = cnt.get("SQ_WAVES")
        lines += ["## Other counters (rocprofv3 --pmc, mean per launch)", "",
                  f"| counter | per launch | per wave | per {a.unit_name} |", "|---|---|---|---|"]
        for c in sorted(cnt):
            pw = f"{cnt[c] / waves:.1f}" if waves else ""
            pu = f"{cnt[c] / a.units:.2f}" if a.units else ""
            lines.append(f"| {c} | {cnt[c]:.4e} | {pw} | {pu} |")
        if a.units:
            lines += ["", f"{a.unit_name}s per launch: {a.units:.4e}."]
        lines.append("")
        if "GRBM_GUI_ACTIVE" in cnt:
            cyc = cnt["GRBM_GUI_ACTIVE"] / a.xcds
            lines += [f"Clock: {cyc:.4e} GPU cycles per launch (GRBM_GUI_ACTIVE / {a.xcds} XCDs), "
                      f"{cyc / (avg_ms * 1e-3) / 1e9:.2f} GHz over the average launch.", ""]
            if "SQ_VALU_MFMA_BUSY_CYCLES" in cnt:
                busy = cnt["SQ_VALU_MFMA_BUSY_CYCLES"] / (cyc * a.simds)
                lines += [f"Matrix pipe busy: SQ_VALU_MFMA_BUSY_CYCLES / (cycles x {a.simds} SIMDs) = {busy:.3f}.", ""]
    if a.mfma_ops:
        tf = a.mfma_ops / (avg_ms * 1e-3) / 1e12
        lines += ["## MFMA roofline", "",
                  f"Executed matrix operations per launch {a.mfma_ops:.4e}; at {avg_ms:.4f} ms that is "
                  f"{tf:.1f} TFLOP/s, {tf / a.mfma_peak:.3f} of the {a.mfma_peak:.0f} TFLOP/s dense peak.", ""]
    lines += ["## Values read by bench.py", "", "```",
              f"kernel = {a.kernel}", f"workload = {a.workload}", f"precision = {a.precision}",
              f"kernel_avg_ms = {avg_ms:.6f}", f"algorithmic_bytes_per_launch = {a.alg_bytes:.0f}"]
    if traffic is not None:
        lines.append(f"traffic_bytes_per_launch = {traffic:.0f}")
    lines += ["```", ""]
    with open(a.output, "w") as f:
        f.write("\n".join(lines))
    print(a.output)


def read_values(path):
    """key = value lines of a summary written by main()."""
    vals = {}
    with open(path) as f:
        for ln in f:
            if " =" in ln and not ln.startswith(("|", "#")):
                k, _, v = ln.strip().partition("=")
                vals[k.strip()] = v.strip()
    return vals


if __name__ == "__main__":
    main()
