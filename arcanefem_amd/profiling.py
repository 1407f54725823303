"""Summaries of rocprofv3 output for the assembly / SpMV / PCG kernels.

Two summaries are produced from the CSV files rocprofv3 writes:

* ``stats``: the per-kernel table of ``--kernel-trace --stats``
  (``*_kernel_stats.csv``) as a short Markdown table;
* ``pmc``: HBM traffic per dispatch from two ``--pmc`` passes, one with
  FETCH_SIZE and one with WRITE_SIZE (KiB per dispatch).  The raw counters
  are scaled by what the same counter reads for a device-to-device copy of
  known size in the same run (``bench.py --calibrate`` issues one of
  256 MiB): factor = copy bytes / counter reading.  The result is written as
  JSON (read by bench.py for ``roofline.traffic``) and as Markdown.

* ``counters``: the per-dispatch median of every counter found in one or
  more ``--pmc`` passes (e.g. SQ_INSTS_VALU, SQ_INSTS_LDS,
  SQ_LDS_BANK_CONFLICT, SQ_WAVES), one row per kernel, as Markdown.

python -m arcanefem_amd.profiling stats --csv K.csv --command "..." --out profiles/X.md
python -m arcanefem_amd.profiling pmc --fetch F.csv --write W.csv --workload tet128 \
    --command "..." --out profiles/X   (writes X.json and X.md)
python -m arcanefem_amd.profiling counters --csv A.csv [B.csv ...] --match REGEX \
    --command "..." --out profiles/X.md
"""

import argparse
import csv
import json
import re
import statistics

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(kernel):
    """'void afem::k_spmv_sell<double, double, 0>(int, ...)' -> 'k_spmv_sell<double, double, 0>'."""
    name = kernel.strip()
    name = re.sub(r"^void\s+", "", name)
    name = name.replace("(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(name):  # cut the argument list at the first top-level '('
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            name = name[:i]
            break
    return name.replace("afem::", "")


def _read(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def counter_per_kernel(path, counter):
    """{short kernel name: [counter value per dispatch]} and the calibration copy reading."""
    per = {}
    copy = None
    for r in _read(path):
        if r["Counter_Name"] != counter:
            continue
        v = float(r["Counter_Value"])
        if r["Kernel_Name"].startswith(COPY_KERNEL):
            grid = int(r["Grid_Size"])
            if copy is None or grid > copy[0]:
                copy = (grid, v)
            continue
        per.setdefault(short_name(r["Kernel_Name"]), []).append(v)
    return per, (copy[1] if copy else None)


def pmc_summary(fetch_csv, write_csv, copy_bytes):
    fetch, fcopy = counter_per_kernel(fetch_csv, "FETCH_SIZE")
    write, wcopy = counter_per_kernel(write_csv, "WRITE_SIZE")
    if not fcopy or not wcopy:
        raise SystemExit("no calibration copy found (run the profiled command with --calibrate)")
    copy_kib = copy_bytes / 1024.0
    ff, wf = copy_kib / fcopy, copy_kib / wcopy
    kernels = {}
    for name in sorted(set(fetch) & set(write)):
        f = statistics.median(fetch[name])
        w = statistics.median(write[name])
        kernels[name] = {
            "dispatches": min(len(fetch[name]), len(write[name])),
            "fetch_kib_raw": f, "write_kib_raw": w,
            "fetch_bytes": f * ff * 1024.0, "write_bytes": w * wf * 1024.0,
            "traffic_bytes": (f * ff + w * wf) * 1024.0,
        }
    return {
        "copy_bytes": copy_bytes,
        "copy_fetch_kib_raw": fcopy, "copy_write_kib_raw": wcopy,
        "fetch_factor": ff, "write_factor": wf,
        "kernels": kernels,
    }


def counter_table(paths, match=None):
    """{kernel: {counter: (median per dispatch, dispatches)}} over the CSV
    files of several --pmc passes; kernels filtered by the regex `match`."""
    acc = {}
    for path in paths:
        for r in _read(path):
            name = short_name(r["Kernel_Name"])
            if r["Kernel_Name"].startswith(COPY_KERNEL) or (match and not re.search(match, name)):
                continue
            acc.setdefault(name, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    return {k: {c: (statistics.median(v), len(v)) for c, v in cs.items()} for k, cs in acc.items()}


def counters_markdown(table, command):
    counters = sorted({c for cs in table.values() for c in cs})
    lines = ["# rocprofv3 counters per dispatch (median)", "", "Commands:", "", "```", command, "```", "",
             "| kernel | dispatches | " + " | ".join(counters) + " |",
             "|---|---:|" + "---:|" * len(counters)]
    for k in sorted(table):
        n = max(v[1] for v in table[k].values())
        cells = [f"{table[k][c][0]:.4g}" if c in table[k] else "" for c in counters]
        lines.append(f"| `{k}` | {n} | " + " | ".join(cells) + " |")
    return "\n".join(lines) + "\n"


def stats_markdown(path, command, top=20):
    rows = _read(path)
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = [
        "# rocprofv3 kernel statistics",
        "",
        "Command:",
        "",
        "```",
        command,
        "```",
        "",
        "| kernel | calls | avg us | min us | max us | total % |",
        "|---|---:|---:|---:|---:|---:|",
    ]
    for r in rows[:top]:
        lines.append(f"| `{short_name(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                     f"{float(r['Percentage']):.1f} |")
    return "\n".join(lines) + "\n"


def grid_markdown(trace_csv, match, top=40):
    """Per (kernel, grid size) durations from a rocprofv3 kernel trace for the
    kernels whose short name matches `match`: one kernel is launched at
    several sizes in one run, and only the row at the benchmark's grid is
    comparable with its event-timed average."""
    import re
    groups = {}
    for r in _read(trace_csv):
        name = short_name(r["Kernel_Name"])
        if not re.search(match, name):
            continue
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        groups.setdefault((name, grid), []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    rows = sorted(groups.items(), key=lambda kv: -sum(kv[1]))[:top]
    lines = ["", f"Per launch size (kernels matching `{match}`):", "",
             "| kernel | grid (work-items) | workgroups | calls | avg us | min us | max us |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    for (name, grid), d in rows:
        lines.append(f"| `{name}` | {grid} | {grid // 256 if grid % 256 == 0 else '-'} | {len(d)} | "
                     f"{sum(d) / len(d):.1f} | {min(d):.1f} | {max(d):.1f} |")
    return "\n".join(lines) + "\n"


def pmc_markdown(summary, command, workload, top=60):
    lines = [
        f"# HBM traffic per dispatch ({workload})",
        "",
        "Two rocprofv3 passes (FETCH_SIZE, then WRITE_SIZE); commands:",
        "",
        "```",
        command,
        "```",
        "",
        f"Calibration: a {summary['copy_bytes']} byte device-to-device copy reads FETCH_SIZE = "
        f"{summary['copy_fetch_kib_raw']:.1f} KiB and WRITE_SIZE = {summary['copy_write_kib_raw']:.1f} KiB, "
        f"so FETCH is scaled by {summary['fetch_factor']:.4f} and WRITE by {summary['write_factor']:.4f}.",
        "",
        "| kernel | dispatches | FETCH raw KiB | WRITE raw KiB | traffic MB (scaled) |",
        "|---|---:|---:|---:|---:|",
    ]
    ks = sorted(summary["kernels"].items(), key=lambda kv: -kv[1]["traffic_bytes"])
    for name, k in ks[:top]:
        lines.append(f"| `{name}` | {k['dispatches']} | {k['fetch_kib_raw']:.1f} | {k['write_kib_raw']:.1f} | "
                     f"{k['traffic_bytes'] / 1e6:.1f} |")
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("--csv", required=True)
    s.add_argument("--command", required=True)
    s.add_argument("--out", required=True)
    s.add_argument("--top", type=int, default=60)
    s.add_argument("--trace", default=None, help="kernel trace CSV for a per-launch-size table")
    s.add_argument("--grid-match", default=None, help="regex on the short kernel names of that table")
    p = sub.add_parser("pmc")
    p.add_argument("--fetch", required=True)
    p.add_argument("--write", required=True)
    p.add_argument("--workload", required=True)
    p.add_argument("--command", required=True)
    p.add_argument("--copy-bytes", type=int, default=256 << 20)
    p.add_argument("--out", required=True, help="output path without extension")
    c = sub.add_parser("counters")
    c.add_argument("--csv", nargs="+", required=True)
    c.add_argument("--match", default=None, help="regex on the short kernel name")
    c.add_argument("--command", required=True)
    c.add_argument("--out", required=True)
    a = ap.parse_args()
    if a.cmd == "counters":
        with open(a.out, "w") as f:
            f.write(counters_markdown(counter_table(a.csv, a.match), a.command))
    elif a.cmd == "stats":
        with open(a.out, "w") as f:
            f.write(stats_markdown(a.csv, a.command, a.top))
            if a.trace and a.grid_match:
                f.write(grid_markdown(a.trace, a.grid_match))
    else:
        summ = pmc_summary(a.fetch, a.write, a.copy_bytes)
        summ["workload"] = a.workload
        summ["command"] = a.command
        with open(a.out + ".json", "w") as f:
            json.dump(summ, f, indent=1, sort_keys=True)
        with open(a.out + ".md", "w") as f:
            f.write(pmc_markdown(summ, a.command, a.workload))


if __name__ == "__main__":
    main()
