"""Summaries of rocprofv3 result databases (rocpd SQLite) as the short text
and JSON files bench.py reads back: a kernel-trace table, calibrated HBM
traffic per dispatch, and the SQ instruction mix of the neighbour passes.

  python3 opencl-sph_amd/sph/profile_summary.py trace DB OUT.md --title T --command C
  python3 opencl-sph_amd/sph/profile_summary.py traffic WORKLOAD FETCH_DB WRITE_DB OUT_PREFIX --copy-mib M --command C
  python3 opencl-sph_amd/sph/profile_summary.py mix WORKLOAD DB OUT_PREFIX --command C [--trace-db TRACE_DB]

Traffic calibration: the profiled program also runs a device-to-device copy
of M MiB (bench.py --copy-probe-mib M); the largest dispatches of the copy
kernel read and write exactly M MiB, so FETCH_SIZE and WRITE_SIZE (KiB per
dispatch, summed over the channels) are each scaled by the factor that
makes the copy's reading equal M MiB.
"""
import argparse
import collections
import json
import os
import sqlite3

COPY_KERNEL = "copyBuffer"


def short(name):
    """Kernel name with its template arguments, without return type and
    parameter list."""
    name = name.replace("sphk::(anonymous namespace)::", "").replace("(anonymous namespace)::", "")
    if name.startswith("void "):
        name = name[5:]
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and i > 0:
            return name[:i]
    return name


def per_dispatch(db, counter):
    """{dispatch_id: (kernel, value summed over components, grid)} for one counter."""
    c = sqlite3.connect(db)
    q = ("select dispatch_id, kernel_name, sum(value), max(grid_size) from counters_collection "
         "where counter_name = ? group by dispatch_id")
    return {d: (short(k), v, g) for d, k, v, g in c.execute(q, (counter,))}


def by_kernel(rows):
    out = collections.defaultdict(list)
    for kernel, value, _ in rows.values():
        out[kernel].append(value)
    return out


def copy_scale(rows, mib):
    """Counter-to-bytes factor from the probe copies: the copy-kernel
    dispatches whose reading is at least half the largest one (every large
    copy launches the same capped grid, so the grid does not tell them
    apart; the probe is by far the largest copy of the run)."""
    copies = [v for k, v, _ in rows.values() if COPY_KERNEL in k]
    if not copies:
        raise SystemExit("no copy-kernel dispatches: run the program with --copy-probe-mib")
    top = max(copies)
    vals = [v for v in copies if v >= 0.5 * top]
    mean = sum(vals) / len(vals)
    return (mib << 20) / (mean * 1024.0), mean


def is_runtime_kernel(name):
    return name.startswith("__amd_rocclr")


def cmd_trace(args):
    c = sqlite3.connect(args.db)
    q = ("select name, count(*), sum(duration), avg(duration), max(vgpr_count), max(accum_vgpr_count), "
         "max(sgpr_count), max(lds_size), max(scratch_size), min(grid_x * grid_y * grid_z), "
         "max(grid_x * grid_y * grid_z), max(workgroup_x * workgroup_y * workgroup_z) "
         "from kernels group by name order by sum(duration) desc")
    rows = list(c.execute(q))
    total = sum(r[2] for r in rows) or 1
    lines = [f"# Kernel trace: {args.title}", "", "Commands (from the repository root on one MI355X):", "", "```",
             *args.command, "```", ""]
    if args.note:
        lines += [args.note, ""]
    lines += ["| kernel | calls | total us | avg us | % | VGPR | AGPR | SGPR | LDS B | scratch B | "
              "grid (work-items, min-max) | wg |", "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for name, n, tot, avg, vg, ag, sg, lds, scr, gmin, gmax, wg in rows[:args.top]:
        grid = f"{gmin}" if gmin == gmax else f"{gmin}-{gmax}"
        lines.append(f"| {short(name)} | {n} | {tot / 1e3:.1f} | {avg / 1e3:.2f} | {100 * tot / total:.2f} | {vg} | "
                     f"{ag} | {sg} | {lds} | {scr} | {grid} | {wg} |")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")


def cmd_traffic(args):
    fetch = per_dispatch(args.fetch_db, "FETCH_SIZE")
    write = per_dispatch(args.write_db, "WRITE_SIZE")
    fs, fcopy = copy_scale(fetch, args.copy_mib)
    ws, wcopy = copy_scale(write, args.copy_mib)
    fk, wk = by_kernel(fetch), by_kernel(write)
    kernels = {}
    for k in fk:
        if is_runtime_kernel(k) or k not in wk:
            continue
        r = sum(fk[k]) / len(fk[k]) * 1024.0 * fs
        w = sum(wk[k]) / len(wk[k]) * 1024.0 * ws
        kernels[k] = {"read_bytes": r, "write_bytes": w, "dispatches": len(fk[k])}
    kernels = dict(sorted(kernels.items(), key=lambda kv: -(kv[1]["read_bytes"] + kv[1]["write_bytes"])))
    summary = {"workload": args.workload, "fetch_scale": fs, "write_scale": ws, "copy_mib": args.copy_mib,
               "kernels": kernels}
    with open(args.out_prefix + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    lines = [f"# HBM traffic per dispatch: {args.workload}", "",
             "Commands (from the repository root on one MI355X; each counter in its own run):", "", "```"]
    lines += args.command + ["```", "",
                             f"Calibration: the {args.copy_mib} MiB device-to-device copy reads FETCH_SIZE = "
                             f"{fcopy:.0f} KiB and WRITE_SIZE = {wcopy:.0f} KiB per dispatch, so FETCH_SIZE is "
                             f"scaled by {fs:.4f} and WRITE_SIZE by {ws:.4f}.", ""]
    if args.note:
        lines += [args.note, ""]
    lines += ["| kernel | dispatches | read bytes/dispatch | write bytes/dispatch | total |", "|---|---|---|---|---|"]
    for k, r in list(kernels.items())[:args.top]:
        lines.append(f"| {k} | {r['dispatches']} | {r['read_bytes']:.4g} | {r['write_bytes']:.4g} | "
                     f"{r['read_bytes'] + r['write_bytes']:.4g} |")
    with open(args.out_prefix + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")


MIX = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_LDS", "SQ_WAVE_CYCLES")
CUS = 256                 # MI355X compute units
CLOCK_HZ = 2.4e9          # MI355X peak engine clock


def mean_durations(db, trace):
    """Mean dispatch duration (ns) per kernel: from a kernel-trace database
    when given (no counter overhead), else from the counter run itself."""
    c = sqlite3.connect(db)
    if trace:
        q = "select name, avg(duration) from kernels group by name"
    else:
        q = ("select kernel_name, avg(d) from (select kernel_name, max(duration) d from counters_collection "
             "group by dispatch_id) group by kernel_name")
    return {short(k): d for k, d in c.execute(q)}


def cmd_mix(args):
    per = {name: by_kernel(per_dispatch(args.db, name)) for name in MIX}
    dur = mean_durations(args.trace_db, True) if args.trace_db else mean_durations(args.db, False)
    table = {}
    for k, waves in per["SQ_WAVES"].items():
        w = sum(waves) / len(waves)
        if w <= 0 or is_runtime_kernel(k):
            continue

        def mean(name):
            v = per[name].get(k)
            return sum(v) / len(v) if v else 0.0

        ns = dur.get(k)
        # a CU issues at most one wave64 VALU instruction per cycle (4 SIMDs,
        # 4 cycles each), so this is the share of the chip's VALU issue slots
        # the kernel used over its duration
        issue = mean("SQ_INSTS_VALU") / (CUS * CLOCK_HZ * ns * 1e-9) if ns else 0.0
        table[k] = {"waves": int(round(w)), "valu_per_wave": round(mean("SQ_INSTS_VALU") / w, 1),
                    "salu_per_wave": round(mean("SQ_INSTS_SALU") / w, 1),
                    "vmem_reads_per_wave": round(mean("SQ_INSTS_VMEM_RD") / w, 1),
                    "lds_per_wave": round(mean("SQ_INSTS_LDS") / w, 1),
                    "wave_cycles": round(mean("SQ_WAVE_CYCLES") / w, 0),
                    "avg_us": round(ns / 1e3, 2) if ns else None,
                    "valu_issue": round(issue, 3)}
    table = dict(sorted(table.items(), key=lambda kv: -kv[1]["valu_per_wave"] * kv[1]["waves"]))
    path = args.out_prefix + ".json"
    merged = json.load(open(path)) if os.path.exists(path) else {}
    merged[args.workload] = table
    commands = merged.pop("_commands", {})
    commands[args.workload] = args.command
    merged["_commands"] = commands
    with open(path, "w") as f:
        json.dump(merged, f, indent=1, sort_keys=True)
    lines = ["# Instruction mix of the neighbour passes (rocprofv3 SQ counters)", "",
             "Per wave: instructions issued by one wave64 on average (SQ_INSTS_* / SQ_WAVES). VALU issue: "
             "SQ_INSTS_VALU / (256 CUs x 2.4 GHz x the kernel's mean duration), the share of the chip's "
             "VALU issue slots used, since a CU issues at most one wave64 VALU instruction per cycle "
             "(4 SIMDs, 4 cycles each). Durations come from the kernel trace named below.", ""]
    for wl, rows in sorted((k, v) for k, v in merged.items() if not k.startswith("_")):
        lines += [f"## {wl}", "", "Commands (from the repository root on one MI355X):", "", "```",
                  *commands.get(wl, []), "```", "",
                  "| kernel | waves | VALU / wave | VMEM reads / wave | LDS / wave | SALU / wave | cycles / wave | "
                  "mean us | VALU issue |", "|---|---|---|---|---|---|---|---|---|"]
        for k, r in list(rows.items())[:args.top]:
            lines.append(f"| {k} | {r['waves']} | {r['valu_per_wave']:.0f} | {r['vmem_reads_per_wave']:.0f} | "
                         f"{r.get('lds_per_wave', 0):.0f} | {r.get('salu_per_wave', 0):.0f} | "
                         f"{r.get('wave_cycles', 0):.0f} | {r.get('avg_us') or 0:.1f} | {r['valu_issue']:.2f} |")
        lines.append("")
    with open(args.out_prefix + ".md", "w") as f:
        f.write("\n".join(lines))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="what", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("out")
    t.add_argument("--title", required=True)
    t.add_argument("--command", action="append", required=True)
    t.add_argument("--note", default="")
    t.add_argument("--top", type=int, default=40)
    t.set_defaults(fn=cmd_trace)
    h = sub.add_parser("traffic")
    h.add_argument("workload")
    h.add_argument("fetch_db")
    h.add_argument("write_db")
    h.add_argument("out_prefix")
    h.add_argument("--copy-mib", type=int, required=True)
    h.add_argument("--command", action="append", required=True)
    h.add_argument("--note", default="")
    h.add_argument("--top", type=int, default=30)
    h.set_defaults(fn=cmd_traffic)
    m = sub.add_parser("mix")
    m.add_argument("workload")
    m.add_argument("db")
    m.add_argument("out_prefix")
    m.add_argument("--command", action="append", required=True)
    m.add_argument("--trace-db", default=None, help="kernel-trace database of the same program, for durations")
    m.add_argument("--top", type=int, default=14)
    m.set_defaults(fn=cmd_mix)
    args = ap.parse_args()
    args.fn(args)


if __name__ == "__main__":
    main()
