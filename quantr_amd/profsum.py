"""Summarise rocprofv3 CSV output into the short text/JSON records kept under profiles/.

Usage:
    python -m quantr_amd.profsum stats  <run_kernel_stats.csv> <out.md> <command...>
    python -m quantr_amd.profsum traffic <FETCH_SIZE csv> <WRITE_SIZE csv> <kernel> <bytes> <out> <command...>
    python -m quantr_amd.profsum valu <counter csv> <kernel> <out.md> <command...>

`traffic` scales both counters with kernels of known traffic that the same process ran:
`k_norm_partial` reads the whole state once (FETCH_SIZE calibration) and `k_init_basis`
writes it once (WRITE_SIZE calibration); the state size in bytes is the `bytes` argument
(one read + one write of it is the tile sweep's algorithmic traffic).
"""
import collections
import csv
import json
import sys


def kernel_kind(name):
    """Compiled stages are named <kind>_<j> (one entry point per group of a module); they are
    grouped by kind."""
    for kind in ("qr_dual_stage", "qr_stage"):
        if name == kind or name.startswith(kind + "_"):
            return kind
    return name


def per_kernel(path, counter=None):
    d = collections.defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if counter is not None and r["Counter_Name"] != counter:
                continue
            d[kernel_kind(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in d.items()}, {k: len(v) for k, v in d.items()}


def stats_from_trace(trace_csv):
    """Per-kind call count, average/min/max duration (ns) and share, from a kernel trace."""
    d = collections.defaultdict(list)
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            d[kernel_kind(r["Kernel_Name"])].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    total = sum(sum(v) for v in d.values()) or 1
    rows = []
    for k, v in sorted(d.items(), key=lambda kv: -sum(kv[1])):
        rows.append({"Name": k, "Calls": len(v), "AverageNs": sum(v) / len(v), "MinNs": min(v), "MaxNs": max(v),
                     "Percentage": "%.2f" % (100.0 * sum(v) / total)})
    return rows


def find(d, sub):
    for k in d:
        if sub in k:
            return k
    raise KeyError(sub)


def stats(csv_path, out, cmd):
    """csv_path: a kernel-trace CSV (grouped here by kind) or a --stats CSV."""
    if csv_path.endswith("kernel_trace.csv"):
        rows = stats_from_trace(csv_path)
    else:
        rows = list(csv.DictReader(open(csv_path)))
    lines = ["# rocprofv3 kernel durations", "", "Command (from the repository root on the GPU host):", "",
             "    " + cmd, "", "| kernel | calls | avg ms | min ms | max ms | share % |", "|---|---|---|---|---|---|"]
    for r in rows:
        name = r["Name"] if len(r["Name"]) < 70 else r["Name"][:67] + "..."
        lines.append("| %s | %s | %.4f | %.4f | %.4f | %s |" % (
            name, r["Calls"], float(r["AverageNs"]) / 1e6, float(r["MinNs"]) / 1e6,
            float(r["MaxNs"]) / 1e6, r["Percentage"]))
    open(out, "w").write("\n".join(lines) + "\n")


def traffic(fetch_csv, write_csv, kernel, state_bytes, out, cmd):
    fd, fn = per_kernel(fetch_csv, "FETCH_SIZE")
    wd, wn = per_kernel(write_csv, "WRITE_SIZE")
    kib = state_bytes / 1024.0
    f_scale = kib / fd[find(fd, "k_norm_partial")]
    w_scale = kib / wd[find(wd, "k_init_basis")]
    kf, kw = find(fd, kernel), find(wd, kernel)
    fetch_b = fd[kf] * f_scale * 1024.0
    write_b = wd[kw] * w_scale * 1024.0
    rec = {
        "kernel": kernel, "command": cmd, "dispatches": fn[kf],
        "fetch_kib_raw": fd[kf], "write_kib_raw": wd[kw],
        "fetch_scale": f_scale, "write_scale": w_scale,
        "fetch_bytes_per_launch": fetch_b, "write_bytes_per_launch": write_b,
        "traffic_bytes_per_launch": fetch_b + write_b,
        "algorithmic_bytes_per_launch": 2.0 * state_bytes,
        "traffic_over_algorithmic": (fetch_b + write_b) / (2.0 * state_bytes),
        "calibration": "FETCH_SIZE scaled by k_norm_partial (reads the state once), "
                       "WRITE_SIZE by k_init_basis (writes the state once), same process",
    }
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
        f.write("\n")
    return rec


def valu(counter_csv, kernel, out, cmd, simds=1024, xcds=8):
    """Per-dispatch VALU counters of one kernel kind. SQ_INSTS_VALU counts wave instructions;
    a wave64 VALU instruction holds a 16-lane SIMD for 4 cycles, so VALU busy = 4 * insts /
    (SIMDs * cycles), with the kernel's cycles from GRBM_GUI_ACTIVE (summed over the XCDs)."""
    d = collections.defaultdict(lambda: collections.defaultdict(list))
    grid = {}
    with open(counter_csv) as f:
        for r in csv.DictReader(f):
            k = kernel_kind(r["Kernel_Name"])
            d[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
            g0, wg0, v0 = grid.get(k, (0, 0, set()))
            v0.add(int(r["VGPR_Count"]))
            grid[k] = (int(r["Grid_Size"]), int(r["Workgroup_Size"]), v0)
    c = d[find(d, kernel)]
    avg = {name: sum(v) / len(v) for name, v in c.items()}
    g, wg, vgpr = grid[find(d, kernel)]
    waves = g / 64.0
    cycles = avg["GRBM_GUI_ACTIVE"] / xcds
    f64 = sum(avg.get(x, 0.0) for x in ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64"))
    lines = ["# rocprofv3 VALU counters: %s" % kernel, "", "Command (from the repository root on the GPU host):", "",
             "    " + cmd, "", "| quantity | per dispatch (average of %d) |" % len(c["SQ_INSTS_VALU"]), "|---|---|"]
    rows = [("grid threads / workgroup", "%d / %d" % (g, wg)),
            ("VGPR_Count column as reported (range over dispatches)", "%d-%d" % (min(vgpr), max(vgpr))),
            ("waves", "%.0f" % waves)]
    for name in sorted(avg):
        rows.append((name, "%.4g" % avg[name]))
    rows += [("VALU instructions per wave", "%.0f" % (avg["SQ_INSTS_VALU"] / waves)),
             ("FP64 share of VALU instructions", "%.2f" % (f64 / avg["SQ_INSTS_VALU"])),
             ("LDS instructions per wave", "%.0f" % (avg.get("SQ_INSTS_LDS", 0.0) / waves)),
             ("kernel cycles per XCD (GRBM_GUI_ACTIVE / %d)" % xcds, "%.4g" % cycles),
             ("VALU busy = 4 * SQ_INSTS_VALU / (%d SIMDs * cycles)" % simds,
              "%.3f" % (4.0 * avg["SQ_INSTS_VALU"] / (simds * cycles)))]
    for a, b in rows:
        lines.append("| %s | %s |" % (a, b))
    open(out, "w").write("\n".join(lines) + "\n")


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(sys.argv[2], sys.argv[3], " ".join(sys.argv[4:]))
    elif sys.argv[1] == "traffic":
        print(json.dumps(traffic(sys.argv[2], sys.argv[3], sys.argv[4], float(sys.argv[5]), sys.argv[6],
                                 " ".join(sys.argv[7:])), indent=1))
    elif sys.argv[1] == "valu":
        valu(sys.argv[2], sys.argv[3], sys.argv[4], " ".join(sys.argv[5:]))
    else:
        raise SystemExit("usage: profsum stats|traffic|valu ...")
