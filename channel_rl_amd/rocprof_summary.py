"""Summaries of rocprofv3 output for the channel step kernels.

    python -m channel_rl_amd.rocprof_summary stats DIR/run_kernel_stats.csv
    python -m channel_rl_amd.rocprof_summary traffic FETCH_DIR/run_counter_collection.csv \
        WRITE_DIR/run_counter_collection.csv

`stats` prints the per-kernel table of `rocprofv3 --kernel-trace --stats`.
`traffic` turns two `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs (KiB per
dispatch) into HBM bytes per launch.  Each counter is scaled by what it reads
for k_stream_copy, whose bytes are known (the same run copies nbytes into a
second buffer): factor = nbytes / (counter KiB x 1024).  On gfx950 FETCH_SIZE
reads half of a streaming copy's loads (factor 2) and WRITE_SIZE all of its
stores (factor 1).

The traffic table's rows `| kernel | launches | fetch KiB | write KiB | bytes |`
are what `load_traffic` reads back.

    python -m channel_rl_amd.rocprof_summary sq PASS1/run_counter_collection.csv \
        PASS2/run_counter_collection.csv

    python -m channel_rl_amd.rocprof_summary stages DIR/run_kernel_stats.csv \
        FETCH_DIR/run_counter_collection.csv WRITE_DIR/run_counter_collection.csv

`stages` joins the two: per kernel, its average duration from the kernel
trace, its scaled HBM bytes per launch, the rate those bytes move at and
that rate as a fraction of the 8 TB/s HBM peak.

    python -m channel_rl_amd.rocprof_summary grids DIR/run_kernel_trace.csv [KERNEL_PREFIX]

`grids` groups the dispatches of a kernel trace by kernel and grid size (a
grouped ensemble launch stacks its members' grids along z, so each group of
channels shows as its own grid): launches, average and total duration; and,
over the whole trace, the sum of the kernels' durations against the union of
their intervals (above 1 where launches overlapped).

`sq` summarises the SQ counter passes listed in SQ_PASSES (medians per
dispatch; each counter is a whole-device sum per dispatch).  Per wave:
VALU and LDS instructions; as fractions of the waves' lifetime
(SQ_WAVE_CYCLES, in quad-cycles like the wait and active counters): time
waiting for anything, waiting to issue, and issuing.  Per unit: the SIMDs'
VALU busy fraction, SQ_ACTIVE_INST_VALU x 4 / (SIMDS x cycles), and the
LDS's, SQ_LDS_IDX_ACTIVE / (CUS x cycles), with cycles = GRBM_GUI_ACTIVE /
XCDS (the counter sums the XCDs' clocks); bank-conflict cycles as a
fraction of SQ_LDS_IDX_ACTIVE.
"""
import csv
import collections
import os
import statistics
import sys

COPY_KERNEL = "k_stream_copy"
COPY_BYTES = 1 << 30  # bench.py copies 1 GiB (read) into 1 GiB (write)


def short_name(name):
    """`void chan::k_rhs_rk<double, 64, 4>(...)` -> `k_rhs_rk<double, 64, 4>`."""
    n = name.replace("(anonymous namespace)::", "").replace("void ", "")
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            n = n[:i]
            break
    return n.replace("chan::", "").strip()


def base_name(name):
    return short_name(name).split("<")[0]


def read_stats(path):
    rows = []
    for r in csv.DictReader(open(path)):
        rows.append((short_name(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]),
                     float(r["AverageNs"]), float(r["Percentage"])))
    return rows


def stats_table(path):
    out = ["| kernel | calls | total ms | average us | % |", "|---|---:|---:|---:|---:|"]
    for n, c, tot, avg, pct in read_stats(path):
        out.append(f"| {n} | {c} | {tot / 1e6:.3f} | {avg / 1e3:.2f} | {pct:.2f} |")
    return "\n".join(out)


def read_counter(path):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        d[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return d


def traffic(fetch_path, write_path):
    """{kernel: (launches, fetch KiB, write KiB, scaled bytes per launch)} and
    the two scale factors."""
    f, w = read_counter(fetch_path), read_counter(write_path)
    copy = next(k for k in f if base_name(k) == COPY_KERNEL)
    # the first copy launch is the warm-up; every launch moves the same bytes
    sf = COPY_BYTES / (statistics.median(f[copy]) * 1024.0)
    sw = COPY_BYTES / (statistics.median(w[copy]) * 1024.0)
    res = {}
    for k in f:
        fk = statistics.mean(f[k])
        wk = statistics.mean(w.get(k, [0.0]))
        res[k] = (len(f[k]), fk, wk, fk * 1024.0 * sf + wk * 1024.0 * sw)
    return res, sf, sw


def traffic_table(fetch_path, write_path):
    res, sf, sw = traffic(fetch_path, write_path)
    out = [f"copy-kernel scale factors: FETCH_SIZE x {sf:.3f}, WRITE_SIZE x {sw:.3f}", "",
           "| kernel | launches | fetch KiB | write KiB | bytes |", "|---|---:|---:|---:|---:|"]
    for k, (n, fk, wk, b) in sorted(res.items(), key=lambda kv: -kv[1][3]):
        out.append(f"| {k} | {n} | {fk:.1f} | {wk:.1f} | {b:.0f} |")
    return "\n".join(out)


HBM_PEAK_GBS = 8000.0


def stages_table(stats_path, fetch_path, write_path):
    res, sf, sw = traffic(fetch_path, write_path)
    out = [f"copy-kernel scale factors: FETCH_SIZE x {sf:.3f}, WRITE_SIZE x {sw:.3f}", "",
           "| kernel | calls | average us | % of time | HBM bytes/launch | GB/s | of HBM peak |",
           "|---|---:|---:|---:|---:|---:|---:|"]
    for n, c, tot, avg, pct in read_stats(stats_path):
        if n not in res or base_name(n) == COPY_KERNEL:
            continue
        b = res[n][3]
        gbs = b / avg  # bytes per ns = GB/s
        out.append(f"| {n} | {c} | {avg / 1e3:.2f} | {pct:.2f} | {b:.0f} | {gbs:.0f} | {gbs / HBM_PEAK_GBS:.3f} |")
    return "\n".join(out)


# the two counter passes (at most 8 SQ and 2 GRBM counters each)
SQ_PASSES = (
    "SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_LDS SQ_WAIT_INST_LDS "
    "SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE SQ_WAIT_ANY GRBM_GUI_ACTIVE",
    "SQ_WAIT_INST_ANY SQ_ACTIVE_INST_ANY SQ_LDS_DATA_FIFO_FULL SQ_LDS_ADDR_CONFLICT SQ_INSTS_VALU "
    "SQ_INSTS_LDS SQ_LDS_UNALIGNED_STALL SQ_INSTS_SALU GRBM_GUI_ACTIVE",
)
XCDS, CUS, SIMDS = 8, 256, 1024  # MI355X


def sq_table(paths):
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    grid = {}
    for p in paths:
        for r in csv.DictReader(open(p)):
            k = short_name(r["Kernel_Name"])
            vals[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
            grid[k] = int(r["Grid_Size"])
    need = ("SQ_WAVE_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_ANY", "SQ_ACTIVE_INST_VALU",
            "SQ_LDS_IDX_ACTIVE", "SQ_LDS_BANK_CONFLICT", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "GRBM_GUI_ACTIVE")
    out = ["| kernel | waves | VALU inst/wave | LDS inst/wave | waiting | waiting to issue | issuing | "
           "VALU busy | LDS busy | LDS bank conflicts |",
           "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    rows = []
    for k, d in vals.items():
        if not k.startswith("k_") or any(c not in d for c in need):
            continue
        m = {c: statistics.median(v) for c, v in d.items()}
        waves = grid[k] / 64.0
        wc = m["SQ_WAVE_CYCLES"]
        cyc = m["GRBM_GUI_ACTIVE"] / XCDS
        valu = m["SQ_ACTIVE_INST_VALU"] * 4.0 / (SIMDS * cyc)
        lds = m["SQ_LDS_IDX_ACTIVE"] / (CUS * cyc)
        bank = m["SQ_LDS_BANK_CONFLICT"] / m["SQ_LDS_IDX_ACTIVE"] if m["SQ_LDS_IDX_ACTIVE"] else 0.0
        rows.append((cyc, f"| {k} | {waves:.0f} | {m['SQ_INSTS_VALU'] / waves:.0f} | {m['SQ_INSTS_LDS'] / waves:.0f} | "
                          f"{m['SQ_WAIT_ANY'] / wc:.2f} | {m['SQ_WAIT_INST_ANY'] / wc:.2f} | "
                          f"{m['SQ_ACTIVE_INST_ANY'] / wc:.2f} | {valu:.2f} | {lds:.2f} | {bank:.2f} |"))
    out += [r for _, r in sorted(rows, key=lambda t: -t[0])]
    return "\n".join(out)


def read_trace(path):
    """[(short name, (grid x, y, z), start ns, end ns)] of a rocprofv3 kernel trace."""
    rows = []
    for r in csv.DictReader(open(path)):
        g = tuple(int(r.get(f"Grid_Size_{a}", r.get(f"Grid_Size{a}", 0)) or 0) for a in "XYZ")
        rows.append((short_name(r["Kernel_Name"]), g, int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    return rows


def grids_table(path, prefix=""):
    rows = [r for r in read_trace(path) if base_name(r[0]).startswith(prefix) or not prefix]
    groups = collections.defaultdict(list)
    for n, g, t0, t1 in rows:
        groups[(n, g)].append(t1 - t0)
    busy = sum(t1 - t0 for _, _, t0, t1 in rows)
    union, end = 0, None
    for _, _, t0, t1 in sorted(rows, key=lambda r: r[2]):
        if end is None or t0 > end:
            union += t1 - t0
            end = t1
        elif t1 > end:
            union += t1 - end
            end = t1
    out = [f"kernel time {busy / 1e6:.3f} ms over {union / 1e6:.3f} ms of intervals "
           f"(overlap factor {busy / max(union, 1):.2f})", "",
           "| kernel | grid (threads) | launches | average us | total ms |", "|---|---|---:|---:|---:|"]
    for (n, g), d in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"| {n} | {g[0]} x {g[1]} x {g[2]} | {len(d)} | {statistics.mean(d) / 1e3:.2f} | "
                   f"{sum(d) / 1e6:.3f} |")
    return "\n".join(out)


def load_traffic(path):
    """Scaled HBM bytes per launch by kernel base name from a traffic table;
    of several instances of a kernel (tile shapes timed by the autotuner),
    the one launched most often: the one the step runs."""
    res, most = {}, {}
    if not os.path.exists(path):
        return res
    for line in open(path):
        p = [s.strip() for s in line.strip().strip("|").split("|")]
        if len(p) == 5 and p[0].startswith("k_"):
            try:
                n, b = int(p[1]), float(p[4])
            except ValueError:
                continue
            k = base_name(p[0])
            if n > most.get(k, -1):
                most[k], res[k] = n, b
    return res


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "stats":
        print(stats_table(sys.argv[2]))
    elif len(sys.argv) >= 4 and sys.argv[1] == "traffic":
        print(traffic_table(sys.argv[2], sys.argv[3]))
    elif len(sys.argv) >= 5 and sys.argv[1] == "stages":
        print(stages_table(sys.argv[2], sys.argv[3], sys.argv[4]))
    elif len(sys.argv) >= 3 and sys.argv[1] == "grids":
        print(grids_table(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else ""))
    elif len(sys.argv) >= 3 and sys.argv[1] == "sq":
        print(sq_table(sys.argv[2:]))
    else:
        print(__doc__)
        sys.exit(2)
