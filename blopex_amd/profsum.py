"""Text summaries of rocprofv3 output (CSV format) for the profiles/ folder.

    python3 -m blopex_amd.profsum stats  <trace-dir> [--title T] [--command C]...
    python3 -m blopex_amd.profsum counters <fetch-dir> <write-dir> [--title T] [--command C]...
    python3 -m blopex_amd.profsum mfma <counter-dir> <trace-dir> [--title T] [--command C]...
    python3 -m blopex_amd.profsum sq <counter-dir> [--min-calls N] [--title T] [--command C]...

`stats` reads `*_kernel_stats.csv` (rocprofv3 --kernel-trace --stats) and
prints the per-kernel table.  `counters` reads the `*_counter_collection.csv`
of one FETCH_SIZE pass and one WRITE_SIZE pass over the same program.  The
program must also run a device-to-device copy of 512 MiB
(`bench.py --roofline-only` does): its counter readings give the factor that
turns each counter's KiB into bytes moved (counter reads x factor =
512 MiB).  For every kernel the summary prints the scaled bytes per launch and
a line `traffic <kernel> <bytes>` that bench.py reads back.  `sq` reads one
pass of SQ counters and prints, for every kernel dispatched at least N times,
the per-dispatch average of each counter and, where the counters are there:
the MFMA pipes' busy share (SQ_VALU_MFMA_BUSY_CYCLES over the SIMD cycles:
SQ_BUSY_CYCLES is summed over the shader arrays, so the kernel's cycles are
SQ_BUSY_CYCLES / arrays, times the SIMD count from the agent info), the share
of wave cycles spent waiting (SQ_WAIT_ANY, SQ_WAIT_INST_LDS over
SQ_WAVE_CYCLES), and bank-conflict cycles per LDS-active cycle.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import os
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"
COPY_BYTES = 512 << 20


def short_name(name: str) -> str:
    """`void bx::(anonymous namespace)::k_dia<float, 7>(long, ...)` -> `k_dia<float, 7>`."""
    s = name.replace("(anonymous namespace)::", "")
    if s.startswith("void "):
        s = s[5:]
    depth = 0
    for i, ch in enumerate(s):
        depth += (ch == "<") - (ch == ">")
        if ch == "(" and depth == 0:
            s = s[:i]
            break
    return s.split("::")[-1] if "<" not in s else s[:s.index("<")].split("::")[-1] + s[s.index("<"):]


def _one(pattern: str) -> str:
    hits = sorted(glob.glob(pattern, recursive=True))
    if not hits:
        raise SystemExit(f"no file matches {pattern}")
    return hits[0]


def stats(trace_dir: str) -> list[str]:
    path = _one(os.path.join(trace_dir, "**", "*kernel_stats.csv"))
    rows = list(csv.DictReader(open(path)))
    out = [f"{'kernel':<48} {'calls':>6} {'avg us':>10} {'min us':>10} {'max us':>10} "
           f"{'total ms':>10} {'%':>6}"]
    for r in rows:
        out.append(f"{short_name(r['Name'])[:48]:<48} {int(r['Calls']):>6} "
                   f"{float(r['AverageNs']) / 1e3:>10.1f} {float(r['MinNs']) / 1e3:>10.1f} "
                   f"{float(r['MaxNs']) / 1e3:>10.1f} {float(r['TotalDurationNs']) / 1e6:>10.3f} "
                   f"{float(r['Percentage']):>6.2f}")
    return out


def _counter(pass_dir: str, counter: str) -> dict[str, list[float]]:
    path = _one(os.path.join(pass_dir, "**", "*counter_collection.csv"))
    vals: dict[str, list[float]] = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            vals[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return vals


def counters(fetch_dir: str, write_dir: str) -> tuple[list[str], dict[str, float]]:
    fetch = _counter(fetch_dir, "FETCH_SIZE")
    write = _counter(write_dir, "WRITE_SIZE")
    if COPY_KERNEL not in fetch or COPY_KERNEL not in write:
        raise SystemExit(f"the passes hold no {COPY_KERNEL} dispatch to calibrate against")
    mean = lambda v: sum(v) / len(v)
    ff = COPY_BYTES / (mean(fetch[COPY_KERNEL]) * 1024.0)
    fw = COPY_BYTES / (mean(write[COPY_KERNEL]) * 1024.0)
    out = [f"calibration: {COPY_KERNEL} of {COPY_BYTES} B reads FETCH_SIZE "
           f"{mean(fetch[COPY_KERNEL]):.0f} KiB -> factor {ff:.3f}; WRITE_SIZE "
           f"{mean(write[COPY_KERNEL]):.0f} KiB -> factor {fw:.3f}",
           "",
           f"{'kernel':<40} {'disp':>5} {'FETCH KiB':>12} {'WRITE KiB':>12} {'read B':>14} "
           f"{'written B':>14} {'traffic B':>14}"]
    traffic = {}
    for name in fetch:
        if name == COPY_KERNEL or name not in write:
            continue
        fk, wk = mean(fetch[name]), mean(write[name])
        rb, wb = fk * 1024.0 * ff, wk * 1024.0 * fw
        traffic[name] = rb + wb
        out.append(f"{name[:40]:<40} {len(fetch[name]):>5} {fk:>12.0f} {wk:>12.0f} {rb:>14.0f} "
                   f"{wb:>14.0f} {rb + wb:>14.0f}")
    out.append("")
    out += [f"traffic {name} {b:.0f}" for name, b in traffic.items()]
    return out, traffic


MFMA_COUNTERS = ("SQ_INSTS_VALU_MFMA_MOPS_F32", "SQ_INSTS_VALU_MFMA_MOPS_F64")
PEAK_TFLOPS = {"F32": 157.3, "F64": 78.6}   # MI355X dense MFMA


def mfma(counter_dir: str, trace_dir: str) -> tuple[list[str], dict[str, tuple[float, float]]]:
    """MFMA flops per dispatch (MOPS x 512, whole GPU) per kernel from one
    counter pass, and achieved TFLOP/s against the kernel-trace durations of
    the same program (FP32 or FP64 peak, whichever the kernel mostly runs)."""
    path = _one(os.path.join(counter_dir, "**", "*counter_collection.csv"))
    per: dict[str, dict[str, list[float]]] = collections.defaultdict(
        lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] in MFMA_COUNTERS:
            per[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    tpath = _one(os.path.join(trace_dir, "**", "*kernel_stats.csv"))
    avg_ns = {short_name(r["Name"]): float(r["AverageNs"]) for r in csv.DictReader(open(tpath))}
    mean = lambda v: sum(v) / len(v) if v else 0.0
    out = [f"{'kernel':<40} {'disp':>5} {'GFLOP/disp':>11} {'avg us':>8} {'TFLOP/s':>8} "
           f"{'peak':>6} {'frac':>6}"]
    res = {}
    for name, c in sorted(per.items()):
        f32 = mean(c["SQ_INSTS_VALU_MFMA_MOPS_F32"]) * 512
        f64 = mean(c["SQ_INSTS_VALU_MFMA_MOPS_F64"]) * 512
        flops = f32 + f64
        if flops == 0 or name not in avg_ns:
            continue
        kind = "F64" if f64 > f32 else "F32"
        tf = flops / avg_ns[name] / 1e3
        if tf < 0.05:
            continue  # a few MFMA tiles in a latency-bound kernel
        res[name] = (tf, PEAK_TFLOPS[kind])
        out.append(f"{name[:40]:<40} {len(c['SQ_INSTS_VALU_MFMA_MOPS_F32']):>5} {flops / 1e9:>11.3f} "
                   f"{avg_ns[name] / 1e3:>8.1f} {tf:>8.1f} {PEAK_TFLOPS[kind]:>6.1f} "
                   f"{tf / PEAK_TFLOPS[kind]:>6.3f}")
    out.append("")
    out += [f"mfma {name} {tf:.2f} {peak:.1f}" for name, (tf, peak) in res.items()]
    return out, res


def sq(counter_dir: str, min_calls: int = 10) -> list[str]:
    path = _one(os.path.join(counter_dir, "**", "*counter_collection.csv"))
    agent = _one(os.path.join(counter_dir, "**", "*agent_info.csv"))
    simds = arrays = 0
    for r in csv.DictReader(open(agent)):
        if int(r.get("Simd_Count", 0) or 0) > 0:
            simds, arrays = int(r["Simd_Count"]), int(r["Array_Count"])
    sums: dict[str, dict[str, float]] = collections.defaultdict(lambda: collections.defaultdict(float))
    disp: dict[str, set] = collections.defaultdict(set)
    names: list[str] = []
    for r in csv.DictReader(open(path)):
        k = short_name(r["Kernel_Name"])
        sums[k][r["Counter_Name"]] += float(r["Counter_Value"])
        disp[k].add(r["Dispatch_Id"])
        if r["Counter_Name"] not in names:
            names.append(r["Counter_Name"])
    out = [f"agent: {simds} SIMDs, {arrays} shader arrays; counters: {' '.join(names)}", ""]
    for k in sorted(sums, key=lambda x: -sums[x].get("SQ_BUSY_CYCLES", 0.0)):
        n = len(disp[k])
        if n < min_calls:
            continue
        a = sums[k]
        out.append(f"{k}  ({n} dispatches; per dispatch)")
        for c in names:
            out.append(f"    {c:<28} {a[c] / n:>14.4g}")
        if a.get("SQ_BUSY_CYCLES") and "SQ_VALU_MFMA_BUSY_CYCLES" in a and simds and arrays:
            share = a["SQ_VALU_MFMA_BUSY_CYCLES"] / (simds * a["SQ_BUSY_CYCLES"] / arrays)
            out.append(f"    MFMA busy share of SIMD cycles  {share:.3f}")
        if a.get("SQ_WAVE_CYCLES"):
            if "SQ_WAIT_ANY" in a:
                out.append(f"    waiting share of wave cycles    {a['SQ_WAIT_ANY'] / a['SQ_WAVE_CYCLES']:.3f}")
            if "SQ_WAIT_INST_LDS" in a:
                out.append(f"    LDS-wait share of wave cycles   {a['SQ_WAIT_INST_LDS'] / a['SQ_WAVE_CYCLES']:.3f}")
        if a.get("SQ_ACTIVE_INST_LDS") and "SQ_LDS_BANK_CONFLICT" in a:
            out.append(f"    bank conflict per LDS-active    {a['SQ_LDS_BANK_CONFLICT'] / a['SQ_ACTIVE_INST_LDS']:.3f}")
        out.append("")
    return out


def read_mfma(path: str, kernel: str) -> tuple[float, float] | None:
    """The `mfma <kernel> <TFLOP/s> <peak>` figures of a committed summary, or None."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("mfma "):
            head, tf, peak = line[5:].rsplit(" ", 2)
            if head == kernel:
                return float(tf), float(peak)
    return None


def read_trace_avg_us(path: str, kernel: str) -> float | None:
    """The `avg us` column of a kernel's row in a committed `stats` summary, or None."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.rsplit(None, 6)
        if len(parts) == 7 and parts[0].strip() == kernel:
            try:
                return float(parts[2])
            except ValueError:
                return None
    return None


def read_traffic(path: str, kernel: str) -> float | None:
    """The `traffic <kernel> <bytes>` figure of a committed counter summary, or None."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic ") and line[8:].rsplit(" ", 1)[0] == kernel:
            return float(line.rsplit(" ", 1)[1])
    return None


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["stats", "counters", "mfma", "sq"])
    ap.add_argument("dirs", nargs="+")
    ap.add_argument("--min-calls", type=int, default=10)
    ap.add_argument("--title", default="")
    ap.add_argument("--command", action="append", default=[])
    a = ap.parse_args(argv)
    head = [a.title] if a.title else []
    if a.command:
        head.append("commands (from the repository root, TMPDIR=/tmp):")
        head += [f"  {c}" for c in a.command]
        extra = f" --min-calls {a.min_calls}" if a.what == "sq" and a.min_calls != 10 else ""
        head.append(f"summary: python3 -m blopex_amd.profsum {a.what} {' '.join(a.dirs)}{extra}")
        head.append("")
    if a.what == "stats":
        body = stats(a.dirs[0])
    elif a.what == "sq":
        body = sq(a.dirs[0], a.min_calls)
    elif a.what == "mfma":
        if len(a.dirs) != 2:
            raise SystemExit("mfma needs <counter-dir> <trace-dir>")
        body, _ = mfma(a.dirs[0], a.dirs[1])
    else:
        if len(a.dirs) != 2:
            raise SystemExit("counters needs <fetch-dir> <write-dir>")
        body, _ = counters(a.dirs[0], a.dirs[1])
    sys.stdout.write("\n".join(head + body) + "\n")


if __name__ == "__main__":
    main()
