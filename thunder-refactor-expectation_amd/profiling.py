"""rocprofv3 helpers for the committed profile summaries.

    python profiling.py copy
        one 1 GiB device-to-device copy, three times: the calibration kernel
        for FETCH_SIZE / WRITE_SIZE (bytes it must move are known);
    python profiling.py summarize --name NAME --trace DIR [--fetch DIR
        --write DIR --cal-fetch DIR --cal-write DIR] --command "..." --out PREFIX
        writes PREFIX_trace.md (kernel durations from --stats) and, with the
        counter runs, PREFIX_traffic.md / PREFIX_traffic.json: per-launch HBM
        read and write bytes of each kernel (average over its dispatches of
        the largest grid), each counter scaled so the copy reads and writes
        exactly its size; with --mfma (a run of the MFMA_COUNTERS below)
        also PREFIX_mfma.md: executed matrix-core flops, rate and busy
        fraction of each kernel that issues MFMA instructions.

Counter directories are rocprofv3 `--output-format csv` runs
(run_counter_collection.csv); the trace directory holds run_kernel_stats.csv.
"""
import argparse
import csv
import json
import os
import re
import sys

COPY_BYTES = 1 << 30


def copy():
    import torch
    n = COPY_BYTES // 4
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {COPY_BYTES} bytes per launch, 3 launches")


def short(name):
    n = name[5:] if name.startswith("void ") else name
    n = n.split("(")[0].strip()
    return n[5:] if n.startswith("thx::") else n


PREC = {"0": "fp32", "1": "bf16", "2": "fp16", "3": "fp8"}


def simple_key(s):
    m = re.match(r"k_score16<(\d), \d+(?:, \d+)?>", s)
    if m:
        return f"k_score16_{PREC[m.group(1)]}"
    m = re.match(r"k_score<(\d), \d+>", s)
    if m:
        return f"k_score_{PREC[m.group(1)]}"
    if s.startswith("k_score_sel<"):
        return "k_score_sel"
    if s.startswith("k_score_local_q<"):
        return "k_score_local_q"
    if s.startswith("k_score_local<"):
        return "k_score_local"
    return None


def counters(d):
    """{(kernel, grid): [values]} from one counter run."""
    out = {}
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for row in csv.DictReader(f):
            key = (short(row["Kernel_Name"]), int(row["Grid_Size"]))
            out.setdefault(key, []).append(float(row["Counter_Value"]))
    return out


def copy_scale(d, per_launch_bytes):
    c = counters(d)
    vals = [v for (k, g), vs in c.items() if "copyBuffer" in k for v in vs]
    kib = max(vals)
    return per_launch_bytes / (kib * 1024.0), kib


# one counter run: 5 SQ counters and 1 GRBM counter (within the per-run limits)
MFMA_COUNTERS = ["SQ_INSTS_VALU_MFMA_MOPS_F32", "SQ_INSTS_VALU_MFMA_MOPS_BF16",
                 "SQ_INSTS_VALU_MFMA_MOPS_F8", "SQ_INSTS_VALU_MFMA_MOPS_F6F4",
                 "SQ_VALU_MFMA_BUSY_CYCLES", "GRBM_GUI_ACTIVE"]
# operand type of each MOPS counter and MI355X's dense peak for it, TFLOP/s
# (the FP8 operands of v_mfma_scale_*_f8f6f4 count under F8 or F6F4)
MFMA_PEAK = {"SQ_INSTS_VALU_MFMA_MOPS_F32": ("fp32", 157.3),
             "SQ_INSTS_VALU_MFMA_MOPS_BF16": ("bf16", 2500.0),
             "SQ_INSTS_VALU_MFMA_MOPS_F8": ("fp8", 5000.0),
             "SQ_INSTS_VALU_MFMA_MOPS_F6F4": ("fp8", 5000.0)}
N_XCC, N_SIMD = 8, 256 * 4


def mfma_summary(d, name, command):
    """Markdown table of the MFMA counters of one run, per (kernel, grid),
    averaged over its dispatches.  MOPS counters count 512 flops each;
    GRBM_GUI_ACTIVE is reported summed over the 8 XCDs, so one XCD's cycles of
    the dispatch are GRBM / 8; MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES (summed over
    the 1024 SIMDs) / (1024 x GRBM / 8)."""
    disp = {}
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for row in csv.DictReader(f):
            e = disp.setdefault(row["Dispatch_Id"], {
                "key": (short(row["Kernel_Name"]), int(row["Grid_Size"])),
                "ns": float(row["End_Timestamp"]) - float(row["Start_Timestamp"]), "c": {}})
            e["c"][row["Counter_Name"]] = e["c"].get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    groups = {}
    for e in disp.values():
        groups.setdefault(e["key"], []).append(e)
    md = [f"# MFMA utilisation per launch: {name}", "",
          "Command (as for the kernel trace):", "", "```", command, "```", "",
          "Per kernel and grid, averaged over its dispatches in this run: executed matrix-core "
          "flops = MOPS x 512 (padding rows and columns included), the dispatch time of the "
          "counter run, the executed rate against MI355X's dense peak for the operand type, "
          "and MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES / (1024 SIMDs x GRBM_GUI_ACTIVE / 8 XCDs); "
          "`clock` = GRBM_GUI_ACTIVE / 8 / time, a check on the normalisation.", "",
          "| Kernel | grid | launches | type | executed TFLOP | ms | TFLOP/s | % of peak | "
          "MFMA busy % | clock GHz |", "|---|---|---|---|---|---|---|---|---|---|"]
    rows = []
    for (k, g), es in groups.items():
        n = len(es)
        avg = {c: sum(e["c"].get(c, 0.0) for e in es) / n for c in MFMA_COUNTERS}
        mops = {c: avg[c] for c in MFMA_PEAK if avg[c] > 0}
        if not mops:
            continue
        c_main = max(mops, key=mops.get)
        typ, peak = MFMA_PEAK[c_main]
        flops = 512.0 * sum(mops.values())
        ns = sum(e["ns"] for e in es) / n
        cyc = avg["GRBM_GUI_ACTIVE"] / N_XCC
        busy = 100.0 * avg["SQ_VALU_MFMA_BUSY_CYCLES"] / (N_SIMD * cyc) if cyc > 0 else float("nan")
        rate = flops / (ns * 1e-9) / 1e12 if ns > 0 else float("nan")
        rows.append((flops, f"| `{k}` | {g} | {n} | {typ} | {flops / 1e12:.3f} | {ns / 1e6:.3f} | "
                            f"{rate:.1f} | {100.0 * rate / peak:.1f} | {busy:.1f} | "
                            f"{cyc / ns:.2f} |"))
    md += [r for _, r in sorted(rows, key=lambda x: -x[0])]
    return "\n".join(md) + "\n"


def summarize(a):
    rows = list(csv.DictReader(open(os.path.join(a.trace, "run_kernel_stats.csv"))))
    lines = [f"# Kernel trace: {a.name}", "", "Command (repository root, `cd /tmp && export "
             "TMPDIR=/tmp` first; rocprofv3 writes its CSV files under the -d directory):", "", "```",
             a.command, "```", "", "| Kernel | Calls | Average ms | Min ms | Max ms | % |",
             "|---|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.3f} | "
                     f"{float(r['MinNs']) / 1e6:.3f} | {float(r['MaxNs']) / 1e6:.3f} | "
                     f"{float(r['Percentage']):.2f} |")
    open(a.out + "_trace.md", "w").write("\n".join(lines) + "\n")
    if a.mfma:
        open(a.out + "_mfma.md", "w").write(mfma_summary(a.mfma, a.name, a.command_mfma))
    if not a.fetch:
        return
    fs, fkib = copy_scale(a.cal_fetch, COPY_BYTES)
    ws, wkib = copy_scale(a.cal_write, COPY_BYTES)
    fc, wc = counters(a.fetch), counters(a.write)
    best = {}
    for (k, g), vs in fc.items():
        if (k, g) not in wc:
            continue
        if k not in best or g > best[k][0]:
            best[k] = (g, sum(vs) / len(vs), sum(wc[(k, g)]) / len(wc[(k, g)]))
    kern = {}
    md = [f"# HBM traffic per launch: {a.name}", "",
          "Commands (as for the kernel trace, one counter per run):", "", "```",
          a.command_fetch, a.command_write, a.command_cal_fetch, a.command_cal_write, "```", "",
          f"Calibration: a {COPY_BYTES >> 20} MiB device copy reads FETCH_SIZE = {fkib:.0f} KiB and "
          f"writes WRITE_SIZE = {wkib:.0f} KiB, so FETCH_SIZE is scaled by {fs:.3f} and "
          f"WRITE_SIZE by {ws:.3f}.  Per kernel: the dispatches of its largest grid, averaged.", "",
          "| Kernel | grid | FETCH_SIZE KiB | WRITE_SIZE KiB | read GB | write GB |",
          "|---|---|---|---|---|---|"]
    # every grid size of every kernel, for lookups of one specific launch
    by_grid = {}
    ambiguous = set()
    for (k, g), vs in fc.items():
        if (k, g) not in wc:
            continue
        e = {"read_GB": sum(vs) / len(vs) * 1024 * fs / 1e9,
             "write_GB": sum(wc[(k, g)]) / len(wc[(k, g)]) * 1024 * ws / 1e9}
        by_grid.setdefault(k, {})[str(g)] = e
        sk = simple_key(k)
        if sk:
            # under the short name a grid two variants share is ambiguous: dropped
            # (the full name still has it)
            seen = by_grid.setdefault(sk, {})
            if str(g) in seen or (sk, str(g)) in ambiguous:
                seen.pop(str(g), None)
                ambiguous.add((sk, str(g)))
            else:
                seen[str(g)] = e
    for k, (g, f, w) in sorted(best.items(), key=lambda x: -(x[1][1] * fs + x[1][2] * ws)):
        e = {"grid": g, "read_GB": f * 1024 * fs / 1e9, "write_GB": w * 1024 * ws / 1e9}
        kern[k] = dict(e, by_grid=by_grid.get(k, {}))
        sk = simple_key(k)
        if sk and (sk not in kern or kern[sk]["grid"] < g):
            kern[sk] = dict(e, by_grid=by_grid.get(sk, {}))
        md.append(f"| `{k}` | {g} | {f:.0f} | {w:.0f} | {e['read_GB']:.3f} | {e['write_GB']:.3f} |")
    json.dump({"workload": a.name, "fetch_scale": fs, "write_scale": ws, "kernels": kern},
              open(a.out + "_traffic.json", "w"), indent=1)
    open(a.out + "_traffic.md", "w").write("\n".join(md) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "summarize"])
    ap.add_argument("--name", default="")
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--cal-fetch")
    ap.add_argument("--cal-write")
    ap.add_argument("--command", default="")
    ap.add_argument("--command-fetch", default="")
    ap.add_argument("--command-write", default="")
    ap.add_argument("--command-cal-fetch", default="")
    ap.add_argument("--command-cal-write", default="")
    ap.add_argument("--mfma")
    ap.add_argument("--command-mfma", default="")
    ap.add_argument("--out")
    a = ap.parse_args()
    if a.what == "copy":
        copy()
    else:
        summarize(a)


if __name__ == "__main__":
    sys.exit(main())
