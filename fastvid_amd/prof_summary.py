"""Summaries of rocprofv3 CSV output as short markdown tables.

    python -m fastvid_amd.prof_summary stats KERNEL_STATS_CSV --title T --cmd C
        per-kernel calls, total and average duration from --kernel-trace --stats

    python -m fastvid_amd.prof_summary traffic FETCH_CSV WRITE_CSV --copy-bytes N --title T --cmd C
        per-kernel FETCH_SIZE / WRITE_SIZE (KiB per dispatch) from two --pmc runs,
        each scaled by what the same counter reads for the largest copy kernel
        dispatch of the run, a device-to-device copy of N bytes (so the table
        reports HBM bytes per dispatch in units of a copy's bytes)

    python -m fastvid_amd.prof_summary counters PMC_CSV [PMC_CSV ...] --title T --cmd C
        per-kernel average per dispatch of every counter in one or more --pmc
        runs, with the derived VALU and LDS instructions per wave, VALU issue
        share of wave cycles and LDS bank-conflict share of LDS cycles

The traffic table ends with one machine-readable line per kernel,
``traffic: <kernel> <fetch_bytes> <write_bytes>``, which bench.py reads to fill
its roofline's ``traffic`` field.
"""

from __future__ import annotations

import argparse
import collections
import csv
import re
import sys


def short_name(name: str) -> str:
    """'void fvh::k_band<true, true, unsigned short>(...)' -> 'k_band<true,true,u16>'."""
    base = name.split("(")[0]
    base = base.replace("void ", "").replace("fvh::", "").strip()
    base = (base.replace("unsigned short", "u16").replace("unsigned int", "u32")
            .replace("unsigned char", "u8").replace(" ", ""))
    return base


def stats_table(path: str) -> list[str]:
    rows = list(csv.DictReader(open(path)))
    out = ["| kernel | calls | total ms | avg us | min us | max us | % |",
           "|---|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        out.append("| %s | %s | %.3f | %.1f | %.1f | %.1f | %s |" % (
            short_name(r["Name"]), r["Calls"], float(r["TotalDurationNs"]) / 1e6,
            float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3,
            r["Percentage"]))
    return out


def per_dispatch(path: str, counter: str):
    """{kernel: [value per dispatch]} and {dispatch: (kernel, grid)}."""
    sums = collections.defaultdict(float)
    info = {}
    for r in csv.DictReader(open(path)):
        if not r["Counter_Name"].startswith(counter):
            continue
        d = r["Dispatch_Id"]
        sums[d] += float(r["Counter_Value"])
        info[d] = (short_name(r["Kernel_Name"]), int(r["Grid_Size"]))
    by_kernel = collections.defaultdict(list)
    for d, v in sums.items():
        by_kernel[info[d][0]].append(v)
    return by_kernel, sums, info


def copy_factor(sums, info, copy_bytes: int) -> float:
    """Counter KiB of the largest copy dispatch divided by the copy's KiB."""
    copies = [(info[d][1], sums[d]) for d in sums if "copy" in info[d][0].lower()]
    if not copies:
        raise SystemExit("no copy kernel dispatch in the run")
    grid, value = max(copies)
    return value / (copy_bytes / 1024.0)


def traffic_table(fetch_csv: str, write_csv: str, copy_bytes: int) -> list[str]:
    fk, fs, fi = per_dispatch(fetch_csv, "FETCH_SIZE")
    wk, ws, wi = per_dispatch(write_csv, "WRITE_SIZE")
    ff = copy_factor(fs, fi, copy_bytes)
    wf = copy_factor(ws, wi, copy_bytes)
    out = [f"Calibration: a {copy_bytes} B device-to-device copy reads {ff:.4f} x and writes "
           f"{wf:.4f} x its size in FETCH_SIZE / WRITE_SIZE KiB; values below are divided by "
           "these factors.", "",
           "| kernel | dispatches | FETCH KiB/dispatch | WRITE KiB/dispatch | HBM MB/dispatch |",
           "|---|---:|---:|---:|---:|"]
    lines = []
    for k in sorted(set(fk) | set(wk)):
        f = sum(fk.get(k, [0])) / max(1, len(fk.get(k, [])))
        w = sum(wk.get(k, [0])) / max(1, len(wk.get(k, [])))
        fb = f * 1024 / ff
        wb = w * 1024 / wf
        out.append("| %s | %d | %.0f | %.0f | %.2f |" % (k, len(fk.get(k, [])), f, w,
                                                          (fb + wb) / 1e6))
        lines.append("traffic: %s %d %d" % (k, int(fb), int(wb)))
    return out + ["", "```"] + lines + ["```"]


def counter_table(paths, cus: int = 256, xcds: int = 8) -> list[str]:
    """Per-kernel mean per dispatch of every counter over the CSVs (passes of
    the same workload, matched by kernel name), then derived shares:
    kernel cycles = GRBM_GUI_ACTIVE / xcds (the counter sums the XCDs);
    VALU busy = SQ_INSTS_VALU / (cus * cycles) (a wave64 VALU instruction
    holds one of a CU's four SIMDs for four cycles); LDS busy =
    SQ_LDS_IDX_ACTIVE / (cus * cycles); bank-conflict share of LDS cycles;
    share of wave cycles spent waiting to issue (SQ_WAIT_INST_ANY /
    SQ_WAVE_CYCLES); L2 miss share of L2 requests."""
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    kernel_of = {}
    for path in paths:
        for r in csv.DictReader(open(path)):
            key = (path, r["Dispatch_Id"])
            kernel_of[key] = short_name(r["Kernel_Name"])
            per[key][r["Counter_Name"]] += float(r["Counter_Value"])
    by_kernel = collections.defaultdict(lambda: collections.defaultdict(list))
    for key, counters in per.items():
        for c, v in counters.items():
            by_kernel[kernel_of[key]][c].append(v)
    names = sorted({c for k in by_kernel.values() for c in k})
    out = ["| kernel | " + " | ".join(names) + " |", "|---|" + "---:|" * len(names)]
    derived = ["", f"Derived ({cus} compute units, {xcds} XCDs):", "",
               "| kernel | cycles | VALU instr / wave | VALU busy | LDS busy | LDS bank-conflict "
               "share | wave wait share | L2 miss share |",
               "|---|---:|---:|---:|---:|---:|---:|---:|"]

    def ratio(mean, num, den, scale=1.0):
        if num in mean and mean.get(den):
            return "%.3g" % (scale * mean[num] / mean[den])
        return "-"

    for k in sorted(by_kernel):
        mean = {c: sum(v) / len(v) for c, v in by_kernel[k].items()}
        out.append("| %s | " % k + " | ".join("%.4g" % mean.get(c, float("nan"))
                                              for c in names) + " |")
        if not mean.get("SQ_WAVES") or not mean.get("GRBM_GUI_ACTIVE"):
            continue
        cycles = mean["GRBM_GUI_ACTIVE"] / xcds
        mean["_cu_cycles"] = cus * cycles
        mean["_hits"] = mean.get("TCC_HIT_sum", 0) + mean.get("TCC_MISS_sum", 0)
        derived.append("| %s | %.0f | %s | %s | %s | %s | %s | %s |" % (
            k, cycles, ratio(mean, "SQ_INSTS_VALU", "SQ_WAVES"),
            ratio(mean, "SQ_INSTS_VALU", "_cu_cycles"), ratio(mean, "SQ_LDS_IDX_ACTIVE", "_cu_cycles"),
            ratio(mean, "SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE"),
            ratio(mean, "SQ_WAIT_INST_ANY", "SQ_WAVE_CYCLES"),
            ratio(mean, "TCC_MISS_sum", "_hits")))
    return out + derived


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--copy-bytes", type=int, required=True)
    c = sub.add_parser("counters")
    c.add_argument("csvs", nargs="+")
    c.add_argument("--cus", type=int, default=256)
    c.add_argument("--xcds", type=int, default=8)
    for p in (s, t, c):
        p.add_argument("--title", default="rocprofv3 summary")
        p.add_argument("--cmd", default="")
    a = ap.parse_args(argv)
    out = [f"# {a.title}", ""]
    if a.cmd:
        out += ["Command:", "", "```", a.cmd, "```", ""]
    if a.mode == "stats":
        out += stats_table(a.csv)
    elif a.mode == "traffic":
        out += traffic_table(a.fetch_csv, a.write_csv, a.copy_bytes)
    else:
        out += counter_table(a.csvs, a.cus, a.xcds)
    print("\n".join(out))
    return 0


def read_traffic(path: str) -> dict:
    """{kernel short name: (fetch_bytes, write_bytes)} from a traffic summary."""
    table = {}
    try:
        for line in open(path):
            m = re.match(r"traffic: (\S+) (\d+) (\d+)", line.strip())
            if m:
                table[m.group(1)] = (int(m.group(2)), int(m.group(3)))
    except OSError:
        pass
    return table


if __name__ == "__main__":
    sys.exit(main())
