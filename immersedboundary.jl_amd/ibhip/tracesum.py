"""Summaries of rocprofv3 output for bench.py runs (the figures under profiles/).

bench.py brackets its timed loop with two one-element torch fill kernels; the dispatches between
them are the timed (graph-replayed) steps, and the contraction launches right after the second
one are its profiled eager pass.

    python -m ibhip.tracesum trace KERNEL_TRACE.csv --steps N [--eager L]
        per-kernel time per step over the timed window, and the average duration of the first
        L contraction launches of the eager pass (the launches bench.py's roofline averages)
    python -m ibhip.tracesum pmc COUNTER_COLLECTION.csv --steps N [--match gemm_eo]
        per-kernel counter averages over the timed window.  With GRBM_GUI_ACTIVE and
        SQ_VALU_MFMA_BUSY_CYCLES: the MFMA-busy fraction, busy cycles over (GRBM_GUI_ACTIVE / 8
        XCDs) x 1024 SIMDs.  With FETCH_SIZE / WRITE_SIZE (KiB): HBM bytes per dispatch, each
        scaled by what the counter reads for bench.py --calib-copy's 256 MiB copy (torch's
        multiply kernel before the timed window) against the 256 MiB it moves.
"""
import argparse
import collections
import csv
import re
import sys

XCDS, SIMDS = 8, 1024
CALIB_BYTES = (1 << 25) * 8


def short(name):
    name = re.sub(r"^void ", "", name)
    return name.split("(")[0]


def window(rows):
    """Indices (a, b) of the two one-element torch fill kernels around the timed loop."""
    def grid(r):
        return int(r.get("Grid_Size") or r.get("Grid_Size_X") or 0)

    marks = [i for i, r in enumerate(rows) if "FillFunctor" in r["Kernel_Name"] and grid(r) <= 256]
    if len(marks) < 2:
        sys.exit(f"no timed window: {len(marks)} marker fill kernels found")
    return marks[0], marks[1]


def trace(path, steps, eager, match):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    a, b = window(rows)
    win = rows[a + 1:b]
    t0 = int(win[0]["Start_Timestamp"])
    t1 = max(int(r["End_Timestamp"]) for r in win)
    busy = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in win)
    agg = collections.defaultdict(lambda: [0, 0])
    for r in win:
        k = short(r["Kernel_Name"])
        agg[k][0] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        agg[k][1] += 1
    print(f"timed window (graph replay): {len(win)} dispatches over {steps} steps; first start "
          f"to last end {(t1 - t0) / 1e3 / steps:.1f} us/step, kernels busy "
          f"{busy / 1e3 / steps:.1f} us/step\n")
    print("| us/step | calls/step | us/call | kernel |\n|---:|---:|---:|---|")
    for k, (ns, n) in sorted(agg.items(), key=lambda kv: -kv[1][0]):
        print(f"| {ns / 1e3 / steps:.1f} | {n / steps:.1f} | {ns / 1e3 / n:.1f} | `{k}` |")
    if eager:
        after = [r for r in rows[b + 1:] if match in r["Kernel_Name"]][:eager]
        per = collections.defaultdict(list)
        for r in after:
            per[short(r["Kernel_Name"])].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        tot = sum(sum(v) for v in per.values())
        print(f"\nprofiled eager pass: the first {len(after)} `{match}` launches after the window "
              f"average {tot / 1e3 / max(len(after), 1):.2f} us\n")
        print("| launches | avg us | kernel |\n|---:|---:|---|")
        for k, v in sorted(per.items(), key=lambda kv: -len(kv[1])):
            print(f"| {len(v)} | {sum(v) / 1e3 / len(v):.1f} | `{k}` |")


def pmc(path, steps, match):
    disp = collections.defaultdict(dict)
    meta = {}
    for r in csv.DictReader(open(path)):
        d = int(r["Dispatch_Id"])
        disp[d][r["Counter_Name"]] = float(r["Counter_Value"])
        meta[d] = r
    order = sorted(disp)
    rows = [meta[d] for d in order]
    a, b = window(rows)
    # the calibration copy: torch's multiply kernel before the window (bench.py --calib-copy)
    pre = [d for d in order[:a] if "MulFunctor" in meta[d]["Kernel_Name"]]
    scale = {}
    for cn in ("FETCH_SIZE", "WRITE_SIZE"):
        vals = [disp[d].get(cn, 0.0) for d in pre]
        if vals and max(vals) > 0:
            scale[cn] = CALIB_BYTES / 1024 / max(vals)
    if scale:
        print("calibration (the 256 MiB copy before the window): " +
              ", ".join(f"{k} x{v:.3f}" for k, v in scale.items()) + "\n")
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    cnt = collections.Counter()
    for d in order[a + 1:b]:
        k = short(meta[d]["Kernel_Name"])
        if match and match not in k:
            continue
        cnt[k] += 1
        for cn, v in disp[d].items():
            agg[k][cn] += v
    cols = sorted({cn for k in agg for cn in agg[k]})
    head = ["dispatches/step", "kernel"] + [f"{c}/dispatch" for c in cols]
    derived = []
    if "GRBM_GUI_ACTIVE" in cols and "SQ_VALU_MFMA_BUSY_CYCLES" in cols:
        derived.append("MFMA busy")
    if "SQ_WAVE_CYCLES" in cols and "SQ_WAIT_INST_ANY" in cols:
        derived.append("wait-inst fraction")
    if "SQ_INSTS_LDS" in cols and "SQ_LDS_BANK_CONFLICT" in cols:
        derived.append("LDS conflict cycles / LDS instr")
    if scale:
        derived.append("HBM MB/dispatch")
    print("| " + " | ".join(head + derived) + " |")
    print("|" + "---:|" * len(head + derived))
    for k, c in sorted(agg.items(), key=lambda kv: -kv[1].get("GRBM_GUI_ACTIVE", 0)):
        n = cnt[k]
        cells = [f"{n / steps:.1f}", f"`{k}`"] + [f"{c[cn] / n:.4g}" for cn in cols]
        if "MFMA busy" in derived:
            g = c["GRBM_GUI_ACTIVE"] / XCDS
            cells.append(f"{c['SQ_VALU_MFMA_BUSY_CYCLES'] / (g * SIMDS):.3f}" if g else "-")
        if "wait-inst fraction" in derived:
            cells.append(f"{c['SQ_WAIT_INST_ANY'] / max(c['SQ_WAVE_CYCLES'], 1):.3f}")
        if "LDS conflict cycles / LDS instr" in derived:
            cells.append(f"{c['SQ_LDS_BANK_CONFLICT'] / max(c['SQ_INSTS_LDS'], 1):.3f}")
        if scale:
            mb = sum(c.get(cn, 0.0) * s for cn, s in scale.items()) * 1024 / n / 1e6
            cells.append(f"{mb:.2f}")
        print("| " + " | ".join(cells) + " |")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("kind", choices=["trace", "pmc"])
    ap.add_argument("csv")
    ap.add_argument("--steps", type=int, required=True, help="timed steps inside the window")
    ap.add_argument("--eager", type=int, default=0, help="contraction launches of the eager pass")
    ap.add_argument("--match", default="", help="kernel-name filter (pmc; trace: eager pass)")
    a = ap.parse_args(argv)
    if a.kind == "trace":
        trace(a.csv, a.steps, a.eager, a.match or "gemm_eo")
    else:
        pmc(a.csv, a.steps, a.match)


if __name__ == "__main__":
    main()
