"""Summarises rocprofv3 CSV output of bench.py into a short Markdown table and
a JSON of calibrated HBM traffic per kernel dispatch.

Inputs: one ``--kernel-trace --stats`` run (``<trace>/<prefix>_kernel_stats.csv``)
and two PMC runs, one with FETCH_SIZE and one with WRITE_SIZE
(``<dir>/<prefix>_counter_collection.csv``).  FETCH_SIZE / WRITE_SIZE (KiB per
dispatch) are scaled by what they read for ``k_copy``, which reads and writes
exactly neq doubles.  An optional third PMC run (``--sq``) with SQ_INSTS_VALU
and SQ_WAVES gives the VALU issue load: a wave64 vector instruction holds a
16-lane SIMD for 4 cycles, so busy = 4 * SQ_INSTS_VALU / (SIMDs * clock *
average duration) at the peak engine clock (``--simds``, ``--clock-ghz``).

Alternatively (``--dram-read`` / ``--dram-write``) the bytes come straight
from the L2's DRAM request counters, which count in 32-byte units (a 64-byte
request counts 2, a 128-byte request 4): TCC_EA0_RDREQ_DRAM_32B_sum and
TCC_EA0_WRREQ_WRITE_DRAM_32B_sum, no scaling; ``k_copy`` then serves as a
check.  On gfx950 FETCH_SIZE counts a 128-byte read request as 64 bytes
(for a copy kernel whose reads are all 128-byte requests it reads exactly
64 bytes per request), so the FETCH_SIZE scaling holds only for kernels
whose reads are all 128-byte requests; the read run therefore also records
the requests by size (TCC_EA0_RDREQ_128B/64B/32B_sum).

  python -m intpointsfem_amd.rocprof_summary --name k_residual_800x800-q2 \\
      --trace /tmp/p_trace --fetch /tmp/p_fetch \\
      --write /tmp/p_write --neq 5126402 --workload "..." --command "..." [--command ...]
"""
import argparse
import collections
import csv
import json
import os


def short(name):
    return name.replace("ipf::(anonymous namespace)::", "").split("(ipf::")[0].split("(long")[0].strip()[:70]


def counters(path, cname):
    agg = collections.defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == cname:
                agg[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}


def roi_stats(trace_dir, prefix, roi):
    """Per-kernel statistics (rows shaped like rocprofv3's kernel_stats.csv)
    of the dispatches that start and end inside the first occurrence of the
    marker range `roi`; the kernel trace is streamed, not loaded."""
    with open(os.path.join(trace_dir, f"{prefix}_marker_api_trace.csv")) as f:
        rng = next(((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in csv.DictReader(f)
                    if r["Function"] == roi), None)
    if rng is None:
        raise SystemExit(f"marker range {roi!r} not found")
    agg = {}
    with open(os.path.join(trace_dir, f"{prefix}_kernel_trace.csv")) as f:
        for r in csv.DictReader(f):
            t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            if t0 < rng[0] or t1 > rng[1]:
                continue
            d = t1 - t0
            a = agg.setdefault(r["Kernel_Name"], [0, 0, d, d])
            a[0] += 1
            a[1] += d
            a[2] = min(a[2], d)
            a[3] = max(a[3], d)
    total = sum(v[1] for v in agg.values()) or 1
    return [{"Name": k, "Calls": v[0], "AverageNs": v[1] / v[0], "MinNs": v[2], "MaxNs": v[3],
             "Percentage": 100.0 * v[1] / total} for k, v in sorted(agg.items(), key=lambda kv: -kv[1][1])]


def whole_stats(trace_dir, prefix, patterns):
    """Per-kernel statistics over the whole kernel trace for the kernels whose
    short name contains one of `patterns` (kernels the run launches outside
    the timed region, e.g. the other residual paths it compares)."""
    agg = {}
    with open(os.path.join(trace_dir, f"{prefix}_kernel_trace.csv")) as f:
        for r in csv.DictReader(f):
            k = short(r["Kernel_Name"])
            if not any(p in k for p in patterns):
                continue
            d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            a = agg.setdefault(r["Kernel_Name"], [0, 0, d, d])
            a[0] += 1
            a[1] += d
            a[2] = min(a[2], d)
            a[3] = max(a[3], d)
    return [{"Name": k, "Calls": v[0], "AverageNs": v[1] / v[0], "MinNs": v[2], "MaxNs": v[3]}
            for k, v in sorted(agg.items(), key=lambda kv: -kv[1][1])]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--name", required=True)
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", default=None, help="FETCH_SIZE run (omit with --write for a trace-only summary)")
    ap.add_argument("--write", default=None)
    ap.add_argument("--dram-read", default=None,
                    help="PMC run with TCC_EA0_RDREQ_DRAM_32B_sum (and the request sizes), instead of --fetch")
    ap.add_argument("--dram-write", default=None,
                    help="PMC run with TCC_EA0_WRREQ_WRITE_DRAM_32B_sum, instead of --write")
    ap.add_argument("--prefix", default="q2")
    ap.add_argument("--neq", type=int, default=0)
    ap.add_argument("--roi", default=None,
                    help="a marker range (rocprofv3 --marker-trace): kernel statistics over the dispatches that run "
                         "inside it, from the kernel trace, instead of the whole-run statistics")
    ap.add_argument("--also", action="append", default=[],
                    help="with --roi: also list, over the whole trace, the kernels whose name contains this text")
    ap.add_argument("--calibrate-from", default=None,
                    help="a traffic JSON of this tool whose k_copy calibration to use when the run has no k_copy")
    ap.add_argument("--sq", default=None, help="PMC run with SQ_INSTS_VALU and SQ_WAVES (optional)")
    ap.add_argument("--simds", type=int, default=1024, help="SIMDs of the GPU (MI355X: 256 CUs x 4)")
    ap.add_argument("--clock-ghz", type=float, default=2.4, help="peak engine clock (MI355X: 2.4 GHz)")
    ap.add_argument("--workload", required=True)
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                  "profiles"))
    a = ap.parse_args()
    if a.roi:
        stats = roi_stats(a.trace, a.prefix, a.roi)
    else:
        with open(os.path.join(a.trace, f"{a.prefix}_kernel_stats.csv")) as f:
            stats = list(csv.DictReader(f))
    lines = [f"# {a.name}: rocprofv3 summary", "", f"Workload: {a.workload}", ""] + (
        [f"Kernel statistics over the dispatches inside the marker range \"{a.roi}\" (the timed region).", ""]
        if a.roi else []) + [
        "Commands (run from /tmp with TMPDIR=/tmp, repository at $R):", "```"] + a.command + [
        "```", "", "## Kernel statistics (kernel trace)", "",
        "| kernel | calls | average µs | min µs | max µs | share |", "|---|---|---|---|---|---|"]
    for r in stats:
        if float(r["Percentage"]) < 0.05:
            continue
        lines.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} % |")
    others = whole_stats(a.trace, a.prefix, a.also) if (a.roi and a.also) else []
    if others:
        lines += ["", "Kernels of the same run outside the timed region (whole trace): "
                  + ", ".join(f"`{p}`" for p in a.also) + ".", "",
                  "| kernel | calls | average µs | min µs | max µs |", "|---|---|---|---|---|"]
        for r in others:
            lines.append(f"| `{short(r['Name'])}` | {r['Calls']} | {r['AverageNs'] / 1e3:.1f} | "
                         f"{r['MinNs'] / 1e3:.1f} | {r['MaxNs'] / 1e3:.1f} |")
    out = None
    if a.fetch and a.write:
        fetch = counters(os.path.join(a.fetch, f"{a.prefix}_counter_collection.csv"), "FETCH_SIZE")
        write = counters(os.path.join(a.write, f"{a.prefix}_counter_collection.csv"), "WRITE_SIZE")
        copyk = next((k for k in fetch if "k_copy" in k), None)
        if copyk is not None:
            true_b = a.neq * 8
            ff = true_b / (fetch[copyk] * 1024)
            fw = true_b / (write[copyk] * 1024)
            cal = (f"Calibration on `k_copy` (reads and writes neq x 8 = {true_b / 1e6:.2f} MB): FETCH_SIZE reads "
                   f"{fetch[copyk] * 1024 / 1e6:.2f} MB -> factor {ff:.3f}; WRITE_SIZE reads "
                   f"{write[copyk] * 1024 / 1e6:.2f} MB -> factor {fw:.3f}.")
        elif a.calibrate_from:
            with open(a.calibrate_from) as f:
                c = json.load(f)["calibration"]
            ff, fw, true_b = c["fetch_factor"], c["write_factor"], c["copy_kernel_bytes"]
            cal = (f"Calibration factors from {os.path.basename(a.calibrate_from)} (its `k_copy` of "
                   f"{true_b / 1e6:.2f} MB, same GPU model and counters): FETCH_SIZE x {ff:.3f}, WRITE_SIZE x {fw:.3f}.")
        else:
            raise SystemExit("no k_copy dispatch in the PMC runs: give --calibrate-from")
        lines += ["", "## HBM traffic per dispatch (PMC)", "", cal, "",
                  "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | read MB (calibrated) | written MB (calibrated) | "
                  "total MB |", "|---|---|---|---|---|---|"]
        out = {"calibration": {"fetch_factor": ff, "write_factor": fw, "copy_kernel_bytes": true_b}, "kernels": {}}
        for k in fetch:
            if k not in write:
                continue
            rb = fetch[k] * 1024 * ff
            wb = write[k] * 1024 * fw
            if rb + wb < 1e6:
                continue
            lines.append(f"| `{k}` | {fetch[k]:.0f} | {write[k]:.0f} | {rb / 1e6:.1f} | {wb / 1e6:.1f} | "
                         f"{(rb + wb) / 1e6:.1f} |")
            out["kernels"][k] = {"read_bytes": rb, "write_bytes": wb, "total_bytes": rb + wb}
    if a.dram_read and a.dram_write:
        rf = os.path.join(a.dram_read, f"{a.prefix}_counter_collection.csv")
        rd = counters(rf, "TCC_EA0_RDREQ_DRAM_32B_sum")
        wr = counters(os.path.join(a.dram_write, f"{a.prefix}_counter_collection.csv"),
                      "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum")
        size = {n: counters(rf, f"TCC_EA0_RDREQ_{n}B_sum") for n in (128, 64, 32)}
        copyk = next((k for k in rd if "k_copy" in k), None)
        chk = ("" if copyk is None or not a.neq else
               f"Check on `k_copy` (reads and writes neq x 8 = {a.neq * 8 / 1e6:.2f} MB): the counters give "
               f"{rd[copyk] * 32 / 1e6:.2f} MB read, {wr.get(copyk, 0) * 32 / 1e6:.2f} MB written.")
        lines += ["", "## HBM traffic per dispatch (PMC)", "",
                  "Bytes = 32 x TCC_EA0_RDREQ_DRAM_32B_sum (read) and 32 x TCC_EA0_WRREQ_WRITE_DRAM_32B_sum "
                  "(written): the L2's requests to DRAM in 32-byte units, unscaled. " + chk, "",
                  "| kernel | read MB | written MB | total MB | read requests 128 B / 64 B / 32 B |",
                  "|---|---|---|---|---|"]
        out = {"calibration": {"counters": "TCC_EA0_RDREQ_DRAM_32B_sum, TCC_EA0_WRREQ_WRITE_DRAM_32B_sum",
                               "fetch_factor": 1.0, "write_factor": 1.0}, "kernels": {}}
        for k in rd:
            if k not in wr:
                continue
            rb, wb = rd[k] * 32, wr[k] * 32
            if rb + wb < 1e6:
                continue
            mix = " / ".join(f"{size[n].get(k, 0.0) / 1e6:.3f} M" for n in (128, 64, 32))
            lines.append(f"| `{k}` | {rb / 1e6:.1f} | {wb / 1e6:.1f} | {(rb + wb) / 1e6:.1f} | {mix} |")
            out["kernels"][k] = {"read_bytes": rb, "write_bytes": wb, "total_bytes": rb + wb,
                                 "read_requests": {f"{n}B": size[n].get(k, 0.0) for n in (128, 64, 32)}}
    if a.sq:
        sqf = os.path.join(a.sq, f"{a.prefix}_counter_collection.csv")
        valu = counters(sqf, "SQ_INSTS_VALU")
        waves = counters(sqf, "SQ_WAVES")
        avg_us = {short(r["Name"]): float(r["AverageNs"]) / 1e3 for r in others}
        avg_us.update({short(r["Name"]): float(r["AverageNs"]) / 1e3 for r in stats})
        lines += ["", "## VALU issue load (PMC)", "",
                  f"busy = 4 cycles x SQ_INSTS_VALU / ({a.simds} SIMDs x {a.clock_ghz} GHz x average duration).", "",
                  "| kernel | SQ_WAVES | SQ_INSTS_VALU | VALU instr / wave | average µs | VALU busy |",
                  "|---|---|---|---|---|---|"]
        if out is None:
            out = {"kernels": {}}
        for k in valu:
            if k not in avg_us or k not in waves or valu[k] < 1e5:
                continue
            busy = 4.0 * valu[k] / (a.simds * a.clock_ghz * 1e3 * avg_us[k])
            lines.append(f"| `{k}` | {waves[k]:.0f} | {valu[k]:.4g} | {valu[k] / waves[k]:.0f} | {avg_us[k]:.1f} | "
                         f"{busy:.2f} |")
            out["kernels"].setdefault(k, {}).update({"valu_insts": valu[k], "waves": waves[k], "valu_busy": busy})
    with open(os.path.join(a.out, f"{a.name}.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    if out is not None:
        with open(os.path.join(a.out, f"{a.name}_traffic.json"), "w") as f:
            json.dump(out, f, indent=1)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
