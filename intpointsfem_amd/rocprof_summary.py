"""Summarises rocprofv3 CSV output of bench.py into a short Markdown table and
a JSON of calibrated HBM traffic per kernel dispatch.

Inputs: one ``--kernel-trace --stats`` run (``<trace>/<prefix>_kernel_stats.csv``)
and two PMC runs, one with FETCH_SIZE and one with WRITE_SIZE
(``<dir>/<prefix>_counter_collection.csv``).  FETCH_SIZE / WRITE_SIZE (KiB per
dispatch) are scaled by what they read for ``k_copy``, which reads and writes
exactly neq doubles.

  python -m intpointsfem_amd.rocprof_summary --name k_residual_800x800-q2 \\
      --trace gpurun_out/prof_trace --fetch gpurun_out/prof_fetch \\
      --write gpurun_out/prof_write --neq 5126402 --workload "..." --command "..." [--command ...]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--name", required=True)
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", default=None, help="FETCH_SIZE run (omit with --write for a trace-only summary)")
    ap.add_argument("--write", default=None)
    ap.add_argument("--prefix", default="q2")
    ap.add_argument("--neq", type=int, default=0)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                  "profiles"))
    a = ap.parse_args()
    with open(os.path.join(a.trace, f"{a.prefix}_kernel_stats.csv")) as f:
        stats = list(csv.DictReader(f))
    lines = [f"# {a.name}: rocprofv3 summary", "", f"Workload: {a.workload}", "",
             "Commands (run from /tmp with TMPDIR=/tmp, repository at $R):", "```"] + a.command + [
        "```", "", "## Kernel statistics (kernel trace)", "",
        "| kernel | calls | average µs | min µs | max µs | share |", "|---|---|---|---|---|---|"]
    for r in stats:
        if float(r["Percentage"]) < 0.05:
            continue
        lines.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} % |")
    out = None
    if a.fetch and a.write:
        fetch = counters(os.path.join(a.fetch, f"{a.prefix}_counter_collection.csv"), "FETCH_SIZE")
        write = counters(os.path.join(a.write, f"{a.prefix}_counter_collection.csv"), "WRITE_SIZE")
        copyk = next(k for k in fetch if "k_copy" in k)
        true_b = a.neq * 8
        ff = true_b / (fetch[copyk] * 1024)
        fw = true_b / (write[copyk] * 1024)
        lines += ["", "## HBM traffic per dispatch (PMC)", "",
                  f"Calibration on `k_copy` (reads and writes neq x 8 = {true_b / 1e6:.2f} MB): FETCH_SIZE reads "
                  f"{fetch[copyk] * 1024 / 1e6:.2f} MB -> factor {ff:.3f}; WRITE_SIZE reads "
                  f"{write[copyk] * 1024 / 1e6:.2f} MB -> factor {fw:.3f}.", "",
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
    with open(os.path.join(a.out, f"{a.name}.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    if out is not None:
        with open(os.path.join(a.out, f"{a.name}_traffic.json"), "w") as f:
            json.dump(out, f, indent=1)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
