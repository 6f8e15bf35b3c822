"""intpointsfem_amd.rocprof_summary on synthetic rocprofv3 CSV files: the
marker-range kernel statistics, the DRAM-counter traffic (32 bytes per
count, unscaled, only kernels present in both PMC runs, the copy-kernel
check line), the FETCH_SIZE / WRITE_SIZE form scaled on the copy kernel, and
the VALU issue fraction."""
import csv
import json
import os
import sys

import pytest

from intpointsfem_amd import rocprof_summary as rs

K = "void ipf::(anonymous namespace)::k_residual_fmt<1, 4, false, true, false>(ipf::GroupDev, void const*)"
C = "ipf::(anonymous namespace)::k_copy(long, double const*, double*)"
X = "void ipf::(anonymous namespace)::k_only_in_read_run(int)"
NEQ = 1000  # the copy kernel moves 8000 bytes each way


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def _counters(path, values):
    """values: [(kernel, counter, value)], one row per dispatch and counter."""
    _write(path, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"],
           [(i, k, c, v) for i, (k, c, v) in enumerate(values)])


def _trace(d):
    # marker range [100, 100000] ns: two dispatches of K inside (10 us, 30 us), one after it
    _write(os.path.join(d, "q2_marker_api_trace.csv"), ["Function", "Start_Timestamp", "End_Timestamp"],
           [("timed residuals", 100, 100000)])
    _write(os.path.join(d, "q2_kernel_trace.csv"), ["Kernel_Name", "Start_Timestamp", "End_Timestamp"],
           [(K, 200, 10200), (K, 20000, 50000), (K, 200000, 200500), (C, 50, 90)])
    _write(os.path.join(d, "q2_kernel_stats.csv"), ["Name", "Calls", "AverageNs", "MinNs", "MaxNs", "Percentage"],
           [(K, 3, 13500, 500, 30000, 100.0)])


def _run(monkeypatch, args):
    monkeypatch.setattr(sys, "argv", ["rocprof_summary"] + args)
    rs.main()


def test_dram_counters_marker_range_and_valu(tmp_path, monkeypatch, capsys):
    d = str(tmp_path)
    _trace(os.path.join(d, "trace"))
    _counters(os.path.join(d, "rd", "q2_counter_collection.csv"), [
        (K, "TCC_EA0_RDREQ_DRAM_32B_sum", 1e6), (K, "TCC_EA0_RDREQ_DRAM_32B_sum", 3e6),
        (K, "TCC_EA0_RDREQ_128B_sum", 5e5), (K, "TCC_EA0_RDREQ_64B_sum", 0), (K, "TCC_EA0_RDREQ_32B_sum", 0),
        (C, "TCC_EA0_RDREQ_DRAM_32B_sum", 250), (C, "TCC_EA0_RDREQ_128B_sum", 62.5),
        (X, "TCC_EA0_RDREQ_DRAM_32B_sum", 1e7)])
    _counters(os.path.join(d, "wr", "q2_counter_collection.csv"), [
        (K, "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum", 5e5), (C, "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum", 250)])
    _counters(os.path.join(d, "sq", "q2_counter_collection.csv"), [
        (K, "SQ_INSTS_VALU", 1.2e6), (K, "SQ_WAVES", 1000)])
    out = os.path.join(d, "out")
    os.makedirs(out)
    _run(monkeypatch, ["--name", "t", "--roi", "timed residuals", "--trace", os.path.join(d, "trace"),
                       "--dram-read", os.path.join(d, "rd"), "--dram-write", os.path.join(d, "wr"),
                       "--sq", os.path.join(d, "sq"), "--neq", str(NEQ), "--workload", "w", "--command", "c",
                       "--out", out])
    md = open(os.path.join(out, "t.md")).read()
    tr = json.load(open(os.path.join(out, "t_traffic.json")))["kernels"]
    short = rs.short(K)
    # marker range: the two dispatches inside, average 20 us
    assert f"| `{short}` | 2 | 20.0 | 10.0 | 30.0 | 100.0 % |" in md
    # 32 bytes per count, averaged over dispatches; written from the write run
    assert tr[short]["read_bytes"] == 32 * 2e6 and tr[short]["write_bytes"] == 32 * 5e5
    assert tr[short]["total_bytes"] == 32 * 2.5e6
    assert tr[short]["read_requests"]["128B"] == 5e5
    # a kernel missing from the write run is left out; the copy check reads 8000 bytes each way
    assert rs.short(X) not in tr
    assert "the counters give 0.01 MB read, 0.01 MB written" in md and "neq x 8 = 0.01 MB" in md
    # VALU busy = 4 * insts / (1024 SIMDs * 2.4 GHz * 20 us)
    assert tr[short]["valu_busy"] == pytest.approx(4 * 1.2e6 / (1024 * 2.4e3 * 20.0))


def test_fetch_size_scaled_on_the_copy_kernel(tmp_path, monkeypatch):
    d = str(tmp_path)
    _trace(os.path.join(d, "trace"))
    # FETCH_SIZE reads half of the copy kernel's 8000 bytes: factor 2 for every kernel
    _counters(os.path.join(d, "f", "q2_counter_collection.csv"), [(K, "FETCH_SIZE", 1000.0), (C, "FETCH_SIZE", 4000 / 1024)])
    _counters(os.path.join(d, "w", "q2_counter_collection.csv"), [(K, "WRITE_SIZE", 500.0), (C, "WRITE_SIZE", 8000 / 1024)])
    out = os.path.join(d, "out")
    os.makedirs(out)
    _run(monkeypatch, ["--name", "t", "--trace", os.path.join(d, "trace"), "--fetch", os.path.join(d, "f"),
                       "--write", os.path.join(d, "w"), "--neq", str(NEQ), "--workload", "w", "--out", out])
    tr = json.load(open(os.path.join(out, "t_traffic.json")))
    assert tr["calibration"]["fetch_factor"] == pytest.approx(2.0)
    assert tr["calibration"]["write_factor"] == pytest.approx(1.0)
    k = tr["kernels"][rs.short(K)]
    assert k["read_bytes"] == pytest.approx(2 * 1000 * 1024) and k["write_bytes"] == pytest.approx(500 * 1024)
