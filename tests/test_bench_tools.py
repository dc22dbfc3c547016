"""Host-side pieces of bench.py: the profiler-counter summary that turns
rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs into HBM bytes per stage."""
import csv
import json
import os

import pytest

import bench

HEADER = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
          "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
          "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def _write_run(d, rows, counter):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as fh:
        w = csv.writer(fh, quoting=csv.QUOTE_ALL)
        w.writerow(HEADER)
        for k, (name, value) in enumerate(rows):
            w.writerow([k, k, "Agent 1", 1, 1, 1, 256, 1, name, 256, 0, 0, 8, 0, 16, counter, value, 0, 1])


def test_pmc_summary_scales_by_the_copy_probe(tmp_path):
    probe_kib = bench.COPY_PROBE_BYTES / 1024
    # the counters read half the probe's bytes on fetch and all of them on write
    _write_run(tmp_path / "cf", [("fill", 0.5), ("copy_kernel", probe_kib / 2)], "FETCH_SIZE")
    _write_run(tmp_path / "cw", [("fill", probe_kib), ("copy_kernel", probe_kib)], "WRITE_SIZE")
    # two steps: k_update once per step, three bounding kernels, one FSD kernel
    run_f = [("svp::k_update(svp::DevMesh)", 100.0), ("svp::k_update(svp::DevMesh)", 100.0),
             ("void svp::k_bnd_flow<8>(svp::DevMesh)", 10.0), ("void svp::k_bnd_apply<8>(x)", 30.0),
             ("svp::k_fsd_group(x)", 7.0)]
    run_w = [("svp::k_update(svp::DevMesh)", 50.0), ("svp::k_update(svp::DevMesh)", 50.0),
             ("void svp::k_bnd_flow<8>(svp::DevMesh)", 4.0), ("void svp::k_bnd_apply<8>(x)", 2.0),
             ("svp::k_fsd_group(x)", 1.0)]
    _write_run(tmp_path / "rf", run_f, "FETCH_SIZE")
    _write_run(tmp_path / "rw", run_w, "WRITE_SIZE")
    out = bench.pmc_summary(str(tmp_path / "cf"), str(tmp_path / "cw"), str(tmp_path / "rf"), str(tmp_path / "rw"))
    assert out["fetch_factor"] == pytest.approx(2.0)
    assert out["write_factor"] == pytest.approx(1.0)
    assert out["steps"] == 2
    b = out["stages"]["bounding"]
    assert b["fetch_bytes"] == pytest.approx(2.0 * 1024 * 40.0 / 2)
    assert b["write_bytes"] == pytest.approx(1.0 * 1024 * 6.0 / 2)
    assert b["bytes"] == pytest.approx(b["fetch_bytes"] + b["write_bytes"])
    u = out["stages"]["update"]
    assert u["bytes"] == pytest.approx(2.0 * 1024 * 100.0 + 1024 * 50.0)
    assert out["stages"]["findSignedDistance"]["bytes"] == pytest.approx((2.0 * 7.0 + 1.0) * 1024 / 2)
    json.dumps(out)


def test_pmc_summary_counts_steps_by_the_templated_update(tmp_path):
    """The update kernel is instantiated per cell size (k_update<6> on
    hexahedra): one launch per step still numbers the steps, and no other
    kernel whose name merely contains k_update counts."""
    probe_kib = bench.COPY_PROBE_BYTES / 1024
    _write_run(tmp_path / "cf", [("copy_kernel", probe_kib)], "FETCH_SIZE")
    _write_run(tmp_path / "cw", [("copy_kernel", probe_kib)], "WRITE_SIZE")
    rows = [("void svp::k_update<6>(svp::DevMesh, svp::DevState)", 8.0)] * 3
    rows += [("svp::k_update_extra(svp::DevMesh)", 1.0)] * 7
    _write_run(tmp_path / "rf", rows, "FETCH_SIZE")
    _write_run(tmp_path / "rw", rows, "WRITE_SIZE")
    out = bench.pmc_summary(str(tmp_path / "cf"), str(tmp_path / "cw"), str(tmp_path / "rf"), str(tmp_path / "rw"))
    assert out["steps"] == 3
    assert out["stages"]["update"]["fetch_bytes"] == pytest.approx(1024 * (3 * 8.0 + 7 * 1.0) / 3)


def test_traffic_of_reads_the_committed_profile():
    path = os.path.join(bench.ROOT, "profiles", "hbm_traffic_hex128p.json")
    if not os.path.exists(path):
        pytest.skip("no committed traffic profile")
    t = bench.traffic_of("hex128p", "bounding")
    assert t is not None and t > 0
    assert bench.traffic_of("no-such-workload", "bounding") is None
