"""bench.py's reading of a rocprofv3 counter-collection file (the in-run
counter pass): per-kernel averages over dispatches, bytes from the 32-byte
request counters, the L2 hit rate and the per-particle-step instruction
counts, on a synthetic file with known values.  No GPU."""
import csv
import sys
from pathlib import Path

import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import bench  # noqa: E402

COLS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
        "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
        "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp",
        "End_Timestamp"]


def _write(path, rows):
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(COLS)
        for did, name, counter, value in rows:
            w.writerow([did, did, "Agent 1", 1, 1, 1, 256, 1, name, 256, 0, 0, 8, 0, 16, counter,
                        float(value), 0, 1])


def test_counter_file_summary(tmp_path):
    walk = "void rp::vwalk_modal_kernel<true>(rp::DGrid const*, rp::StepArgs)"
    adv = "void rp::advect_kernel<rp::Cur, true>(rp::DGrid const*, rp::StepArgs)"
    other = "__amd_rocclr_fillBufferAligned"
    per = {walk: dict(TCC_HIT_sum=600, TCC_MISS_sum=400, TCC_EA0_RDREQ_DRAM_32B_sum=1000,
                      TCC_EA0_WRREQ_WRITE_DRAM_32B_sum=500, SQ_INSTS_VALU=300, SQ_INSTS_VALU_FMA_F64=50,
                      SQ_INSTS_VALU_MUL_F64=20, SQ_INSTS_VALU_ADD_F64=10, SQ_INSTS_MFMA=16),
           adv: dict(TCC_HIT_sum=300, TCC_MISS_sum=100, TCC_EA0_RDREQ_DRAM_32B_sum=200,
                     TCC_EA0_WRREQ_WRITE_DRAM_32B_sum=100, SQ_INSTS_VALU=100, SQ_INSTS_VALU_FMA_F64=10,
                     SQ_INSTS_VALU_MUL_F64=0, SQ_INSTS_VALU_ADD_F64=4, SQ_INSTS_MFMA=0)}
    rows = []
    did = 0
    for rep in range(3):                      # three dispatches of each, the same counts each time
        for name, cs in per.items():
            did += 1
            rows += [(did, name, c, v) for c, v in cs.items()]
    did += 1
    rows += [(did, other, c, 1e9) for c in per[walk]]   # not a step kernel: ignored
    f = tmp_path / "pmc_counter_collection.csv"
    _write(f, rows)
    n = 128.0                                  # active particles per dispatch
    got = bench.summarize_counters(f, n)
    w, a = got["by_kernel"]["vwalk_modal_kernel"], got["by_kernel"]["advect_kernel"]
    assert w["dispatches"] == 3 and a["dispatches"] == 3
    assert w["read_bytes_per_dispatch"] == 32.0 * 1000 and w["write_bytes_per_dispatch"] == 32.0 * 500
    assert got["bytes_per_launch"] == 32.0 * (1000 + 500 + 200 + 100)
    assert got["bytes_per_particle_step"] == pytest.approx(32.0 * 1800 / n)
    assert got["l2_hit_rate"] == pytest.approx(900 / 1400)
    assert w["valu_lane_instr_per_particle_step"] == pytest.approx(64 * 300 / n)
    assert got["valu_lane_instr_per_particle_step"] == pytest.approx(64 * 400 / n)
    assert got["fp64_flop_per_particle_step"] == pytest.approx(64 * (2 * 50 + 20 + 10 + 2 * 10 + 4) / n)
    assert got["mfma_flop_per_particle_step"] == pytest.approx(16 / n * 4096)
    assert "fillBuffer" not in str(got["by_kernel"])


def test_counter_file_without_step_kernels_is_an_error(tmp_path):
    f = tmp_path / "pmc_counter_collection.csv"
    _write(f, [(1, "__amd_rocclr_copyBuffer", "TCC_HIT_sum", 5)])
    with pytest.raises(ValueError):
        bench.summarize_counters(f, 10.0)
