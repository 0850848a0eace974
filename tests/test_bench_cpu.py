"""Host logic of bench.py's queue-kernel roofline table (no GPU): kernel names of a rocprofv3
summary reduced to their short form, kernel groups summed, and achieved bandwidth computed as
algorithmic bytes over summed kernel time."""
import contextlib
import io
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_short_kernel_names():
    assert bench._short_kernel("void gpq::(anonymous namespace)::merge_tasks_kernel<256>(gpq::MergeTask const*)") \
        == "merge_tasks_kernel<256>"
    assert bench._short_kernel("gpq::(anonymous namespace)::select_scan_kernel(unsigned int const*, unsigned int*)") \
        == "select_scan_kernel"
    assert bench._short_kernel("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"
    assert bench._short_kernel("void gpq::(anonymous namespace)::phold_step_kernel<gpq::GaqView>(gpq::GaqView, "
                                "gpq::(anonymous namespace)::PholdDev)") == "phold_step_kernel<gpq::GaqView>"


def test_roofline_table_joins_time_and_bytes(tmp_path):
    stats = tmp_path / "kt_kernel_stats.csv"
    stats.write_text(
        '"Name","Calls","TotalDurationNs","AverageNs","Percentage","MinNs","MaxNs","StdDev"\n'
        '"void gpq::(anonymous namespace)::merge_tasks_kernel<256>(gpq::MergeTask const*)",10,2000000,200000,50,1,1,0\n'
        '"void gpq::(anonymous namespace)::tile_sort_kernel<256>(gpq::Event const*, gpq::Event*, unsigned int const*)"'
        ',10,500000,50000,12,1,1,0\n'
        '"void gpq::(anonymous namespace)::merge_pass_kernel<256>(gpq::Event const*)",20,500000,25000,12,1,1,0\n'
        '"gpq::(anonymous namespace)::select_count_kernel(gpq::Event const* const*)",10,1000000,100000,25,1,1,0\n'
        '"gpq::(anonymous namespace)::gaq_plan_kernel(int)",10,100,10,0,1,1,0\n')
    counts = tmp_path / "counts.json"
    groups = {"sort": {"alg_bytes": 3.2e9, "kernels": ["tile_sort_kernel", "merge_pass_kernel"], "model": "s"},
              "merge": {"alg_bytes": 8.0e9, "kernels": ["merge_tasks_kernel"], "model": "m"},
              "select": {"alg_bytes": 1.0e9, "kernels": ["select_"], "model": "c"}}
    counts.write_text("[bench] log line\n" + json.dumps({"queue_roofline": "gaq", "groups": groups}) + "\n")
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        bench.roofline_table(str(stats), str(counts))
    rows = {ln.split("|")[1].strip(): [c.strip() for c in ln.split("|")[1:-1]]
            for ln in buf.getvalue().splitlines() if ln.startswith("| ") and not ln.startswith("| group")}
    # merge: 8e9 B over 2 ms = 4000 GB/s = 50% of 8 TB/s
    assert rows["merge"][2] == "10" and rows["merge"][6] == "4000" and rows["merge"][7] == "50.0%"
    # sort: tile sort and merge passes together, 3.2e9 B over 1 ms
    assert rows["sort"][2] == "30" and rows["sort"][6] == "3200"
    assert rows["select"][6] == "1000"
    assert "gaq_plan_kernel" in buf.getvalue()
