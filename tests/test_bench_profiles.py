"""CPU: the bench line's `roofline.traffic` and `fp64` objects come from the
committed rocprof summaries (profiles/k_rhs_hbm_traffic.md and
profiles/k_rhs_fp64_valu.md).  Every workload the bench measures has its
entries there, and the ratio the bench reads is the one the summary's table
derives from its measured bytes (read + write per point over the
algorithmic bytes per point), so that an edit of a summary cannot leave the
bench reporting a figure the table does not hold."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

WORKLOADS = ("batched_scalar", "brill", "plain_512x1024")


def _table_rows(path):
    rows = {}
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0].split(" ")[0] in WORKLOADS:
            rows[cells[0].split(" ")[0]] = [float(c) for c in cells[2:]]
    return rows


def test_traffic_ratios_are_the_tables():
    import bench
    rows = _table_rows(os.path.join(ROOT, "profiles", "k_rhs_hbm_traffic.md"))
    assert sorted(rows) == sorted(WORKLOADS)
    for w in WORKLOADS:
        read, write, traffic, alg, ratio = rows[w]
        assert abs(read + write - traffic) <= 0.11, w
        assert abs(traffic / alg - ratio) <= 5e-4, w
        got = bench.measured_traffic(w, 1.0e6)
        assert got is not None and abs(got / 1.0e6 - ratio) < 1e-12, (w, got)
        assert 1.0 <= ratio < 2.0, w
    assert bench.measured_traffic("no_such_workload", 1.0) is None


def test_fp64_rates_come_from_the_counter_summary():
    import bench
    for w in ("batched_scalar", "plain_512x1024"):
        f = bench.fp64_roofline(w, 1.0e6, 1.0)
        assert f is not None and f["flop_per_point"] > 1000.0
        assert abs(f["achieved"] - f["flop_per_point"] * 1.0e6 / 1.0e-3 / 1.0e12) < 1e-9
        assert abs(f["frac"] - f["achieved"] / f["peak"]) < 1e-15
    assert bench.fp64_roofline("batched_scalar", 1.0e6, 0.0) is None
