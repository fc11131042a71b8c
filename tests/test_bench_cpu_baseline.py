"""CPU: the cpu_baseline leg of bench.py (the oracle timed on the host),
one thread and several threads each stepping an instance of its own."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def test_cpu_baseline_one_and_several_threads(oracle_mod):
    import bench
    from oaxis import workloads
    texts = [workloads.schwarzschild(1.0, 32, 64, 0), workloads.schwarzschild(1.5, 32, 64, 0)]
    v1, s1, n1 = bench.cpu_baseline(texts, 0.3, 1)
    assert v1 > 0.0 and s1 >= 0.3 and n1 >= 1
    v2, s2, n2 = bench.cpu_baseline(texts, 0.3, 2)
    assert v2 > 0.0 and s2 >= 0.3 and n2 >= 2       # each thread steps at least once
    assert 1 <= bench.host_threads() <= 16
    rec = bench.cpu_baseline_record(texts, 0.2, "two Schwarzschild instances")
    assert rec["kind"] == "port" and rec["cores"] == bench.host_threads()
    assert rec["value"] > 0.0 and rec["one_core"]["value"] > 0.0


def test_cpu_baseline_reports_a_setup_error_instead_of_waiting(oracle_mod):
    """A worker thread whose instance cannot be built releases the others
    (they start together) and the error reaches the caller."""
    import pytest
    import bench
    import oaxis
    from oaxis import workloads
    texts = [workloads.schwarzschild(1.0, 32, 64, 0), "no_such_key = 1\n"]
    with pytest.raises(oaxis.OaxisError):
        bench.cpu_baseline(texts, 0.2, 2)
