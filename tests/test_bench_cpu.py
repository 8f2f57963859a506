"""Host logic of bench.py: the roofline fields it reads from the committed profiles (the copy-calibrated
traffic table and the SQ counter table) are found under the forward kernel's current instance name."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_committed_traffic_and_counters_cover_the_forward_kernel():
    bench = load_bench()
    name = bench.FWD_KERNEL
    traffic = bench.measured_traffic(name)
    # at least the bins written (1.2 GB per launch), well under 10x the algorithmic bytes
    assert traffic is not None and 1.0e9 < traffic < 1.2e10
    insts = bench.measured_counter(name, "SQ_INSTS_LDS")
    assert insts is not None and insts > 0
    lds = bench.lds_roofline(2.5)
    assert lds is not None and 0.0 < lds["frac"] < 1.0
    assert abs(lds["achieved"] - insts * 256.0 / 2.5e-3 / 1e9) < 1e-6 * lds["achieved"]
    valu = bench.measured_counter(name, "SQ_INSTS_VALU")
    issue = bench.valu_issue(2.5)
    assert issue is not None and 0.0 < issue["frac"] < 1.0
    assert abs(issue["achieved"] - valu / 2.5e-3 / 1e9) < 1e-6 * issue["achieved"]


def test_committed_tables_are_from_the_current_kernels():
    """The tables record the SHA-256 of the fast kernels' source they were measured with."""
    bench = load_bench()
    assert bench.table_current("traffic_panel128.json")
    assert bench.table_current("counters_panel128.json")


def test_unknown_kernels_give_no_figures():
    bench = load_bench()
    assert bench.measured_traffic("no_such_kernel") is None
    assert bench.measured_counter("no_such_kernel", "SQ_INSTS_LDS") is None
