"""The per-launch figures bench.py quotes come from the committed rocprof
summaries: every kernel shape the benchmark can choose has its measured
HBM traffic (and, for the batched shapes the autotuner picks, its VALU
instruction count) on record, and the parsed values are the ones written."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_recorded_for_every_fused_shape():
    b = _bench()
    for shape in ("512,4", "512,2", "1024,1", "1024,2", "256,2", "256,4", "256,8", "256,16"):
        t = b.measured_traffic(f"cullFusedKernel<{shape}>")
        assert t is not None and 1.3e9 < t < 1.7e9, shape  # about 1.05x the 1.39 GB algorithmic
    assert b.measured_traffic("cullKernel<256,8>") == 9009600000
    assert b.measured_traffic("cullKernel<512,4>") == 8917600000
    assert b.measured_traffic("processDrawablesKernel<3>") == 2330500000
    assert b.measured_traffic("noSuchKernel<1,1>") is None


def test_valu_recorded_for_the_batched_shapes():
    b = _bench()
    assert b.measured_per_launch("valu_wave_instructions", "cullKernel<256,8>") == 647785236
    assert b.measured_per_launch("valu_wave_instructions", "cullKernel<512,4>") == 687555878
    v = b.valu_issue("cullKernel<256,8>", 1.81, 256)
    assert abs(v["issue_frac"] - 647785236 / (256 * 2.4e9 * 1.81e-3)) < 1e-12
    assert b.valu_issue("cullKernel<128,8>", 2.0, 256) is None
