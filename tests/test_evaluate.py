"""Host logic of the evaluator-equivalent harness: the gate rules of the
reference evaluator (scripts/evaluate.py:740-755) and the sample geometry."""

from fastvid_amd.evaluate import GATE_SAMPLES, distribution, performance_failures


def _result(fmt, width, height, frames, mode, enc, dec):
    perf = {"mode": mode, "encode": {}, "decode": {}}
    key = "gpps" if mode == "throughput" else "median_ms"
    perf["encode"][key] = enc
    perf["decode"][key] = dec
    return {"format": fmt, "width": width, "height": height, "frame_count": frames,
            "performance": perf}


def test_throughput_gates():
    assert performance_failures(_result("yuv422", 3840, 2160, 24, "throughput", 2.0, 3.0)) == []
    assert performance_failures(_result("yuv422", 3840, 2160, 24, "throughput", 1.99, 3.0)) == [
        "encode throughput < 2.0 GP/s"]
    assert performance_failures(_result("rgb444", 3840, 2160, 24, "throughput", 1.5, 1.9)) == [
        "decode throughput < 2.0 GP/s"]
    # gates apply to 24-frame 4K samples only
    assert performance_failures(_result("rgb444", 3840, 2160, 23, "throughput", 0.1, 0.1)) == []


def test_latency_gates():
    assert performance_failures(_result("rgb444", 1920, 1080, 1, "latency", 0.99, 0.49)) == []
    assert performance_failures(_result("rgb444", 1920, 1080, 1, "latency", 1.0, 0.5)) == [
        "encode latency >= 1.0 ms", "decode latency >= 0.5 ms"]
    assert performance_failures(_result("yuv422", 1920, 1080, 1, "latency", 9.0, 9.0)) == []


def test_distribution_and_samples():
    d = distribution([0.003, 0.001, 0.002, 0.004] + [0.002] * 16)
    assert d["median_seconds"] == 0.002 and d["min_seconds"] == 0.001
    assert d["max_seconds"] == 0.004 and d["gate_statistic"] == "median"
    ids = {s.id: s for s in GATE_SAMPLES}
    assert ids["4k-yuv422-10-x24"].plane_shapes == ((2160, 3840), (2160, 1920), (2160, 1920))
    assert ids["4k-rgb444-10-x24"].batch_frames == 24
    assert ids["1080p-rgb444-10"].raw_bytes_per_frame == 1920 * 1080 * 3 * 2


def test_evaluator_frames_have_the_sample_shapes():
    import numpy as np

    from fastvid_amd import synthetic
    for fmt, bd in (("yuv422", 10), ("rgb444", 10), ("rgb444", 8), ("gray", 12)):
        planes = synthetic.frame(7, 65, 33, fmt, bd)
        cw = 33 if fmt == "yuv422" else 65
        shapes = [(33, 65)] + ([(33, cw)] * 2 if fmt != "gray" else [])
        assert [p.shape for p in planes] == shapes
        assert all(p.dtype == np.uint16 and int(p.max()) < (1 << bd) for p in planes)
        again = synthetic.frame(7, 65, 33, fmt, bd)
        assert all(np.array_equal(a, b) for a, b in zip(planes, again))
