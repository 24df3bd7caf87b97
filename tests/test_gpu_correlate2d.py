"""correlate2d (direct) on the HIP library: bit-identical to the oracle (whose
restatement reproduces the reference's correlate2d digests) in every element
type, mode and boundary, and to the reference's recorded digests on its own
probe inputs."""
import hashlib
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

NP = {"FP32": np.float32, "FP16": np.float16, "INT32": np.int32, "INT16": np.int16, "INT8": np.int8}
HERE = os.path.dirname(os.path.abspath(__file__))


def _inputs(name, shape, seed):
    rng = np.random.default_rng(seed)
    if orc.INTEGRAL[name]:
        hi = {"INT32": 1 << 20, "INT16": 3000, "INT8": 100}[name]
        return rng.integers(-hi, hi, shape).astype(NP[name])
    return (rng.standard_normal(shape) * 3).astype(NP[name])


def _dominates(s1, s2):
    return (s1[0] >= s2[0] and s1[1] >= s2[1]) or (s2[0] >= s1[0] and s2[1] >= s1[1])


@pytest.mark.parametrize("name", list(NP))
@pytest.mark.parametrize("s1,s2", [((13, 17), (3, 5)), ((40, 100), (7, 9)), ((70, 130), (1, 1)), ((5, 4), (8, 9)),
                                   ((33, 65), (4, 6)), ((20, 30), (60, 70)), ((50, 600), (5, 7))])
@pytest.mark.parametrize("mode", ["full", "same", "valid"])
@pytest.mark.parametrize("boundary", ["fill", "wrap", "symm"])
def test_correlate2d_bit_identical(name, s1, s2, mode, boundary):
    """The three dispatch forms: the wide kernel (8 x 256 outputs per block,
    2 x 4 per thread, window in at most 48 KiB of LDS) with partial last
    tiles and several column blocks, the 16 x 64 tile kernel for kernels
    whose wide window exceeds that budget (60 x 70), and the per-term direct
    form beyond both; in2 larger than in1, even kernel sizes ('same' origin
    r2/2, c2/2), a nonzero fill value."""
    from zkx_dev_amd import convolution
    if mode == "valid" and not _dominates(s1, s2):
        pytest.skip("valid needs one input to dominate")
    print(f"correlate2d {name} {s1} {s2} {mode} {boundary}", flush=True)
    a, b = _inputs(name, s1, sum(s1)), _inputs(name, s2, sum(s2) + 1)
    fill = -3.6 if orc.INTEGRAL[name] else 0.75
    y = convolution.correlate2d(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode, boundary,
                                fill).cpu().numpy()
    ref = orc.correlate2d(a, b, mode, boundary, fill, name)
    assert y.dtype == ref.dtype and y.shape == ref.shape
    assert np.array_equal(y.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("name", list(NP))
def test_correlate2d_large_kernel_direct_path(name):
    """in2 of 100 x 200: its staged window would exceed the LDS tile, so the
    per-term global form runs."""
    from zkx_dev_amd import convolution
    a, b = _inputs(name, (30, 30), 5), _inputs(name, (100, 200), 6)
    for boundary in ("fill", "symm"):
        y = convolution.correlate2d(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), "same", boundary,
                                    2).cpu().numpy()
        ref = orc.correlate2d(a, b, "same", boundary, 2, name)
        assert np.array_equal(y.view(np.uint8), ref.view(np.uint8))


def test_correlate2d_reference_probe_digests():
    """The reference's recorded correlate2d probe cases: the device output's
    sha256 equals the reference CPU (and GPU) output digest."""
    from zkx_dev_amd import convolution
    cases = [c for c in json.load(open(os.path.join(HERE, "golden", "probe_digests.json"), encoding="utf-8"))
             if c["op"] == "correlate2d"]
    assert len(cases) == 15
    for c in cases:
        a = c["args"]
        dt = a["dtype"]
        r1, c1, r2, c2 = (int(a[k]) for k in ("rows1", "cols1", "rows2", "cols2"))
        x = orc.correlate2d_probe_input(r1 * c1, int(a["seed"]), 3, dt).reshape(r1, c1)
        h = orc.correlate2d_probe_input(r2 * c2, int(a["seed"]), 7, dt).reshape(r2, c2)
        y = convolution.correlate2d(torch.from_numpy(x).cuda(), torch.from_numpy(h).cuda(), a["mode"],
                                    a["boundary"], float(a["fillvalue"])).cpu().numpy()
        assert hashlib.sha256(np.ascontiguousarray(y).tobytes()).hexdigest() == c["records"][0]["cpu_output_digest"], a


def test_correlate2d_validation():
    from zkx_dev_amd import convolution
    a = torch.ones((8, 8), device="cuda")
    with pytest.raises(ValueError):
        convolution.correlate2d(a, a[:3, :3], "middle")
    with pytest.raises(ValueError):
        convolution.correlate2d(a, a[:3, :3], "same", "reflect")
    with pytest.raises(ValueError):
        convolution.correlate2d(a[:4, :8], a[:8, :4], "valid")  # neither dominates
    with pytest.raises(ValueError):
        convolution.correlate2d(a.to(torch.int16), a[:3, :3].to(torch.int16), "same", "fill", float("nan"))
    with pytest.raises(ValueError):
        convolution.correlate2d(a, a[:3, :3], "same", out=a)  # output aliases an input
    with pytest.raises(ValueError):
        convolution.correlate2d(a.reshape(-1), a[:3, :3])
    assert tuple(convolution.correlate2d(a[:0], a[:3, :3]).shape) == (0, 0)
    # the integer NaN fill is ignored where no fill is used
    y = convolution.correlate2d(a.to(torch.int16), a[:3, :3].to(torch.int16), "same", "wrap", float("nan"))
    assert int(y.sum()) == 9 * 64
