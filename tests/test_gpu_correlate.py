"""correlate (1-D, direct) on the HIP library against the oracle, whose
restatement reproduces the reference's correlate digests bit for bit: the
device result must be bit-identical for every element type and mode."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

NP = {"FP32": np.float32, "FP16": np.float16, "INT32": np.int32, "INT16": np.int16, "INT8": np.int8}


def _inputs(name, n, seed):
    rng = np.random.default_rng(seed)
    if orc.INTEGRAL[name]:
        hi = {"INT32": 1 << 20, "INT16": 3000, "INT8": 100}[name]
        return rng.integers(-hi, hi, n).astype(NP[name])
    return (rng.standard_normal(n) * 3).astype(NP[name])


@pytest.mark.parametrize("name", list(NP))
@pytest.mark.parametrize("l1,l2", [(1000, 31), (128, 17), (31, 1000), (1, 1), (5, 5), (4097, 255), (3, 40),
                                   (20000, 700), (9000, 256), (70000, 3), (6000, 1)])
@pytest.mark.parametrize("mode", ["full", "same", "valid"])
def test_correlate_bit_identical(name, l1, l2, mode):
    """Shapes that put blocks of 2048 outputs wholly inside in1 (LDS form,
    in2 staged 256 taps at a time: 700 = 2 x 256 + 188) beside edge blocks."""
    from zkx_dev_amd import convolution
    print(f"correlate {name} l1={l1} l2={l2} {mode}", flush=True)
    a, b = _inputs(name, l1, l1), _inputs(name, l2, l2 + 1)
    y = convolution.correlate(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode).cpu().numpy()
    ref = orc.correlate(a, b, mode, name)
    assert y.dtype == ref.dtype and y.shape == ref.shape
    assert np.array_equal(y.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("name", list(NP))
def test_correlate_probe_cases(name):
    """The reference probe's signals ('same', direct) for counts 128/1024/10000."""
    from zkx_dev_amd import convolution
    for n, m in ((128, 17), (1024, 31), (10000, 31)):
        x = orc.correlate_probe_signal(n, 41, name)
        r = orc.correlate_probe_signal(m, 7, name)
        y = convolution.correlate(torch.from_numpy(x).cuda(), torch.from_numpy(r).cuda(), "same").cpu().numpy()
        assert np.array_equal(y.view(np.uint8), orc.correlate(x, r, "same", name).view(np.uint8))


def test_correlate_float_vs_numpy():
    """FP32 results against numpy's FP64 correlate (mode names as numpy)."""
    from zkx_dev_amd import convolution
    a, b = _inputs("FP32", 3000, 1), _inputs("FP32", 77, 2)
    for mode in ("full", "same", "valid"):
        y = convolution.correlate(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode).cpu().numpy()
        ref = np.correlate(a.astype(np.float64), b.astype(np.float64), mode)
        assert orc.rel_l2(y, ref) < 1e-6


def test_correlate_validation():
    from zkx_dev_amd import convolution
    a = torch.ones(16, device="cuda")
    with pytest.raises(ValueError):
        convolution.correlate(a, a[:4], "middle")
    with pytest.raises(ValueError):
        convolution.correlate(a, torch.ones(4, dtype=torch.int32, device="cuda"))
    with pytest.raises(ValueError):
        convolution.correlate(a, a[:4], "same", out=a)  # output aliases an input
    assert convolution.correlate(a[:0], a[:4]).numel() == 0


def _small_ints(name, n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(-9, 10, n).astype(NP[name])


@pytest.mark.parametrize("name", list(NP))
@pytest.mark.parametrize("l1,l2", [(1000, 31), (31, 1000), (20000, 700), (9000, 4096), (1, 1), (7, 7)])
@pytest.mark.parametrize("mode", ["full", "same", "valid"])
def test_correlate_fft_method(name, l1, l2, mode):
    """method='fft' (complex FP32 FFT, the reference's workspace overload)
    against the direct sums of the oracle: floats within FP32 FFT round-off;
    integers with sums well inside FP32 precision are exact after rounding,
    including INT8's wrap modulo 2^8."""
    from zkx_dev_amd import convolution
    print(f"correlate fft {name} l1={l1} l2={l2} {mode}", flush=True)
    if orc.INTEGRAL[name]:
        a, b = _small_ints(name, l1, l1), _small_ints(name, l2, l2 + 3)
    else:
        a, b = _inputs(name, l1, l1), _inputs(name, l2, l2 + 1)
    y = convolution.correlate(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode, "fft").cpu().numpy()
    ref = orc.correlate(a, b, mode, name)
    assert y.dtype == ref.dtype and y.shape == ref.shape
    if orc.INTEGRAL[name]:
        assert np.array_equal(y, ref)
    else:
        tol = 1e-5 if name == "FP32" else 2e-3  # FP16 output rounding
        assert orc.rel_l2(y.astype(np.float64), ref.astype(np.float64)) < tol


@pytest.mark.parametrize("mode", ["full", "same", "valid"])
def test_correlate_fft_both_long(mode):
    """Both inputs longer than 4096 samples: one whole-length transform pair
    (the reference pads to the next power of two and takes any length)."""
    from zkx_dev_amd import convolution
    a, b = _inputs("FP32", 30000, 5), _inputs("FP32", 5001, 6)
    y = convolution.correlate(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode, "fft").cpu().numpy()
    ref = np.correlate(a.astype(np.float64), b.astype(np.float64), mode)
    assert orc.rel_l2(y.astype(np.float64), ref) < 1e-5
    ai, bi = _small_ints("INT16", 9000, 7), _small_ints("INT16", 4500, 8)
    yi = convolution.correlate(torch.from_numpy(ai).cuda(), torch.from_numpy(bi).cuda(), mode, "fft").cpu().numpy()
    assert np.array_equal(yi, orc.correlate(ai, bi, mode, "INT16"))


def test_correlate_fft_two_streams_one_thread():
    """Two fft-method calls from one host thread on two streams, back to back:
    the per-thread workspace is handed from one stream to the other in order,
    so both results match the oracle."""
    from zkx_dev_amd import convolution
    a1, b1 = _inputs("FP32", 200000, 11), _inputs("FP32", 3000, 12)
    a2, b2 = _inputs("FP32", 150000, 13), _inputs("FP32", 2500, 14)
    t = [torch.from_numpy(v).cuda() for v in (a1, b1, a2, b2)]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        y1 = convolution.correlate(t[0], t[1], "same", "fft", stream=s1)
        y2 = convolution.correlate(t[2], t[3], "same", "fft", stream=s2)
        torch.cuda.synchronize()
        assert orc.rel_l2(y1.cpu().numpy().astype(np.float64),
                          np.correlate(a1.astype(np.float64), b1.astype(np.float64), "same")) < 1e-5
        assert orc.rel_l2(y2.cpu().numpy().astype(np.float64),
                          np.correlate(a2.astype(np.float64), b2.astype(np.float64), "same")) < 1e-5


def test_correlate_fft_limits():
    from zkx_dev_amd import convolution
    a = torch.ones(5000, device="cuda")
    with pytest.raises(ValueError):
        convolution.correlate(a, a[:4], "same", "fft", out=a)  # aliasing
    assert convolution.correlate(a[:0], a[:4], "full", "fft").numel() == 0
