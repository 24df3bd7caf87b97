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


def test_correlate_fft_limits():
    from zkx_dev_amd import convolution
    from zkx_dev_amd._lib import ZkxError
    a = torch.ones(5000, device="cuda")
    with pytest.raises(ZkxError):
        convolution.correlate(a, a[:4097], "full", "fft")  # both inputs longer than 4096
    with pytest.raises(ValueError):
        convolution.correlate(a, a[:4], "same", "fft", out=a)  # aliasing
    assert convolution.correlate(a[:0], a[:4], "full", "fft").numel() == 0
