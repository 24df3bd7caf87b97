"""The operator API for the reference's other element types (__half, int32,
int16, int8 beside float; radartools_typed.cpp:810-811,
filtering_typed.cu:683+) and the complex filtering / fftconvolve extensions,
through tests/cpp/typed_ops.cpp (g++, linked to libzkx_radar.so).

The reference loads every sample to float and runs its FP32 kernels
(operator_compute_policy.h:42-45,64-67), so each typed result must equal,
bit for bit, the float path of this library on the converted samples; on top
of that the type-dependent rules are checked against the oracle: the output
storage dtype tables (radartools_typed.cpp:271-278, filtering_typed.cpp:62-101),
the integer CFAR detection (radartools_kernels.cuh:23-37; orc.ca_cfar), the
resample_poly coefficient scaling in the coefficients' type
(filtering_kernels.cuh:473-497; orc.scaled_custom_device) and the FP32
device cfar_alpha (radartools_kernels.cuh:136-148; orc.cfar_alpha_f32)."""
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "typed_ops")
P, R, T, ROWS, N, TAPS, UP, DOWN = 16, 256, 16, 3, 1000, 31, 3, 2
DTYPES = {"FP32": (0, np.float32), "FP16": (1, np.float16), "INT32": (2, np.int32), "INT16": (3, np.int16),
          "INT8": (4, np.int8)}
# ResamplePolyOutputDtype {fp16, fp32, fp64, int32, int16, int8}
RP_CODE = {"FP16": 0, "FP32": 1, "FP64": 2, "INT32": 3, "INT16": 4, "INT8": 5}


def _inputs(name, rng):
    dt = DTYPES[name][1]
    if name in ("FP32", "FP16"):
        def sig(n, scale=1.0):
            return (rng.standard_normal(n) * scale).astype(dt)
        cfar = rng.exponential(1.0, P * R).astype(dt)
        cfar[rng.integers(0, P * R, 8)] = dt(60.0)
        zi = sig(ROWS * (TAPS - 1))
        coeffs = (np.hanning(TAPS + 2)[1:-1] / 8).astype(dt)
    else:
        hi = {"INT32": 1 << 28, "INT16": 2000, "INT8": 100}[name]

        def sig(n, scale=1.0):
            return rng.integers(-hi, hi + 1, n).astype(dt)
        cfar = rng.integers(0, min(hi, 40) + 1, P * R).astype(dt)
        cfar[rng.integers(0, P * R, 8)] = dt(min(hi, 120))
        zi = sig(ROWS * (TAPS - 1))
        coeffs = rng.integers(-5, 6, TAPS).astype(dt)
        coeffs[0] = dt(100)  # wraps when scaled in int8
    x = sig(P * R * 2)
    tmpl = sig(T * 2)
    xf = sig(ROWS * N)
    h = sig(TAPS) if name in ("FP32", "FP16") else rng.integers(-3, 4, TAPS).astype(dt)
    return dict(x=x, tmpl=tmpl, cfar=cfar, xf=xf, h=h, zi=zi, coeffs=coeffs)


def _run(name, inp, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    hdr = np.array([DTYPES[name][0], P, R, T, ROWS, N, TAPS, UP, DOWN], np.int32)
    blob = hdr.tobytes() + b"".join(inp[k].tobytes() for k in ("x", "tmpl", "cfar", "xf", "h", "zi", "coeffs"))
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    src.write_bytes(blob)
    print(f"running typed_ops {name}", flush=True)
    r = subprocess.run([EXE, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out, off, blocks = dst.read_bytes(), 0, []
    kinds = {0: np.float32, 1: np.float64, 2: np.uint8, 3: DTYPES[name][1], 4: np.int32}
    while off < len(out):
        kind = int(np.frombuffer(out, np.int32, 1, off)[0])
        nbytes = int(np.frombuffer(out, np.int64, 1, off + 4)[0])
        off += 12
        blocks.append(np.frombuffer(out, kinds[kind], nbytes // np.dtype(kinds[kind]).itemsize, off))
        off += nbytes
    return blocks


def _c(a):
    return a.view(np.complex128) if a.dtype == np.float64 else a.view(np.complex64)


@pytest.mark.parametrize("name", list(DTYPES))
def test_typed_operators(name, tmp_path):
    from zkx_dev_amd import convolution, filtering, radar
    from zkx_dev_amd._lib import check, lib
    rng = np.random.default_rng(20261203 + DTYPES[name][0])
    inp = _inputs(name, rng)
    b = iter(_run(name, inp, tmp_path))
    integral = name.startswith("INT")
    f32 = {k: v.astype(np.float32) for k, v in inp.items()}  # the reference's load to float
    dev = {k: torch.from_numpy(v).cuda() for k, v in f32.items()}
    x = dev["x"].view(torch.complex64).reshape(P, R)
    tm = dev["tmpl"].view(torch.complex64)

    # pulse compression: FP64 storage for integer types without a window, always with one
    for win, norm in (("none", False), ("hamming", True)):
        flag, val = next(b)[0], _c(next(b))
        assert flag == (1 if (integral or win != "none") else 0)
        ref = radar.pulse_compression(x, tm, normalize=norm, window=win, widen=False).cpu().numpy()
        assert np.array_equal(val.reshape(P, R), ref.astype(val.dtype)), (name, win)
    # pulse Doppler (Hann: FP64 storage) and the preallocated FP32 overload
    flag, val = next(b)[0], _c(next(b))
    assert flag == 1
    ref = radar.pulse_doppler(x, window="hann", widen=False).cpu().numpy()
    assert np.array_equal(val.reshape(P, R), ref.astype(np.complex128))
    val = _c(next(b))
    assert np.array_equal(val.reshape(P, R), radar.pulse_doppler(x, widen=False).cpu().numpy())
    # cfar_alpha: CPU form FP64 of the loaded pfa; device form FP32
    pfa = 1.0 if integral else float(np.float32(DTYPES[name][1](1e-3)))
    a_cpu, a_dev = next(b)[0], next(b)[0]
    assert a_cpu == orc.cfar_alpha(pfa, 24)
    a32 = orc.cfar_alpha_f32(pfa, 24)
    assert a_dev == float(np.float32(a_dev))  # an FP32 value in FP64 storage
    assert abs(a_dev - float(a32)) <= orc.cfar_alpha_f32_tolerance(pfa, 24)
    # CA-CFAR on the typed map: thresholds and detections as the reference CPU form
    thr, det = next(b), next(b)
    cmap = inp["cfar"].reshape(P, R)
    othr, odet = orc.ca_cfar(cmap, (2, 2), (4, 4), 1e-3)
    assert np.array_equal(thr.reshape(P, R).view(np.uint32), othr.view(np.uint32))
    assert np.array_equal(det.reshape(P, R), odet)
    fthr, _ = orc.ca_cfar(cmap.astype(np.float32), (2, 2), (4, 4), 1e-3)
    assert np.array_equal(fthr, othr)  # thresholds come from the float samples
    # upfirdn: FP64 storage for int32
    flag, val = next(b)[0], next(b)
    assert flag == (1 if name == "INT32" else 0)
    xf = dev["xf"].reshape(ROWS, N)
    ref = filtering.upfirdn(dev["h"], xf, UP, DOWN, axis=1).cpu().numpy()
    assert np.array_equal(val.reshape(ref.shape), ref.astype(val.dtype))
    # resample_poly, default design: FP64
    flag, val = next(b)[0], next(b)
    assert flag == RP_CODE["FP64"]
    ref = filtering.resample_poly(xf, UP, DOWN, axis=1, widen=False).cpu().numpy()
    assert np.array_equal(val.reshape(ref.shape), ref.astype(np.float64))
    # resample_poly, custom coefficients scaled in their own type
    flag, val = next(b)[0], next(b)
    want = "FP64" if name == "INT32" else "FP32"
    assert flag == RP_CODE[want]
    taps = np.array([orc.scaled_custom_device(v, UP, name) for v in inp["coeffs"].tolist()], np.float32)
    if name == "INT8":
        assert taps[0] == 44.0  # 100 * 3 wraps in int8
    td = torch.from_numpy(taps).cuda()
    y = torch.empty((ROWS, (N * UP + DOWN - 1) // DOWN), dtype=torch.float32, device="cuda")
    check(lib().zkx_resample_poly_f32(td.data_ptr(), TAPS, xf.data_ptr(), ROWS, N, 1, UP, DOWN, y.data_ptr(), None))
    assert np.array_equal(val.reshape(y.shape), y.cpu().numpy().astype(val.dtype))
    # identity: a copy in the input's own type
    flag, val = next(b)[0], next(b)
    assert flag == RP_CODE[name]
    assert np.array_equal(val.view(np.uint8), inp["xf"].view(np.uint8))
    # firfilter with zi: float results
    fy, fzf = next(b), next(b)
    zi = dev["zi"].reshape(ROWS, TAPS - 1)
    ry, rzf = filtering.firfilter(dev["h"], xf, axis=1, zi=zi)
    assert np.array_equal(fy, ry.cpu().numpy().reshape(-1))
    assert np.array_equal(fzf, rzf.cpu().numpy().reshape(-1))
    # fftconvolve (same) of the pulses with the template
    cv = _c(next(b))
    ref = filtering.fftconvolve(x, tm, mode="same").cpu().numpy()
    assert np.array_equal(cv.reshape(ref.shape), ref)
    # decimate by down+1 along axis 1: designed filter -> FP64 storage; custom
    # coefficients in the input type -> FP32 (FP64 for int32); causal order 24
    qd = DOWN + 1
    flag, val = next(b)[0], next(b)
    assert flag == 1
    ref = filtering.decimate(xf, qd, axis=1, widen=False).cpu().numpy()
    assert np.array_equal(val.reshape(ref.shape), ref.astype(np.float64))
    flag, val = next(b)[0], next(b)
    assert flag == (1 if name == "INT32" else 0)
    ref = filtering.decimate(xf, qd, dev["h"], axis=1).cpu().numpy()
    assert np.array_equal(val.reshape(ref.shape), ref.astype(val.dtype))
    val = next(b)
    ref = filtering.decimate(xf, qd, 24, axis=1, zero_phase=False, widen=False).cpu().numpy()
    assert np.array_equal(val.reshape(ref.shape), ref.astype(np.float64))
    # channelize_poly, 4 channels: complex64 always
    val = _c(next(b))
    ref = filtering.channelize_poly(dev["xf"], dev["h"], 4).cpu().numpy()
    assert np.array_equal(val.reshape(ref.shape), ref)
    # correlate (direct) of the first line with the taps: the reference's
    # arithmetic in the input type, bit for bit (the oracle reproduces its digests)
    for mode in ("full", "same", "valid"):
        val = next(b)
        ref = orc.correlate(inp["xf"][:N], inp["h"], mode, name)
        assert np.array_equal(val.view(np.uint8), ref.view(np.uint8)), (name, mode)
    # the FFT method (workspace overload) equals the Python layer's FFT method
    val = next(b)
    ref = convolution.correlate(torch.from_numpy(np.ascontiguousarray(inp["xf"][:N])).cuda(),
                                torch.from_numpy(inp["h"]).cuda(), "same", "fft").cpu().numpy()
    assert np.array_equal(val.view(np.uint8), ref.view(np.uint8)), (name, "fft")
    # ambgfun of the template: 2-D FP32 for float/half, FP64 storage for
    # integers (radartools_typed.cpp:119-130); the cuts are FP64 storage
    flag, val = next(b)[0], next(b)
    assert flag == (1 if integral else 0)
    ref = radar.ambgfun(tm, 1000.0, 100.0).cpu().numpy()
    assert np.array_equal(val.reshape(ref.shape), ref.astype(val.dtype), equal_nan=True)
    val = next(b)
    ref = radar.ambgfun(tm, 1000.0, 100.0, cut="delay", cutValue=0.002).cpu().numpy()
    assert val.dtype == np.float64 and np.array_equal(val, ref, equal_nan=True)
    val = next(b)
    ref = radar.ambgfun(tm, 1000.0, 100.0, cut="doppler", cutValue=37.0).cpu().numpy()
    assert val.dtype == np.float64 and np.array_equal(val, ref, equal_nan=True)
    # correlate2d of the filtering samples as 30 x 100 with 30 taps as 5 x 6
    for mode, boundary in (("same", "symm"), ("full", "fill"), ("valid", "wrap")):
        val = next(b)
        ref = orc.correlate2d(inp["xf"].reshape(30, 100), inp["h"][:30].reshape(5, 6), mode, boundary, 2.0, name)
        assert np.array_equal(val.view(np.uint8), ref.reshape(-1).view(np.uint8)), (name, mode, boundary)
    # N-D correlate (FFT) of the same pair, and the Python binding's result
    for mode in ("full", "same", "valid"):
        val = next(b)
        k = inp["h"][:30].reshape(5, 6)
        ref = orc.correlate_nd(inp["xf"].reshape(30, 100), k, mode, name)
        py = convolution.correlate(torch.from_numpy(inp["xf"].reshape(30, 100)).cuda(),
                                   torch.from_numpy(np.ascontiguousarray(k)).cuda(), mode).cpu().numpy()
        assert np.array_equal(val.view(np.uint8), py.reshape(-1).view(np.uint8)), (name, mode)
        if name == "INT32":
            pass  # sums of 2^28-sized samples: FP32 transforms are not exact there; the C++ and Python paths agree
        elif integral:
            assert np.array_equal(val, ref.reshape(-1)), (name, mode)
        elif name == "FP16":
            assert np.allclose(val.astype(np.float64), ref.reshape(-1).astype(np.float64), rtol=5e-3, atol=5e-3)
        else:
            assert orc.rel_l2(val, ref.reshape(-1)) < 1e-5, (name, mode)
    if name == "FP32":
        # complex extensions: real taps act on each component
        uc = _c(next(b))
        ref = filtering.upfirdn(dev["h"], x, UP, DOWN, axis=1).cpu().numpy()
        assert np.array_equal(uc.reshape(ref.shape), ref)
        rc = _c(next(b))
        ref = filtering.resample_poly(x, UP, DOWN, axis=1, coefficients=dev["coeffs"]).cpu().numpy()
        assert np.array_equal(rc.reshape(ref.shape), ref)
        fc = _c(next(b)).reshape(P, R)
        re = filtering.firfilter(dev["h"], x.real.contiguous(), axis=1).cpu().numpy()
        im = filtering.firfilter(dev["h"], x.imag.contiguous(), axis=1).cpu().numpy()
        assert np.array_equal(fc.real, re) and np.array_equal(fc.imag, im)
        # N-D fftconvolve through the C++ overload: bit-identical to the
        # Python binding, within 1e-5 of the FP64 oracle
        k = inp["h"][:30].reshape(5, 6)
        r1 = next(b)
        py = filtering.fftconvolve(dev["xf"].reshape(30, 100), torch.from_numpy(np.ascontiguousarray(k)).cuda(),
                                   "same").cpu().numpy()
        assert np.array_equal(r1.view(np.uint32), py.reshape(-1).view(np.uint32))
        assert orc.rel_l2(r1, orc.fftconvolve_nd(inp["xf"].reshape(30, 100), k, "same").reshape(-1)) < 1e-5
        r2 = _c(next(b))
        tm2 = tm.reshape(1, -1)
        py = filtering.fftconvolve(x, tm2, "full", [-1]).cpu().numpy()
        assert np.array_equal(r2.view(np.uint64), py.reshape(-1).view(np.uint64))
        ref = orc.fftconvolve_nd(x.cpu().numpy(), tm2.cpu().numpy(), "full", [-1])
        assert orc.rel_l2(r2, ref.reshape(-1)) < 1e-5
    assert next(b, None) is None
