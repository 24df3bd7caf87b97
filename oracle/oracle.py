"""Python face of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker: the product
library never routes through it.

Two layers:
  * ctypes wrappers over ``oracle/_build/libzkx_oracle.so`` (``zkx_oracle.c``),
    the faithful restatement of the reference CPU functions, pinned bit-for-bit
    against the reference's recorded probe digests;
  * numpy FP64 FFT forms of the same maths (``*_fft64``), used as the fast FP64
    ground truth at full sizes, pinned against the C direct sums at small sizes.

Reference functions followed are cited at each wrapper (paths relative to the
reference repository).
"""
from __future__ import annotations

import ctypes
import hashlib
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libzkx_oracle.so")

WINDOWS = {"none": 0, "hann": 1, "hamming": 2, "explicit_fp32": 3, "explicit_fp64": 4}

_lib = None


def build():
    """Compile the C oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i, d, sz, u64 = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_uint64
        sig = {
            "orc_probe_complex_input": (None, [sz, i, i, P]),
            "orc_probe_real_input": (None, [sz, i, i, P]),
            "orc_upfirdn_probe_input": (None, [sz, u64, i, P]),
            "orc_resample_probe_input": (None, [sz, u64, i, P]),
            "orc_firfilter_probe_signal": (None, [i, i, i, P]),
            "orc_window": (None, [i, i, i, P]),
            "orc_pulse_compression": (i, [P, i, i, P, i, i, i, P, i, i, P]),
            "orc_pulse_doppler": (i, [P, i, i, i, i, P, i, P]),
            "orc_ambgfun": (i, [P, sz, i, P, sz, i, i, d, d, P]),
            "orc_firwin_lowpass_hamming": (i, [i, d, P]),
            "orc_decimate": (i, [P, i, i, i, i, P, i, i, i, P]),
            "orc_channelize_poly": (i, [P, ctypes.c_longlong, P, ctypes.c_longlong, i, P]),
            "orc_correlate_len": (ctypes.c_longlong, [ctypes.c_longlong, ctypes.c_longlong, i]),
            "orc_correlate_direct": (i, [P, ctypes.c_longlong, P, ctypes.c_longlong, i, i, P]),
            "orc_correlate2d_shape": (i, [ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, i, P, P]),
            "orc_correlate2d": (i, [P, ctypes.c_longlong, ctypes.c_longlong, P, ctypes.c_longlong, ctypes.c_longlong, i, i, ctypes.c_double, i, P]),
            "orc_cfar_alpha": (d, [d, i]),
            "orc_ca_cfar_count": (i, [i, i, i, i, i]),
            "orc_ca_cfar_alpha_f": (ctypes.c_float, [i, i, i, i, i, d]),
            "orc_ca_cfar": (i, [P, P, d, d, i, i, i, i, i, i, i, d, P, P]),
            "orc_ca_cfar_exact": (i, [P, i, i, i, i, i, i, d, P]),
            "orc_upfirdn_out_len": (ctypes.c_longlong, [ctypes.c_longlong, ctypes.c_longlong, i, i]),
            "orc_upfirdn": (i, [P, i, P, i, i, i, i, i, i, P]),
            "orc_resample_plan_fill": (i, [i, i, i, i, i, P]),
            "orc_resample_bessel_i0": (d, [d]),
            "orc_resample_design": (i, [i, i, i, d, P, P]),
            "orc_resample_poly": (i, [P, i, P, i, i, i, i, i, i, P]),
            "orc_firfilter": (i, [P, i, P, i, i, i, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def sha256(arr) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


# --------------------------------------------------------------------------
# probe generators (reference probe programs, see zkx_oracle.c header)
# --------------------------------------------------------------------------
INTEGRAL = {"FP32": False, "FP16": False, "INT32": True, "INT16": True, "INT8": True}
NP_DTYPE = {"FP32": np.float32, "FP16": np.float16, "INT32": np.int32,
            "INT16": np.int16, "INT8": np.int8}


def probe_complex_input(n, seed, dtype):
    out = np.empty(2 * n, np.float32)
    lib().orc_probe_complex_input(n, seed, INTEGRAL[dtype], _p(out))
    return out.astype(NP_DTYPE[dtype])  # interleaved (re, im) of the typed dtype


def probe_real_input(n, seed, dtype):
    out = np.empty(n, np.float32)
    lib().orc_probe_real_input(n, seed, INTEGRAL[dtype], _p(out))
    return out.astype(NP_DTYPE[dtype])


def upfirdn_probe_input(n, seed, dtype):
    out = np.empty(n, np.float32)
    lib().orc_upfirdn_probe_input(n, seed, INTEGRAL[dtype], _p(out))
    return out.astype(NP_DTYPE[dtype])


def resample_probe_input(n, seed, dtype):
    out = np.empty(n, np.float32)
    lib().orc_resample_probe_input(n, seed, INTEGRAL[dtype], _p(out))
    return out.astype(NP_DTYPE[dtype])


def firfilter_probe_signal(n, seed, dtype):
    out = np.empty(n, np.float64)
    lib().orc_firfilter_probe_signal(n, seed, INTEGRAL[dtype], _p(out))
    if dtype == "FP16":  # task2_filter_probe.cu cv<__half>: double -> float -> half
        return out.astype(np.float32).astype(np.float16)
    return out.astype(NP_DTYPE[dtype])


# --------------------------------------------------------------------------
# radar ops
# --------------------------------------------------------------------------
def window(kind, length, exact=False):
    out = np.empty(length, np.float64)
    lib().orc_window(WINDOWS[kind], length, int(exact), _p(out))
    return out


def _as_c64_f32(x):
    """complex array -> interleaved float32 (values the reference loads)."""
    x = np.asarray(x)
    if np.iscomplexobj(x):
        v = np.empty(x.shape + (2,), np.float32)
        v[..., 0] = x.real
        v[..., 1] = x.imag
        return np.ascontiguousarray(v.reshape(-1))
    return np.ascontiguousarray(x.astype(np.float32).reshape(-1))


def pulse_compression(x, tmpl, nfft=-1, window="none", normalize=False, exact=False,
                      explicit_window=None):
    """radartools_typed.cpp:618-675 pulse_compression_typed_cpu (direct circular
    correlation in FP64).  x: complex [P, S]; tmpl: complex [T]. -> complex128 [P, nfft]."""
    x = np.asarray(x)
    P, S = x.shape
    T = np.asarray(tmpl).shape[0]
    n = S if nfft == -1 else nfft
    out = np.empty(2 * P * n, np.float64)
    ew = None if explicit_window is None else np.ascontiguousarray(explicit_window, np.float32)
    rc = lib().orc_pulse_compression(_p(_as_c64_f32(x)), P, S, _p(_as_c64_f32(tmpl)), T, nfft,
                                     WINDOWS[window], _p(ew), int(normalize), int(exact), _p(out))
    if rc:
        raise ValueError(f"pulse_compression rc={rc}")
    return out.view(np.complex128).reshape(P, n)


def pulse_doppler(x, nfft=-1, window="none", exact=False, explicit_window=None):
    """radartools_typed.cpp:714-759 pulse_doppler_typed_cpu (direct DFT along
    pulses; faithful mode keeps the float pi).  x: complex [P, R] -> complex128 [nfft, R]."""
    x = np.asarray(x)
    P, R = x.shape
    n = P if nfft == -1 else nfft
    out = np.empty(2 * n * R, np.float64)
    ew = None if explicit_window is None else np.ascontiguousarray(explicit_window, np.float32)
    rc = lib().orc_pulse_doppler(_p(_as_c64_f32(x)), P, R, nfft, WINDOWS[window], _p(ew),
                                 int(exact), _p(out))
    if rc:
        raise ValueError(f"pulse_doppler rc={rc}")
    return out.view(np.complex128).reshape(n, R)


def decimate_probe_input(n, seed, dtype):
    """ZKX/test_all/operators/cases/filtering/decimate/operator_probe.cu:29:
    r = (i*17 + seed) % 23 - 11, times 0.25 for floating types, cast to dtype."""
    i = np.arange(n, dtype=np.uint64)
    r = ((i * np.uint64(17) + np.uint64(seed)) % np.uint64(23)).astype(np.int64) - 11
    if INTEGRAL[dtype]:
        return r.astype(NP_DTYPE[dtype])
    return (r.astype(np.float32) * np.float32(0.25)).astype(NP_DTYPE[dtype])


def channelize_probe_input(count, seed, taps, dtype):
    """ZKX/test_all/operators/cases/filtering/channelize_poly/operator_probe.cu:122-135:
    x: (i*17 + seed) % 23 - 11 (x0.25 for floating types); taps:
    (i*7 + 3*seed) % 9 - 4 (x0.125)."""
    i = np.arange(count, dtype=np.uint64)
    if taps:
        r = ((i * np.uint64(7) + np.uint64(3 * seed)) % np.uint64(9)).astype(np.int64) - 4
        scale = np.float32(0.125)
    else:
        r = ((i * np.uint64(17) + np.uint64(seed)) % np.uint64(23)).astype(np.int64) - 11
        scale = np.float32(0.25)
    if INTEGRAL[dtype]:
        return r.astype(NP_DTYPE[dtype])
    return (r.astype(np.float32) * scale).astype(NP_DTYPE[dtype])


def channelize_poly(x, h, n_chans):
    """channelize_poly_typed_cpu (filtering_typed.cpp:443-466 and the CPU form):
    FP64 polyphase sums and +i DFT over phases, complex64 [n_chans, N // n_chans]."""
    xs = np.ascontiguousarray(np.asarray(x).reshape(-1).astype(np.float64))
    hs = np.ascontiguousarray(np.asarray(h).reshape(-1).astype(np.float64))
    M = int(n_chans)
    if M <= 0:
        raise ValueError("channelize_poly n_chans")
    P = xs.size // M
    out = np.empty(2 * M * P, np.float32)
    rc = lib().orc_channelize_poly(_p(xs), xs.size, _p(hs), hs.size, M, _p(out))
    if rc:
        raise ValueError(f"channelize_poly rc={rc}")
    return out.view(np.complex64).reshape(M, P)


CORRELATE_MODES = {"full": 0, "same": 1, "valid": 2}
_INT_BITS = {"INT32": 32, "INT16": 16, "INT8": 8}


def correlate_probe_signal(count, salt, dtype):
    """ZKX/test_all/operators/suites/task2_features/task2_features_probe.cu:120-131:
    4 sin(0.031 i) + 2 cos(0.071 i) + ((17 i + salt) % 13 - 6)/8, rounded half
    away from zero for integer types, then cast (libm sin/cos through Python's
    math module)."""
    import math
    raw = np.array([4.0 * math.sin(0.031 * i) + 2.0 * math.cos(0.071 * i) + ((i * 17 + salt) % 13 - 6) * 0.125
                    for i in range(count)], np.float64)
    if INTEGRAL[dtype]:
        return np.where(raw >= 0, np.floor(raw + 0.5), -np.floor(-raw + 0.5)).astype(NP_DTYPE[dtype])
    if dtype == "FP16":
        return raw.astype(np.float32).astype(np.float16)
    return raw.astype(np.float32)


def correlate(in1, in2, mode="full", dtype="FP32"):
    """correlate_typed_cpu direct (convolution_typed.cpp:285-334) for 1-D
    inputs of element type `dtype`: returns the output in that type (float32,
    float16 rounded from the FP32 sum, or the integer wrapped modulo 2^bits)."""
    a = np.ascontiguousarray(np.asarray(in1).reshape(-1).astype(np.float64))
    b = np.ascontiguousarray(np.asarray(in2).reshape(-1).astype(np.float64))
    m = CORRELATE_MODES[mode]
    n = lib().orc_correlate_len(a.size, b.size, m)
    out = np.empty(max(n, 0), np.float64)
    rc = lib().orc_correlate_direct(_p(a), a.size, _p(b), b.size, m, _INT_BITS.get(dtype, 0), _p(out))
    if rc:
        raise ValueError(f"correlate rc={rc}")
    if INTEGRAL[dtype]:
        return out.astype(NP_DTYPE[dtype])
    if dtype == "FP16":
        return out.astype(np.float32).astype(np.float16)
    return out.astype(np.float32)


CORRELATE2D_BOUNDARIES = {"fill": 0, "wrap": 1, "symm": 2}


def correlate2d_probe_input(count, seed, salt, dtype):
    """ZKX/test_all/operators/cases/convolution/correlate2d/operator_probe.cu:26-27:
    r = (17 i + seed + salt) mod 11 - 5 (unsigned 64-bit arithmetic); integers
    store r, floating types r / 4 (float, then half rounded to nearest)."""
    i = np.arange(count, dtype=np.uint64)
    r = ((i * np.uint64(17) + np.uint64(seed) + np.uint64(salt)) % np.uint64(11)).astype(np.int64) - 5
    if INTEGRAL[dtype]:
        return r.astype(NP_DTYPE[dtype])
    x = r.astype(np.float32) * np.float32(0.25)
    return x.astype(np.float16) if dtype == "FP16" else x


def correlate2d_fill(fillvalue, dtype):
    """convolution_typed.cpp:116-131 / 28-37: the fill value in the element
    type: integers truncate toward zero and wrap modulo 2^bits (a non-finite
    value is an error), half rounds the float value to nearest, float rounds."""
    if INTEGRAL[dtype]:
        if not math.isfinite(fillvalue):
            raise ValueError("typed correlate2d integral fillvalue must be finite")
        bits = _INT_BITS[dtype]
        w = math.fmod(math.trunc(fillvalue), 2.0 ** bits)
        if w < 0:
            w += 2.0 ** bits
        w = int(w)
        return float(w - (1 << bits) if w >= (1 << (bits - 1)) else w)
    if dtype == "FP16":
        return float(np.float16(np.float32(fillvalue)))
    return float(np.float32(fillvalue))


def correlate2d(in1, in2, mode="full", boundary="fill", fillvalue=0.0, dtype="FP32"):
    """correlate2d_typed_cpu (convolution_typed.cpp:416-470) for 2-D inputs of
    element type `dtype`; the result in that type (float32, float16 rounded
    from the FP32 sum, or the integer wrapped modulo 2^bits)."""
    a = np.ascontiguousarray(np.asarray(in1, np.float64))
    b = np.ascontiguousarray(np.asarray(in2, np.float64))
    if a.ndim != 2 or b.ndim != 2:
        raise ValueError("correlate2d inputs must be 2-D")
    m = CORRELATE_MODES[mode]
    if boundary not in CORRELATE2D_BOUNDARIES:
        raise ValueError("typed correlate2d boundary must be fill, wrap, or symm")
    fill = 0.0 if mode == "valid" else (correlate2d_fill(fillvalue, dtype) if boundary == "fill" else 0.0)
    rows, cols = ctypes.c_longlong(0), ctypes.c_longlong(0)
    rc = lib().orc_correlate2d_shape(a.shape[0], a.shape[1], b.shape[0], b.shape[1], m, ctypes.byref(rows),
                                     ctypes.byref(cols))
    if rc:
        raise ValueError(f"correlate2d shape rc={rc}")
    out = np.empty((rows.value, cols.value), np.float64)
    if out.size:
        rc = lib().orc_correlate2d(_p(a), a.shape[0], a.shape[1], _p(b), b.shape[0], b.shape[1], m,
                                   CORRELATE2D_BOUNDARIES[boundary], fill, _INT_BITS.get(dtype, 0), _p(out))
        if rc:
            raise ValueError(f"correlate2d rc={rc}")
    if INTEGRAL[dtype]:
        return out.astype(NP_DTYPE[dtype])
    if dtype == "FP16":
        return out.astype(np.float32).astype(np.float16)
    return out.astype(np.float32)


def firwin_lowpass_hamming(numtaps, cutoff):
    """filter_design_typed.cpp:268-296 firwin_typed_cpu<float>(numtaps, {cutoff},
    Hamming, pass_zero, scale, fs = 2): FP64 taps; cutoff is loaded as float."""
    out = np.empty(int(numtaps), np.float64)
    rc = lib().orc_firwin_lowpass_hamming(int(numtaps), float(np.float32(cutoff)), _p(out))
    if rc:
        raise ValueError("firwin cutoff")
    return out


def decimate_design(q, filter_order=None):
    """design_decimate_filter (filtering_typed.cpp:318-336): order 20*q unless
    given, cutoff 1/q in float."""
    order = 20 * q if filter_order is None else int(filter_order)
    if order < 0:
        raise ValueError("decimate filter order")
    return firwin_lowpass_hamming(order + 1, np.float32(1.0) / np.float32(q))


def decimate(x, q, axis=-1, filter_order=None, zero_phase=True, taps=None, fp32=None):
    """decimate_typed_cpu (filtering_typed.cpp:2229-2317).  x: loaded samples
    (any real dtype, converted to float64 exactly); taps: custom coefficients
    (loaded), or None for the default design.  fp32 (default: custom taps)
    selects the FP32-accumulating path; the result is float64 storage."""
    x = np.asarray(x)
    xs = np.ascontiguousarray(x.astype(np.float64))
    ax = axis % xs.ndim
    shape = xs.shape
    outer = int(np.prod(shape[:ax], dtype=np.int64))
    n = shape[ax]
    inner = int(np.prod(shape[ax + 1:], dtype=np.int64))
    h = decimate_design(q, filter_order) if taps is None else np.ascontiguousarray(taps, np.float64)
    if fp32 is None:
        fp32 = taps is not None
    out_n = 0 if n == 0 else 1 + (n - 1) // q
    out = np.empty(outer * out_n * inner, np.float64)
    rc = lib().orc_decimate(_p(xs), outer, n, inner, int(q), _p(h), h.size, int(bool(zero_phase)), int(bool(fp32)),
                            _p(out))
    if rc:
        raise ValueError(f"decimate rc={rc}")
    return out.reshape(shape[:ax] + (out_n,) + shape[ax + 1:])


AMBGFUN_CUTS = {"2d": 0, "delay": 1, "doppler": 2}


def ambgfun(x, fs=1.0, prf=1.0, y=None, cut="2d", cut_value=0.0, nx=None, ny=None,
            output_fp32=True):
    """radartools_typed.cpp:329-409 ambgfun_typed_cpu (FP64 direct sums).
    x, y: complex 1-D arrays of loaded samples (y=None means y = x); nx/ny the
    first extents when a higher-rank input is flattened (the norm covers all
    samples).  The 2-D cut of a float/half input is returned as FP32 (the
    reference narrows it); the cuts and integer inputs stay FP64.  prf does not
    enter the result (as in the reference)."""
    del prf
    xf = _as_c64_f32(np.asarray(x).reshape(-1))
    xc = xf.size // 2
    nx = xc if nx is None else int(nx)
    if y is None:
        yf, yc, ny_ = None, 0, 0
    else:
        yf = _as_c64_f32(np.asarray(y).reshape(-1))
        yc = yf.size // 2
        ny_ = yc if ny is None else int(ny)
    n_y = nx if y is None else ny_
    rows = nx + n_y - 1
    nfreq = 1 << max(0, (rows - 1).bit_length())
    code = AMBGFUN_CUTS[cut]
    size = rows * nfreq if code == 0 else (nfreq if code == 1 else rows)
    out = np.empty(size, np.float64)
    rc = lib().orc_ambgfun(_p(xf), xc, nx, _p(yf), yc, ny_, code, float(fs), float(cut_value), _p(out))
    if rc:
        raise ValueError(f"ambgfun rc={rc}")
    if code == 0:
        out = out.reshape(rows, nfreq)
        if output_fp32:
            out = out.astype(np.float32)
    return out


def ambgfun_fft64_row(xn, yn, row, nfreq):
    """One delay row of cuSignal's 2-D cut (radartools.py:127-143,207-213) for
    normalised waveforms: nfreq * |fftshift(ifft(yn[c] conj(xn[c-(nx-1)+row]), nfreq))|."""
    nx, ny = xn.size, yn.size
    c = np.arange(ny)
    xc = c - (nx - 1) + row
    live = (xc >= 0) & (xc < nx)
    d = np.zeros(ny, np.complex128)
    d[live] = yn[live] * np.conj(xn[xc[live]])
    return nfreq * np.abs(np.fft.fftshift(np.fft.ifft(d, nfreq)))


def ambgfun_fft64(x, fs=1.0, y=None, cut="2d", cut_value=0.0):
    """The cuSignal formulation (cusignal/radartools/radartools.py:185-234)
    with numpy's FP64 FFTs: an independent computation of the same function,
    used to pin the direct-sum restatement above."""
    x = np.asarray(x, np.complex128)
    xn = x / np.linalg.norm(x)
    yn = xn if y is None else np.asarray(y, np.complex128) / np.linalg.norm(y)
    nx, ny = xn.size, yn.size
    rows = nx + ny - 1
    nfreq = 1 << max(0, (rows - 1).bit_length())
    if cut == "2d":
        return np.stack([ambgfun_fft64_row(xn, yn, row, nfreq) for row in range(rows)])
    if cut == "delay":
        fd = -fs / 2 + fs * np.arange(nfreq) / nfreq
        xs = np.fft.ifft(np.fft.fft(xn, nfreq) * np.exp(1j * 2 * np.pi * fd * cut_value))
        ypad = np.zeros(nfreq, np.complex128)
        ypad[:ny] = yn
        return nfreq * np.abs(np.fft.ifftshift(np.fft.ifft(ypad * np.conj(xs), nfreq)))
    t = np.arange(nx) / fs
    fy = np.fft.fft(yn, rows)
    fx = np.fft.fft(xn * np.exp(1j * 2 * np.pi * cut_value * t), rows)
    return np.abs(np.fft.fftshift(np.fft.ifft(fy * np.conj(fx))))


# ---------------------------------------------------------------------------
# echo source (the step before the radar path)
# ---------------------------------------------------------------------------

def echo_noise_bits(index, seed):
    """Task1 step1 noise hash (ZKX/Task1/accuracy/validation/step1_validation.cu:8-16):
    a 32-bit integer mix of (index, seed); index is a uint64 array."""
    m = np.uint64(0xFFFFFFFF)
    i = np.asarray(index, np.uint64)
    v = (np.uint64(seed) ^ ((i * np.uint64(747796405) + np.uint64(2891336453)) & m)) & m
    v ^= v >> np.uint64(16)
    v = (v * np.uint64(2246822519)) & m
    v ^= v >> np.uint64(13)
    v = (v * np.uint64(3266489917)) & m
    return v ^ (v >> np.uint64(16))


def echo_uniform_noise(index, seed, noise_std):
    """step1_validation.cu:18-22: ((bits & 0xffff) / 32767.5 - 1) * std, in FP32."""
    b = (echo_noise_bits(index, seed) & np.uint64(0xFFFF)).astype(np.float32)
    return (b / np.float32(32767.5) - np.float32(1.0)) * np.float32(noise_std)


def echo_source(num_pulses, samples_per_pulse, pulse_samples, sample_rate_hz, prf_hz,
                bandwidth_hz, target_delay_samples, doppler_hz, target_amplitude, noise_std,
                noise_seed):
    """Task1 step1 CPU reference (step1_validation.cu:24-80): LFM waveform with
    FP64 phase, one delayed Doppler-shifted echo (FP64 phase), FP32 hash noise.
    Float parameters are taken as FP32 numbers, as the reference's config holds
    them.  Returns (waveform [T], noiseless [P,S], noise [P,S], echo [P,S]),
    all complex64."""
    f32 = lambda v: float(np.float32(v))
    fs, prf, bw = f32(sample_rate_hz), f32(prf_hz), f32(bandwidth_hz)
    amp, fd = f32(target_amplitude), f32(doppler_hz)
    P, S, T, d = num_pulses, samples_per_pulse, pulse_samples, target_delay_samples
    pi = 3.14159265358979323846
    width = T / fs
    slope = bw / width
    n = np.arange(T, dtype=np.float64)
    t = n / fs - 0.5 * width
    ph = pi * slope * t * t
    w = (np.cos(ph).astype(np.float32) + 1j * np.sin(ph).astype(np.float32)).astype(np.complex64)
    idx = np.arange(P * S, dtype=np.int64)
    pulse, sample = idx // S, idx % S
    src = sample - d
    live = (src >= 0) & (src < T)
    time = pulse.astype(np.float64) / prf + sample.astype(np.float64) / fs
    phase = 2.0 * pi * fd * time
    c, s_ = np.cos(phase), np.sin(phase)
    wr = np.where(live, w.real[np.clip(src, 0, T - 1)].astype(np.float64), 0.0)
    wi = np.where(live, w.imag[np.clip(src, 0, T - 1)].astype(np.float64), 0.0)
    nl_re = np.where(live, (amp * (wr * c - wi * s_)), 0.0).astype(np.float32)
    nl_im = np.where(live, (amp * (wr * s_ + wi * c)), 0.0).astype(np.float32)
    base = 2 * idx.astype(np.uint64)
    nz_re = echo_uniform_noise(base, noise_seed, noise_std)
    nz_im = echo_uniform_noise(base + np.uint64(1), noise_seed, noise_std)
    noiseless = (nl_re + 1j * nl_im).astype(np.complex64).reshape(P, S)
    noise = (nz_re + 1j * nz_im).astype(np.complex64).reshape(P, S)
    echo = ((nl_re + nz_re) + 1j * (nl_im + nz_im)).astype(np.complex64).reshape(P, S)
    return w, noiseless, noise, echo


def cfar_alpha(pfa, reference):
    """radartools_typed.cpp:592-603 cfar_alpha_typed_cpu (FP64)."""
    return lib().orc_cfar_alpha(float(pfa), int(reference))


def ca_cfar_count(guard, ref):
    if len(guard) == 1:
        return 2 * ref[0]
    return lib().orc_ca_cfar_count(2, guard[0], guard[1], ref[0], ref[1])


def ca_cfar_alpha_f(guard, ref, pfa):
    rank = len(guard)
    g1 = guard[1] if rank == 2 else 0
    r1 = ref[1] if rank == 2 else 0
    return np.float32(lib().orc_ca_cfar_alpha_f(rank, guard[0], g1, ref[0], r1, float(pfa)))


def ca_cfar(x, guard, ref, pfa):
    """radartools_typed.cpp:465-555 ca_cfar_typed_cpu: FP32 integral image
    (axis 0 then axis 1), border threshold 0, strict '>' detection.
    x: real array, rank 1 or 2 (integer dtypes use the integer detection rule)."""
    x = np.asarray(x)
    rank = x.ndim
    rows = x.shape[0]
    cols = x.shape[1] if rank == 2 else 1
    ints = None
    lo = hi = 0.0
    if np.issubdtype(x.dtype, np.integer):
        ints = np.ascontiguousarray(x.astype(np.int64).reshape(-1))
        info = np.iinfo(x.dtype)
        lo, hi = float(info.min), float(info.max)
    xf = np.ascontiguousarray(x.astype(np.float32).reshape(-1))
    thr = np.empty(xf.size, np.float32)
    det = np.empty(xf.size, np.uint8)
    g1 = guard[1] if rank == 2 else 0
    r1 = ref[1] if rank == 2 else 0
    rc = lib().orc_ca_cfar(_p(xf), _p(ints), lo, hi, rank, rows, cols, guard[0], g1, ref[0],
                           r1, float(pfa), _p(thr), _p(det))
    if rc:
        raise ValueError(f"ca_cfar rc={rc}")
    return thr.reshape(x.shape), det.reshape(x.shape)


def ca_cfar_exact(x, guard, ref, pfa):
    """FP64 direct box sums with an FP64 alpha (the exact statistic)."""
    x = np.ascontiguousarray(np.asarray(x, np.float32))
    rows, cols = x.shape
    thr = np.empty(x.size, np.float64)
    lib().orc_ca_cfar_exact(_p(x), rows, cols, guard[0], guard[1], ref[0], ref[1],
                            float(pfa), _p(thr))
    return thr.reshape(rows, cols)


# --------------------------------------------------------------------------
# filtering ops
# --------------------------------------------------------------------------
def _lines(shape, axis):
    shape = list(shape)
    axis = axis + len(shape) if axis < 0 else axis
    outer = int(np.prod(shape[:axis])) if axis > 0 else 1
    inner = int(np.prod(shape[axis + 1:])) if axis + 1 < len(shape) else 1
    return outer, shape[axis], inner, axis


def upfirdn(h, x, up=1, down=1, axis=-1, fp64_acc=False):
    """filtering_typed.cpp:2548-2616 upfirdn_typed_cpu.  Float path sums float
    products in float; fp64_acc is the INT32 path.  Returns float64 values."""
    x = np.asarray(x)
    h = np.ascontiguousarray(np.asarray(h, np.float64).reshape(-1))
    outer, n, inner, ax = _lines(x.shape, axis)
    out_n = lib().orc_upfirdn_out_len(n, h.size, up, down)
    shape = list(x.shape)
    shape[ax] = out_n
    out = np.empty(int(np.prod(shape)), np.float64)
    rc = lib().orc_upfirdn(_p(h), h.size, _p(np.ascontiguousarray(x.astype(np.float64).reshape(-1))),
                           outer, n, inner, up, down, int(fp64_acc), _p(out))
    if rc:
        raise ValueError(f"upfirdn rc={rc}")
    return out.reshape(shape)


def resample_plan(n, up, down, custom_count=None):
    f = np.empty(8, np.int32)
    rc = lib().orc_resample_plan_fill(n, up, down, int(custom_count is not None),
                                      custom_count or 0, _p(f))
    if rc:
        raise ValueError(f"resample_poly rc={rc}")
    keys = ["up", "down", "identity", "half_length", "coefficient_count", "prefix_zeros",
            "output_start", "output_count"]
    return dict(zip(keys, map(int, f)))


def resample_design(up, down, window_mode="kaiser", beta=5.0, design_window=None):
    """filtering_typed.cpp:176-218 design_resample_poly_coefficients (FP64)."""
    plan = resample_plan(1, up, down)
    out = np.empty(plan["coefficient_count"], np.float64)
    dw = None if design_window is None else np.ascontiguousarray(design_window, np.float64)
    mode = {"kaiser": 0, "hamming": 1, "explicit": 2}[window_mode]
    lib().orc_resample_design(up, down, mode, float(beta), _p(dw), _p(out))
    return out


def resample_poly_core(taps, x, up, down, axis, fp64_acc):
    """Polyphase loop of filtering_typed.cpp:2650-2712 given scaled taps."""
    x = np.asarray(x)
    taps = np.ascontiguousarray(np.asarray(taps, np.float64))
    outer, n, inner, ax = _lines(x.shape, axis)
    plan = resample_plan(n, up, down, taps.size)
    shape = list(x.shape)
    shape[ax] = plan["output_count"]
    out = np.empty(int(np.prod(shape)), np.float64)
    rc = lib().orc_resample_poly(_p(taps), taps.size,
                                 _p(np.ascontiguousarray(x.astype(np.float64).reshape(-1))),
                                 outer, n, inner, up, down, int(fp64_acc), _p(out))
    if rc:
        raise ValueError(f"resample_poly rc={rc}")
    return out.reshape(shape)


def resample_poly_f32(x, up, down, axis=0, taps=None, window_mode="kaiser", beta=5.0):
    """resample_poly_typed_cpu for float32 input.  Custom taps: taps*up in
    float, FP32 polyphase sum, FP32 result.  Designed taps: FP64 design and
    FP64 sum (FP64 result).  Returns (values, dtype_name)."""
    x = np.asarray(x, np.float32)
    outer, n, inner, ax = _lines(x.shape, axis)
    plan = resample_plan(n, up, down, None if taps is None else len(taps))
    if plan["identity"]:
        return x.astype(np.float64), "FP32"
    if taps is not None:
        t = np.asarray(taps, np.float32)
        if plan["up"] < 65536:
            scaled = (t * np.float32(plan["up"])).astype(np.float64)
        else:
            scaled = t.astype(np.float64) * plan["up"]
        return resample_poly_core(scaled, x, up, down, axis, False), "FP32"
    coeffs = resample_design(up, down, window_mode, beta)
    return resample_poly_core(coeffs, x, up, down, axis, True), "FP64"


def _scaled_custom(value, up, dtype):
    """filtering_typed.cpp:129-157 resample_poly_scaled_custom_value."""
    def wrap(v, bits):
        v = int(v) & ((1 << bits) - 1)
        return v - (1 << bits) if v >= 1 << (bits - 1) else v
    if dtype == "FP32":
        return float(np.float32(value) * np.float32(up)) if up < 65536 else float(value) * up
    if dtype == "FP16":
        loaded = np.float32(np.float16(value))
        if up <= 255:
            return float(np.float32(np.float16(loaded * np.float32(up))))
        return float(loaded * np.float32(up)) if up < 65536 else float(loaded) * up
    v = int(value)
    if dtype == "INT32":
        return float(wrap(v * up, 32))
    if dtype == "INT16":
        return float(wrap(v * up, 16 if up < 32768 else 32))
    if up <= 127:
        return float(wrap(v * up, 8))
    return float(wrap(v * up, 16 if up < 32768 else 32))


def scaled_custom_device(value, up, dtype):
    """resample_poly_scaled_custom_value_device (filtering_kernels.cuh:473-497),
    the device form the GPU path uses: as _scaled_custom, except that float
    and half products stay in FP32 at every up, and the result is a float."""
    if dtype == "FP32":
        return np.float32(np.float32(value) * np.float32(up))
    if dtype == "FP16":
        v = np.float32(np.float16(value)) * np.float32(up)
        return np.float32(np.float16(v)) if up <= 255 else np.float32(v)
    return np.float32(_scaled_custom(value, up, dtype))


def cfar_alpha_f32(pfa, reference):
    """cfar_alpha_kernel (radartools_kernels.cuh:136-148): alpha in FP32 from
    the pfa loaded to float, n * (powf(p, -1/n) - 1)."""
    n = np.float32(reference)
    p = np.float32(pfa)
    return np.float32(n * (np.power(p, np.float32(-1.0) / n) - np.float32(1.0)))


def cfar_alpha_f32_tolerance(pfa, reference):
    """How far two FP32 evaluations of cfar_alpha_f32 may differ when their
    powf results differ by up to 2 ulp (powf is not correctly rounded on every
    platform): n * 2 ulp(p^(-1/n)) carried through the subtraction, plus 2 ulp
    of the result."""
    n = np.float32(reference)
    pw = np.power(np.float32(pfa), np.float32(-1.0) / n)
    a = cfar_alpha_f32(pfa, reference)
    return float(n) * 2 * float(np.spacing(pw)) + 2 * float(np.spacing(np.float32(max(abs(a), 1e-30))))


def resample_poly_typed(x, dtype, up, down, axis=0, taps=None, window_mode="kaiser", beta=5.0):
    """resample_poly_typed_cpu (filtering_typed.cpp:2617-2715) for any of the
    five input dtypes.  Returns (values as float64, output dtype name)."""
    x = np.asarray(x)
    outer, n, inner, ax = _lines(x.shape, axis)
    plan = resample_plan(n, up, down, None if taps is None else len(taps))
    if plan["identity"]:
        return x.astype(np.float64), dtype
    if taps is None:
        return resample_poly_core(resample_design(up, down, window_mode, beta), x, up, down,
                                  axis, True), "FP64"
    scaled = np.array([_scaled_custom(v, plan["up"], dtype) for v in np.asarray(taps).tolist()])
    if dtype == "INT32":
        fp32 = False
    elif dtype in ("INT16", "INT8"):
        fp32 = plan["up"] < 32768
    else:
        fp32 = plan["up"] < 65536
    if fp32:
        scaled = scaled.astype(np.float32).astype(np.float64)
    return resample_poly_core(scaled, x, up, down, axis, not fp32), "FP32" if fp32 else "FP64"


def firfilter(b, x, axis=-1, zi=None):
    """filtering_typed.cpp:1536-1592 firfilter_typed_cpu (FP64 accumulate -> f32).
    zi, if given, must already have the state shape (axis length K-1).
    Returns (y float32, zf float32 or None)."""
    x = np.asarray(x)
    b = np.ascontiguousarray(np.asarray(b, np.float64).reshape(-1))
    outer, n, inner, ax = _lines(x.shape, axis)
    y = np.empty(x.size, np.float32)
    zf = None
    zi_arr = None
    if zi is not None:
        zshape = list(x.shape)
        zshape[ax] = b.size - 1
        zi_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(zi, np.float64), zshape).reshape(-1))
        zf = np.empty(int(np.prod(zshape)), np.float32)
    rc = lib().orc_firfilter(_p(b), b.size, _p(np.ascontiguousarray(x.astype(np.float64).reshape(-1))),
                             outer, n, inner, _p(zi_arr), _p(y), _p(zf))
    if rc:
        raise ValueError(f"firfilter rc={rc}")
    if zf is not None:
        zshape = list(x.shape)
        zshape[ax] = b.size - 1
        zf = zf.reshape(zshape)
    return y.reshape(x.shape), zf


# --------------------------------------------------------------------------
# FP64 FFT forms (numpy), the fast exact ground truth at full sizes
# --------------------------------------------------------------------------
def pulse_compression_fft64(x, tmpl, nfft=-1, window="none", normalize=False):
    """Exact-mode pulse compression via FP64 FFTs: circular correlation of the
    zero-padded/truncated pulse with the windowed, normalised template."""
    x = np.asarray(x, np.complex128)
    P, S = x.shape
    n = S if nfft == -1 else nfft
    t = np.asarray(tmpl, np.complex128) * window_exact(window, len(tmpl))
    if normalize:
        t = t / math.sqrt(float(np.sum(np.abs(t) ** 2)))
    t = t[:n]
    xp = np.zeros((P, n), np.complex128)
    m = min(S, n)
    xp[:, :m] = x[:, :m]
    H = np.fft.fft(t, n)
    return np.fft.ifft(np.fft.fft(xp, axis=1) * np.conj(H)[None, :], axis=1)


def pulse_doppler_fft64(x, nfft=-1, window="none"):
    """Exact-mode Doppler processing via an FP64 FFT along pulses."""
    x = np.asarray(x, np.complex128)
    P, R = x.shape
    n = P if nfft == -1 else nfft
    w = window_exact(window, P)
    xw = x * w[:, None]
    used = min(P, n)
    return np.fft.fft(xw[:used], n, axis=0)


def window_exact(kind, length):
    return window(kind, length, exact=True)


def rel_l2(a, ref, floor=1e-12):
    """ZKX/test_all/support/metrics/accuracy.cpp:7-23: ||a-r||_2 / max(||r||_2, floor)."""
    a = np.asarray(a).astype(np.complex128).reshape(-1)
    r = np.asarray(ref).astype(np.complex128).reshape(-1)
    den = max(float(np.linalg.norm(r)), floor)
    return float(np.linalg.norm(a - r)) / den


# ---------------------------------------------------------------------------
# N-D convolution and correlation (FP64 ground truth)
# ---------------------------------------------------------------------------
def fft_axes(axes, ndim):
    """cuSignal's axes handling for the N-D transforms
    (cusignal/utils/fftpack_helper.py:279-360, convolve.py:257-266): None is
    every axis, negatives count from the end, out-of-range / repeated / empty
    given axes raise ValueError; returned sorted."""
    if axes is None:
        return list(range(ndim))
    ax = np.atleast_1d(np.asarray(axes))
    ax = [int(a) + ndim if int(a) < 0 else int(a) for a in ax]
    if any(a < 0 or a >= ndim for a in ax):
        raise ValueError("axes exceeds dimensionality of input")
    if len(set(ax)) != len(ax):
        raise ValueError("all axes must be unique")
    if not ax:
        raise ValueError("when provided, axes cannot be empty")
    return sorted(ax)


def _direct_conv_full(a, b, axes):
    """Full linear convolution over `axes` by direct FP64 sums (tiny inputs):
    every pair of elements whose non-axis coordinates broadcast together adds
    a[i] * b[j] at i + j on the convolved axes."""
    nd = a.ndim
    shape = [a.shape[d] + b.shape[d] - 1 if d in axes else max(a.shape[d], b.shape[d]) for d in range(nd)]
    out = np.zeros(shape, np.complex128)
    for i in np.ndindex(*a.shape):
        for j in np.ndindex(*b.shape):
            k, ok = [], True
            for d in range(nd):
                if d in axes:
                    k.append(i[d] + j[d])
                else:
                    if a.shape[d] != 1 and b.shape[d] != 1 and i[d] != j[d]:
                        ok = False
                        break
                    k.append(max(i[d] if a.shape[d] > 1 else 0, j[d] if b.shape[d] > 1 else 0))
            if ok:
                out[tuple(k)] += complex(a[i]) * complex(b[j])
    return out


def fftconvolve_nd(in1, in2, mode="full", axes=None, direct=False):
    """cuSignal fftconvolve (convolve.py:159-320, convolution_utils.py:40-100):
    rank and broadcast checks, 'valid' operand swap decided over every
    dimension, full extent s1 + s2 - 1 on the convolved axes and max(s1, s2)
    elsewhere, centred crop (_centered) to the mode's extents, a real result
    for two real inputs.  The full convolution is computed in complex128 by
    numpy FFTs of the full length (direct=True: by direct sums)."""
    a = np.asarray(in1)
    b = np.asarray(in2)
    if a.ndim == b.ndim == 0:
        return a * b
    if a.ndim != b.ndim:
        raise ValueError("in1 and in2 should have the same dimensionality")
    if a.size == 0 or b.size == 0:
        return np.array([])
    ax = fft_axes(axes, a.ndim)
    s1, s2 = np.array(a.shape), np.array(b.shape)
    other = [d for d in range(a.ndim) if d not in ax]
    for d in other:
        if not (s1[d] == s2[d] or s1[d] == 1 or s2[d] == 1):
            raise ValueError("incompatible shapes for in1 and in2")
    complex_result = np.iscomplexobj(a) or np.iscomplexobj(b)
    shape = np.maximum(s1, s2)
    shape[ax] = s1[ax] + s2[ax] - 1
    if mode == "valid":
        ok1 = bool(np.all(s1 >= s2))
        ok2 = bool(np.all(s2 >= s1))
        if not (ok1 or ok2):
            raise ValueError("For 'valid' mode, one must be at least as large as the other in every dimension")
        if not ok1:
            a, b, s1, s2 = b, a, s2, s1
    elif mode not in ("full", "same"):
        raise ValueError("acceptable mode flags are 'valid', 'same', or 'full'")
    a64 = a.astype(np.complex128)
    b64 = b.astype(np.complex128)
    if direct:
        ret = _direct_conv_full(a64, b64, ax)
    else:
        fshape = [int(shape[d]) for d in ax]
        ret = np.fft.ifftn(np.fft.fftn(a64, fshape, axes=ax) * np.fft.fftn(b64, fshape, axes=ax), axes=ax)
        ret = ret[tuple(slice(int(n)) for n in shape)]
    if not complex_result:
        ret = ret.real
    if mode == "full":
        keep = shape
    elif mode == "same":
        keep = s1
    else:
        keep = shape.copy()
        keep[ax] = s1[ax] - s2[ax] + 1
    start = (np.array(ret.shape) - keep) // 2
    return ret[tuple(slice(int(s), int(s + k)) for s, k in zip(start, keep))]


def correlate_nd(in1, in2, mode="full", dtype="FP32"):
    """correlate_typed_cpu for rank 0 / >= 2 (convolution_typed.cpp:153-199,
    349-410) with the FFT method's storage: extents full s1 + s2 - 1, same
    s1, valid |s1 - s2| + 1 per dimension ('valid' needs one input to
    dominate in every dimension); the full correlation
    z[f] = sum_k in1[f - (s2 - 1) + k] * in2[k] in FP64, the centred crop
    start (full - out) // 2, stored in `dtype` (integers rounded half to
    even and wrapped modulo 2^bits, half rounded to nearest)."""
    a = np.asarray(in1).astype(np.float64)
    b = np.asarray(in2).astype(np.float64)
    if a.ndim != b.ndim:
        raise ValueError("typed correlate inputs must have the same rank")
    if a.ndim == 0:
        full = np.asarray(np.float32(np.float32(a) * np.float32(b)), np.float64)
        return _store_typed(full, dtype)
    if a.size == 0 or b.size == 0:
        return np.zeros(0, NP_DTYPE[dtype])
    s1, s2 = np.array(a.shape), np.array(b.shape)
    if mode == "full":
        out = s1 + s2 - 1
    elif mode == "same":
        out = s1.copy()
    elif mode == "valid":
        if not (np.all(s1 >= s2) or np.all(s2 >= s1)):
            raise ValueError("typed correlate valid mode requires one input to dominate in every dimension")
        out = np.abs(s1 - s2) + 1
    else:
        raise ValueError("typed correlate mode must be full, same, or valid")
    full_shape = s1 + s2 - 1
    # correlation = convolution with in2 reversed in every dimension
    rb = b[tuple(slice(None, None, -1) for _ in range(b.ndim))]
    dims = list(range(a.ndim))
    full = np.real(np.fft.ifftn(np.fft.fftn(a, full_shape, axes=dims) * np.fft.fftn(rb, full_shape, axes=dims),
                                axes=dims))
    start = (full_shape - out) // 2
    return _store_typed(full[tuple(slice(int(s), int(s + k)) for s, k in zip(start, out))], dtype)


def _store_typed(v, dtype):
    v = np.asarray(v, np.float64)
    if INTEGRAL[dtype]:
        bits = _INT_BITS[dtype]
        w = np.mod(np.rint(v.astype(np.float32).astype(np.float64)), 2.0 ** bits).astype(np.uint64)
        return w.astype(np.dtype(f"uint{bits}")).view(NP_DTYPE[dtype])
    if dtype == "FP16":
        return v.astype(np.float32).astype(np.float16)
    return v.astype(np.float32)
