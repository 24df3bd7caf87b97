"""Host-side mirror of the reference's correlate API on torch device tensors.

Follows ZKX/cusignal_cpp/include/cusignal/operators/convolution/
convolution_typed.h (correlate_device(in1, in2, out, mode), the direct and
FFT methods in 1-D, the N-D FFT method, and correlate2d_device) in
cusignal's call forms: inputs of one element type (float32, float16, int32,
int16, int8), output in that type.  The direct
method keeps the reference's arithmetic bit for bit (an FP32 fma chain for
floating types, sums modulo 2^bits for integers).  All compute runs in
libzkx_radar.so.
"""
from __future__ import annotations

import torch

from ._check import out_tensor
from ._lib import INVALID_ARGUMENT, INVALID_DTYPE, INVALID_SHAPE, ArgumentError, check, lib
from .radar import _ptr, _stream

_DTYPES = {torch.float32: 0, torch.float16: 1, torch.int32: 2, torch.int16: 3, torch.int8: 4}
_MODES = {"full": 0, "same": 1, "valid": 2}


def correlate(in1, in2, mode="full", method="auto", *, out=None, stream=None):
    """Cross-correlation of two device tensors of one dtype and one rank.

    Rank 1: method "direct" (the reference's direct sums, bit for bit), "fft"
    (the reference's workspace overload: complex FP32 FFT, integers rounded
    half to even and wrapped) or "auto" (direct, as the reference's overload
    without a workspace).  Rank 0 or >= 2 (correlate_device with a
    CorrelateNDWorkspace, convolution_typed.cu:150-241): "fft" or "auto"
    (FFT over every dimension); "direct" is refused as the reference refuses
    it."""
    import ctypes

    if mode not in _MODES:
        raise ArgumentError(INVALID_ARGUMENT, "typed correlate mode must be full, same, or valid")
    if method not in ("direct", "auto", "fft"):
        raise ArgumentError(INVALID_ARGUMENT, "typed correlate method must be direct, fft, or auto")
    for t, name in ((in1, "in1"), (in2, "in2")):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise ArgumentError(INVALID_ARGUMENT, f"{name} must be a device tensor")
    if in1.dtype != in2.dtype or in1.dtype not in _DTYPES:
        raise ArgumentError(INVALID_DTYPE, "correlate inputs must share one of float32, float16, int32, int16, int8")
    if in1.dim() != in2.dim():
        raise ArgumentError(INVALID_SHAPE, "typed correlate inputs must have the same rank")
    a, b = in1.contiguous(), in2.contiguous()
    if a.dim() != 1:
        if method == "direct" and a.dim() > 1:
            raise ArgumentError(INVALID_ARGUMENT, "typed correlate direct method is only implemented for 1D")
        rank = a.dim()
        s1 = (ctypes.c_int * max(rank, 1))(*a.shape)
        s2 = (ctypes.c_int * max(rank, 1))(*b.shape)
        dims = (ctypes.c_int * max(rank, 1))()
        count = ctypes.c_longlong(0)
        check(lib().zkx_correlate_nd_out_shape(s1, s2, rank, _MODES[mode], dims, ctypes.byref(count)))
        shape = () if rank == 0 else ((0,) if count.value == 0 else tuple(int(v) for v in dims[:rank]))
        if out is None:
            out = torch.empty(shape, dtype=a.dtype, device=a.device)
        elif out.numel() != count.value or out.dtype != a.dtype:
            raise ArgumentError(INVALID_SHAPE, "typed correlate output shape mismatch")
        out_tensor(out, "out", a, a.dtype, count.value)
        check(lib().zkx_correlate_nd(_ptr(a), s1, _ptr(b), s2, rank, _DTYPES[a.dtype], _MODES[mode], _ptr(out),
                                     _stream(stream)))
        return out
    n = lib().zkx_correlate_out_len(a.numel(), b.numel(), _MODES[mode])
    if out is None:
        out = torch.empty(max(n, 0), dtype=a.dtype, device=a.device)
    elif out.numel() != n or out.dtype != a.dtype:
        raise ArgumentError(INVALID_SHAPE, "typed correlate output shape mismatch")
    out_tensor(out, "out", a, a.dtype, max(n, 0))
    fn = lib().zkx_correlate_fft if method == "fft" else lib().zkx_correlate_direct
    check(fn(_ptr(a), a.numel(), _ptr(b), b.numel(), _DTYPES[a.dtype], _MODES[mode], _ptr(out), _stream(stream)))
    return out


_BOUNDARIES = {"fill": 0, "wrap": 1, "symm": 2}


def correlate2d(in1, in2, mode="full", boundary="fill", fillvalue=0, *, out=None, stream=None):
    """2-D cross-correlation of two device matrices of the same dtype
    (cusignal.correlate2d's call form; correlate2d_device,
    convolution_typed.h:374-379).  boundary 'fill' pads with `fillvalue`
    stored in the element type, 'wrap' is circular, 'symm' mirrors with the
    edge repeated; 'valid' ignores both."""
    import ctypes

    if mode not in _MODES:
        raise ValueError("typed correlate2d mode must be full, same, or valid")
    if boundary not in _BOUNDARIES:
        raise ValueError("typed correlate2d boundary must be fill, wrap, or symm")
    for t, name in ((in1, "in1"), (in2, "in2")):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise ValueError(f"{name} must be a device tensor")
    if in1.dtype != in2.dtype or in1.dtype not in _DTYPES:
        raise ValueError("correlate2d inputs must share one of float32, float16, int32, int16, int8")
    if in1.dim() != 2 or in2.dim() != 2:
        raise ValueError("correlate2d inputs must both be 2-D arrays")
    a, b = in1.contiguous(), in2.contiguous()
    rows, cols = ctypes.c_longlong(0), ctypes.c_longlong(0)
    code = lib().zkx_correlate2d_out_shape(a.shape[0], a.shape[1], b.shape[0], b.shape[1], _MODES[mode],
                                           ctypes.byref(rows), ctypes.byref(cols))
    if code in (1001, 1002):
        raise ValueError(lib().zkx_last_error_message().decode())
    check(code)
    shape = (rows.value, cols.value)
    if out is None:
        out = torch.empty(shape, dtype=a.dtype, device=a.device)
    elif tuple(out.shape) != shape or out.dtype != a.dtype or not out.is_contiguous():
        raise ValueError("typed correlate2d output shape mismatch")
    out_tensor(out, "out", a, a.dtype, shape[0] * shape[1])
    code = lib().zkx_correlate2d(_ptr(a), a.shape[0], a.shape[1], _ptr(b), b.shape[0], b.shape[1],
                                 _DTYPES[a.dtype], _MODES[mode], _BOUNDARIES[boundary], float(fillvalue),
                                 _ptr(out), _stream(stream))
    if code in (1001, 1002):
        raise ValueError(lib().zkx_last_error_message().decode())
    check(code)
    return out


__all__ = ["correlate", "correlate2d"]
