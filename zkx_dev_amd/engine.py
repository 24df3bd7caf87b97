"""Heterogeneous dwell batch engine (C ABI zkx_engine_*) on torch device tensors.

Many dwells of different pulse counts P (powers of two, 16..256), range
lengths R (any; powers of two up to 8192 use one circular FFT per pulse,
others overlap-save segments) and resident chirp templates run as one batch:
two persistent launches replayed from a HIP graph.  Per dwell the result is
the reference chain pulse_compression (nfft = R) -> pulse_doppler (nfft = P)
-> |.|^2 -> CA-CFAR (FP64 windowed sums), as in Task1 steps 2-4
(Task1/pipeline/src/steps/step{2,3,4}).

Dwells are packed back to back: `pack()` lays out a list of [P, R] cubes in
one flat complex64 buffer and returns the descriptor table with input and
output offsets.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib
from .radar import _ptr, _stream

WINDOWS = {"none": 0, "hann": 1, "hamming": 2}

DESC_DTYPE = np.dtype([("num_pulses", np.int32), ("range_bins", np.int32), ("template_id", np.int32),
                       ("reserved", np.int32), ("input_offset", np.int64), ("output_offset", np.int64)])
assert DESC_DTYPE.itemsize == 32


def make_descs(shapes, template_ids):
    """Descriptor table for dwells of shapes [(P, R), ...] packed back to back
    (inputs and outputs both at the running sum of P*R)."""
    d = np.zeros(len(shapes), DESC_DTYPE)
    off = 0
    for i, ((p, r), t) in enumerate(zip(shapes, template_ids)):
        d[i] = (p, r, t, 0, off, off)
        off += p * r
    return d, off


class DwellEngine:
    """zkx_engine: resident templates, per-lane job tables, graph and workspace."""

    def __init__(self, pc_window="hamming", normalize=True, doppler_window="hann", guard=(2, 2),
                 reference=(8, 8), pfa=1e-6, lanes=1):
        h = ctypes.c_void_p()
        check(lib().zkx_engine_create(ctypes.byref(h), WINDOWS[pc_window], int(bool(normalize)),
                                      WINDOWS[doppler_window], guard[0], guard[1], reference[0],
                                      reference[1], float(pfa), int(lanes)))
        self._h = h
        self.lanes = lanes
        self.guard, self.reference, self.pfa = tuple(guard), tuple(reference), pfa
        self.pc_window, self.doppler_window, self.normalize = pc_window, doppler_window, normalize
        self._templates = []

    def close(self):
        if self._h:
            check(lib().zkx_engine_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_template(self, h):
        """Register a complex64 template (device or host); returns its id."""
        t = torch.as_tensor(h)
        if t.dtype != torch.complex64:
            t = t.to(torch.complex64)
        t = t.contiguous()
        if not t.is_cuda:
            t = t.cuda()
        tid = ctypes.c_int()
        check(lib().zkx_engine_add_template(self._h, _ptr(t), t.numel(), ctypes.byref(tid)))
        torch.cuda.synchronize()
        self._templates.append(t.numel())
        return tid.value

    def set_options(self, use_graph=True, pc_grid=0, dc_grid=0):
        check(lib().zkx_engine_set_options(self._h, int(bool(use_graph)), int(pc_grid), int(dc_grid)))

    @staticmethod
    def _descs(descs):
        d = np.ascontiguousarray(descs, DESC_DTYPE)
        return d, ctypes.c_void_p(d.ctypes.data)

    def run(self, x, descs, threshold, detections, rd_map=None, lane=0, stream=None):
        """x: flat complex64 device buffer; threshold float32 / detections
        uint8 (and rd_map complex64) flat device buffers covering every
        dwell's output range."""
        d, dp = self._descs(descs)
        for t, dt in ((x, torch.complex64), (threshold, torch.float32), (detections, torch.uint8)):
            if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != dt or not t.is_contiguous():
                raise ValueError("engine buffers must be contiguous device tensors of the documented dtypes")
        if rd_map is not None and (rd_map.dtype != torch.complex64 or not rd_map.is_cuda or
                                   not rd_map.is_contiguous() or rd_map.device != x.device):
            raise ValueError("rd_map must be a contiguous complex64 tensor on the input's device")
        if threshold.device != x.device or detections.device != x.device:
            raise ValueError("engine outputs must be on the input's device")
        if len(d):
            need_in = int(np.max(d["input_offset"] + d["num_pulses"].astype(np.int64) * d["range_bins"]))
            need_out = int(np.max(d["output_offset"] + d["num_pulses"].astype(np.int64) * d["range_bins"]))
            if x.numel() < need_in or threshold.numel() < need_out or detections.numel() < need_out or \
                    (rd_map is not None and rd_map.numel() < need_out):
                raise ValueError("engine buffers are smaller than the dwell table needs")
        check(lib().zkx_engine_run(self._h, int(lane), _ptr(x), dp, len(d), _ptr(threshold), _ptr(detections),
                                   _ptr(rd_map), _stream(stream)))

    def plan(self, descs):
        d, dp = self._descs(descs)
        n = np.zeros(len(d), np.int32)
        s = np.zeros(len(d), np.int32)
        check(lib().zkx_engine_plan(self._h, dp, len(d), ctypes.c_void_p(n.ctypes.data),
                                    ctypes.c_void_p(s.ctypes.data)))
        return n, s

    def autotune(self, x, descs, threshold, detections, lane=0, stream=None):
        d, dp = self._descs(descs)
        check(lib().zkx_engine_autotune(self._h, int(lane), _ptr(x), dp, len(d), _ptr(threshold),
                                        _ptr(detections), _stream(stream)))

    def tuning(self):
        n = ctypes.c_int()
        check(lib().zkx_engine_tuning_export(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        check(lib().zkx_engine_tuning_export(self._h, buf, n.value + 1, ctypes.byref(n)))
        return buf.value.decode()

    def load_tuning(self, text):
        check(lib().zkx_engine_tuning_import(self._h, text.encode()))


__all__ = ["DwellEngine", "make_descs", "DESC_DTYPE"]
