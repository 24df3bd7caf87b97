"""Validation of caller-supplied output tensors, shared by the operator
wrappers: every output a kernel writes is checked before any launch, so a
host, undersized, expanded or foreign-device tensor is a ValueError instead of
an out-of-bounds device write."""
import torch


def out_tensor(t, name, like, dtype, numel):
    """Return `t` if it is a contiguous device tensor on `like`'s device with
    the given dtype and exactly `numel` elements; None passes through."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.device != like.device:
        raise ValueError(f"{name} must be on {like.device}, got {t.device}")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if t.numel() != numel:
        raise ValueError(f"{name} must have {numel} elements, got {t.numel()}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t
