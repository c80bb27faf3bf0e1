"""Device-array plumbing: torch tensors hold device memory, the C ABI sees raw
pointers.  Grid fields are Julia-layout (column-major, y index fastest): a
field of reference shape (ny, nx) is a torch tensor of shape (ny, nx) whose
strides are (1, ny), i.e. the transpose of a C-contiguous (nx, ny) tensor.
"""
from __future__ import annotations

import numpy as np
import torch

_DTYPES = {
    np.dtype(np.float64): torch.float64,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.int8): torch.int8,
    np.dtype(np.int16): torch.int16,
    np.dtype(np.int64): torch.int64,
}


def device(index: int = 0) -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the EarthBox GPU path needs a HIP device")
    return torch.device("cuda", index)


def to_device(a: np.ndarray, dev: torch.device | None = None) -> torch.Tensor:
    """Copy a numpy vector or matrix to the device in Julia (column-major) layout."""
    dev = dev or device()
    a = np.asarray(a)
    if a.ndim == 1:
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    if a.ndim == 2:
        t = torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)
        return t.T
    if a.ndim == 3:
        t = torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).to(dev)
        return t.permute(2, 1, 0)
    raise ValueError("only vectors, matrices and 3-D arrays are supported")


def empty(shape, dtype, dev: torch.device | None = None) -> torch.Tensor:
    dev = dev or device()
    tdt = _DTYPES[np.dtype(dtype)]
    if isinstance(shape, int) or len(shape) == 1:
        return torch.empty(shape, dtype=tdt, device=dev)
    if len(shape) == 3:
        ny, nx, nz = shape
        return torch.empty((nz, nx, ny), dtype=tdt, device=dev).permute(2, 1, 0)
    ny, nx = shape
    return torch.empty((nx, ny), dtype=tdt, device=dev).T


def zeros(shape, dtype, dev: torch.device | None = None) -> torch.Tensor:
    t = empty(shape, dtype, dev)
    t.zero_()
    return t


def to_host(t: torch.Tensor) -> np.ndarray:
    """Device tensor -> numpy array with the reference's (ny, nx) indexing."""
    if t.dim() == 1:
        return t.cpu().numpy()
    if t.dim() == 3:
        return t.permute(2, 1, 0).contiguous().cpu().numpy().transpose(2, 1, 0)
    return np.asfortranarray(t.T.contiguous().cpu().numpy().T)


def ptr(t: torch.Tensor | None, dtype=None, numel: int | None = None) -> int | None:
    """Raw device pointer of a Julia-layout tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise TypeError("expected a device tensor")
    if dtype is not None and t.dtype != _DTYPES[np.dtype(dtype)]:
        raise TypeError(f"expected dtype {dtype}, got {t.dtype}")
    if t.dim() == 2:
        if not t.T.is_contiguous():
            raise ValueError("2-D fields must be column-major (ny, nx) views")
    elif t.dim() == 3:
        # column-major grid fields (ny, nx, nz) or row-major images (H, W, C)
        if not (t.permute(2, 1, 0).is_contiguous() or t.is_contiguous()):
            raise ValueError("3-D arrays must be dense column-major fields or row-major images")
    elif not t.is_contiguous():
        raise ValueError("vectors must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"expected {numel} elements, got {t.numel()}")
    return t.data_ptr()
