"""Field: the Taichi-field surface the reference's step loop uses, over a GPU tensor.

The reference allocates `ti.field(float, shape)` and `ti.Vector.field(n, float,
shape)` (run_edge.py:174-254) and calls `copy_from`, `fill`, `to_numpy`,
`from_numpy`, `shape` and `[None]` on them (run_edge.py:253,559-581,599-601,
621-624,644,694-707; sim/mgpcg.py:178,183-184).  A Field keeps the same
methods and meaning; the storage differs: a vector field of n components is
n contiguous planes (SoA, `tensor[c, i, j, k]`) so that the kernels read it
coalesced, while `to_numpy`/`from_numpy` keep Taichi's AoS `(..., n)` view.

Every write through this module or a kernel wrapper bumps a per-storage
version, which is how derived data (the Hermite tables) knows when to rebuild.
Code that writes `Field.data` directly must call `Field.touch()` afterwards.
"""
import itertools

import numpy as np
import torch

from . import _lib

# storage address -> version; versions come from one global counter, so a new
# allocation at a recycled address never matches a version seen before
_VERSIONS = {}
_COUNTER = itertools.count(1)


def touch_ptr(ptr):
    _VERSIONS[ptr] = next(_COUNTER)


def version_of(ptr):
    return _VERSIONS.get(ptr, 0)


def stream_handle():
    return torch.cuda.current_stream().cuda_stream


def _torch_dtype(dtype):
    if dtype in (None, float, "f32", np.float32, torch.float32):
        return torch.float32
    if dtype in (int, "i32", np.int32, torch.int32):
        return torch.int32
    if dtype in (np.float64, torch.float64, "f64"):
        return torch.float64
    raise TypeError(f"unsupported field dtype {dtype!r}")


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("edge_hip needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU path")


class Field:
    """A scalar (n=None) or n-vector field on a grid `shape`, stored on the GPU.

    Layout: an n-vector field is n planes (data shape (n,) + shape), except a
    4-vector (the smoke's RGBA), whose channels sit together per cell (data
    shape shape + (4,)): the smoke kernels gather a 16-byte tap."""

    def __init__(self, n, dtype, shape, device="cuda"):
        require_gpu()
        _lib.lib()  # fail loudly now if the HIP library is missing
        if isinstance(shape, int):
            shape = (shape,)
        self.shape = tuple(int(s) for s in shape)
        self.n = n
        self.dtype = _torch_dtype(dtype)
        self.channels_last = n == 4
        grid = self.shape if self.shape else (1,)
        if not n:
            tshape = grid
        elif self.channels_last:
            tshape = grid + (n,)
        else:
            tshape = (n,) + grid
        self.data = torch.zeros(tshape, dtype=self.dtype, device=device)
        self.touch()

    # -- identity / versions -------------------------------------------------
    @property
    def ptr(self):
        return self.data.data_ptr()

    def touch(self):
        touch_ptr(self.ptr)

    @property
    def version(self):
        return version_of(self.ptr)

    @property
    def numel(self):
        return self.data.numel()

    def swap_storage(self, other):
        """Exchange the storage of two same-shaped fields (a pointer exchange)."""
        assert self.data.shape == other.data.shape and self.dtype == other.dtype
        self.data, other.data = other.data, self.data

    # -- Taichi field methods ---------------------------------------------------
    def copy_from(self, other):
        assert other.data.shape == self.data.shape, (other.data.shape, self.data.shape)
        if self.dtype == torch.float32:
            _lib.call("edge_copy", other.ptr, self.ptr, self.numel, stream_handle())
        else:  # integer masks: setup data, moved as bytes
            self.data.copy_(other.data)
        self.touch()

    def fill(self, value):
        if self.dtype == torch.float32:
            _lib.call("edge_fill", self.ptr, float(value), self.numel, stream_handle())
        else:
            self.data.fill_(value)
        self.touch()

    def to_numpy(self):
        a = self.data.detach().cpu().numpy()
        if not self.shape:
            return a.reshape(())
        if self.n and not self.channels_last:
            return np.ascontiguousarray(np.moveaxis(a, 0, -1))
        return a.copy()

    def from_numpy(self, arr):
        arr = np.asarray(arr)
        if self.n:
            assert arr.shape == self.shape + (self.n,), (arr.shape, self.shape, self.n)
            if not self.channels_last:
                arr = np.moveaxis(arr, -1, 0)
        else:
            assert arr.shape == self.shape, (arr.shape, self.shape)
        t = torch.from_numpy(np.ascontiguousarray(arr)).to(self.dtype)
        self.data.copy_(t.reshape(self.data.shape))
        self.touch()

    def __getitem__(self, key):
        if key is None and not self.shape:
            return self.data.reshape(-1)[0].item()
        raise TypeError("only 0-d fields support [None]; index the numpy view instead")

    def __setitem__(self, key, value):
        if key is None and not self.shape:
            self.data.fill_(value)
            self.touch()
            return
        raise TypeError("only 0-d fields support [None] assignment")

    # -- SoA planes (kernels) -----------------------------------------------------
    def plane_ptr(self, c):
        if self.channels_last:
            raise TypeError("a 4-vector field keeps its channels per cell: it has no planes")
        return self.data[c].data_ptr()

    def __repr__(self):
        return f"Field(n={self.n}, shape={self.shape}, dtype={self.dtype})"


def as_ptr(x):
    """Device pointer of a Field or tensor (None -> NULL)."""
    if x is None:
        return None
    if isinstance(x, Field):
        return x.ptr
    if isinstance(x, torch.Tensor):
        return x.data_ptr()
    raise TypeError(f"expected Field or tensor, got {type(x)}")


def touch(*xs):
    for x in xs:
        if x is None:
            continue
        touch_ptr(as_ptr(x))


class _VectorNS:
    @staticmethod
    def field(n, dtype, shape):
        return Field(n, dtype, shape)


class ti:
    """The subset of the taichi namespace the step loop touches."""
    f32 = torch.float32
    f64 = torch.float64
    i32 = torch.int32
    Vector = _VectorNS

    @staticmethod
    def field(dtype, shape):
        return Field(None, dtype, shape)

    @staticmethod
    def sync():
        _lib.call("edge_stream_sync", stream_handle())

    @staticmethod
    def init(**kwargs):
        require_gpu()
        _lib.lib()
