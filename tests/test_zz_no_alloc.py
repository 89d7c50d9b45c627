"""The allocation-free particle scatter, edge_lp_p2g_noalloc, against the
allocating edge_lp_p2g through the C-ABI.

Two calls in a row, the second on a smaller grid with fewer particles (so the
shared device array holds the first call's keys and ranges when the second
runs): each result equals the allocating scatter of the same input (the same
kernels and the same stable sort order; the FP32 atomics of the tile boxes and
across tiles may add in another order, so 1e-6 of scale) and the oracle's
sequential scatter (1e-5, as the allocating scatter's own parity test).  Sizes
beyond the entry point's maximum are rejected on the host, before any launch.
"""
import numpy as np
import pytest
import torch


def scale_err(gpu, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(np.asarray(gpu, np.float64) - ref)) / max(np.max(np.abs(ref)), 1e-30))


def case(seed, dims, dx, n, num):
    """Particles clustered around a few centres (many per cell) plus uniform ones."""
    rng = np.random.default_rng(seed)
    L = np.array(dims) * dx
    k = n // 2
    centres = rng.uniform(0.3, 0.7, size=(3, 5)) * L[:, None]
    pick = rng.integers(0, 5, k)
    clustered = centres[:, pick] + rng.normal(scale=2.0 * dx, size=(3, k))
    uniform = rng.uniform(0, 1, size=(3, n - k)) * L[:, None]
    pos = np.clip(np.concatenate([clustered, uniform], 1), 0, L[:, None]).astype(np.float32)
    diff = rng.normal(size=(3, n)).astype(np.float32)
    rr = rng.uniform(0.5, 1.5, n).astype(np.float32)
    return pos, diff, rr, num


def scatter(name, dev, dims, dx, pos, diff, rr, num, radius=0.01):
    from edge_hip import _lib
    from edge_hip.field import stream_handle
    nx, ny, nz = dims
    n = pos.shape[1]
    P = torch.from_numpy(np.ascontiguousarray(pos)).to(dev)
    D = torch.from_numpy(np.ascontiguousarray(diff)).to(dev)
    R = torch.from_numpy(rr).to(dev)
    shapes = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    drag = [torch.full(s, 7.0, device=dev) for s in shapes]   # must come back zeroed where nothing lands
    _lib.call(name, *[d.data_ptr() for d in drag], D.data_ptr(), P.data_ptr(), R.data_ptr(), n, num, dx, radius,
              nx * dx, ny * dx, nz * dx, nx, ny, nz, stream_handle())
    torch.cuda.synchronize()
    return [d.cpu().numpy() for d in drag]


CALLS = [((64, 48, 40), 1.0 / 48, 300000, 290000, 21),   # first call
         ((32, 24, 20), 1.0 / 24, 40000, 39000, 22)]      # second, smaller


@pytest.mark.gpu
def test_noalloc_scatter_equals_the_allocating_one(orc32, dev):
    for dims, dx, n, num, seed in CALLS:
        pos, diff, rr, num = case(seed, dims, dx, n, num)
        print(f"p2g no-alloc: grid {dims}, {num} of {n} particles", flush=True)
        got = scatter("edge_lp_p2g_noalloc", dev, dims, dx, pos, diff, rr, num)
        ref = scatter("edge_lp_p2g", dev, dims, dx, pos, diff, rr, num)
        want = orc32.lp_p2g(diff, pos, rr, num, dx, 0.01, tuple(np.array(dims) * dx), dims)
        for a in range(3):
            assert np.abs(want[a]).max() > 0
            e_alloc, e_orc = scale_err(got[a], ref[a]), scale_err(got[a], want[a])
            print(f"  grid {a}: vs edge_lp_p2g {e_alloc:.3g}, vs the oracle {e_orc:.3g}")
            assert e_alloc < 1e-6
            assert e_orc < 1e-5


@pytest.mark.gpu
def test_noalloc_scatter_rejects_sizes_beyond_its_maximum(dev):
    """Over the maximum: an error, and the outputs untouched (every buffer is
    sized for its call, so nothing would be out of bounds either way)."""
    from edge_hip import _lib
    from edge_hip.field import stream_handle
    dx = 1.0 / 24
    # 2^20 + 1 particles: one more than the entry point holds
    n = (1 << 20) + 1
    dims = (32, 24, 20)
    P = torch.full((3, n), 0.5, device=dev)
    D = torch.zeros((3, n), device=dev)
    R = torch.ones(n, device=dev)
    drag = [torch.zeros(s, device=dev) for s in [(33, 24, 20), (32, 25, 20), (32, 24, 21)]]
    with pytest.raises(RuntimeError, match="exceed its maximum"):
        _lib.call("edge_lp_p2g_noalloc", *[d.data_ptr() for d in drag], D.data_ptr(), P.data_ptr(), R.data_ptr(), n,
                  n, dx, 0.01, 32 * dx, 24 * dx, 20 * dx, *dims, stream_handle())
    # a grid above 4,194,304 tile-padded cells (264 x 128 x 128 pads to 272 x 128 x 128)
    big = [torch.zeros(s, device=dev) for s in [(265, 128, 128), (264, 129, 128), (264, 128, 129)]]
    with pytest.raises(RuntimeError, match="exceed its maximum"):
        _lib.call("edge_lp_p2g_noalloc", *[d.data_ptr() for d in big], D.data_ptr(), P.data_ptr(), R.data_ptr(), 10,
                  10, dx, 0.01, 264 * dx, 128 * dx, 128 * dx, 264, 128, 128, stream_handle())
    torch.cuda.synchronize()
    assert all(float(d.abs().max()) == 0.0 for d in drag + big)


@pytest.mark.parametrize("n,dims", [((1 << 20) + 1, (32, 24, 20)), (10, (264, 128, 128))])
def test_noalloc_limits_checked_on_the_host(n, dims):
    """The same rejections without a GPU: the checks run before any HIP call
    (placeholder addresses, never dereferenced)."""
    from edge_hip import _lib
    import os
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libedge_hip.so not built")
    dx = 1.0 / 24
    fake = [64 * (i + 1) for i in range(6)]
    with pytest.raises(RuntimeError, match="exceed its maximum"):
        _lib.call("edge_lp_p2g_noalloc", *fake, n, n, dx, 0.01, dims[0] * dx, dims[1] * dx, dims[2] * dx, *dims,
                  None)
