"""The allocation-free encoder transform (pacmap_mlp_transform_noalloc)
against the allocating one (pacmap_mlp_transform_dev), through the C ABI.

Both run the same kernels on the same inputs, so the bar is bit-identical
outputs.  Two calls in a row, the second with a smaller input (fewer points,
a narrower input, other weights and output width), check that the static
workspace keeps nothing from one call that the next depends on.  Sizes past
the workspace are rejected before any device work.
"""
import numpy as np
import pytest

import pacmap_mi355x as pm


def _noalloc(x, p, dims):
    xd, pd = pm.DeviceArray(x), pm.DeviceArray(p)
    yd = pm.DeviceArray(np.zeros((x.shape[0], dims), np.float32))
    pm.mlp.transform_noalloc(xd, pd, yd, hidden=256, dims=dims)
    return yd.numpy()


@pytest.mark.gpu
def test_noalloc_transform_matches_allocating_entry_point():
    rng = np.random.default_rng(77)
    for n, d, dims in ((20000, 784, 3), (1000, 37, 2)):
        print(f"noalloc transform n={n} d={d} dims={dims}", flush=True)
        x = rng.normal(size=(n, d)).astype(np.float32)
        p = pm.mlp.init(d, 256, dims, seed=n)
        p[-dims:] = rng.normal(size=dims).astype(np.float32)
        ref = pm.mlp.transform(pm.DeviceArray(x), p, hidden=256, dims=dims)
        got = _noalloc(x, p, dims)
        assert np.isfinite(ref).all()
        assert np.array_equal(got, ref)


@pytest.mark.gpu
def test_noalloc_transform_rejects_sizes_past_its_workspace():
    x = pm.DeviceArray(np.zeros((4, 2049), np.float32))
    p = pm.DeviceArray(np.zeros(pm.mlp.param_count(2049, 256, 2), np.float32))
    y = pm.DeviceArray(np.zeros((4, 2), np.float32))
    with pytest.raises(pm.Unsupported) as e:
        pm.mlp.transform_noalloc(x, p, y, hidden=256, dims=2)
    assert "2048" in str(e.value)
    x2 = pm.DeviceArray(np.zeros((4, 16), np.float32))
    p2 = pm.DeviceArray(np.zeros(pm.mlp.param_count(16, 128, 2), np.float32))
    with pytest.raises(pm.Unsupported):
        pm.mlp.transform_noalloc(x2, p2, y, hidden=128, dims=2)
