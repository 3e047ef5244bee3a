"""Parametric encoder: the FP64 oracle pinned on the CPU, the HIP kernels
against it on the GPU.

Bars:
* oracle: central finite differences (relative 1e-6) and the reference's
  pair-gradient restatement (oracle.pacmap_grad, gradient.rs:98-151, 1e-5);
* GPU gradient and transform vs the oracle with the GPU's bf16 rounding
  points emulated: 2e-3 norm-wise (FP32 vs FP64 accumulation and rare
  one-ulp bf16 rounding flips), and 5e-2 vs the pure FP64 gradient;
* training: deterministic (bit-identical parameters on a re-run) and the
  epoch loss falls.
"""
import numpy as np
import pytest

import mlp_oracle as MO
import oracle as O

D, H, DIMS = 20, 64, 2


def _pairs(rng, n, counts=(300, 150, 600)):
    out = []
    for c in counts:
        p = rng.integers(0, n, size=(c, 2)).astype(np.uint32)
        p[:5, 1] = p[:5, 0]  # self pairs contribute nothing (gradient.rs:116-118)
        out.append(p)
    return out


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n = 40
    x = rng.normal(size=(n, D)).astype(np.float32)
    p = rng.normal(scale=0.3, size=MO.param_count(D, H, DIMS))
    nn, mn, fp = _pairs(rng, n, (30, 15, 60))
    w = (2.0, 3.0, 1.0)
    f = lambda q: MO.loss_and_grad(q, x, nn, mn, fp, w, D, H, DIMS)[0]  # noqa: E731
    _, g = MO.loss_and_grad(p, x, nn, mn, fp, w, D, H, DIMS)
    idx = rng.choice(p.size, 60, replace=False)
    idx = np.concatenate([idx, [p.size - 1, p.size - 3, H * D, H * D + H + 5]])  # biases too
    eps = 1e-6
    fd = np.array([(f(p + eps * np.eye(1, p.size, i)[0]) - f(p - eps * np.eye(1, p.size, i)[0])) / (2 * eps)
                   for i in idx])
    err = np.abs(fd - g[idx]).max() / np.abs(g[idx]).max()
    assert err < 1e-6, err


def test_oracle_pair_gradient_matches_reference_restatement():
    """dL/dY of the encoder loss, summed per point, is gradient.rs's gradient."""
    rng = np.random.default_rng(1)
    n = 50
    y = rng.normal(size=(n, DIMS)).astype(np.float32)
    nn, mn, fp = _pairs(rng, n, (40, 20, 80))
    w = (2.0, 3.0, 1.0)
    pairs, kind = MO.typed_pairs(nn, mn, fp)
    loss, gi = MO.pair_terms(y[pairs[:, 0]].astype(np.float64), y[pairs[:, 1]].astype(np.float64),
                             kind, pairs[:, 0] == pairs[:, 1], np.array(w))
    g = np.zeros((n, DIMS))
    np.add.at(g, pairs[:, 0], gi)
    np.add.at(g, pairs[:, 1], -gi)
    ref = O.pacmap_grad(y, nn, mn, fp, *w).astype(np.float64)
    assert np.abs(g - ref[:n]).max() <= 1e-5 * np.abs(ref[:n]).max()
    assert abs(loss.sum() - ref[n, 0]) <= 1e-5 * abs(ref[n, 0])


def test_oracle_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-39], np.float32)
    r = MO.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0 + 2 ** -6                # 1 + 3/256 -> 1 + 4/256
    assert r[3] == -2.5
    assert np.abs(r[4] - 3.0e-39) < 2 ** -133   # subnormal kept to bf16 precision


@pytest.mark.gpu
def test_mlp_gradient_matches_oracle():
    import pacmap_mi355x as pm
    rng = np.random.default_rng(2)
    n, d, h = 500, 100, 256
    x = rng.normal(size=(n, d)).astype(np.float32)
    p = pm.mlp.init(d, h, DIMS, seed=5)
    p[-DIMS:] = 0.0
    nn, mn, fp = _pairs(rng, n)
    w = np.array([2.0, 3.0, 1.0], np.float32)
    g, loss = pm.mlp.grad(x, p, nn, mn, fp, w, hidden=h, dims=DIMS)
    lo, go = MO.loss_and_grad(p, x, nn, mn, fp, w, d, h, DIMS, emulate=True)
    l64, g64 = MO.loss_and_grad(p, x, nn, mn, fp, w, d, h, DIMS, emulate=False)
    blocks = np.cumsum([0, h * d, h, h * h, h, DIMS * h, DIMS])
    # the loss depends on y_i - y_j only, so its b3 gradient is exactly zero:
    # the GPU's is rounding noise of the sum of +-dY
    gb3 = g[blocks[5]:]
    assert np.abs(gb3).max() <= 1e-4 * np.abs(g[blocks[4]:blocks[5]]).max()
    assert np.abs(go[blocks[5]:]).max() <= 1e-12 * np.abs(go[blocks[4]:blocks[5]]).max()
    for a, b, name in zip(blocks[:-2], blocks[1:-1], ["W1", "b1", "W2", "b2", "W3"]):
        e = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        e64 = np.linalg.norm(g[a:b] - g64[a:b]) / np.linalg.norm(g64[a:b])
        print(f"{name}: rel err vs bf16-emulating oracle {e:.2e}, vs FP64 {e64:.2e}", flush=True)
        assert e < 2e-3, name
        assert e64 < 5e-2, name
    assert abs(loss - lo) <= 1e-3 * abs(lo)


@pytest.mark.gpu
def test_mlp_transform_matches_oracle():
    import pacmap_mi355x as pm
    rng = np.random.default_rng(3)
    n, d, h = 1000, 37, 128
    x = rng.normal(size=(n, d)).astype(np.float32)
    p = pm.mlp.init(d, h, 3, seed=9)
    y = pm.mlp.transform(x, p, hidden=h, dims=3)
    ref = MO.transform(p, x, d, h, 3, emulate=True)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"transform max err {err:.2e}", flush=True)
    assert err < 2e-3


@pytest.mark.gpu
@pytest.mark.parametrize("n,d,dims", [
    (1, 100, 2), (33, 100, 1), (1000, 100, 2), (1000, 37, 3), (33, 101, 8),
    (1000, 2176, 8),   # widest input whose block still fits in LDS with 8 outputs
    (300, 3000, 2),    # block LDS above 160 KiB: the unfused per-layer kernels
])
def test_mlp_transform_hidden256_matches_oracle(n, d, dims):
    """The default hidden width (256) runs the fused transform kernel while its
    LDS fits, the per-layer kernels otherwise; both against the bf16-emulating
    oracle (2e-3 of the largest output), and host and device inputs agree
    bit for bit."""
    import pacmap_mi355x as pm
    rng = np.random.default_rng(n + d + dims)
    x = rng.normal(size=(n, d)).astype(np.float32)
    p = pm.mlp.init(d, 256, dims, seed=d)
    p[-dims:] = rng.normal(size=dims).astype(np.float32)  # a nonzero output bias
    print(f"transform n={n} d={d} dims={dims}", flush=True)
    y = pm.mlp.transform(x, p, hidden=256, dims=dims)
    xd = pm.DeviceArray(x)
    yd = pm.mlp.transform(xd, p, hidden=256, dims=dims)
    ref = MO.transform(p, x, d, 256, dims, emulate=True)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"  max err {err:.2e}", flush=True)
    assert err < 2e-3
    assert np.array_equal(y, yd)


@pytest.mark.gpu
def test_mlp_fit_deterministic_and_loss_falls():
    import pacmap_mi355x as pm
    rng = np.random.default_rng(4)
    n, d = 3000, 20
    centres = rng.normal(scale=3.0, size=(5, d))
    x = (centres[rng.integers(0, 5, n)] + rng.normal(size=(n, d))).astype(np.float32)
    nn, mn, fp, xp = pm.generate_pairs(x, pm.Configuration(seed=1))
    p0 = pm.mlp.init(xp.shape[1], 64, 2, seed=1)
    kw = dict(hidden=64, epochs=4, batch_pairs=2048, learning_rate=1e-3, seed=3)
    p1, l1 = pm.mlp.fit(xp, nn, mn, fp, p0, **kw)
    p2, l2 = pm.mlp.fit(xp, nn, mn, fp, p0, **kw)
    print("epoch loss", l1, flush=True)
    assert np.array_equal(p1, p2) and np.array_equal(l1, l2)
    assert np.isfinite(p1).all()
    assert l1[-1] < l1[0]


@pytest.mark.gpu
def test_mlp_transform_kept_weights_follow_parameter_changes():
    """The fused transform keeps its converted weights between calls while
    the parameters stay byte-equal: a change made in place to the same array
    (a trained step), a different shape, and a release of the cached memory
    must each give the result of a first call on those parameters."""
    import pacmap_mi355x as pm
    rng = np.random.default_rng(21)
    n, d = 777, 100
    x = pm.DeviceArray(rng.normal(size=(n, d)).astype(np.float32))
    p = pm.mlp.init(d, 256, 2, seed=4)
    y1 = pm.mlp.transform(x, p)
    assert np.array_equal(pm.mlp.transform(x, p), y1)  # reused weights, same bits
    p[5] += 0.25  # W1 entry, changed in place
    p[-1] += 1.0  # output bias
    y2 = pm.mlp.transform(x, p)
    assert not np.array_equal(y2, y1)
    pm.release_cached_memory()
    assert np.array_equal(pm.mlp.transform(x, p), y2)  # rebuilt from scratch: same bits
    q = pm.mlp.init(d, 256, 3, seed=4)  # another shape
    y3 = pm.mlp.transform(x, q, dims=3)
    assert y3.shape == (n, 3)
    assert np.array_equal(pm.mlp.transform(x, p), y2)


@pytest.mark.gpu
def test_generate_pairs_match_fit_transform_pipeline():
    """generate_pairs returns the pairs fit_transform optimises: with them as
    AllProvided (and the same Value init) the two embeddings are identical."""
    import pacmap_mi355x as pm
    rng = np.random.default_rng(6)
    n, d = 1500, 30
    x = rng.normal(size=(n, d)).astype(np.float32)
    y0 = rng.normal(size=(n, 2)).astype(np.float32)
    cfg = pm.Configuration(seed=11, initialization=pm.Initialization.Value(y0), num_iters=(10, 10, 10))
    nn, mn, fp, xp = pm.generate_pairs(x, cfg)
    k, n_mn, n_fp = 10, 5, 20
    assert nn.shape == (n * k, 2) and mn.shape == (n * n_mn, 2) and fp.shape == (n * n_fp, 2)
    assert xp.shape == (n, d)
    ya, _ = pm.fit_transform(x, cfg)
    cfg2 = pm.Configuration(seed=11, initialization=pm.Initialization.Value(y0), num_iters=(10, 10, 10),
                            pair_configuration=pm.PairConfiguration.AllProvided(nn, mn, fp))
    yb, _ = pm.fit_transform(x, cfg2)
    assert np.array_equal(ya, yb)
