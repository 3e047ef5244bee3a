"""Parity of the HIP path against the CPU oracle, through the C ABI.

Bars (written in each test):
* kNN indices and distances: bit-exact to the oracle (the reference's exact
  algorithm with the 8-lane distance order, ties by neighbour index);
* sigma, scaled distances and NN pair picks: bit-exact;
* gradient, Adam, 10-iteration embeddings and losses: 1e-5 relative in FP32
  (the reference's own chunk reduction order is not fixed, gradient.rs:53-70).
"""
import numpy as np
import pytest

import oracle as O
import pacmap_mi355x as pm

pytestmark = pytest.mark.gpu

REL = 1e-5  # relative tolerance for floating-point parity (FP32)


def order_spread(fn, nn, mn, fp):
    """Largest element change between two valid summation orders of the
    reference algorithm: `fn(nn, mn, fp)` with each pair array in given and in
    reversed row order (the same pairs, so the same mathematical result)."""
    a = np.asarray(fn(nn, mn, fp), np.float64)
    b = np.asarray(fn(np.ascontiguousarray(nn[::-1]), np.ascontiguousarray(mn[::-1]),
                      np.ascontiguousarray(fp[::-1])), np.float64)
    return float(np.abs(a - b).max())


def assert_parity(got, ref, spread, rel=REL):
    """FP32 parity of a multi-iteration result: norm-wise relative error
    <= rel, and per element |got - ref| <= rel * max(|ref|, rms) + 4 * spread,
    where `spread` is the largest element difference between two valid
    summation orders of the oracle (Adam amplifies reordering noise in
    components whose gradient nearly cancels, in the reference as well)."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    nerr = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
    rms = np.sqrt(np.mean(ref ** 2))
    bound = rel * np.maximum(np.abs(ref), rms) + 4.0 * spread + 1e-30
    worst = (np.abs(got - ref) / bound).max()
    print(f"parity: norm-wise rel err {nerr:.2e}, oracle order spread {spread:.2e}, "
          f"worst element {worst:.3f} of bound", flush=True)
    assert nerr <= rel, f"norm-wise relative error {nerr:.3e} > {rel}"
    assert worst <= 1.0, f"element error {worst:.2f}x the bound"


def other_orders(y0, nn, mn, fp, it, nthreads=1):
    """The oracle's 10-iteration result under two other valid summation
    orders of the same pairs: every list reversed, and the far pairs
    permuted (the reference's rayon reduce fixes no order, gradient.rs:53-70)."""
    perm = np.random.default_rng(9).permutation(len(fp))
    return [O.optimize(y0, np.ascontiguousarray(nn[::-1]), np.ascontiguousarray(mn[::-1]),
                       np.ascontiguousarray(fp[::-1]), 1.0, it, nthreads=nthreads)[0],
            O.optimize(y0, nn, mn, np.ascontiguousarray(fp[perm]), 1.0, it, nthreads=nthreads)[0]]


def assert_parity_large(got, ref, alts, rel=REL, norm=True):
    """10-iteration parity at sizes where valid summation orders of the
    reference itself disagree near the tolerance.  From a Value init (std
    1e-4) Adam's first steps are about lr * sign(g), so a gradient component
    within rounding noise of zero takes either sign depending on the order;
    `alts` are the oracle under other valid orders.  Bars: norm-wise error
    at most max(rel, 3x the largest norm-wise difference between oracle
    orders), and no more elements off by more than rel * max(|ref|, rms)
    than 3x the most such elements of another oracle order, plus 10.
    norm=False keeps only the element count (for sizes where single chaotic
    rows dominate the norm; assert_shadowing then bounds every step)."""
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    nrm = np.linalg.norm(ref)
    rms = np.sqrt(np.mean(ref ** 2))
    tol = rel * np.maximum(np.abs(ref), rms)
    a_err = max(np.linalg.norm(np.asarray(a, np.float64) - ref) / nrm for a in alts)
    a_out = max(int((np.abs(np.asarray(a, np.float64) - ref) > tol).sum()) for a in alts)
    g_err = np.linalg.norm(got - ref) / nrm
    g_out = int((np.abs(got - ref) > tol).sum())
    print(f"parity: norm-wise rel err {g_err:.2e} (oracle orders {a_err:.2e}), elements beyond "
          f"{rel:g} relative {g_out} (oracle orders {a_out}) of {ref.size}", flush=True)
    if norm:
        assert g_err <= max(rel, 3.0 * a_err)
    assert g_out <= 3 * a_out + 10


def assert_shadowing(snaps, y0, nn, mn, fp, lr, it, nthreads=1, rel=REL):
    """Stepwise parity of an optimiser trajectory (snaps[k] = Y after
    iteration k): every GPU iterate is compared with one oracle iteration
    (gradient.rs:35-151, adam.rs:35-60, weights.rs:52-81) applied to the
    GPU's previous iterate, with the Adam moments carried along the GPU's
    trajectory.  Each step is held to 1e-5 without the chaotic growth that
    end points accumulate over iterations: norm-wise error at most rel, and
    no more elements beyond rel * max(|y|, rms) than 3x those of the oracle
    under another valid order (every pair list reversed) on the same step,
    plus 10 (Adam's first steps are about lr * sign(g), so a component
    within rounding noise of zero takes either sign)."""
    rev = [np.ascontiguousarray(a[::-1]) for a in (nn, mn, fp)]
    prev = np.asarray(y0, np.float32)
    m = np.zeros_like(prev)
    v = np.zeros_like(prev)
    for itr in range(int(sum(it))):
        w = O.find_weights(1000.0, itr, it[0], it[1])
        preds = []
        for lists in ((nn, mn, fp), rev):
            g = O.pacmap_grad(prev, *lists, *w, nthreads=nthreads)
            p, mm, vv = prev.copy(), m.copy(), v.copy()
            O.update_embedding_adam(p, g, mm, vv, np.float32(0.9), np.float32(0.999), lr, itr)
            preds.append((p, mm, vv))
        pred, m, v = preds[0]
        ref = pred.astype(np.float64)
        tol = rel * np.maximum(np.abs(ref), np.sqrt(np.mean(ref ** 2)))
        got = np.asarray(snaps[itr], np.float64)
        nerr = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        out = int((np.abs(got - ref) > tol).sum())
        alt = int((np.abs(preds[1][0].astype(np.float64) - ref) > tol).sum())
        print(f"  step {itr}: norm-wise rel err {nerr:.2e}, elements beyond {rel:g} {out} "
              f"(oracle reversed order {alt})", flush=True)
        assert nerr <= rel
        assert out <= 3 * alt + 10
        prev = np.asarray(snaps[itr], np.float32)


def oracle_loss64(y0, nn, mn, fp, lr, iters):
    """FP64 loss of every iteration along the oracle's FP32 trajectory
    (lib.rs:803-823).  The reference's own f32 loss sums 131072-term chunks
    sequentially and carries ~1e-5 relative error of its own."""
    y = np.array(y0, np.float32, copy=True)
    m = np.zeros_like(y)
    v = np.zeros_like(y)
    out = []
    for itr in range(int(sum(iters))):
        w = O.find_weights(1000.0, itr, iters[0], iters[1])
        g = O.pacmap_grad(y, nn, mn, fp, w[0], w[1], w[2])
        out.append(O.pacmap_grad_f64(y.astype(np.float64), nn, mn, fp, float(w[0]), float(w[1]),
                                     float(w[2]))[-1, 0])
        O.update_embedding_adam(y, g, m, v, np.float32(0.9), np.float32(0.999), lr, itr)
    return np.array(out)


def assert_loss(got, ref64, rel=REL):
    """Every iteration's loss within rel of the FP64 loss."""
    err = np.abs(np.asarray(got, np.float64) - ref64) / np.abs(ref64)
    print(f"loss: worst relative error vs FP64 {err.max():.2e}", flush=True)
    assert err.max() <= rel


def assert_rel(got, ref, rel=REL):
    """FP32 parity: norm-wise relative error <= rel, and per element
    |got - ref| <= rel * max(|ref|, rms(ref)).

    The per-element floor sits at the RMS magnitude because a component that
    nearly cancels carries the absolute reordering error of its summands, in
    the reference as much as here (gradient.rs:53-70 has no fixed order)."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape
    if ref.size == 0:
        return
    nerr = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
    rms = np.sqrt(np.mean(ref ** 2))
    bound = rel * np.maximum(np.abs(ref), rms) + 1e-30
    worst = (np.abs(got - ref) / bound).max()
    print(f"parity: norm-wise rel err {nerr:.2e}, worst element {worst:.3f} of bound", flush=True)
    assert nerr <= rel, f"norm-wise relative error {nerr:.3e} > {rel}"
    assert worst <= 1.0, f"element error {worst:.2f}x the bound"


def _pairs(rng, n, rows):
    return rng.integers(0, n, size=(rows, 2)).astype(np.uint32)


# ------------------------------------------------------------------- kNN --
KNN_CASES = [
    # (n, d, k, kind, seed)
    (500, 100, 60, "normal", 1),
    (700, 784, 60, "uniform", 2),
    (300, 1000, 10, "sym", 3),
    (400, 3, 60, "normal", 4),
    (200, 8, 5, "normal", 5),
    (150, 7, 149, "normal", 6),
    (333, 17, 100, "normal", 7),
    (300, 16, 200, "normal", 8),
    (520, 33, 448, "uniform", 9),
    (500, 4, 60, "grid", 10),     # many exact distance ties
    (256, 12, 30, "dups", 11),    # duplicated points: zero distances
    (1100, 128, 60, "normal", 12),
]


def _data(kind, n, d, seed):
    rng = np.random.default_rng(seed)
    if kind == "normal":
        return rng.normal(size=(n, d)).astype(np.float32)
    if kind == "uniform":
        return rng.uniform(0, 1, size=(n, d)).astype(np.float32)
    if kind == "sym":
        return rng.uniform(-1, 1, size=(n, d)).astype(np.float32)
    if kind == "grid":
        return rng.integers(0, 3, size=(n, d)).astype(np.float32)
    if kind == "dups":
        base = rng.normal(size=(n // 4, d)).astype(np.float32)
        return np.repeat(base, 4, axis=0)[:n].copy()
    if kind == "three":
        base = rng.normal(size=(3, d)).astype(np.float32)
        return base[rng.integers(0, 3, size=n)].copy()
    if kind == "clusters":  # 12 well-separated clusters, rows in random order
        base = rng.normal(scale=12.0, size=(12, d)).astype(np.float32)
        return (base[rng.integers(0, 12, size=n)] + rng.normal(size=(n, d))).astype(np.float32)
    raise ValueError(kind)


@pytest.mark.parametrize("n,d,k,kind,seed", KNN_CASES)
def test_knn_exact_bitwise(n, d, k, kind, seed):
    x = _data(kind, n, d, seed)
    print(f"knn case n={n} d={d} k={k} {kind}", flush=True)
    idx, dist = pm.knn.find_k_nearest_neighbors(x, k)
    ref_i, ref_d = O.knn_rowwise(x, k, nthreads=8)
    assert idx.shape == ref_i.shape == (n, min(k, n - 1))
    assert np.array_equal(dist.view(np.uint32), ref_d.view(np.uint32)), "distance bits differ"
    assert np.array_equal(idx, ref_i), "neighbour indices differ"


MFMA_CASES = [
    # (n, d, k, kind, seed): large enough for the certified MFMA candidate path
    (4000, 100, 60, "normal", 101),
    (3000, 128, 10, "uniform", 102),
    (5000, 256, 60, "sym", 103),
    (2600, 33, 100, "normal", 104),
    (3000, 12, 30, "dups", 105),
    (4000, 4, 60, "grid", 106),
    # three distinct points: a tile epilogue lists up to 1,024 pairs, more
    # than a wave's 128-entry staging area, so it flushes mid-epilogue
    (2048, 8, 60, "three", 107),
]


@pytest.mark.parametrize("n,d,k,kind,seed", MFMA_CASES)
def test_knn_mfma_paths_bitwise(n, d, k, kind, seed):
    """Every kNN path (split-bf16 MFMA, bf16 MFMA, direct) equals the oracle bit for bit."""
    x = _data(kind, n, d, seed)
    ref_i, ref_d = O.knn_rowwise(x, k, nthreads=16)
    for mode in ("fp32", "bf16", "direct"):
        print(f"knn mode={mode} n={n} d={d} k={k} {kind}", flush=True)
        idx, dist = pm.knn.find_k_nearest_neighbors(x, k, precision=mode)
        t = pm.last_timings()
        print(f"  path={t['knn_path']} rescanned={t['knn_rescanned_rows']} "
              f"kernel={t['knn_kernel_ms']:.3f} ms rerank={t['knn_rerank_ms']:.3f} ms", flush=True)
        assert np.array_equal(dist.view(np.uint32), ref_d.view(np.uint32)), mode
        assert np.array_equal(idx, ref_i), mode


@pytest.mark.parametrize("n,d,k,kind,seed", [(4000, 100, 60, "normal", 121), (3000, 12, 30, "dups", 122)])
def test_knn_split_pass_bitwise(monkeypatch, n, d, k, kind, seed):
    """The candidate pass that scans every (query, candidate) pair in split
    ranges (PACMAP_KNN_SYM=0) instead of the upper triangle once equals the
    oracle bit for bit too."""
    x = _data(kind, n, d, seed)
    ref_i, ref_d = O.knn_rowwise(x, k, nthreads=16)
    monkeypatch.setenv("PACMAP_KNN_SYM", "0")
    idx, dist = pm.knn.find_k_nearest_neighbors(x, k)
    assert np.array_equal(dist.view(np.uint32), ref_d.view(np.uint32))
    assert np.array_equal(idx, ref_i)


@pytest.mark.parametrize("n,d,k,seed", [(12000, 100, 60, 131), (9000, 40, 30, 132), (5000, 2, 10, 133)])
def test_knn_pruned_pass_bitwise(monkeypatch, n, d, k, seed):
    """The symmetric pass over rows in Morton order of their first six
    coordinates skips the tile pairs whose boxes lie farther apart than both
    rows' taus: on well-separated clusters most pairs are skipped, and the
    neighbours equal the oracle's and the unpruned pass's bit for bit."""
    x = _data("clusters", n, d, seed)
    ref_i, ref_d = O.knn_rowwise(x, k, nthreads=16)
    idx, dist = pm.knn.find_k_nearest_neighbors(x, k)
    t = pm.last_timings()
    print(f"pruned: {t['knn_tile_pairs_skipped']} of {t['knn_tile_pairs']} tile pairs, "
          f"rescanned {t['knn_rescanned_rows']}", flush=True)
    assert t["knn_tile_pairs"] > 0 and t["knn_tile_pairs_skipped"] > t["knn_tile_pairs"] // 2
    assert np.array_equal(dist.view(np.uint32), ref_d.view(np.uint32))
    assert np.array_equal(idx, ref_i)
    monkeypatch.setenv("PACMAP_KNN_PRUNE", "0")
    idx0, dist0 = pm.knn.find_k_nearest_neighbors(x, k)
    assert pm.last_timings()["knn_tile_pairs_skipped"] == 0
    assert np.array_equal(dist0.view(np.uint32), dist.view(np.uint32)) and np.array_equal(idx0, idx)


def test_knn_tau_sample_on_a_manifold():
    """On 3-D data in 100 dimensions the Morton order puts a row's true
    neighbours in its own and nearby tiles.  The tau sample is drawn from rows
    spread over the input, not from tiles of that order, so tau still lists
    about 3k candidates per row and almost no row falls back to the exact
    rescan (a sample of every 12th tile of the order left 4.4% of a 100,000
    point Swiss roll short of k); the neighbours equal the oracle's bit for
    bit."""
    rng = np.random.default_rng(141)
    n, d, k = 12000, 100, 60
    basis, _ = np.linalg.qr(rng.standard_normal((d, 3)))
    t = rng.uniform(1.5 * np.pi, 4.5 * np.pi, n)
    p3 = np.stack([t * np.cos(t), rng.uniform(0.0, 21.0, n), t * np.sin(t)], 1) + rng.normal(0.0, 0.3, (n, 3))
    x = (p3 @ basis.T + rng.normal(0.0, 0.01, (n, d))).astype(np.float32)
    ref_i, ref_d = O.knn_rowwise(x, k, nthreads=16)
    idx, dist = pm.knn.find_k_nearest_neighbors(x, k)
    tm = pm.last_timings()
    print(f"manifold: rescanned {tm['knn_rescanned_rows']} of {n}, skipped {tm['knn_tile_pairs_skipped']} of "
          f"{tm['knn_tile_pairs']} tile pairs", flush=True)
    assert tm["knn_rescanned_rows"] <= n // 200
    assert np.array_equal(dist.view(np.uint32), ref_d.view(np.uint32))
    assert np.array_equal(idx, ref_i)


@pytest.mark.parametrize("n,d,k,seed", [(6000, 100, 60, 111), (5000, 256, 30, 112), (4000, 20, 10, 113)])
def test_knn_approximate_modes(n, d, k, seed):
    """FP8 (per-row scaled E4M3 candidates) and the uncertified bf16 modes:
    the returned neighbours carry their exact FP32 distances (the oracle's,
    bit for bit, wherever the index matches) and recall stays high; split
    bf16 without the certificate still finds every neighbour."""
    rng = np.random.default_rng(seed)
    c = rng.normal(scale=1.5, size=(20, d))
    x = (c[rng.integers(0, 20, n)] + rng.normal(size=(n, d)) * rng.uniform(0.2, 1.0, d)).astype(np.float32)
    ref_i, ref_d = O.knn_rowwise(x, k, nthreads=16)
    floor = {("fp8", False): 0.9, ("bf16", False): 0.99, ("fp32", False): 0.9999}
    for (mode, cert), lo in floor.items():
        idx, dist = pm.knn.find_k_nearest_neighbors(x, k, precision=mode, certify=cert)
        t = pm.last_timings()
        recall = np.mean([len(np.intersect1d(a, b)) for a, b in zip(idx, ref_i)]) / k
        print(f"knn {mode} certify={cert}: path={t['knn_path']} certified={t['knn_certified']} "
              f"recall={recall:.5f} kernel={t['knn_kernel_ms']:.3f} ms", flush=True)
        assert t["knn_certified"] == 0
        assert recall >= lo, (mode, recall)
        same = idx == ref_i
        assert np.array_equal(dist[same].view(np.uint32), ref_d[same].view(np.uint32))
        assert (np.diff(dist, axis=1) >= 0).all()


def test_knn_certificate_covers_most_rows():
    """On clustered data the split-bf16 certificate leaves few rows to rescan."""
    rng = np.random.default_rng(107)
    n, d = 8000, 100
    c = rng.normal(scale=2.0, size=(10, d))
    x = (c[rng.integers(0, 10, n)] + rng.normal(size=(n, d))).astype(np.float32)
    pm.knn.find_k_nearest_neighbors(x, 60)
    t = pm.last_timings()
    assert t["knn_path"] == 1
    assert t["knn_rescanned_rows"] <= 0.01 * n


@pytest.mark.parametrize("kind", ["clustered", "grid"])
def test_knn_rerank_prunes_by_upper_bounds(kind):
    """The certified split-bf16 rerank gives exact distances only to listed
    candidates whose lower bound is within the k-th smallest upper bound:
    far fewer than are listed, with the neighbour sets still equal to the
    oracle's bit for bit (the grid has many exactly tied distances)."""
    rng = np.random.default_rng(108)
    n, d, k = 8000, 100, 60
    if kind == "clustered":
        c = rng.normal(scale=2.0, size=(10, d))
        x = (c[rng.integers(0, 10, n)] + rng.normal(size=(n, d))).astype(np.float32)
    else:
        x = _data("grid", n, 24, 109)
    idx, dist = pm.knn.find_k_nearest_neighbors(x, k)
    t = pm.last_timings()
    ref_i, ref_d = O.knn_rowwise(x, k, nthreads=16)
    rows = n - t["knn_rescanned_rows"]
    print(f"knn {kind}: listed {t['knn_listed'] / max(rows, 1):.1f} per row, reranked "
          f"{t['knn_reranked'] / max(rows, 1):.1f}, rescanned rows {t['knn_rescanned_rows']}", flush=True)
    assert t["knn_path"] == 1 and t["knn_certified"] == 1
    assert np.array_equal(dist.view(np.uint32), ref_d.view(np.uint32))
    assert np.array_equal(idx, ref_i)
    assert t["knn_reranked"] >= k * rows
    if kind == "clustered":
        assert t["knn_reranked"] <= 0.6 * t["knn_listed"]


def test_knn_matches_triple_sort_reference_algorithm():
    x = _data("grid", 300, 5, 21)
    idx, dist = pm.knn.find_k_nearest_neighbors(x, 60)
    ref_i, ref_d = O.knn_triple_sort(x, 60, nthreads=4)
    assert np.array_equal(idx, ref_i) and np.array_equal(dist, ref_d)


def test_knn_golden_kats(golden):
    for case in golden("knn.json")["cases"]:
        if "embeddings" in case:
            x = np.array(case["embeddings"], np.float32).reshape(len(case["embeddings"]), -1)
        else:
            x = np.zeros(case["embeddings_shape"], np.float32)
        idx, dist = pm.knn.find_k_nearest_neighbors(x, case["k"])
        assert list(idx.shape) == case["shape"]
        for point, exp in case.get("expected", []):
            got = sorted(zip(idx[point].tolist(), dist[point].tolist()))
            want = sorted((int(i), float(v)) for i, v in exp)
            assert [g[0] for g in got] == [w[0] for w in want]
            for (_, gd), (_, wd) in zip(got, want):
                assert abs(gd - wd) <= max(1e-4, 1e-4 * abs(wd))


def test_knn_zero_features_ties_by_index():
    x = np.zeros((70, 0), np.float32)
    idx, dist = pm.knn.find_k_nearest_neighbors(x, 5)
    assert (dist == 0).all()
    exp = np.array([[j for j in range(70) if j != i][:5] for i in range(70)], np.uint32)
    assert np.array_equal(idx, exp)


# -------------------------------------------------------- sigma / scale --
def test_sigma_scale_and_nn_pairs_bitwise():
    x = _data("normal", 600, 20, 31)
    kp, knn = 60, 10
    idx, dist = O.knn_rowwise(x, kp, nthreads=8)
    sig, sc, pairs = pm.sigma_scale(dist, idx, knn)
    ref_sig = O.sigma(dist)
    ref_sc = O.scale_dist(dist, ref_sig, idx)
    ref_pairs = O.sample_neighbors_pair(ref_sc, idx, knn)
    assert np.array_equal(sig.view(np.uint32), ref_sig.view(np.uint32))
    assert np.array_equal(sc.view(np.uint32), ref_sc.view(np.uint32))
    assert np.array_equal(pairs, ref_pairs)


@pytest.mark.parametrize("which", [0, 1])
def test_scale_dist_golden_on_gpu(golden, which):
    """distance.rs:120-190: the reference's expected scaled distances for its
    given sigma (exact for the 2x2 case, 1e-6 for the 10x5 one), and the
    oracle's bits."""
    c = golden("scale_dist.json")["cases"][which]
    dist = np.array(c["knn_distances"], np.float32)
    sig = np.array(c["sig"], np.float32)
    idx = np.array(c["neighbors"], np.uint32)
    got = pm.scale_dist(dist, sig, idx)
    np.testing.assert_allclose(got, np.array(c["expected"], np.float32), rtol=0, atol=c["tolerance"])
    assert np.array_equal(got.view(np.uint32), O.scale_dist(dist, sig, idx).view(np.uint32))
    # sigma from the distances themselves (neighbors.rs:79-84) through the other seam
    sig2, sc2, _ = pm.sigma_scale(dist, idx, 0)
    assert np.array_equal(sig2, O.sigma(dist))
    assert np.array_equal(sc2, O.scale_dist(dist, O.sigma(dist), idx))


# --------------------------------------------------------------- sampling --
@pytest.mark.parametrize("n,d,knn,n_mn,n_fp,seed", [
    (2000, 100, 10, 10, 20, 42),
    (777, 13, 15, 7, 30, 7),
    (9, 4, 3, 4, 10, 3),       # fewer points than draws: short samples, (0, 0) rows
    (5000, 33, 10, 5, 20, 2024),
    (300, 20, 10, 40, 5, 11),  # more than 32 MN pairs a point: picks read back from the pair rows
    (12, 5, 3, 12, 4, 5),      # many rejected draws: the MN streams run past their precomputed values
])
def test_sample_pairs_match_oracle(n, d, knn, n_mn, n_fp, seed):
    """The GPU samplers restate the same RNG streams as the oracle: identical pairs."""
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, d)).astype(np.float32)
    nn = np.stack([np.repeat(np.arange(n), knn),
                   rng.integers(0, n, n * knn)], 1).astype(np.uint32)
    mn, fp = pm.sample_pairs(x, nn, knn, n_mn, n_fp, seed)
    assert np.array_equal(mn, O.sample_mn_pair_deterministic(x, n_mn, seed))
    assert np.array_equal(fp, O.sample_fp_pair_deterministic(n, nn, knn, n_fp, seed))


# --------------------------------------------------------------- gradient --
def test_grad_golden(golden):
    g = golden("gradient.json")
    out = pm.pacmap_grad(np.array(g["y_test"], np.float32), g["pair_neighbors"], g["pair_mn"],
                         g["pair_fp"], g["w_neighbors"], g["w_mn"], g["w_fp"])
    np.testing.assert_allclose(out, np.array(g["grad_python"], np.float32), rtol=0, atol=1e-6)


@pytest.mark.parametrize("dims", [2, 3, 1])
def test_grad_matches_oracle(dims):
    rng = np.random.default_rng(40 + dims)
    n = 3000
    y = rng.normal(size=(n, dims)).astype(np.float32)
    nn, mn, fp = _pairs(rng, n, 30000), _pairs(rng, n, 15000), _pairs(rng, n, 60000)
    got = pm.pacmap_grad(y, nn, mn, fp, 2.0, 730.5, 1.0)
    ref = O.pacmap_grad(y, nn, mn, fp, 2.0, 730.5, 1.0)
    ref64 = O.pacmap_grad_f64(y.astype(np.float64), nn, mn, fp, 2.0, 730.5, 1.0)
    assert_rel(got[:-1], ref[:-1])
    assert abs(got[-1, 0] - ref64[-1, 0]) <= REL * abs(ref64[-1, 0])
    assert np.abs(got[:-1].astype(np.float64).sum(0)).max() < 1e-3  # zero-sum invariant
    # no less accurate than the oracle's own (reference) summation order
    e_gpu = np.linalg.norm(got[:-1] - ref64[:-1])
    e_ref = np.linalg.norm(ref[:-1] - ref64[:-1])
    assert e_gpu <= 2.0 * e_ref + 1e-7 * np.linalg.norm(ref64[:-1])


@pytest.mark.parametrize("n", [3000, 20000])
def test_grad_hub_rows_sorted_on_the_gpu(n):
    """Hub rows longer than a wavefront sorts (2048 entries): one point in
    6,000 NN pairs (sorted in LDS) and another in 40,000 FP pairs (sorted in
    place in global memory).  Every CSR row is sorted by entry value, so the
    gradient is bit-identical for any order of the input pair lists; it
    matches the oracle at 1e-5, and 10 iterations match the oracle's.
    n = 20000 runs the LDS-staged gradient, n = 3000 the one-pass kernel."""
    rng = np.random.default_rng(n)
    y = rng.normal(size=(n, 2)).astype(np.float32)
    hub_a = np.stack([np.zeros(6000, np.int64), rng.integers(1, n, 6000)], 1)
    hub_b = np.stack([rng.integers(0, n, 40000), np.ones(40000, np.int64)], 1)
    nn = np.concatenate([_pairs(rng, n, n * 5), hub_a]).astype(np.uint32)
    mn = _pairs(rng, n, n * 3)
    fp = np.concatenate([_pairs(rng, n, n * 10), hub_b]).astype(np.uint32)
    w = (2.0, 120.0, 1.0)
    g1 = pm.pacmap_grad(y, nn, mn, fp, *w)
    perm = [rng.permutation(len(a)) for a in (nn, mn, fp)]
    g2 = pm.pacmap_grad(y, nn[perm[0]], mn[perm[1]], fp[perm[2]], *w)
    assert np.array_equal(g1[:-1].view(np.uint32), g2[:-1].view(np.uint32))
    assert_rel(g1[:-1], O.pacmap_grad(y, nn, mn, fp, *w)[:-1])
    it = (4, 3, 3)
    y0 = O.value_init(y)
    got, _, _ = pm.optimize(y0, nn, mn, fp, 1.0, it)
    assert pm.last_timings()["grad_sell"] == (1 if n >= 16384 else 0)
    ry, _, _ = O.optimize(y0, nn, mn, fp, 1.0, it)
    assert_parity_large(got, ry, other_orders(y0, nn, mn, fp, it))


def test_grad_rejects_out_of_range_pairs():
    y = np.zeros((10, 2), np.float32)
    with pytest.raises(pm.InvalidPairIndex):
        pm.pacmap_grad(y, np.array([[0, 10]]), np.zeros((0, 2)), np.zeros((0, 2)), 1, 1, 1)


# ------------------------------------------------------------------- Adam --
def test_adam_golden(golden):
    a = golden("adam.json")
    y = np.array(a["y"], np.float32)
    m = np.zeros_like(y)
    v = np.zeros_like(y)
    pm.update_embedding_adam(y, np.array(a["grad"], np.float32), m, v, a["beta1"], a["beta2"],
                             a["lr"], int(a["itr"]))
    for got, key in ((y, "y_expected"), (m, "m_expected"), (v, "v_expected")):
        np.testing.assert_allclose(got, np.array(a[key], np.float32), rtol=0, atol=a["epsilon"])


def test_adam_matches_oracle():
    rng = np.random.default_rng(50)
    y = rng.normal(size=(5000, 2)).astype(np.float32)
    g = rng.normal(size=(5000, 2)).astype(np.float32) * np.float32(1e-3)
    m = rng.normal(size=(5000, 2)).astype(np.float32) * np.float32(1e-3)
    v = np.abs(rng.normal(size=(5000, 2))).astype(np.float32) * np.float32(1e-6)
    y2, m2, v2 = y.copy(), m.copy(), v.copy()
    pm.update_embedding_adam(y, g, m, v, np.float32(0.9), np.float32(0.999), 1.0, 17)
    O.update_embedding_adam(y2, g, m2, v2, np.float32(0.9), np.float32(0.999), 1.0, 17)
    assert_rel(y, y2, rel=1e-6)
    assert_rel(m, m2, rel=1e-6)
    assert_rel(v, v2, rel=1e-6)


# -------------------------------------------------------------- optimiser --
def _problem(seed, n=1500):
    rng = np.random.default_rng(seed)
    y0 = O.value_init(rng.normal(size=(n, 2)).astype(np.float32))
    return y0, _pairs(rng, n, n * 10), _pairs(rng, n, n * 10), _pairs(rng, n, n * 20)


def test_optimize_10_iterations_matches_oracle():
    y0, nn, mn, fp = _problem(60)
    it = (4, 3, 3)
    y, _, loss = pm.optimize(y0, nn, mn, fp, 1.0, it)
    ry, _, rloss = O.optimize(y0, nn, mn, fp, 1.0, it)
    assert_parity(y, ry, order_spread(lambda a, b, c: O.optimize(y0, a, b, c, 1.0, it)[0], nn, mn, fp))
    assert_loss(loss, oracle_loss64(y0, nn, mn, fp, 1.0, it))


@pytest.mark.parametrize("paged", [False, True])
@pytest.mark.parametrize("dims", [2, 3])
def test_optimize_target_blocked_matches_oracle(monkeypatch, dims, paged):
    """The two-pass target-blocked gradient (used when Y outgrows an XCD's L2)
    forced on at a small n, in its gather form and in its paged form (pages
    of 100 points, so each range is walked in 4 pages): same bar against the
    oracle, same loss, and deterministic."""
    y0, nn, mn, fp = _problem(62, n=3000)
    if dims != 2:
        y0 = np.random.default_rng(63).normal(size=(y0.shape[0], dims)).astype(np.float32)
    it = (4, 3, 3)
    monkeypatch.setenv("PACMAP_GRAD_TARGET_BLOCKS", "1")
    monkeypatch.setenv("PACMAP_GRAD_PAGED", "1" if paged else "0")
    monkeypatch.setenv("PACMAP_GRAD_PAGE_POINTS", "100")
    y, _, loss = pm.optimize(y0, nn, mn, fp, 1.0, it)
    t = pm.last_timings()
    assert t["grad_target_blocks"] == 8 and t["grad_paged"] == int(paged)
    if paged:
        assert t["grad_page_pages"] == 4
    y2, _, loss2 = pm.optimize(y0, nn, mn, fp, 1.0, it)
    assert np.array_equal(y, y2) and np.array_equal(loss, loss2)
    ry, _, _ = O.optimize(y0, nn, mn, fp, 1.0, it)
    assert_parity(y, ry, order_spread(lambda a, b, c: O.optimize(y0, a, b, c, 1.0, it)[0], nn, mn, fp))
    assert_loss(loss, oracle_loss64(y0, nn, mn, fp, 1.0, it))
    monkeypatch.setenv("PACMAP_GRAD_TARGET_BLOCKS", "0")
    pm.optimize(y0, nn, mn, fp, 1.0, it)
    assert pm.last_timings()["grad_target_blocks"] == 0


@pytest.mark.parametrize("n,dims", [(3000, 2), (3000, 3), (777, 1), (20000, 2), (120000, 2)])
def test_optimize_lds_sliced_matches_target_blocked(monkeypatch, n, dims):
    """The LDS-staged sliced-ELL gradient restages the target-blocked kernel:
    with 8 target ranges (the same ranges, per-lane entry order, butterfly
    and range-order combine) the embeddings are bit-identical; with 4 merged
    ranges the same bar against the oracle holds.  Both are deterministic.
    n = 3000 leaves a partial last slice; 20000 takes the default path;
    120000 needs 8 ranges and chunks capped at 4096 rows."""
    y0, nn, mn, fp = _problem(64, n=n)
    if dims != 2:
        y0 = np.random.default_rng(65).normal(size=(n, dims)).astype(np.float32)
    it = (4, 3, 3)
    monkeypatch.setenv("PACMAP_GRAD_TARGET_BLOCKS", "1")
    monkeypatch.setenv("PACMAP_GRAD_PAGED", "0")
    yt, _, _ = pm.optimize(y0, nn, mn, fp, 1.0, it)
    assert pm.last_timings()["grad_target_blocks"] == 8 and pm.last_timings()["grad_paged"] == 0
    monkeypatch.setenv("PACMAP_GRAD_TARGET_BLOCKS", "0")
    if n < 16384:
        monkeypatch.setenv("PACMAP_GRAD_SELL", "1")
    spread = None
    for ranges in ((8,) if n > 80000 else (8, 4)):
        monkeypatch.setenv("PACMAP_GRAD_SELL_RANGES", str(ranges))
        y, _, loss = pm.optimize(y0, nn, mn, fp, 1.0, it)
        t = pm.last_timings()
        assert t["grad_sell"] == 1 and t["grad_target_blocks"] == 0 and t["grad_sell_ranges"] == ranges
        assert t["grad_sell_entries"] >= 2 * (len(nn) + len(mn) + len(fp)) - 2 * n
        if ranges == 8:
            assert np.array_equal(y, yt)
        y2, _, loss2 = pm.optimize(y0, nn, mn, fp, 1.0, it)
        assert np.array_equal(y, y2) and np.array_equal(loss, loss2)
        assert_loss(loss, oracle_loss64(y0, nn, mn, fp, 1.0, it))
        if n <= 3000:
            ry, _, _ = O.optimize(y0, nn, mn, fp, 1.0, it)
            if spread is None:
                spread = order_spread(lambda a, b, c: O.optimize(y0, a, b, c, 1.0, it)[0], nn, mn, fp)
            assert_parity(y, ry, spread)
    monkeypatch.delenv("PACMAP_GRAD_SELL_RANGES")
    monkeypatch.setenv("PACMAP_GRAD_SELL", "0")
    pm.optimize(y0, nn, mn, fp, 1.0, it)
    assert pm.last_timings()["grad_sell"] == 0


@pytest.mark.parametrize("n,dims,rows,points", [
    (3000, 2, 0, 64), (3000, 3, 96, 100), (777, 1, 0, 0), (20000, 2, 512, 300), (150000, 2, 0, 0)])
def test_paged_gradient_matches_oracle_and_target_blocked(monkeypatch, n, dims, rows, points):
    """The paged target-blocked gradient (rows' Y and sums in LDS, the target
    range walked in pages of Y staged through LDS; the default wherever the
    target-blocked kernel runs), with chunks and pages forced small so that
    slices are partial and each range takes several pages: the gradient
    matches the oracle at 1e-5, the loss the FP64 loss, the gather form of
    the kernel within 1e-5, and it is bit-identical from run to run.  The
    layout holds every entry once plus padding."""
    rng = np.random.default_rng(70 + n + dims)
    y = (rng.normal(size=(n, dims)) * 3).astype(np.float32)
    nn, mn, fp = _pairs(rng, n, n * 10), _pairs(rng, n, n * 5), _pairs(rng, n, n * 20)
    w = (2.0, 730.5, 1.0)
    monkeypatch.setenv("PACMAP_GRAD_TARGET_BLOCKS", "1")
    monkeypatch.setenv("PACMAP_GRAD_PAGED", "1")
    if rows:
        monkeypatch.setenv("PACMAP_GRAD_PAGE_ROWS", str(rows))
    if points:
        monkeypatch.setenv("PACMAP_GRAD_PAGE_POINTS", str(points))
    print(f"paged gradient n={n} dims={dims} rows<={rows or 'plan'} points<={points or 'plan'}", flush=True)
    g = pm.pacmap_grad(y, nn, mn, fp, *w)
    t = pm.last_timings()
    assert t["grad_paged"] == 1 and t["grad_target_blocks"] == 8
    if points:
        assert t["grad_page_pages"] >= 2
    entries = 2 * int((nn[:, 0] != nn[:, 1]).sum() + (mn[:, 0] != mn[:, 1]).sum() + (fp[:, 0] != fp[:, 1]).sum())
    assert t["grad_page_words"] >= entries
    print(f"  pages {t['grad_page_pages']}, words {t['grad_page_words']} for {entries} entries", flush=True)
    g2 = pm.pacmap_grad(y, nn, mn, fp, *w)
    assert np.array_equal(g.view(np.uint32), g2.view(np.uint32))
    ref = O.pacmap_grad(y, nn, mn, fp, *w)
    assert_rel(g[:-1], ref[:-1])
    ref64 = O.pacmap_grad_f64(y.astype(np.float64), nn, mn, fp, *w)
    assert abs(g[-1, 0] - ref64[-1, 0]) <= REL * abs(ref64[-1, 0])
    monkeypatch.setenv("PACMAP_GRAD_PAGED", "0")
    gt = pm.pacmap_grad(y, nn, mn, fp, *w)
    assert pm.last_timings()["grad_paged"] == 0
    assert_rel(g[:-1], gt[:-1])


def test_paged_gradient_hub_rows(monkeypatch):
    """Hub rows in the paged layout: one point in 6,000 NN pairs and another
    in 40,000 FP pairs; their entries fill the widest slices of every page.
    Bit-identical for any order of the pair lists, the oracle's gradient at
    1e-5, and 10 iterations within the oracle's order spread."""
    n = 20000
    rng = np.random.default_rng(77)
    y = rng.normal(size=(n, 2)).astype(np.float32)
    hub_a = np.stack([np.zeros(6000, np.int64), rng.integers(1, n, 6000)], 1)
    hub_b = np.stack([rng.integers(0, n, 40000), np.ones(40000, np.int64)], 1)
    nn = np.concatenate([_pairs(rng, n, n * 5), hub_a]).astype(np.uint32)
    mn = _pairs(rng, n, n * 3)
    fp = np.concatenate([_pairs(rng, n, n * 10), hub_b]).astype(np.uint32)
    w = (2.0, 120.0, 1.0)
    monkeypatch.setenv("PACMAP_GRAD_TARGET_BLOCKS", "1")
    monkeypatch.setenv("PACMAP_GRAD_PAGED", "1")
    monkeypatch.setenv("PACMAP_GRAD_PAGE_POINTS", "400")
    g1 = pm.pacmap_grad(y, nn, mn, fp, *w)
    assert pm.last_timings()["grad_paged"] == 1
    perm = [rng.permutation(len(a)) for a in (nn, mn, fp)]
    g2 = pm.pacmap_grad(y, nn[perm[0]], mn[perm[1]], fp[perm[2]], *w)
    assert np.array_equal(g1[:-1].view(np.uint32), g2[:-1].view(np.uint32))
    assert_rel(g1[:-1], O.pacmap_grad(y, nn, mn, fp, *w)[:-1])
    it = (4, 3, 3)
    y0 = O.value_init(y)
    got, _, loss = pm.optimize(y0, nn, mn, fp, 1.0, it)
    assert pm.last_timings()["grad_paged"] == 1
    ry, _, _ = O.optimize(y0, nn, mn, fp, 1.0, it)
    assert_parity_large(got, ry, other_orders(y0, nn, mn, fp, it))
    assert_loss(loss, oracle_loss64(y0, nn, mn, fp, 1.0, it))


def _device_free_bytes():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_thread_caches_released_at_thread_exit():
    """Calls from short-lived threads: each thread gets its own stream, block
    cache, graph and pinned buffers, computes the main thread's embedding bit
    for bit, and returns what it held when it exits (free device memory after
    eight threads have run and exited is within 64 MiB of before)."""
    import threading
    y0, nn, mn, fp = _problem(80, n=20000)
    it = (4, 3, 3)
    ref = pm.optimize(y0, nn, mn, fp, 1.0, it)[0]
    before = _device_free_bytes()
    for i in range(8):
        out = {}

        def work():
            try:
                out["y"] = pm.optimize(y0, nn, mn, fp, 1.0, it)[0]
            except Exception as e:  # noqa: BLE001 - re-raised on the main thread
                out["err"] = e

        th = threading.Thread(target=work)
        th.start()
        th.join()
        assert "err" not in out, out.get("err")
        assert np.array_equal(out["y"], ref)
    after = _device_free_bytes()
    print(f"  device free before {before >> 20} MiB, after 8 threads {after >> 20} MiB", flush=True)
    assert before - after < (64 << 20)


def test_optimizer_graph_replayed_only_for_identical_launches():
    """A call whose launches match the previous call's replays its graph; the
    results equal a freshly captured run, and any change recaptures."""
    y0, nn, mn, fp = _problem(66, n=2500)
    it = (5, 5, 5)
    pm.release_cached_memory()
    a = pm.optimize(y0, nn, mn, fp, 1.0, it)
    assert pm.last_timings()["graph_reused"] == 0
    b = pm.optimize(y0, nn, mn, fp, 1.0, it)
    assert pm.last_timings()["graph_reused"] == 1
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])
    c = pm.optimize(y0, nn, mn, fp, 0.5, it)  # another learning rate: other Adam scalars
    assert pm.last_timings()["graph_reused"] == 0
    assert not np.array_equal(c[0], a[0])
    y1 = (y0 * 2.0).astype(np.float32)  # same launches, other inputs
    d = pm.optimize(y1, nn, mn, fp, 0.5, it)
    assert pm.last_timings()["graph_reused"] == 1
    pm.release_cached_memory()
    e = pm.optimize(y1, nn, mn, fp, 0.5, it)
    assert pm.last_timings()["graph_reused"] == 0
    assert np.array_equal(d[0], e[0])


def test_optimize_full_schedule_deterministic():
    y0, nn, mn, fp = _problem(61, n=2000)
    a = pm.optimize(y0, nn, mn, fp, 1.0, (100, 100, 250))
    b = pm.optimize(y0, nn, mn, fp, 1.0, (100, 100, 250))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])


def test_optimize_snapshots_follow_reference_slots():
    y0, nn, mn, fp = _problem(62, n=500)
    snaps = [3, 6, 6, 0, 99]
    it = (3, 2, 2)
    y, s, _ = pm.optimize(y0, nn, mn, fp, 1.0, it, snapshots=snaps)
    ry, rs, _ = O.optimize(y0, nn, mn, fp, 1.0, it, snapshots=snaps)
    spread = order_spread(lambda a, b, c: O.optimize(y0, a, b, c, 1.0, it, snapshots=snaps)[1],
                          nn, mn, fp)
    assert_parity(s[0], rs[0], spread)
    assert_parity(s[1], rs[1], spread)
    assert not s[2].any() and not s[4].any()
    # slot 3 (value 0) is never written; only slot 0 starts from the initial Y
    assert not s[3].any()


def test_optimize_snapshot_slot_zero_keeps_initial_state():
    y0, nn, mn, fp = _problem(63, n=300)
    _, s, _ = pm.optimize(y0, nn, mn, fp, 1.0, (2, 1, 1), snapshots=[0, 2])
    assert np.array_equal(s[0], y0)


# ---------------------------------------------------------- fit_transform --
def test_fit_transform_value_allprovided_matches_oracle():
    rng = np.random.default_rng(70)
    n, d = 1200, 30
    x = rng.normal(size=(n, d)).astype(np.float32)
    v = rng.normal(size=(n, 2)).astype(np.float32)
    nn, mn, fp = _pairs(rng, n, n * 10), _pairs(rng, n, n * 5), _pairs(rng, n, n * 20)
    it = (4, 3, 3)
    cfg = pm.Configuration(initialization=pm.Initialization.Value(v),
                           pair_configuration=pm.PairConfiguration.AllProvided(nn, mn, fp),
                           num_iters=it)
    y, snaps, loss = pm.fit_transform(x, cfg, return_loss=True)
    y0 = O.value_init(v)
    ry, _, rloss = O.optimize(y0, nn, mn, fp, 1.0, it)
    assert snaps is None
    assert_parity(y, ry, order_spread(lambda a, b, c: O.optimize(y0, a, b, c, 1.0, it)[0], nn, mn, fp))
    assert_loss(loss, oracle_loss64(y0, nn, mn, fp, 1.0, it))
    # Without a loss buffer the optimiser runs its loss-free kernels: same embedding.
    y2, _ = pm.fit_transform(x, cfg)
    assert np.array_equal(y2, y)


def _oracle_pairs_generate(x, k, n_mn, n_fp, seed):
    """Oracle pipeline of generate_pairs (neighbors.rs:56-108) on normalised x."""
    xn = O.normalize(x)
    n = x.shape[0]
    kp = min(k + 50, n - 1)
    idx, dist = O.knn_rowwise(xn, kp, nthreads=8)
    sig = O.sigma(dist)
    sc = O.scale_dist(dist, sig, idx)
    nn = O.sample_neighbors_pair(sc, idx, min(k, n - 1))
    mn = O.sample_mn_pair_deterministic(xn, n_mn, seed)
    fp = O.sample_fp_pair_deterministic(n, nn, min(k, n - 1), n_fp, seed)
    return nn, mn, fp


def test_fit_transform_generate_matches_oracle_pipeline():
    """Generate + seed + Value init: kNN, sigma, NN/MN/FP pairs and 10 iterations."""
    rng = np.random.default_rng(71)
    n, d = 900, 40
    centers = rng.normal(scale=3.0, size=(6, d))
    x = (centers[rng.integers(0, 6, n)] + rng.normal(size=(n, d))).astype(np.float32)
    v = rng.normal(size=(n, 2)).astype(np.float32)
    it = (5, 3, 2)
    cfg = pm.Configuration(initialization=pm.Initialization.Value(v), seed=42, num_iters=it,
                           override_neighbors=10, mid_near_ratio=1.0)
    y, _, loss = pm.fit_transform(x, cfg, return_loss=True)
    nn, mn, fp = _oracle_pairs_generate(x, 10, 10, 20, 42)
    y0 = O.value_init(v)
    ry, _, rloss = O.optimize(y0, nn, mn, fp, 1.0, it)
    assert_parity(y, ry, order_spread(lambda a, b, c: O.optimize(y0, a, b, c, 1.0, it)[0], nn, mn, fp))
    assert_loss(loss, oracle_loss64(y0, nn, mn, fp, 1.0, it))


def test_fit_transform_neighbors_provided_matches_oracle():
    """PairConfiguration::NeighborsProvided (lib.rs:651-676): the given NN
    pairs are kept, MN and FP pairs come from the seeded samplers on the
    preprocessed data with the given NN pairs as the far-pair reject lists
    (neighbors.rs:127-145); 10 iterations from a Value init at 1e-5."""
    rng = np.random.default_rng(72)
    n, d, k = 1500, 30, 10
    x = rng.normal(size=(n, d)).astype(np.float32)
    v = rng.normal(size=(n, 2)).astype(np.float32)
    nn = np.stack([np.repeat(np.arange(n), k), rng.integers(0, n, n * k)], 1).astype(np.uint32)
    it = (4, 3, 3)
    cfg = pm.Configuration(initialization=pm.Initialization.Value(v), seed=7, num_iters=it,
                           pair_configuration=pm.PairConfiguration.NeighborsProvided(nn))
    y, _, loss = pm.fit_transform(x, cfg, return_loss=True)
    xn = O.normalize(x)
    mn = O.sample_mn_pair_deterministic(xn, 5, 7)
    fp = O.sample_fp_pair_deterministic(n, nn, k, 20, 7)
    y0 = O.value_init(v)
    ry, _, _ = O.optimize(y0, nn, mn, fp, 1.0, it)
    assert_parity(y, ry, order_spread(lambda a, b, c: O.optimize(y0, a, b, c, 1.0, it)[0], nn, mn, fp))
    assert_loss(loss, oracle_loss64(y0, nn, mn, fp, 1.0, it))


def test_random_init_matches_oracle_draw():
    """Initialization::Random(Some(seed)) (lib.rs:769-778): with no iterations
    fit_transform returns the initial Y, which is the oracle's rand_distr
    StandardNormal draw (ziggurat) from SmallRng::seed_from_u64(seed), bit for
    bit, in row-major order (parity unpinned against rand_distr itself)."""
    x = np.random.default_rng(73).normal(size=(500, 12)).astype(np.float32)
    for dims, seed in ((2, 42), (3, 7)):
        cfg = pm.Configuration(initialization=pm.Initialization.Random(seed), embedding_dimensions=dims,
                               num_iters=(0, 0, 0), seed=1)
        y, _ = pm.fit_transform(x, cfg)
        assert np.array_equal(y.view(np.uint32), O.random_init(500, dims, seed).view(np.uint32))


def _reference_test_data():
    """tests.rs:24-37: 1000 x 50 U(-1, 1) plus a per-cluster offset (5 clusters)."""
    rng = np.random.default_rng(80)
    x = rng.uniform(-1, 1, size=(1000, 50)).astype(np.float32)
    x += (np.arange(1000) // 200)[:, None].astype(np.float32)
    return x


def _structure_checks(y, snaps, cfg, n=1000):
    """tests.rs:79-147."""
    assert y.shape == (n, cfg.embedding_dimensions)
    assert np.isfinite(y).all()
    assert (np.abs(y.mean(0)) < 1.0).all()
    assert (y.std(0) > 1e-6).all()
    lab = np.arange(n) // 200
    dd = np.sqrt(((y[:, None, :].astype(np.float64) - y[None, :, :]) ** 2).sum(-1))
    iu = np.triu_indices(n, 1)
    same = (lab[:, None] == lab[None, :])[iu]
    assert dd[iu][same].mean() < dd[iu][~same].mean()
    if cfg.snapshots is not None:
        assert snaps.shape == (len(cfg.snapshots), n, cfg.embedding_dimensions)


@pytest.mark.parametrize("which", [0, 1, 2])
def test_fit_transform_reference_configurations(which):
    """The three configurations of tests.rs:40-71."""
    x = _reference_test_data()
    cfgs = [
        pm.Configuration(),
        pm.Configuration(initialization=pm.Initialization.Pca(), override_neighbors=15, seed=42,
                         snapshots=[100, 200, 300]),
        pm.Configuration(initialization=pm.Initialization.Random(42), mid_near_ratio=0.3,
                         far_pair_ratio=1.5, override_neighbors=10, seed=42, learning_rate=0.8,
                         num_iters=(50, 50, 100)),
    ]
    y, snaps = pm.fit_transform(x, cfgs[which])
    _structure_checks(y, snaps, cfgs[which])


def test_fit_transform_pca_path_high_dimensional():
    """d > 100 with PCA init: randomized-PCA path of lib.rs:456-484 (parity unpinned)."""
    rng = np.random.default_rng(90)
    n, d = 3000, 200
    centers = rng.normal(scale=2.0, size=(8, d))
    lab = rng.integers(0, 8, n)
    x = (centers[lab] + rng.normal(size=(n, d))).astype(np.float32)
    y, _ = pm.fit_transform(x, pm.Configuration(seed=7))
    assert np.isfinite(y).all() and y.shape == (n, 2)
    # 10-NN preservation against the input space is well above chance (the
    # isotropic 200-d noise makes within-cluster neighbours nearly random)
    from_in = O.knn_rowwise(x, 10, nthreads=8)[0]
    from_out = O.knn_rowwise(y, 10, nthreads=8)[0]
    pres = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(from_in, from_out)])
    print(f"10-NN preservation {pres:.3f} (chance {10 / n:.4f})", flush=True)
    assert pres > 5 * 10 / n
    # clusters separate: same-label neighbours dominate in the embedding
    assert (lab[from_out] == lab[:, None]).mean() > 0.9


def _pca_init_expected(x, comps):
    """FP64 numpy restatement of the PCA initialisation: Y0 = (x - mean) V[:, :2]
    * 0.01 (lib.rs:767-768), V the leading covariance eigenvectors, each signed
    so that its largest-magnitude loading is positive."""
    xd = x.astype(np.float64)
    xc = xd - xd.mean(axis=0)
    w, v = np.linalg.eigh(xc.T @ xc)
    order = np.argsort(-w, kind="stable")[:comps]
    v = v[:, order]
    v = v * np.sign(v[np.abs(v).argmax(axis=0), np.arange(comps)])
    return xc @ v[:, :2] * 0.01, w[order]


@pytest.mark.parametrize("n,d", [(4000, 60), (4000, 120), (3000, 300),
                                 (4000, 61), (3000, 105)])  # odd sizes: a dummy index per Jacobi round
def test_pca_initialisation_matches_fp64_eigensolve(n, d):
    """With 0 iterations fit_transform returns its PCA initialisation.  d <= 100
    normalises and takes an exact 2-component PCA (lib.rs:489-519); d > 100
    projects onto the leading 100 axes (lib.rs:456-484) and keeps the first two
    columns.  Both must equal the FP64 eigenvectors of the covariance (parity
    unpinned against petal-decomposition; this pins the axes themselves).
    Spectrum 10 * 0.85^k keeps every eigenvalue gap wide."""
    rng = np.random.default_rng(n + d)
    basis, _ = np.linalg.qr(rng.standard_normal((d, d)))
    x = (rng.standard_normal((n, d)) * (10.0 * 0.85 ** np.arange(d))) @ basis.T + 3.0
    x = x.astype(np.float32)
    cfg = pm.Configuration(seed=3, num_iters=(0, 0, 0))
    y, _ = pm.fit_transform(x, cfg)
    if d <= 100:
        xn = x.astype(np.float32)
        xn = (xn - xn.min()) / xn.max()  # lib.rs:500-503 (global min and max)
        xn = xn - xn.astype(np.float64).mean(axis=0)
        exp, _ = _pca_init_expected(xn, 2)
    else:
        exp, _ = _pca_init_expected(x, 100)
    err = np.abs(y - exp).max() / np.abs(exp).max()
    print(f"PCA init: max error {err:.2e} of max |Y0|", flush=True)
    assert err < 1e-4


def test_pca_host_upload_in_chunks_matches_device_input():
    """A host X on the PCA path is uploaded in 32 MiB row chunks whose FP64
    sums run while the next chunk is in flight; X already in device memory
    takes the same chunks in one launch.  The PCA initialisation, the pairs
    and the embedding must be the same bits either way (here 3 chunks, the
    last one short), and the axes must still be the FP64 ones, also for data
    far from the origin (the sums are taken about the first chunk's mean)."""
    rng = np.random.default_rng(77)
    n, d = 60000, 320  # 32 MiB holds 26,214 rows of 320 floats
    basis, _ = np.linalg.qr(rng.standard_normal((d, d)))
    x = ((rng.standard_normal((n, d)) * (10.0 * 0.85 ** np.arange(d))) @ basis.T + 1000.0).astype(np.float32)
    cfg = pm.Configuration(seed=3, num_iters=(0, 0, 0))
    y_host, _ = pm.fit_transform(x, cfg)
    y_dev, _ = pm.fit_transform(pm.DeviceArray(x), cfg)
    assert np.array_equal(y_host.view(np.uint32), y_dev.view(np.uint32))
    exp, _ = _pca_init_expected(x, 100)
    err = np.abs(y_host - exp).max() / np.abs(exp).max()
    print(f"chunked PCA init (mean 1000): max error {err:.2e}", flush=True)
    assert err < 1e-4
    cfg = pm.Configuration(seed=3, num_iters=(2, 2, 2))
    y_host, _ = pm.fit_transform(x, cfg)
    y_dev, _ = pm.fit_transform(pm.DeviceArray(x), cfg)
    assert np.array_equal(y_host.view(np.uint32), y_dev.view(np.uint32))


def test_pca_initialisation_rank_deficient():
    """n < d: the covariance has rank n - 1 < 100 requested axes; the
    orthonormalisation takes its shifted retry and the leading axes are still
    the FP64 ones."""
    rng = np.random.default_rng(5)
    n, d = 60, 300
    x = (rng.standard_normal((n, d)) * np.linspace(3.0, 0.5, d)).astype(np.float32)
    y, _ = pm.fit_transform(x, pm.Configuration(seed=3, num_iters=(0, 0, 0)))
    exp, w = _pca_init_expected(x, min(100, n))
    assert np.isfinite(y).all()
    err = np.abs(y - exp).max() / np.abs(exp).max()
    print(f"rank-deficient PCA init: max error {err:.2e}", flush=True)
    assert err < 1e-3


def test_pointer_kinds_checked_at_the_boundary():
    """A device allocation where host memory is expected, or host memory where
    device memory is, returns InvalidArgument instead of faulting."""
    import ctypes as C
    x = np.random.default_rng(74).normal(size=(300, 8)).astype(np.float32)
    xd = pm.DeviceArray(x)
    c = pm.Configuration(num_iters=(1, 1, 1))._to_c([])
    y = np.zeros((300, 2), np.float32)
    st = pm.Status()
    rc = pm.lib().pacmap_fit_transform(xd.ptr, 300, 8, C.byref(c), y.ctypes.data, None, None, C.byref(st))
    assert rc == pm.InvalidArgument.code and b"device memory" in st.message
    rc = pm.lib().pacmap_fit_transform_dev(x.ctypes.data, 300, 8, C.byref(c), y.ctypes.data, None, None,
                                           C.byref(st))
    assert rc == pm.InvalidArgument.code and b"not device memory" in st.message
    yd = pm.DeviceArray(y)
    rc = pm.lib().pacmap_fit_transform_dev(xd.ptr, 300, 8, C.byref(c), yd.ptr, None, None, C.byref(st))
    assert rc == pm.InvalidArgument.code
    # the rejected calls leave no error behind: a valid call still succeeds
    out, _ = pm.fit_transform(xd, pm.Configuration(num_iters=(1, 1, 1)))
    assert np.isfinite(out).all()
