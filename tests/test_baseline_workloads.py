"""Parity at the benchmark workloads, through the C ABI, against the oracle.

Each workload is built by bench.py's own generator, so these tests check the
data the benchmark measures.  The oracle cannot run a full exact kNN at these
sizes in seconds, so the checks are split by stage:

* kNN: every row the certificate sent to the exact rescan plus a seeded sample
  of rows, brute-forced by the oracle in the reference's distance order
  (distance.rs:50-75, knn.rs:137-194): indices and distance bits identical;
* sigma, scaled distances and NN picks: the oracle's functions applied to the
  GPU's (sample-verified) neighbour table must give the GPU's NN pairs bit for
  bit, over all rows;
* MN and FP pairs: the oracle's seeded samplers (sampling.rs:77-142) on the
  GPU's preprocessed data, bit for bit (all points at 70,000, a sample of
  points at 1,000,000);
* gradient: one evaluation at full size through the optimiser's own kernel
  (LDS-staged sliced ELL at 70,000, paged target-blocked at 1,000,000) against the
  oracle at 1e-5 relative, the loss against FP64;
* optimiser: 10 iterations from a Value initialisation on the full generated
  pair set against the oracle, with a bar set by how far other valid
  summation orders of the oracle itself land (see assert_parity_large); at
  1,000,000 points every step against one oracle step from the GPU's
  previous iterate (see assert_shadowing);
* the full 450-iteration run: 10-NN preservation, same-cluster neighbour share
  and random-triplet accuracy of the GPU embedding within 0.01 of the
  oracle's embedding from the same initial Y and pairs.

Workloads: 70,000 x 784 GMM (PCA -> 100, k = 10, 10 MN, 20 FP); 1,000,000 x
128 Zipf GMM (k = 10, 5 MN, 20 FP); 1,000 x 1,000 U(-1, 1) with Random(42);
200,000 x 256 anisotropic GMM precision sweep; the 150,000 / 100,000-point
Swiss-roll encoder.
"""
import os
import sys
import time

import numpy as np
import pytest

import oracle as O
import pacmap_mi355x as pm

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (the benchmark's data generators)

from test_gpu_parity import (REL, assert_parity, assert_parity_large, assert_rel, assert_shadowing,  # noqa: E402
                             order_spread, other_orders)

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _gradient_parity(y, nn, mn, fp, itr):
    """One gradient at full size through the optimiser's kernel in gradient
    mode, against the oracle at 1e-5 (per element floored at the RMS), with
    the loss against FP64."""
    w = O.find_weights(1000.0, itr, 100, 100)
    got = pm.pacmap_grad(y, nn, mn, fp, *w)
    ref = O.pacmap_grad(y, nn, mn, fp, *w, nthreads=THREADS)
    assert_rel(got[:-1], ref[:-1])
    ref64 = O.pacmap_grad_f64(y.astype(np.float64), nn, mn, fp, *(float(v) for v in w))
    print(f"  loss {got[-1, 0]:.6e}, FP64 {ref64[-1, 0]:.6e}", flush=True)
    assert abs(got[-1, 0] - ref64[-1, 0]) <= REL * abs(ref64[-1, 0])


def _check_knn_rows(x, idx, dist, rows, k):
    """GPU neighbour rows vs the oracle's brute force of the same rows."""
    ri, rd = O.knn_rows(x, k, rows, nthreads=THREADS)
    bad = np.nonzero(~(ri == idx[rows]).all(1) | ~(rd.view(np.uint32) == dist[rows].view(np.uint32)).all(1))[0]
    print(f"  kNN rows checked {len(rows)}, mismatching {len(bad)}", flush=True)
    assert bad.size == 0, f"rows {rows[bad[:10]]} differ from the oracle"


def _nn_from_knn(idx, dist, k):
    """sigma, scale_dist and sample_neighbors_pair (neighbors.rs:79-89) of the oracle."""
    return O.sample_neighbors_pair(O.scale_dist(dist, O.sigma(dist), idx), idx, k)


def _purity(y, lab, k=10):
    """Share of each point's k nearest embedding neighbours with its own label."""
    b, _ = pm.knn.find_k_nearest_neighbors(np.ascontiguousarray(y, np.float32), k)
    return float((lab[b] == lab[:, None]).mean())


def _preservation(x_in, y, k=10):
    a, _ = pm.knn.find_k_nearest_neighbors(x_in, k)
    b, _ = pm.knn.find_k_nearest_neighbors(np.ascontiguousarray(y, np.float32), k)
    a.sort(axis=1)
    b.sort(axis=1)
    return float(np.mean([np.intersect1d(p, q, assume_unique=True).size for p, q in zip(a, b)]) / k)


# ------------------------------------------------------------ 70,000 x 784 --
@pytest.fixture(scope="module")
def c2():
    print("C2: generating 70000 x 784 and its pairs", flush=True)
    x, lab = bench.make_data()
    cfg = bench.config(pm)
    nn, mn, fp, xp = pm.generate_pairs(x, cfg)
    resc = pm.last_rescanned_rows()
    t = pm.last_timings()
    print(f"C2: pairs {len(nn)} / {len(mn)} / {len(fp)}, kNN path {t['knn_path']}, "
          f"rescanned {len(resc)}", flush=True)
    return dict(x=x, lab=lab, cfg=cfg, nn=nn, mn=mn, fp=fp, xp=xp, resc=resc, knn_path=t["knn_path"],
                certified=t["knn_certified"])


def test_c2_knn_rows_bitwise(c2):
    """kNN of the 70,000 x 100 PCA projection: the certified split-bf16 path,
    every rescanned row and 2,000 sampled rows equal the oracle bit for bit."""
    xp = c2["xp"]
    assert xp.shape == (70000, 100) and c2["knn_path"] == 1 and c2["certified"] == 1
    idx, dist = pm.knn.find_k_nearest_neighbors(xp, 60)
    resc = pm.last_rescanned_rows()
    assert np.array_equal(resc, c2["resc"])  # same input, same certificate decisions
    rows = np.union1d(np.random.default_rng(1).choice(70000, 2000, replace=False), resc).astype(np.uint32)
    _check_knn_rows(xp, idx, dist, rows, 60)
    assert (np.diff(dist, axis=1) >= 0).all()
    assert not (idx == np.arange(70000, dtype=np.uint32)[:, None]).any()
    c2["idx"], c2["dist"] = idx, dist


def test_c2_nn_mn_fp_pairs_bitwise(c2):
    """NN pairs from the GPU's neighbour table through the oracle's sigma,
    scale_dist and sort (all 70,000 rows); MN and FP pairs from the oracle's
    seeded samplers on the same preprocessed data (all points)."""
    if "idx" not in c2:
        c2["idx"], c2["dist"] = pm.knn.find_k_nearest_neighbors(c2["xp"], 60)
    nn_ref = _nn_from_knn(c2["idx"], c2["dist"], 10)
    assert np.array_equal(c2["nn"], nn_ref)
    t0 = time.perf_counter()
    assert np.array_equal(c2["mn"], O.sample_mn_pair_deterministic(c2["xp"], 10, 42))
    assert np.array_equal(c2["fp"], O.sample_fp_pair_deterministic(70000, c2["nn"], 10, 20, 42))
    print(f"  oracle samplers {time.perf_counter() - t0:.1f} s", flush=True)


def test_c2_optimizer_10_iterations(c2):
    """Value init + AllProvided with the full 2.8M generated pairs through
    fit_transform (LDS-staged sliced-ELL gradient at this size): 1e-5 against
    the oracle, loss-free and loss-recording runs identical."""
    n = 70000
    v = np.random.default_rng(2).normal(size=(n, 2)).astype(np.float32)
    it = (4, 3, 3)
    allp = pm.PairConfiguration.AllProvided(c2["nn"], c2["mn"], c2["fp"])
    cfg = pm.Configuration(initialization=pm.Initialization.Value(v), pair_configuration=allp,
                           num_iters=it, override_neighbors=10, mid_near_ratio=1.0, seed=42)
    y, _, loss = pm.fit_transform(c2["x"], cfg, return_loss=True)
    t = pm.last_timings()
    assert t["grad_sell"] == 1
    y0 = O.value_init(v)
    ry, _, _ = O.optimize(y0, c2["nn"], c2["mn"], c2["fp"], 1.0, it, nthreads=THREADS)
    assert_parity_large(y, ry, other_orders(y0, c2["nn"], c2["mn"], c2["fp"], it, THREADS))
    y2, _ = pm.fit_transform(c2["x"], cfg)
    assert np.array_equal(y, y2)
    assert pm.pacmap_grad(y, c2["nn"], c2["mn"], c2["fp"], 2.0, 500.0, 1.0) is not None
    assert pm.last_timings()["grad_sell"] == 1
    _gradient_parity(y, c2["nn"], c2["mn"], c2["fp"], 5)
    _gradient_parity(y, c2["nn"], c2["mn"], c2["fp"], 300)


def test_c2_full_run_matches_oracle_quality(c2):
    """The benchmark's own call (PCA init, Generate, 450 iterations) against
    the oracle's 450 iterations from the same initial Y (x_pre[:, :2] * 0.01,
    lib.rs:767) on the same pairs: 10-NN preservation within 0.01 (near chance
    for both: isotropic noise in 784 dimensions leaves no within-cluster
    neighbourhood a 2-D map can keep), the same-cluster share of each point's
    10 embedding neighbours within 0.01, and random-triplet accuracy within
    0.01."""
    xd = pm.DeviceArray(c2["x"])
    y, _ = pm.fit_transform(xd, c2["cfg"])
    y0 = (c2["xp"][:, :2] * np.float32(0.01)).astype(np.float32)
    t0 = time.perf_counter()
    ry, _, _ = O.optimize(y0, c2["nn"], c2["mn"], c2["fp"], 1.0, (100, 100, 250), nthreads=THREADS)
    print(f"  oracle 450 iterations {time.perf_counter() - t0:.1f} s", flush=True)
    p_gpu, p_ref = _preservation(c2["xp"], y), _preservation(c2["xp"], ry)
    u_gpu, u_ref = _purity(y, c2["lab"]), _purity(ry, c2["lab"])
    t_gpu = bench.triplet_accuracy(c2["xp"], y, np.random.default_rng(7))
    t_ref = bench.triplet_accuracy(c2["xp"], ry, np.random.default_rng(7))
    print(f"  10-NN preservation: GPU {p_gpu:.4f}, oracle {p_ref:.4f}; same-cluster neighbours: GPU "
          f"{u_gpu:.4f}, oracle {u_ref:.4f}; triplet accuracy: GPU {t_gpu:.4f}, oracle {t_ref:.4f}", flush=True)
    assert np.isfinite(y).all()
    assert abs(p_gpu - p_ref) <= 0.01 and abs(u_gpu - u_ref) <= 0.01 and abs(t_gpu - t_ref) <= 0.01
    assert u_gpu > 0.99


# --------------------------------------------------------- 1,000,000 x 128 --
@pytest.fixture(scope="module")
def c5():
    print("C5: generating 1000000 x 128 and its pairs", flush=True)
    x, _ = bench.make_large()
    cfg = bench.config_large(pm)
    nn, mn, fp, xp = pm.generate_pairs(x, cfg)
    resc = pm.last_rescanned_rows()
    print(f"C5: pairs {len(nn)} / {len(mn)} / {len(fp)}, rescanned {len(resc)}", flush=True)
    return dict(x=x, nn=nn, mn=mn, fp=fp, xp=xp, resc=resc)


def test_c5_knn_nn_and_sampled_pairs_bitwise(c5):
    """1,000,000 x 100 (PCA of 128): every rescanned row and 300 sampled rows
    of the kNN bit for bit; NN pairs of all rows from the GPU table through
    the oracle; MN and FP pairs of 2,000 sampled points from the oracle's
    seeded samplers."""
    n, xp = 1_000_000, c5["xp"]
    assert xp.shape == (n, 100)
    idx, dist = pm.knn.find_k_nearest_neighbors(xp, 60)
    assert np.array_equal(pm.last_rescanned_rows(), c5["resc"])
    rng = np.random.default_rng(3)
    rows = np.union1d(rng.choice(n, 300, replace=False), c5["resc"]).astype(np.uint32)
    _check_knn_rows(xp, idx, dist, rows, 60)
    assert np.array_equal(c5["nn"], _nn_from_knn(idx, dist, 10))
    del idx, dist
    pts = np.sort(rng.choice(n, 2000, replace=False)).astype(np.uint32)
    mn = O.sample_mn_rows(xp, 5, 42, pts).reshape(-1, 5, 2)
    fp = O.sample_fp_rows(n, c5["nn"], 10, 20, 42, pts).reshape(-1, 20, 2)
    assert np.array_equal(c5["mn"].reshape(n, 5, 2)[pts], mn)
    assert np.array_equal(c5["fp"].reshape(n, 20, 2)[pts], fp)


def test_c5_optimizer_10_iterations_target_blocked(c5):
    """10 iterations of the target-blocked gradient (the path at 1M points,
    in its paged form) on all 35M generated pairs from a Value
    initialisation.  Every step is held to 1e-5 against one oracle step from
    the GPU's previous iterate (assert_shadowing), and the end point has no
    more elements beyond 1e-5 than 3x another oracle order.  The end point's
    norm is only reported: at a million points it is dominated by single
    chaotic rows (measured: row 994139 is 9e-5 off after 3 iterations under
    any order and 0.16 to 0.75 off after 10, under the gather and the paged
    forms of the kernel), which says nothing about the kernel."""
    n = 1_000_000
    y0 = O.value_init(np.random.default_rng(4).normal(size=(n, 2)).astype(np.float32))
    it = (4, 3, 3)
    y, snaps, _ = pm.optimize(y0, c5["nn"], c5["mn"], c5["fp"], 1.0, it, snapshots=list(range(1, 11)),
                              with_loss=False)
    t = pm.last_timings()
    assert t["grad_target_blocks"] == 8 and t["grad_paged"] == 1
    assert np.array_equal(snaps[-1], y)
    assert_shadowing(snaps, y0, c5["nn"], c5["mn"], c5["fp"], 1.0, it, THREADS)
    ry, _, _ = O.optimize(y0, c5["nn"], c5["mn"], c5["fp"], 1.0, it, nthreads=THREADS)
    assert_parity_large(y, ry, other_orders(y0, c5["nn"], c5["mn"], c5["fp"], it, THREADS), norm=False)
    assert pm.pacmap_grad(y, c5["nn"], c5["mn"], c5["fp"], 2.0, 500.0, 1.0) is not None
    assert pm.last_timings()["grad_paged"] == 1
    _gradient_parity(y, c5["nn"], c5["mn"], c5["fp"], 5)
    _gradient_parity(y, c5["nn"], c5["mn"], c5["fp"], 300)


# ----------------------------------------------------- 1,000 x 1,000, Random --
def test_c1_random_init_pipeline():
    """1,000 x 1,000 U(-1, 1) from seed 42, Random(Some(42)) with seed 42,
    exact kNN, default ratios and schedule: the preprocessed data, all pairs
    and the initial Y equal the oracle's (the Random draw is rand_distr's
    ziggurat on the restated SmallRng, parity unpinned against rand_distr);
    10 iterations within 1e-5; the full run's 10-NN preservation and
    random-triplet accuracy within 0.01 of the oracle's (uniform data in
    1,000 dimensions: both near chance, so the embedding's quality is not
    what this checks, only that the two runs agree)."""
    x = np.random.default_rng(42).uniform(-1, 1, size=(1000, 1000)).astype(np.float32)
    cfg = pm.Configuration(initialization=pm.Initialization.Random(42), seed=42)
    nn, mn, fp, xp = pm.generate_pairs(x, cfg)
    xn = O.normalize(x)
    assert np.array_equal(xp.view(np.uint32), xn.view(np.uint32))
    idx, dist = O.knn_rowwise(xn, 60, nthreads=THREADS)
    assert np.array_equal(nn, _nn_from_knn(idx, dist, 10))
    assert np.array_equal(mn, O.sample_mn_pair_deterministic(xn, 5, 42))
    assert np.array_equal(fp, O.sample_fp_pair_deterministic(1000, nn, 10, 20, 42))
    y0 = O.random_init(1000, 2, 42)
    yi, _ = pm.fit_transform(x, pm.Configuration(initialization=pm.Initialization.Random(42), seed=42,
                                                 num_iters=(0, 0, 0)))
    assert np.array_equal(yi, y0)
    it = (4, 3, 3)
    y10, _ = pm.fit_transform(x, pm.Configuration(initialization=pm.Initialization.Random(42), seed=42,
                                                  num_iters=it))
    ry, _, _ = O.optimize(y0, nn, mn, fp, 1.0, it)
    assert_parity(y10, ry, order_spread(lambda a, b, c: O.optimize(y0, a, b, c, 1.0, it)[0], nn, mn, fp))
    y, _ = pm.fit_transform(x, cfg)
    ry, _, _ = O.optimize(y0, nn, mn, fp, 1.0, (100, 100, 250), nthreads=THREADS)
    p_gpu, p_ref = _preservation(xn, y), _preservation(xn, ry)
    t_gpu = bench.triplet_accuracy(xn, y, np.random.default_rng(7))
    t_ref = bench.triplet_accuracy(xn, ry, np.random.default_rng(7))
    print(f"  10-NN preservation: GPU {p_gpu:.4f}, oracle {p_ref:.4f}; triplet accuracy: GPU {t_gpu:.4f}, "
          f"oracle {t_ref:.4f}", flush=True)
    assert abs(p_gpu - p_ref) <= 0.01 and abs(t_gpu - t_ref) <= 0.01


# ------------------------------------------------- 200,000 x 256 precisions --
def test_c3_precision_sweep():
    """Every candidate precision on 200,000 x 256: the certified modes (split
    bf16, bf16, direct) equal the oracle bit for bit on 500 sampled rows and
    every rescanned row; the uncertified modes keep recall@60 against FP64
    above their floors, and every returned distance is the exact FP32 one
    (the oracle's bits where the neighbour matches, within 2e-6 relative of
    FP64 everywhere)."""
    x = bench.make_sweep()
    n, k = x.shape[0], 60
    rows = np.sort(np.random.default_rng(5).choice(n, 500, replace=False)).astype(np.uint32)
    ti, td = bench.fp64_top(x, rows, k)
    ri, rd = O.knn_rows(x, k, rows, nthreads=THREADS)
    floors = {("fp32", True): None, ("bf16", True): None, ("direct", True): None,
              ("fp32", False): 0.9999, ("bf16", False): 0.999, ("fp8", False): 0.99}
    for (mode, cert), floor in floors.items():
        idx, dist = pm.knn.find_k_nearest_neighbors(x, k, precision=mode, certify=cert)
        t = pm.last_timings()
        recall = np.mean([np.intersect1d(idx[q], ti[j]).size for j, q in enumerate(rows)]) / k
        x64 = x.astype(np.float64)
        d64 = np.sqrt(((x64[idx[rows]] - x64[rows][:, None, :]) ** 2).sum(2))
        rel = float((np.abs(dist[rows] - d64) / np.maximum(d64, 1e-30)).max())
        print(f"  {mode} certify={cert}: recall@60 {recall:.6f}, max rel dist err vs FP64 {rel:.2e}, "
              f"rescanned {t['knn_rescanned_rows']}, knn {t['knn_ms']:.1f} ms", flush=True)
        assert rel <= 2e-6
        if floor is None:
            assert t["knn_certified"] == 1
            resc = pm.last_rescanned_rows()
            extra = np.setdiff1d(resc, rows)[:200].astype(np.uint32)
            assert np.array_equal(idx[rows], ri) and np.array_equal(dist[rows].view(np.uint32), rd.view(np.uint32))
            if extra.size:
                _check_knn_rows(x, idx, dist, extra, k)
        else:
            assert t["knn_certified"] == 0
            assert recall >= floor
            same = idx[rows] == ri
            assert np.array_equal(dist[rows][same].view(np.uint32), rd[same].view(np.uint32))


# --------------------------------------------------- Swiss-roll encoder ----
def test_c4_encoder_gain_over_plain_method():
    """What the trained network gains over the method without it, on the
    held-out Swiss roll: embedding 100,000 new points takes one encoder
    transform instead of a whole fit_transform (at least 100x faster, both
    with the points resident in HBM), and its 10-NN preservation is no more
    than 0.03 below the plain fit's.  Triplet accuracy is reported.  The
    training set's pair counts are the n > 10,000 defaults (lib.rs:572-578):
    28 NN, 14 MN and 56 FP pairs per point at 150,000 points."""
    rng = np.random.default_rng(42)
    basis, _ = np.linalg.qr(rng.standard_normal((100, 3)))
    x_tr, _ = bench.make_swiss_roll(150_000, rng, basis)
    x_te, _ = bench.make_swiss_roll(100_000, rng, basis)
    cfg = pm.Configuration(seed=42)
    nn, mn, fp, xp_tr = pm.generate_pairs(x_tr, cfg)
    assert (len(nn), len(mn), len(fp)) == (150_000 * 28, 150_000 * 14, 150_000 * 56)
    lo, hi = float(x_tr.min()), float(x_tr.max())
    mean = ((x_tr - np.float32(lo)) / np.float32(hi)).mean(axis=0, dtype=np.float64).astype(np.float32)
    xp_te = np.ascontiguousarray((x_te - np.float32(lo)) / np.float32(hi) - mean, np.float32)
    params, eloss = pm.mlp.fit(xp_tr, nn, mn, fp, pm.mlp.init(100, 256, 2, seed=0), epochs=20,
                               batch_pairs=8192, seed=1)
    assert eloss[-1] < eloss[0]
    xd_te, xd_raw = pm.DeviceArray(xp_te), pm.DeviceArray(x_te)

    def best(fn, reps):
        out = fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t0)
        return out, min(ts)

    # the 0.2 ms transform is timed over more repetitions than the 20 ms fit
    y_enc, t_enc = best(lambda: pm.mlp.transform(xd_te, params), 21)
    (y_plain, _), t_plain = best(lambda: pm.fit_transform(xd_raw, cfg), 3)
    p_enc, p_plain = _preservation(xp_te, y_enc), _preservation(xp_te, y_plain)
    tr_enc = bench.triplet_accuracy(xp_te, y_enc, np.random.default_rng(7))
    tr_plain = bench.triplet_accuracy(xp_te, y_plain, np.random.default_rng(7))
    print(f"  transform {t_enc * 1e3:.3f} ms, plain fit {t_plain * 1e3:.2f} ms ({t_plain / t_enc:.1f}x); "
          f"10-NN preservation encoder {p_enc:.4f} plain {p_plain:.4f}; triplet accuracy encoder "
          f"{tr_enc:.4f} plain {tr_plain:.4f}", flush=True)
    assert t_plain / t_enc >= 100.0
    assert p_enc >= p_plain - 0.03
