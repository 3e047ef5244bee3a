"""CPU checks of the C ABI: the library loads, exports every declared symbol,
lays out its structs like the Python mirror, and rejects invalid input on the
host before any GPU work (lib.rs:294-367 error order)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
import pacmap_mi355x as pm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "pacmap.h"), encoding="utf-8").read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pacmap_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = pm.lib()
    names = _declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert pm.version().startswith("pacmap-mi355x")


def test_config_defaults_match_reference():
    """Configuration::default (lib.rs:208-224) through the C struct layout."""
    c = pm.default_config_c()
    assert c.embedding_dimensions == 2
    assert c.init_kind == pm.INIT_PCA
    assert c.mid_near_ratio == np.float32(0.5) and c.far_pair_ratio == np.float32(2.0)
    assert c.has_override_neighbors == 0 and c.has_seed == 0
    assert c.pair_kind == pm.PAIRS_GENERATE
    assert c.learning_rate == np.float32(1.0)
    assert list(c.num_iters) == [100, 100, 250]
    assert c.has_snapshots == 0 and c.approx_threshold == 8000
    assert c.knn_precision == 0 and c.record_loss == 1
    py = pm.Configuration()
    assert (py.embedding_dimensions, py.mid_near_ratio, py.far_pair_ratio, py.learning_rate,
            py.num_iters, py.approx_threshold) == (2, 0.5, 2.0, 1.0, (100, 100, 250), 8000)


def test_weights_seam_matches_oracle_on_host():
    for itr in range(0, 460, 7):
        w_mn, w_nn, w_fp = pm.find_weights(1000.0, itr, 100, 100)
        o = O.find_weights(1000.0, itr, 100, 100)
        assert (w_nn, w_mn, w_fp) == (o[0], o[1], o[2])


def test_sample_size_error():
    with pytest.raises(pm.SampleSize):
        pm.fit_transform(np.zeros((1, 5), np.float32))
    with pytest.raises(pm.SampleSize):
        pm.fit_transform(np.zeros((0, 5), np.float32))


def test_neighbor_and_far_count_errors():
    x = np.random.default_rng(0).normal(size=(50, 5)).astype(np.float32)
    with pytest.raises(pm.InvalidNeighborCount):
        pm.fit_transform(x, pm.Configuration(override_neighbors=0))
    with pytest.raises(pm.InvalidFarPointCount):
        pm.fit_transform(x, pm.Configuration(far_pair_ratio=0.0))


def test_nearest_neighbor_shape_error():
    """lib.rs:651-658: expected [n * n_neighbors, 2]."""
    x = np.random.default_rng(0).normal(size=(50, 5)).astype(np.float32)
    bad = np.zeros((499, 2), np.uint32)
    with pytest.raises(pm.InvalidNearestNeighborShape) as ei:
        pm.fit_transform(x, pm.Configuration(
            pair_configuration=pm.PairConfiguration.NeighborsProvided(bad)))
    st = ei.value.status
    assert list(st.expected) == [500, 2] and list(st.actual) == [499, 2]


def test_empty_features_error():
    with pytest.raises(pm.EmptyArrayMinMax):
        pm.fit_transform(np.zeros((10, 0), np.float32), pm.Configuration(
            initialization=pm.Initialization.Random(1)))


def test_invalid_value_init_shape():
    x = np.zeros((10, 3), np.float32)
    with pytest.raises(pm.InvalidArgument):
        pm.fit_transform(x, pm.Configuration(initialization=pm.Initialization.Value(
            np.zeros((9, 2), np.float32))))


def test_knn_empty_input_without_gpu():
    """knn.rs:141-143 returns (0, 0) arrays before any device work."""
    idx, dist = pm.knn.find_k_nearest_neighbors(np.zeros((0, 128), np.float32), 5)
    assert idx.shape == (0, 0) and dist.shape == (0, 0)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(pm, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(pm, "_lib", None)
    with pytest.raises(ImportError):
        pm.lib()


@pytest.mark.parametrize("n,counts", [(70_000, (23, 12, 46)), (150_000, (28, 14, 56)),
                                      (1_000_000, (40, 20, 80)), (10_000, (10, 5, 20))])
def test_generate_pairs_counts_without_override(n, counts):
    """lib.rs:565-597 as the library applies it: above 10,000 points k is
    round(10 + 15 (log10 n - 4)), MN and FP counts round(k * 0.5) and
    round(k * 2.0).  The count query (NULL outputs) runs no device work."""
    import ctypes as C
    x = np.zeros((n, 3), np.float32)
    c = pm.Configuration()._to_c([])
    out = np.zeros(4, np.uint64)
    st = pm.Status()
    rc = pm.lib().pacmap_generate_pairs(x.ctypes.data, n, 3, C.byref(c), None, None, None, None,
                                        out.ctypes.data, C.byref(st))
    assert rc == 0, st.message
    k, mn, fp = counts
    assert tuple(int(v) for v in out) == (n * k, n * mn, n * fp, 3)
    assert O.decide_num_pairs(n) == (0, counts)
