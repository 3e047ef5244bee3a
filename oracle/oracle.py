"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module.  It is the checker, never the product: the product path
is the HIP library in ``pacmap-rs_amd`` and never loads this file.

Every wrapper mirrors one function of beamform/pacmap-rs 0.2.6; the C file
(``pacmap_oracle.c``) cites the reference file:line it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u64 = C.c_uint64


def build() -> None:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_find_weights.argtypes = [C.c_float, _u64, _u64, _u64, _f32p]
        L.orc_grad.argtypes = [_f32p, _u64, _u64, _u32p, _u64, _u32p, _u64, _u32p, _u64,
                               C.c_float, C.c_float, C.c_float, _f32p, C.c_int]
        L.orc_grad_f64.argtypes = [_f64p, _u64, _u64, _u32p, _u64, _u32p, _u64, _u32p, _u64,
                                   C.c_double, C.c_double, C.c_double, _f64p]
        L.orc_powi.argtypes = [C.c_float, C.c_int32]
        L.orc_powi.restype = C.c_float
        L.orc_adam_lr_t.argtypes = [C.c_float, C.c_float, C.c_float, _u64]
        L.orc_adam_lr_t.restype = C.c_float
        L.orc_adam.argtypes = [_f32p, _f32p, _f32p, _f32p, _u64, _u64, C.c_float, C.c_float,
                               C.c_float, _u64]
        L.orc_euclid.argtypes = [_f32p, _f32p, _u64]
        L.orc_euclid.restype = C.c_float
        L.orc_euclid_f64.argtypes = [_f32p, _f32p, _u64]
        L.orc_euclid_f64.restype = C.c_double
        L.orc_scale_dist.argtypes = [_f32p, _f32p, _u32p, _u64, _u64, _f32p]
        L.orc_sigma.argtypes = [_f32p, _u64, _u64, _f32p]
        L.orc_knn_triple_sort.argtypes = [_f32p, _u64, _u64, _u64, _u32p, _f32p, C.c_int]
        L.orc_knn_rowwise.argtypes = [_f32p, _u64, _u64, _u64, _u32p, _f32p, C.c_int]
        L.orc_decide_num_pairs.argtypes = [_u64, C.c_int, _u64, C.c_float, C.c_float, _u64p]
        L.orc_rng_seed_from_u64.argtypes = [C.c_void_p, _u64]
        L.orc_rng_next_u64.argtypes = [C.c_void_p]
        L.orc_rng_next_u64.restype = C.c_uint64
        L.orc_rng_gen_range_u32.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.orc_rng_gen_range_u32.restype = C.c_uint32
        L.orc_sample_fp.argtypes = [_u64, C.c_uint32, _u32p, _u64, _u64, C.c_uint32, C.c_void_p,
                                    _u32p]
        L.orc_sample_fp.restype = C.c_uint64
        L.orc_sample_fp_pair_det.argtypes = [_u64, _u32p, _u64, _u64, _u64, _u32p]
        L.orc_sample_mn_pair_det.argtypes = [_f32p, _u64, _u64, _u64, _u64, _u32p]
        L.orc_sample_neighbors_pair.argtypes = [_f32p, _u32p, _u64, _u64, _u64, _u32p]
        L.orc_value_init.argtypes = [_f32p, _u64, _u64, _f32p]
        L.orc_normalize.argtypes = [_f32p, _u64, _u64, _f32p]
        L.orc_knn_rows.argtypes = [_f32p, _u64, _u64, _u64, _u32p, _u64, _u32p, _f32p, C.c_int]
        L.orc_sample_mn_rows.argtypes = [_f32p, _u64, _u64, _u64, _u64, _u32p, _u64, _u32p]
        L.orc_sample_fp_rows.argtypes = [_u64, _u32p, _u64, _u64, _u64, _u32p, _u64, _u32p]
        L.orc_rng_seed_from_bytes.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_zig_tables.argtypes = [_f64p, _f64p]
        L.orc_standard_normal.argtypes = [C.c_void_p]
        L.orc_standard_normal.restype = C.c_double
        L.orc_random_init.argtypes = [_u64, _u64, _u64, _f32p]
        L.orc_optimize.argtypes = [_f32p, _u64, _u64, _u32p, _u64, _u32p, _u64, _u32p, _u64,
                                   C.c_float, _u64p, C.c_void_p, _u64, C.c_void_p, C.c_void_p,
                                   C.c_int]
        _lib = L
    return _lib


def _pairs(p):
    p = np.ascontiguousarray(p, dtype=np.uint32).reshape(-1, 2)
    return p, p.shape[0]


def find_weights(w_mn_init, itr, p1, p2):
    out = np.zeros(3, np.float32)
    lib().orc_find_weights(w_mn_init, itr, p1, p2, out)
    return out  # w_nn, w_mn, w_fp


def pacmap_grad(y, nn, mn, fp, w_nn, w_mn, w_fp, nthreads=1):
    """gradient.rs:35-74 -> (n+1) x dims array, loss at [n, 0]."""
    y = np.ascontiguousarray(y, np.float32)
    n, dims = y.shape
    nn, a = _pairs(nn)
    mn, b = _pairs(mn)
    fp, c = _pairs(fp)
    out = np.zeros((n + 1, dims), np.float32)
    rc = lib().orc_grad(y, n, dims, nn, a, mn, b, fp, c, w_nn, w_mn, w_fp, out, nthreads)
    if rc:
        raise ValueError(f"orc_grad failed ({rc})")
    return out


def pacmap_grad_f64(y, nn, mn, fp, w_nn, w_mn, w_fp):
    y = np.ascontiguousarray(y, np.float64)
    n, dims = y.shape
    nn, a = _pairs(nn)
    mn, b = _pairs(mn)
    fp, c = _pairs(fp)
    out = np.zeros((n + 1, dims), np.float64)
    rc = lib().orc_grad_f64(y, n, dims, nn, a, mn, b, fp, c, w_nn, w_mn, w_fp, out)
    if rc:
        raise ValueError(f"orc_grad_f64 failed ({rc})")
    return out


def powi(a, b):
    return lib().orc_powi(a, b)


def adam_lr_t(lr, beta1, beta2, itr):
    return lib().orc_adam_lr_t(lr, beta1, beta2, itr)


def update_embedding_adam(y, grad, m, v, beta1, beta2, lr, itr):
    """adam.rs:35-60, in place on float32 arrays y, m, v."""
    n, dims = y.shape
    g = np.ascontiguousarray(grad[:n], np.float32)
    lib().orc_adam(y, g, m, v, n, dims, beta1, beta2, lr, itr)


def euclid(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().orc_euclid(a, b, a.size)


def euclid_f64(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().orc_euclid_f64(a, b, a.size)


def scale_dist(knn_dist, sig, nbrs):
    knn_dist = np.ascontiguousarray(knn_dist, np.float32)
    n, k = knn_dist.shape
    out = np.zeros((n, k), np.float32)
    lib().orc_scale_dist(knn_dist, np.ascontiguousarray(sig, np.float32),
                         np.ascontiguousarray(nbrs, np.uint32), n, k, out)
    return out


def sigma(knn_dist):
    knn_dist = np.ascontiguousarray(knn_dist, np.float32)
    n, k = knn_dist.shape
    out = np.zeros(n, np.float32)
    lib().orc_sigma(knn_dist, n, k, out)
    return out


def knn_triple_sort(x, k, nthreads=1):
    """knn.rs:137-194 (ties by neighbour index)."""
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    if n == 0:
        return np.zeros((0, 0), np.uint32), np.zeros((0, 0), np.float32)
    kk = min(k, n - 1)
    idx = np.zeros((n, kk), np.uint32)
    dist = np.zeros((n, kk), np.float32)
    if lib().orc_knn_triple_sort(x, n, d, k, idx, dist, nthreads):
        raise MemoryError("triple array allocation failed")
    return idx, dist


def knn_rowwise(x, k, nthreads=1):
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    if n == 0:
        return np.zeros((0, 0), np.uint32), np.zeros((0, 0), np.float32)
    kk = min(k, n - 1)
    idx = np.zeros((n, kk), np.uint32)
    dist = np.zeros((n, kk), np.float32)
    lib().orc_knn_rowwise(x, n, d, k, idx, dist, nthreads)
    return idx, dist


def knn_rows(x, k, rows, nthreads=1):
    """knn_rowwise for the listed query rows only: (len(rows) x min(k, n-1))."""
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    rows = np.ascontiguousarray(rows, np.uint32)
    kk = min(k, n - 1)
    idx = np.zeros((rows.size, kk), np.uint32)
    dist = np.zeros((rows.size, kk), np.float32)
    if rows.size and lib().orc_knn_rows(x, n, d, k, rows, rows.size, idx, dist, nthreads):
        raise ValueError("row index out of range")
    return idx, dist


def decide_num_pairs(n, override=None, mn_ratio=0.5, fp_ratio=2.0):
    out = np.zeros(3, np.uint64)
    rc = lib().orc_decide_num_pairs(n, override is not None, override or 0, mn_ratio, fp_ratio,
                                    out)
    return rc, tuple(int(v) for v in out)


class Rng:
    """rand 0.8.5 SmallRng::seed_from_u64 restatement (parity unpinned)."""

    def __init__(self, seed):
        self._buf = (C.c_uint64 * 4)()
        lib().orc_rng_seed_from_u64(C.addressof(self._buf), seed)

    @property
    def ptr(self):
        return C.addressof(self._buf)

    @classmethod
    def from_seed(cls, seed_bytes):
        """SeedableRng::from_seed with 32 bytes."""
        r = cls(0)
        lib().orc_rng_seed_from_bytes(r.ptr, bytes(seed_bytes))
        return r

    def standard_normal(self):
        return lib().orc_standard_normal(self.ptr)

    def next_u64(self):
        return lib().orc_rng_next_u64(self.ptr)

    def gen_range(self, low, high):
        return lib().orc_rng_gen_range_u32(self.ptr, low, high)


def sample_fp(n_samples, maximum, reject, self_ind, rng):
    reject = np.ascontiguousarray(reject, np.uint32)
    out = np.zeros(max(n_samples, 1), np.uint32)
    got = lib().orc_sample_fp(n_samples, maximum, reject if reject.size else np.zeros(1, np.uint32),
                              reject.size, 1, self_ind, rng.ptr, out)
    return out[:got]


def sample_fp_pair_deterministic(n, pair_neighbors, n_neighbors, n_fp, seed):
    pn = np.ascontiguousarray(pair_neighbors, np.uint32)
    out = np.zeros((n * n_fp, 2), np.uint32)
    lib().orc_sample_fp_pair_det(n, pn, n_neighbors, n_fp, seed, out)
    return out


def sample_mn_pair_deterministic(x, n_mn, seed):
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    out = np.zeros((n * n_mn, 2), np.uint32)
    lib().orc_sample_mn_pair_det(x, n, d, n_mn, seed, out)
    return out


def sample_mn_rows(x, n_mn, seed, rows):
    """sample_mn_pair_deterministic's rows for the listed points only."""
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    rows = np.ascontiguousarray(rows, np.uint32)
    out = np.zeros((rows.size * n_mn, 2), np.uint32)
    lib().orc_sample_mn_rows(x, n, d, n_mn, seed, rows, rows.size, out)
    return out


def sample_fp_rows(n, pair_neighbors, n_neighbors, n_fp, seed, rows):
    """sample_fp_pair_deterministic's rows for the listed points only."""
    pn = np.ascontiguousarray(pair_neighbors, np.uint32)
    rows = np.ascontiguousarray(rows, np.uint32)
    out = np.zeros((rows.size * n_fp, 2), np.uint32)
    lib().orc_sample_fp_rows(n, pn, n_neighbors, n_fp, seed, rows, rows.size, out)
    return out


def zig_tables():
    x = np.zeros(257)
    f = np.zeros(257)
    lib().orc_zig_tables(x, f)
    return x, f


def random_init(n, dims, seed):
    """lib.rs:769-778 Random(Some(seed)): N(0, 1) f32 draws, row-major."""
    out = np.zeros((n, dims), np.float32)
    lib().orc_random_init(n, dims, seed, out)
    return out


def sample_neighbors_pair(scaled, nbrs, n_neighbors):
    scaled = np.ascontiguousarray(scaled, np.float32)
    n, kp = scaled.shape
    out = np.zeros((n * n_neighbors, 2), np.uint32)
    lib().orc_sample_neighbors_pair(scaled, np.ascontiguousarray(nbrs, np.uint32), n, kp,
                                    n_neighbors, out)
    return out


def value_init(y):
    y = np.ascontiguousarray(y, np.float32)
    out = np.zeros_like(y)
    lib().orc_value_init(y, y.shape[0], y.shape[1], out)
    return out


def normalize(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    rc = lib().orc_normalize(x, x.shape[0], x.shape[1], out)
    if rc:
        raise ValueError("EmptyArrayMinMax")
    return out


def optimize(y0, nn, mn, fp, lr=1.0, iters=(100, 100, 250), snapshots=None, nthreads=1):
    """lib.rs:781-837 from an already-initialised Y. Returns (y, snaps, loss)."""
    y = np.array(y0, np.float32, copy=True, order="C")
    n, dims = y.shape
    nn, a = _pairs(nn)
    mn, b = _pairs(mn)
    fp, c = _pairs(fp)
    it = np.array(iters, np.uint64)
    total = int(it.sum())
    loss = np.zeros(max(total, 1), np.float32)
    snaps = None
    sidx = None
    ns = 0
    if snapshots is not None:
        sidx = np.ascontiguousarray(snapshots, np.uint64)
        ns = sidx.size
        snaps = np.zeros((ns, n, dims), np.float32)
    rc = lib().orc_optimize(y, n, dims, nn, a, mn, b, fp, c, lr, it,
                            sidx.ctypes.data if sidx is not None and ns else None, ns,
                            snaps.ctypes.data if snaps is not None and ns else None,
                            loss.ctypes.data, nthreads)
    if rc:
        raise ValueError(f"orc_optimize failed ({rc})")
    return y, snaps, loss[:total]
