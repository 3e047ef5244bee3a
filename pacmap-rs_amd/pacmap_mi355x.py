"""Python mirror of pacmap-rs's public interface over libpacmap_mi355x.so.

The reference is a Rust crate whose public surface is

* ``pacmap::fit_transform(x, Configuration)`` (src/lib.rs:294-367),
* ``Configuration`` / ``Initialization`` / ``PairConfiguration`` (src/lib.rs:163-267),
* ``pacmap::knn::find_k_nearest_neighbors(data, k)`` (src/knn.rs:137),
* ``PaCMapError`` (src/lib.rs:870-922).

This module keeps those names, argument meanings and error variants and
calls the MI355X library through its C ABI (include/pacmap.h).  There is no
CPU fallback: importing it without the built HIP library raises.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpacmap_mi355x.so")

# --------------------------------------------------------------------------
# C ABI structures (include/pacmap.h)
# --------------------------------------------------------------------------


class Status(C.Structure):
    _fields_ = [("code", C.c_int32), ("gpu_error", C.c_int32),
                ("expected", C.c_uint64 * 2), ("actual", C.c_uint64 * 2),
                ("message", C.c_char * 256)]


class _Config(C.Structure):
    _fields_ = [
        ("embedding_dimensions", C.c_uint64),
        ("init_kind", C.c_int32),
        ("init_random_has_seed", C.c_int32),
        ("init_random_seed", C.c_uint64),
        ("init_value", C.c_void_p),
        ("init_value_rows", C.c_uint64),
        ("init_value_cols", C.c_uint64),
        ("mid_near_ratio", C.c_float),
        ("far_pair_ratio", C.c_float),
        ("has_override_neighbors", C.c_int32),
        ("override_neighbors", C.c_uint64),
        ("has_seed", C.c_int32),
        ("seed", C.c_uint64),
        ("pair_kind", C.c_int32),
        ("pair_neighbors", C.c_void_p),
        ("pair_neighbors_rows", C.c_uint64),
        ("pair_neighbors_cols", C.c_uint64),
        ("pair_mn", C.c_void_p),
        ("pair_mn_rows", C.c_uint64),
        ("pair_fp", C.c_void_p),
        ("pair_fp_rows", C.c_uint64),
        ("learning_rate", C.c_float),
        ("num_iters", C.c_uint64 * 3),
        ("has_snapshots", C.c_int32),
        ("snapshots", C.c_void_p),
        ("num_snapshots", C.c_uint64),
        ("approx_threshold", C.c_uint64),
        ("knn_precision", C.c_int32),
        ("record_loss", C.c_int32),
    ]


class Timings(C.Structure):
    _fields_ = [(name, C.c_double) for name in (
        "total_ms", "upload_ms", "preprocess_ms", "pca_cov_ms", "pca_eig_ms", "knn_ms",
        "knn_kernel_ms")] + [("knn_kernel_launches", C.c_uint64)] + [
        (name, C.c_double) for name in ("sigma_ms", "sampling_ms", "csr_ms", "graph_build_ms",
                                        "optimize_ms")] + [
        (name, C.c_uint64) for name in ("iterations", "csr_entries", "knn_rescanned_rows",
                                        "knn_candidates")] + [
        ("knn_rerank_ms", C.c_double), ("knn_fallback_ms", C.c_double), ("knn_path", C.c_uint64),
        ("knn_sample_ms", C.c_double), ("knn_certified", C.c_uint64),
        ("grad_target_blocks", C.c_uint64), ("grad_sell", C.c_uint64),
        ("grad_sell_entries", C.c_uint64), ("grad_sell_ranges", C.c_uint64), ("graph_reused", C.c_uint64),
        ("knn_listed", C.c_uint64), ("knn_reranked", C.c_uint64), ("grad_paged", C.c_uint64),
        ("grad_page_words", C.c_uint64), ("grad_page_pages", C.c_uint64), ("knn_tile_pairs", C.c_uint64),
        ("knn_tile_pairs_skipped", C.c_uint64)]


INIT_PCA, INIT_VALUE, INIT_RANDOM = 0, 1, 2
PAIRS_GENERATE, PAIRS_NEIGHBORS_PROVIDED, PAIRS_ALL_PROVIDED = 0, 1, 2
KNN_PRECISIONS = {"fp32": 0, "bf16": 1, "fp8": 2, "direct": 3}
KNN_APPROX = 0x100  # drop the certificate and the rescan (approximate neighbours)

_P = C.c_void_p
_u64 = C.c_uint64
_lib = None


class MlpConfig(C.Structure):
    _fields_ = [
        ("hidden", C.c_uint64),
        ("epochs", C.c_uint64),
        ("batch_pairs", C.c_uint64),
        ("learning_rate", C.c_float),
        ("seed", C.c_uint64),
        ("schedule_iters", C.c_uint64 * 3),
    ]


def lib():
    """Load the HIP library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build the HIP library first "
                              "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        L.pacmap_config_default.argtypes = [C.POINTER(_Config)]
        L.pacmap_fit_transform.argtypes = [_P, _u64, _u64, C.POINTER(_Config), _P, _P, _P,
                                           C.POINTER(Status)]
        L.pacmap_fit_transform_dev.argtypes = L.pacmap_fit_transform.argtypes
        L.pacmap_knn_exact.argtypes = [_P, _u64, _u64, _u64, _P, _P, C.POINTER(Status)]
        L.pacmap_knn.argtypes = [_P, _u64, _u64, _u64, C.c_int32, _P, _P, _P, C.POINTER(Status)]
        L.pacmap_grad.argtypes = [_P, _u64, _u64, _P, _u64, _P, _u64, _P, _u64, C.c_float,
                                  C.c_float, C.c_float, _P, C.POINTER(Status)]
        L.pacmap_adam_step.argtypes = [_P, _P, _P, _P, _u64, _u64, C.c_float, C.c_float,
                                       C.c_float, _u64, C.POINTER(Status)]
        L.pacmap_weights.argtypes = [C.c_float, _u64, _u64, _u64, _P]
        L.pacmap_weights.restype = None
        L.pacmap_sigma_scale.argtypes = [_P, _P, _u64, _u64, _u64, _P, _P, _P, C.POINTER(Status)]
        L.pacmap_scale_dist.argtypes = [_P, _P, _P, _u64, _u64, _P, C.POINTER(Status)]
        L.pacmap_sample_pairs.argtypes = [_P, _u64, _u64, _P, _u64, _u64, _u64, _u64, _P, _P,
                                          C.POINTER(Status)]
        L.pacmap_optimize.argtypes = [_P, _u64, _u64, _P, _u64, _P, _u64, _P, _u64, C.c_float,
                                      _P, _P, _u64, _P, _P, C.POINTER(Status)]
        L.pacmap_generate_pairs.argtypes = [_P, _u64, _u64, C.POINTER(_Config), _P, _P, _P, _P, _P,
                                            C.POINTER(Status)]
        L.pacmap_last_rescanned_rows.argtypes = [_P, _u64]
        L.pacmap_last_rescanned_rows.restype = _u64
        L.pacmap_mlp_config_default.argtypes = [C.POINTER(MlpConfig)]
        L.pacmap_mlp_config_default.restype = None
        L.pacmap_mlp_param_count.argtypes = [_u64, _u64, _u64]
        L.pacmap_mlp_param_count.restype = _u64
        L.pacmap_mlp_init.argtypes = [_u64, _u64, _u64, _u64, _P, C.POINTER(Status)]
        L.pacmap_mlp_fit.argtypes = [_P, _u64, _u64, _u64, _P, _u64, _P, _u64, _P, _u64,
                                     C.POINTER(MlpConfig), _P, _P, C.POINTER(Status)]
        L.pacmap_mlp_transform.argtypes = [_P, _u64, _u64, _u64, _u64, _P, _P, C.POINTER(Status)]
        L.pacmap_mlp_transform_dev.argtypes = L.pacmap_mlp_transform.argtypes
        L.pacmap_mlp_transform_noalloc.argtypes = L.pacmap_mlp_transform.argtypes
        L.pacmap_mlp_grad.argtypes = [_P, _u64, _u64, _u64, _u64, _P, _P, _u64, _P, _u64, _P, _u64,
                                      _P, _P, _P, C.POINTER(Status)]
        L.pacmap_dev_alloc.argtypes = [_u64, C.POINTER(C.c_void_p), C.POINTER(Status)]
        L.pacmap_dev_free.argtypes = [_P]
        L.pacmap_dev_upload.argtypes = [_P, _P, _u64, C.POINTER(Status)]
        L.pacmap_dev_download.argtypes = [_P, _P, _u64, C.POINTER(Status)]
        L.pacmap_release_cached_memory.argtypes = []
        L.pacmap_release_cached_memory.restype = None
        L.pacmap_dev_copy.argtypes = [_P, _P, _u64, C.c_uint32, C.POINTER(C.c_double), C.POINTER(Status)]
        L.pacmap_last_timings.argtypes = [C.POINTER(Timings)]
        L.pacmap_last_timings.restype = None
        L.pacmap_device_info.argtypes = [C.c_char_p, _u64, C.POINTER(C.c_int32)]
        L.pacmap_version.restype = C.c_char_p
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# errors (lib.rs:870-922)
# --------------------------------------------------------------------------
class PaCMapError(Exception):
    code = -1

    def __init__(self, message="", status: Optional[Status] = None):
        super().__init__(message)
        self.status = status


def _err(name, code):
    return type(name, (PaCMapError,), {"code": code})


SampleSize = _err("SampleSize", 1)
InvalidNearestNeighborShape = _err("InvalidNearestNeighborShape", 2)
EmptyArrayMean = _err("EmptyArrayMean", 3)
Normal = _err("Normal", 4)
InvalidNeighborCount = _err("InvalidNeighborCount", 5)
InvalidFarPointCount = _err("InvalidFarPointCount", 6)
EmptyArrayMinMax = _err("EmptyArrayMinMax", 7)
ZeroRange = _err("ZeroRange", 8)
Pca = _err("Pca", 9)
Neighbors = _err("Neighbors", 10)
MissingTransform = _err("MissingTransform", 11)
InvalidPairIndex = _err("InvalidPairIndex", 12)
InvalidArgument = _err("InvalidArgument", 13)
GpuError = _err("GpuError", 14)
Unsupported = _err("Unsupported", 15)
_ERRORS = {c.code: c for c in (SampleSize, InvalidNearestNeighborShape, EmptyArrayMean, Normal,
                               InvalidNeighborCount, InvalidFarPointCount, EmptyArrayMinMax,
                               ZeroRange, Pca, Neighbors, MissingTransform, InvalidPairIndex,
                               InvalidArgument, GpuError, Unsupported)}


def _check(rc, st: Status):
    if rc != 0:
        cls = _ERRORS.get(rc, PaCMapError)
        msg = st.message.decode(errors="replace")
        if rc == InvalidNearestNeighborShape.code:
            msg = (f"Invalid shape for nearest neighbor pairs. Expected {list(st.expected)}, "
                   f"got {list(st.actual)}")
        raise cls(msg, st)


# --------------------------------------------------------------------------
# configuration (lib.rs:163-267)
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Initialization:
    """Initialization::{Pca, Value(Array2<f32>), Random(Option<u64>)} (lib.rs:229-239)."""
    kind: int = INIT_PCA
    value: Optional[np.ndarray] = None
    seed: Optional[int] = None

    @staticmethod
    def Pca():
        return Initialization(INIT_PCA)

    @staticmethod
    def Value(value):
        return Initialization(INIT_VALUE, value=np.ascontiguousarray(value, np.float32))

    @staticmethod
    def Random(seed: Optional[int] = None):
        return Initialization(INIT_RANDOM, seed=seed)


@dataclasses.dataclass
class PairConfiguration:
    """PairConfiguration::{Generate, NeighborsProvided, AllProvided} (lib.rs:244-267)."""
    kind: int = PAIRS_GENERATE
    pair_neighbors: Optional[np.ndarray] = None
    pair_mn: Optional[np.ndarray] = None
    pair_fp: Optional[np.ndarray] = None

    @staticmethod
    def Generate():
        return PairConfiguration(PAIRS_GENERATE)

    @staticmethod
    def NeighborsProvided(pair_neighbors):
        return PairConfiguration(PAIRS_NEIGHBORS_PROVIDED,
                                 pair_neighbors=np.ascontiguousarray(pair_neighbors, np.uint32))

    @staticmethod
    def AllProvided(pair_neighbors, pair_mn, pair_fp):
        f = lambda p: np.ascontiguousarray(p, np.uint32).reshape(-1, 2)  # noqa: E731
        return PairConfiguration(PAIRS_ALL_PROVIDED, f(pair_neighbors), f(pair_mn), f(pair_fp))


@dataclasses.dataclass
class Configuration:
    """Configuration with the reference's defaults (lib.rs:163-224)."""
    embedding_dimensions: int = 2
    initialization: Initialization = dataclasses.field(default_factory=Initialization.Pca)
    mid_near_ratio: float = 0.5
    far_pair_ratio: float = 2.0
    override_neighbors: Optional[int] = None
    seed: Optional[int] = None
    pair_configuration: PairConfiguration = dataclasses.field(
        default_factory=PairConfiguration.Generate)
    learning_rate: float = 1.0
    num_iters: Tuple[int, int, int] = (100, 100, 250)
    snapshots: Optional[Sequence[int]] = None
    approx_threshold: int = 8000
    # MI355X extensions
    knn_precision: str = "fp32"
    record_loss: bool = True

    def _to_c(self, keep):
        c = _Config()
        lib().pacmap_config_default(C.byref(c))
        c.embedding_dimensions = self.embedding_dimensions
        ini = self.initialization
        c.init_kind = ini.kind
        if ini.kind == INIT_VALUE:
            v = np.ascontiguousarray(ini.value, np.float32)
            keep.append(v)
            c.init_value = v.ctypes.data
            c.init_value_rows, c.init_value_cols = (v.shape + (1,))[:2] if v.ndim else (0, 0)
        if ini.kind == INIT_RANDOM and ini.seed is not None:
            c.init_random_has_seed, c.init_random_seed = 1, ini.seed
        c.mid_near_ratio = self.mid_near_ratio
        c.far_pair_ratio = self.far_pair_ratio
        if self.override_neighbors is not None:
            c.has_override_neighbors, c.override_neighbors = 1, self.override_neighbors
        if self.seed is not None:
            c.has_seed, c.seed = 1, self.seed
        pc = self.pair_configuration
        c.pair_kind = pc.kind
        if pc.pair_neighbors is not None:
            p = np.ascontiguousarray(pc.pair_neighbors, np.uint32)
            keep.append(p)
            c.pair_neighbors = p.ctypes.data
            shp = p.shape if p.ndim == 2 else (p.size, 1)
            c.pair_neighbors_rows, c.pair_neighbors_cols = shp
        for name in ("pair_mn", "pair_fp"):
            arr = getattr(pc, name)
            if arr is not None:
                p = np.ascontiguousarray(arr, np.uint32).reshape(-1, 2)
                keep.append(p)
                setattr(c, name, p.ctypes.data)
                setattr(c, name + "_rows", p.shape[0])
        c.learning_rate = self.learning_rate
        for i in range(3):
            c.num_iters[i] = self.num_iters[i]
        if self.snapshots is not None:
            s = np.ascontiguousarray(self.snapshots, np.uint64)
            keep.append(s)
            c.has_snapshots, c.num_snapshots = 1, s.size
            c.snapshots = s.ctypes.data if s.size else None
        c.approx_threshold = self.approx_threshold
        c.knn_precision = KNN_PRECISIONS[self.knn_precision]
        c.record_loss = 1 if self.record_loss else 0
        return c


def default_config_c() -> _Config:
    c = _Config()
    lib().pacmap_config_default(C.byref(c))
    return c


# --------------------------------------------------------------------------
# entry points
# --------------------------------------------------------------------------
class DeviceArray:
    """A float32 matrix resident in device memory (HBM)."""

    def __init__(self, x):
        x = np.ascontiguousarray(x, np.float32)
        self.shape = x.shape
        self.ptr = C.c_void_p()
        st = Status()
        _check(lib().pacmap_dev_alloc(x.nbytes, C.byref(self.ptr), C.byref(st)), st)
        _check(lib().pacmap_dev_upload(self.ptr, x.ctypes.data, x.nbytes, C.byref(st)), st)

    def numpy(self):
        """Copy of the device matrix on the host."""
        out = np.empty(self.shape, np.float32)
        st = Status()
        _check(lib().pacmap_dev_download(out.ctypes.data, self.ptr, out.nbytes, C.byref(st)), st)
        return out

    def free(self):
        if self.ptr:
            lib().pacmap_dev_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fit_transform(x, config: Optional[Configuration] = None, return_loss: bool = False):
    """fit_transform (lib.rs:294-367): returns (embedding, snapshots or None)[, loss]."""
    config = config or Configuration()
    dev = isinstance(x, DeviceArray)
    if not dev:
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 2:
            raise ValueError("x must be 2-D")
    n, d = x.shape
    keep = []
    c = config._to_c(keep)
    dims = config.embedding_dimensions
    y = np.zeros((n, dims), np.float32)
    snaps = None
    if config.snapshots is not None:
        snaps = np.zeros((len(config.snapshots), n, dims), np.float32)
    total = int(sum(config.num_iters))
    # The reference computes the loss inside every gradient call but never
    # returns it (lib.rs:806); it is only produced here when asked for.
    loss = np.zeros(max(total, 1), np.float32) if return_loss else None
    st = Status()
    fn = lib().pacmap_fit_transform_dev if dev else lib().pacmap_fit_transform
    rc = fn(x.ptr if dev else x.ctypes.data, n, d, C.byref(c), y.ctypes.data,
            snaps.ctypes.data if snaps is not None and snaps.size else None,
            loss.ctypes.data if loss is not None else None, C.byref(st))
    _check(rc, st)
    if return_loss:
        return y, snaps, loss[:total]
    return y, snaps


class knn:  # noqa: N801 -- mirrors the `pacmap::knn` module
    @staticmethod
    def find_k_nearest_neighbors(data, k: int, precision: str = "fp32", certify: bool = True):
        """knn.rs:137-194: (n x min(k, n-1) u32 indices, f32 distances), ascending.

        precision "fp32" (split bf16), "bf16" and "direct" are exact; "fp8"
        and certify=False return the exact top-k of the MFMA candidates
        (approximate neighbours)."""
        data = np.ascontiguousarray(data, np.float32)
        if data.ndim != 2:
            raise ValueError("data must be 2-D")
        n, d = data.shape
        if n == 0:
            return np.zeros((0, 0), np.uint32), np.zeros((0, 0), np.float32)
        kk = min(k, n - 1)
        idx = np.zeros((n, kk), np.uint32)
        dist = np.zeros((n, kk), np.float32)
        stats = np.zeros(2, np.uint64)
        st = Status()
        prec = KNN_PRECISIONS[precision] | (0 if certify else KNN_APPROX)
        rc = lib().pacmap_knn(data.ctypes.data, n, d, k, prec,
                              idx.ctypes.data, dist.ctypes.data, stats.ctypes.data, C.byref(st))
        _check(rc, st)
        return idx, dist


def _pairs(p):
    p = np.ascontiguousarray(p, np.uint32).reshape(-1, 2)
    return p, p.shape[0]


def pacmap_grad(y, pair_neighbors, pair_mn, pair_fp, w_neighbors, w_mn, w_fp):
    """Seam for gradient.rs:35-74: (n+1) x dims gradient, loss at [n, 0]."""
    y = np.ascontiguousarray(y, np.float32)
    n, dims = y.shape
    a, na = _pairs(pair_neighbors)
    b, nb = _pairs(pair_mn)
    c, nc = _pairs(pair_fp)
    out = np.zeros((n + 1, dims), np.float32)
    st = Status()
    rc = lib().pacmap_grad(y.ctypes.data, n, dims, a.ctypes.data, na, b.ctypes.data, nb,
                           c.ctypes.data, nc, w_neighbors, w_mn, w_fp, out.ctypes.data,
                           C.byref(st))
    _check(rc, st)
    return out


def update_embedding_adam(y, grad, m, v, beta1, beta2, lr, itr):
    """Seam for adam.rs:35-60; updates the float32 arrays y, m, v in place."""
    for arr in (y, m, v):
        if arr.dtype != np.float32 or not arr.flags.c_contiguous:
            raise TypeError("y, m, v must be C-contiguous float32")
    n, dims = y.shape
    g = np.ascontiguousarray(np.asarray(grad)[:n], np.float32)
    st = Status()
    rc = lib().pacmap_adam_step(y.ctypes.data, g.ctypes.data, m.ctypes.data, v.ctypes.data, n,
                                dims, beta1, beta2, lr, itr, C.byref(st))
    _check(rc, st)


def find_weights(w_mn_init, itr, phase_1_iters, phase_2_iters):
    """Seam for weights.rs:52-81: (w_mn, w_neighbors, w_fp)."""
    out = np.zeros(3, np.float32)
    lib().pacmap_weights(w_mn_init, itr, phase_1_iters, phase_2_iters, out.ctypes.data)
    return float(out[1]), float(out[0]), float(out[2])


def sigma_scale(knn_dist, knn_idx, n_neighbors=0):
    """Seam for neighbors.rs:77-89: (sigma, scaled distances, NN pairs)."""
    knn_dist = np.ascontiguousarray(knn_dist, np.float32)
    knn_idx = np.ascontiguousarray(knn_idx, np.uint32)
    n, kp = knn_dist.shape
    sig = np.zeros(n, np.float32)
    sc = np.zeros((n, kp), np.float32)
    pairs = np.zeros((n * n_neighbors, 2), np.uint32)
    st = Status()
    rc = lib().pacmap_sigma_scale(knn_dist.ctypes.data, knn_idx.ctypes.data, n, kp, n_neighbors,
                                  sig.ctypes.data, sc.ctypes.data,
                                  pairs.ctypes.data if n_neighbors else None, C.byref(st))
    _check(rc, st)
    return sig, sc, pairs


def scale_dist(knn_distances, sig, neighbors):
    """Seam for distance.rs:27-37: knn_distances^2 / (sig_i * sig_neighbor)."""
    dist = np.ascontiguousarray(knn_distances, np.float32)
    sig = np.ascontiguousarray(sig, np.float32)
    nbrs = np.ascontiguousarray(neighbors, np.uint32)
    n, k = dist.shape
    out = np.zeros((n, k), np.float32)
    st = Status()
    _check(lib().pacmap_scale_dist(dist.ctypes.data, sig.ctypes.data, nbrs.ctypes.data, n, k,
                                   out.ctypes.data, C.byref(st)), st)
    return out


def sample_pairs(x, pair_neighbors, n_neighbors, n_mn, n_fp, seed):
    """Seam for the seeded MN/FP samplers (sampling.rs:77-142): (pair_mn, pair_fp)."""
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    pn = np.ascontiguousarray(pair_neighbors, np.uint32)
    mn = np.zeros((n * n_mn, 2), np.uint32)
    fp = np.zeros((n * n_fp, 2), np.uint32)
    st = Status()
    rc = lib().pacmap_sample_pairs(x.ctypes.data, n, d, pn.ctypes.data, n_neighbors, n_mn, n_fp,
                                   seed, mn.ctypes.data, fp.ctypes.data, C.byref(st))
    _check(rc, st)
    return mn, fp


def optimize(y0, pair_neighbors, pair_mn, pair_fp, lr=1.0, num_iters=(100, 100, 250),
             snapshots=None, with_loss=True):
    """Seam for the optimiser loop of pacmap() (lib.rs:781-837) from a given Y.

    Returns (y, snapshots or None, per-iteration loss or None when with_loss
    is False, which runs the loss-free kernels)."""
    y = np.array(y0, np.float32, copy=True, order="C")
    n, dims = y.shape
    a, na = _pairs(pair_neighbors)
    b, nb = _pairs(pair_mn)
    c, nc = _pairs(pair_fp)
    it = np.ascontiguousarray(num_iters, np.uint64)
    total = int(it.sum())
    loss = np.zeros(max(total, 1), np.float32) if with_loss else None
    snaps = sidx = None
    ns = 0
    if snapshots is not None:
        sidx = np.ascontiguousarray(snapshots, np.uint64)
        ns = sidx.size
        snaps = np.zeros((ns, n, dims), np.float32)
    st = Status()
    rc = lib().pacmap_optimize(y.ctypes.data, n, dims, a.ctypes.data, na, b.ctypes.data, nb,
                               c.ctypes.data, nc, lr, it.ctypes.data,
                               sidx.ctypes.data if ns else None, ns,
                               snaps.ctypes.data if ns else None,
                               loss.ctypes.data if loss is not None else None, C.byref(st))
    _check(rc, st)
    return y, snaps, loss[:total] if loss is not None else None


def generate_pairs(x, config: Optional[Configuration] = None):
    """PairConfiguration::Generate as fit_transform runs it (lib.rs:629-650,
    neighbors.rs:56-108) after the same preprocessing (lib.rs:442-532).

    Returns (pair_neighbors, pair_mn, pair_fp, x_preprocessed)."""
    config = config or Configuration()
    x = np.ascontiguousarray(x, np.float32)
    if x.ndim != 2:
        raise ValueError("x must be 2-D")
    n, d = x.shape
    keep = []
    c = config._to_c(keep)
    counts = np.zeros(4, np.uint64)
    st = Status()
    L = lib()
    _check(L.pacmap_generate_pairs(x.ctypes.data, n, d, C.byref(c), None, None, None, None,
                                   counts.ctypes.data, C.byref(st)), st)
    nn = np.zeros((int(counts[0]), 2), np.uint32)
    mn = np.zeros((int(counts[1]), 2), np.uint32)
    fp = np.zeros((int(counts[2]), 2), np.uint32)
    xp = np.zeros((n, int(counts[3])), np.float32)
    _check(L.pacmap_generate_pairs(x.ctypes.data, n, d, C.byref(c), nn.ctypes.data,
                                   mn.ctypes.data, fp.ctypes.data, xp.ctypes.data,
                                   counts.ctypes.data, C.byref(st)), st)
    return nn, mn, fp, xp


class mlp:  # noqa: N801 -- namespace of the parametric encoder entry points
    """Encoder d -> hidden -> hidden -> dims (GELU) trained with the PaCMAP pair
    loss; flat FP32 parameters W1, b1, W2, b2, W3, b3 (row-major, out x in)."""

    @staticmethod
    def param_count(d, hidden=256, dims=2):
        return int(lib().pacmap_mlp_param_count(d, hidden, dims))

    @staticmethod
    def init(d, hidden=256, dims=2, seed=0):
        p = np.zeros(mlp.param_count(d, hidden, dims), np.float32)
        st = Status()
        _check(lib().pacmap_mlp_init(d, hidden, dims, seed, p.ctypes.data, C.byref(st)), st)
        return p

    @staticmethod
    def fit(x, pair_neighbors, pair_mn, pair_fp, params, dims=2, hidden=256, epochs=20,
            batch_pairs=8192, learning_rate=1e-3, seed=0, schedule=(100, 100, 250)):
        """Trains a copy of params; returns (trained params, mean loss per epoch)."""
        x = np.ascontiguousarray(x, np.float32)
        n, d = x.shape
        p = np.array(params, np.float32, copy=True)
        if p.size != mlp.param_count(d, hidden, dims):
            raise ValueError("params has the wrong size for (d, hidden, dims)")
        a, na = _pairs(pair_neighbors)
        b, nb = _pairs(pair_mn)
        c, nc = _pairs(pair_fp)
        cfg = MlpConfig()
        lib().pacmap_mlp_config_default(C.byref(cfg))
        cfg.hidden, cfg.epochs, cfg.batch_pairs = hidden, epochs, batch_pairs
        cfg.learning_rate, cfg.seed = learning_rate, seed
        for i in range(3):
            cfg.schedule_iters[i] = schedule[i]
        loss = np.zeros(max(epochs, 1), np.float32)
        st = Status()
        _check(lib().pacmap_mlp_fit(x.ctypes.data, n, d, dims, a.ctypes.data, na, b.ctypes.data, nb,
                                    c.ctypes.data, nc, C.byref(cfg), p.ctypes.data, loss.ctypes.data,
                                    C.byref(st)), st)
        return p, loss[:epochs]

    @staticmethod
    def transform(x, params, hidden=256, dims=2):
        dev = isinstance(x, DeviceArray)
        if not dev:
            x = np.ascontiguousarray(x, np.float32)
        n, d = x.shape
        p = np.ascontiguousarray(params, np.float32)
        y = np.empty((n, dims), np.float32)  # every element is written
        st = Status()
        fn = lib().pacmap_mlp_transform_dev if dev else lib().pacmap_mlp_transform
        _check(fn(x.ptr if dev else x.ctypes.data, n, d, hidden, dims, p.ctypes.data, y.ctypes.data,
                  C.byref(st)), st)
        return y

    @staticmethod
    def transform_noalloc(x_dev, params_dev, y_dev, hidden=256, dims=2):
        """Allocation-free transform: x_dev (n x d), params_dev and y_dev
        (n x dims) are DeviceArrays; y_dev receives encoder(x)."""
        n, d = x_dev.shape
        if tuple(y_dev.shape) != (n, dims):
            raise ValueError("y_dev must be n x dims")
        st = Status()
        _check(lib().pacmap_mlp_transform_noalloc(x_dev.ptr, n, d, hidden, dims, params_dev.ptr, y_dev.ptr,
                                                  C.byref(st)), st)

    @staticmethod
    def grad(x, params, pair_neighbors, pair_mn, pair_fp, weights, hidden=256, dims=2):
        """Mean pair loss and its parameter gradient over one batch of pairs."""
        x = np.ascontiguousarray(x, np.float32)
        n, d = x.shape
        p = np.ascontiguousarray(params, np.float32)
        a, na = _pairs(pair_neighbors)
        b, nb = _pairs(pair_mn)
        c, nc = _pairs(pair_fp)
        w = np.ascontiguousarray(weights, np.float32)
        g = np.zeros_like(p)
        loss = np.zeros(1, np.float32)
        st = Status()
        _check(lib().pacmap_mlp_grad(x.ctypes.data, n, d, hidden, dims, p.ctypes.data,
                                     a.ctypes.data, na, b.ctypes.data, nb, c.ctypes.data, nc,
                                     w.ctypes.data, g.ctypes.data, loss.ctypes.data, C.byref(st)), st)
        return g, float(loss[0])


def release_cached_memory() -> None:
    """Free the device blocks the library cached for the calling thread."""
    lib().pacmap_release_cached_memory()


def hbm_copy(nbytes: int, repeats: int = 10) -> float:
    """Streaming device copy of nbytes (HBM reference); returns ms per copy."""
    st = Status()
    a, b = C.c_void_p(), C.c_void_p()
    L = lib()
    _check(L.pacmap_dev_alloc(nbytes, C.byref(a), C.byref(st)), st)
    try:
        _check(L.pacmap_dev_alloc(nbytes, C.byref(b), C.byref(st)), st)
        try:
            ms = C.c_double()
            _check(L.pacmap_dev_copy(b, a, nbytes, repeats, C.byref(ms), C.byref(st)), st)
            return ms.value
        finally:
            L.pacmap_dev_free(b)
    finally:
        L.pacmap_dev_free(a)


def last_timings() -> dict:
    t = Timings()
    lib().pacmap_last_timings(C.byref(t))
    return {name: getattr(t, name) for name, _ in Timings._fields_}


def last_rescanned_rows():
    """Rows the last kNN on this thread sent to the exact rescan (ascending)."""
    L = lib()
    count = int(L.pacmap_last_rescanned_rows(None, 0))
    out = np.zeros(count, np.uint32)
    if count:
        L.pacmap_last_rescanned_rows(out.ctypes.data, count)
    return out


def device_info():
    name = C.create_string_buffer(64)
    cus = C.c_int32()
    rc = lib().pacmap_device_info(name, 64, C.byref(cus))
    return (name.value.decode(), cus.value) if rc == 0 else (None, 0)


def version() -> str:
    return lib().pacmap_version().decode()
