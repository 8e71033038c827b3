"""ORACLE — test infrastructure only.

numpy wrappers over liblpa_oracle.so (our CPU restatement of the reference,
lpa_oracle.c) and, where it has been built from the reference's own sources,
oracle/_ref/libref_v0_0.so / libref_v1_0.so.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liblpa_oracle.so")
REF_LIBS = {"v0_0": os.path.join(HERE, "_ref", "libref_v0_0.so"),
            "v1_0": os.path.join(HERE, "_ref", "libref_v1_0.so")}

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_oracle = None
_refs = {}


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} is not built; run `make -C {HERE}`")
        lib = ctypes.CDLL(ORACLE_LIB)
        common = [_P, _P, _I, _I, _I, _I, _P, _P, _P, _P]
        lib.oracle_v0_0.argtypes = common + [_P, _F, _P, _P]
        lib.oracle_v1_0.argtypes = common + [_P, _F]
        lib.oracle_walk_f32.argtypes = common + [_P, _F, _P, _P, _P, _I]
        lib.oracle_walk_f64.argtypes = common + [_D, _D, _P, _P, _P, _I]
        lib.oracle_walk_q.argtypes = common + [_P, _F, _I, _F, _P, _P, _I]
        lib.oracle_adjoint_f64.argtypes = [_P, _P, _I, _I, _I, _I, _P, _P, _P, _P, _D, _P, _P, _P]
        lib.oracle_adjoint_f64_planted.argtypes = [_P, _P, _I, _I, _I, _I, _P, _P, _P, _P, _D, _P, _P, _P, _I]
        lib.oracle_useful_taps.argtypes = [_P, ctypes.c_size_t, _I, _I, _P]
        lib.oracle_useful_taps.restype = ctypes.c_longlong
        for f in ("oracle_v0_0", "oracle_v1_0", "oracle_walk_f32", "oracle_walk_q", "oracle_walk_f64",
                  "oracle_adjoint_f64", "oracle_adjoint_f64_planted"):
            getattr(lib, f).restype = _I
        _oracle = lib
    return _oracle


def _kargs(ks):
    return (_ptr(ks.weights), _ptr(ks.offsets), _ptr(ks.rows), _ptr(ks.cols))


def _img(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return x, x.shape[1], x.shape[0]


def v0_0(x, ks, variances, gamma, want_maps=False):
    """Restated v0_0 CPU path.  Returns y, or (y, scale_maps, margins)."""
    x, W, H = _img(x)
    y = np.zeros((H, W), np.float32)
    var = np.ascontiguousarray(variances, np.float32)
    maps = np.zeros((ks.D, H, W), np.uint8) if want_maps else None
    marg = np.zeros((ks.D, H, W), np.float32) if want_maps else None
    if _lib().oracle_v0_0(_ptr(x), _ptr(y), W, H, ks.D, ks.S, *_kargs(ks), _ptr(var), float(gamma), _ptr(maps),
                          _ptr(marg)):
        raise MemoryError
    return (y, maps, marg) if want_maps else y


def v1_0(x, ks, variances, gamma):
    x, W, H = _img(x)
    y = np.zeros((H, W), np.float32)
    var = np.ascontiguousarray(variances, np.float32)
    if _lib().oracle_v1_0(_ptr(x), _ptr(y), W, H, ks.D, ks.S, *_kargs(ks), _ptr(var), float(gamma)):
        raise MemoryError
    return y


def walk_f32(x, ks, variances, gamma, threads=None, want_est=False):
    """Early-exit walk, the same bits as v0_0.  Returns (y, scale_maps, margins[, est])."""
    x, W, H = _img(x)
    y = np.zeros((H, W), np.float32)
    var = np.ascontiguousarray(variances, np.float32)
    maps = np.zeros((ks.D, H, W), np.uint8)
    marg = np.zeros((ks.D, H, W), np.float32)
    est = np.zeros((ks.D, H, W), np.float32) if want_est else None
    threads = threads or min(16, os.cpu_count() or 1)
    if _lib().oracle_walk_f32(_ptr(x), _ptr(y), W, H, ks.D, ks.S, *_kargs(ks), _ptr(var), float(gamma), _ptr(maps),
                              _ptr(marg), _ptr(est), int(threads)):
        raise MemoryError
    return (y, maps, marg, est) if want_est else (y, maps, marg)


def walk_q(x, ks, variances, gamma, frac_bits, qrange, threads=None, want_est=False):
    """Fixed-point ring-sum walk (uniform-weight kernels).  Returns (y, scale_maps[, est])."""
    x, W, H = _img(x)
    y = np.zeros((H, W), np.float32)
    var = np.ascontiguousarray(variances, np.float32)
    maps = np.zeros((ks.D, H, W), np.uint8)
    est = np.zeros((ks.D, H, W), np.float32) if want_est else None
    threads = threads or min(16, os.cpu_count() or 1)
    rc = _lib().oracle_walk_q(_ptr(x), _ptr(y), W, H, ks.D, ks.S, *_kargs(ks), _ptr(var), float(gamma), int(frac_bits),
                              float(qrange), _ptr(maps), _ptr(est), int(threads))
    if rc == -2:
        raise ValueError("a value lies outside the fixed-point range")
    if rc == -3:
        raise ValueError("kernel weights are not uniform")
    if rc:
        raise MemoryError
    return (y, maps, est) if want_est else (y, maps)


def walk_f64(x, ks, sigma, gamma, frozen=None, threads=None, want_est=False):
    """FP64 walk (variances sigma^2*sum(w^2) in double).  frozen: optional
    uint8 [D,H,W] scales to use instead of walking.  Returns (y, scale_maps[, est])."""
    x, W, H = _img(x)
    y = np.zeros((H, W), np.float64)
    maps = np.zeros((ks.D, H, W), np.uint8)
    est = np.zeros((ks.D, H, W), np.float64) if want_est else None
    fr = None if frozen is None else np.ascontiguousarray(frozen, np.uint8)
    threads = threads or min(16, os.cpu_count() or 1)
    if _lib().oracle_walk_f64(_ptr(x), _ptr(y), W, H, ks.D, ks.S, *_kargs(ks), float(sigma), float(gamma), _ptr(fr),
                              _ptr(maps), _ptr(est), int(threads)):
        raise MemoryError
    return (y, maps, est) if want_est else (y, maps)


def adjoint_f64(x, dy, ks, sigma, scales, planted_mistake=0):
    """FP64 dL/dx (H x W) and dL/dk (dense layout, like ks.weights).
    planted_mistake 1..5 selects one of the oracle's planted mistakes (for the
    test that shows the finite-difference pin catches them)."""
    x, W, H = _img(x)
    dy = np.ascontiguousarray(dy, np.float32)
    sc = np.ascontiguousarray(scales, np.uint8)
    dx = np.zeros((H, W), np.float64)
    dk = np.zeros(ks.weights.size, np.float64)
    args = (_ptr(x), _ptr(dy), W, H, ks.D, ks.S, *_kargs(ks), float(sigma), _ptr(sc), _ptr(dx), _ptr(dk))
    rc = (_lib().oracle_adjoint_f64_planted(*args, int(planted_mistake)) if planted_mistake
          else _lib().oracle_adjoint_f64(*args))
    if rc:
        raise MemoryError
    return dx, dk


def useful_taps(scale_maps, nnz):
    """Nonzero taps the walk evaluates: scales 0..min(s*+1, S-1) per pixel and direction."""
    m = np.ascontiguousarray(scale_maps, np.uint8)
    D = m.shape[0]
    nnz = np.ascontiguousarray(nnz, np.int32)
    S = nnz.size // D
    return int(_lib().oracle_useful_taps(_ptr(m), m[0].size, D, S, _ptr(nnz)))


# ---------------------------------------------------------------------------
# The reference's own CPU code, built from its sources into oracle/_ref
# ---------------------------------------------------------------------------

def ref_available(version="v0_0"):
    return os.path.exists(REF_LIBS[version])


def ref(version="v0_0"):
    if version not in _refs:
        path = REF_LIBS[version]
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C {HERE} ref, where the reference tree exists)")
        lib = ctypes.CDLL(path)
        lib.ref_load_kernels.argtypes = [ctypes.c_char_p, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]
        lib.ref_load_kernels.restype = _I
        lib.ref_kernels_copy.argtypes = [_P, _P, _P, _P]
        lib.ref_variances.argtypes = [_F, _P]
        lib.ref_noisy.argtypes = [_P, _P, _I, _F, ctypes.c_uint]
        lib.ref_psnr.argtypes = [_P, _P, _I]
        lib.ref_psnr.restype = _D
        lib.ref_denoise.argtypes = [_P, _P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _F]
        lib.ref_loadf.argtypes = [ctypes.c_char_p, ctypes.POINTER(_I), ctypes.POINTER(_I), _P, _I]
        lib.ref_loadf.restype = _I
        _refs[version] = lib
    return _refs[version]


def ref_load_kernels(path, version="v0_0"):
    """(D, S, weights, offsets, rows, cols) from the reference's loader, or None."""
    lib = ref(version)
    D, S, n = _I(), _I(), _I()
    if not lib.ref_load_kernels(path.encode(), ctypes.byref(D), ctypes.byref(S), ctypes.byref(n)):
        return None
    K = D.value * S.value
    w = np.zeros(n.value, np.float32)
    off, rows, cols = (np.zeros(K, np.int32) for _ in range(3))
    lib.ref_kernels_copy(_ptr(w), _ptr(off), _ptr(rows), _ptr(cols))
    return D.value, S.value, w, off, rows, cols


def ref_variances(sigma, K, version="v0_0"):
    out = np.zeros(K, np.float32)
    ref(version).ref_variances(float(sigma), _ptr(out))
    return out


def ref_noisy(clean, sigma, seed, version="v0_0"):
    clean = np.ascontiguousarray(clean, np.float32)
    out = np.zeros_like(clean)
    ref(version).ref_noisy(_ptr(clean), _ptr(out), clean.size, float(sigma), int(seed))
    return out


def ref_psnr(a, b, version="v0_0"):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return ref(version).ref_psnr(_ptr(a), _ptr(b), a.size)


def ref_denoise(x, ks, variances, gamma, version="v0_0"):
    x, W, H = _img(x)
    y = np.zeros((H, W), np.float32)
    var = np.ascontiguousarray(variances, np.float32)
    ref(version).ref_denoise(_ptr(x), _ptr(y), W, H, ks.D, ks.S, *_kargs(ks), _ptr(var), float(gamma))
    return y


def ref_loadf(path, version="v0_0"):
    lib = ref(version)
    W, H = _I(), _I()
    if not lib.ref_loadf(path.encode(), ctypes.byref(W), ctypes.byref(H), None, 0):
        return None
    out = np.zeros((H.value, W.value), np.float32)
    lib.ref_loadf(path.encode(), ctypes.byref(W), ctypes.byref(H), _ptr(out), out.size)
    return out
