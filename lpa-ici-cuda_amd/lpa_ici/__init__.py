"""Python face of the MI355X LPA-ICI denoiser.

Host helpers (kernel files, noise, PSNR, synthetic images, image I/O) take
and return numpy arrays; device calls take torch tensors on the GPU and pass
their pointers and the current HIP stream to the C ABI.  Every numeric path
runs in liblpa_ici.so; nothing here computes a result itself.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (DEFAULT_KERNELS, FWD_EXACT, FWD_GRAPH, FWD_PLAIN, FWD_RING_FLOAT, FWD_WEIGHTED, GAUSSIAN_M20_KERNELS,
                   KERNEL_DIR, LIB_PATH, LpaError, LpaImage, check, lib)

__all__ = [
    "DEFAULT_KERNELS", "GAUSSIAN_M20_KERNELS", "KERNEL_DIR", "LIB_PATH", "FWD_PLAIN", "FWD_EXACT", "FWD_GRAPH",
    "FWD_WEIGHTED", "FWD_RING_FLOAT",
    "LpaError", "KernelSet", "Plan", "load_kernels", "wedge_kernels", "make_kernels", "write_kernels", "kernel_variances",
    "make_noisy", "psnr", "synth_image",
    "load_image", "write_hdr", "denoise", "estimates", "backward", "image_struct", "same_layout", "tile_configs", "lib",
    "workspace_peak",
]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class KernelSet:
    """A kernel set: the dense cropped kernels on the host (as the reference's
    loader leaves them) and, once `.handle` is first used, the device-resident
    run-length layout."""

    def __init__(self, num_dirs, num_scales, weights, offsets, rows, cols, path=None):
        self.D, self.S = int(num_dirs), int(num_scales)
        self.weights, self.offsets = _f32(weights), _i32(offsets)
        self.rows, self.cols = _i32(rows), _i32(cols)
        self.path = path
        self._handle = None

    @property
    def K(self):
        return self.D * self.S

    def kernel(self, d, s):
        k = d * self.S + s
        o, r, c = self.offsets[k], self.rows[k], self.cols[k]
        return self.weights[o:o + r * c].reshape(r, c)

    def variances(self, sigma):
        return kernel_variances(self, sigma)

    @property
    def handle(self):
        if self._handle is None:
            h = ctypes.c_void_p()
            check(lib().lpa_kset_create(self.D, self.S, _ptr(self.weights), _ptr(self.offsets), _ptr(self.rows),
                                        _ptr(self.cols), ctypes.byref(h)))
            self._handle = h
        return self._handle

    def info(self):
        i = _lib.KsetInfo()
        check(lib().lpa_kset_get_info(self.handle, ctypes.byref(i)))
        return {n: getattr(i, n) for n, _ in _lib.KsetInfo._fields_}

    def runs(self):
        """Host copy of the run-length layout as numpy arrays."""
        info = self.info()
        R, NW, K = info["num_runs"], info["num_weights"], self.K
        out = {n: np.zeros(R, np.int32) for n in ("run_row", "run_col", "run_len", "run_woff")}
        out.update({n: np.zeros(K, np.int32) for n in ("kernel_first_run", "kernel_num_runs", "nnz")})
        out["weights"] = np.zeros(max(NW, 1), np.float32)
        out["sumsq"] = np.zeros(K, np.float32)
        order = ["run_row", "run_col", "run_len", "run_woff", "kernel_first_run", "kernel_num_runs", "weights",
                 "nnz", "sumsq"]
        check(lib().lpa_kset_get_runs(self.handle, *[_ptr(out[n]) for n in order]))
        out["weights"] = out["weights"][:NW]
        return out

    def ring_taps_per_kernel(self):
        """Taps of each kernel's ring (ring form only): the new support of
        scale s over scale s-1, so a walk to scale s adds rings 0..s."""
        out = np.zeros(self.K, np.int64)
        for d in range(self.D):
            prev = set()
            for s in range(self.S):
                k = self.kernel(d, s)
                pos = {(r - k.shape[0] // 2, c - k.shape[1] // 2) for r, c in zip(*np.nonzero(k))}
                out[d * self.S + s] = len(pos - prev)
                prev = pos
        return out

    def ring_runs_per_kernel(self):
        """Row runs of each kernel's ring (ring form only), cut at 32 taps as
        the packer cuts them: the fixed-point walk does two integer adds per
        pixel per run."""
        out = np.zeros(self.K, np.int64)
        for d in range(self.D):
            prev = set()
            for s in range(self.S):
                k = self.kernel(d, s)
                pos = {(r - k.shape[0] // 2, c - k.shape[1] // 2) for r, c in zip(*np.nonzero(k))}
                ring = sorted(pos - prev)
                n, i = 0, 0
                while i < len(ring):
                    j = i + 1
                    while j < len(ring) and ring[j][0] == ring[i][0] and ring[j][1] == ring[j - 1][1] + 1 and j - i < 32:
                        j += 1
                    n += 1
                    i = j
                out[d * self.S + s] = n
                prev = pos
        return out

    def set_tile_config(self, index):
        check(lib().lpa_kset_set_tile_config(self.handle, int(index)))

    def tile_config(self):
        return lib().lpa_kset_get_tile_config(self.handle)

    def close(self):
        if self._handle is not None:
            lib().lpa_kset_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_kernels(path=DEFAULT_KERNELS):
    """load_LPA_kernels_standardized semantics (see include/lpa_ici.h)."""
    D, S, nw = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    p = path.encode()
    check(lib().lpa_load_kernels(p, ctypes.byref(D), ctypes.byref(S), ctypes.byref(nw), None, None, None, None))
    w = np.zeros(nw.value, np.float32)
    K = D.value * S.value
    off, rows, cols = (np.zeros(K, np.int32) for _ in range(3))
    check(lib().lpa_load_kernels(p, ctypes.byref(D), ctypes.byref(S), ctypes.byref(nw), _ptr(w), _ptr(off),
                                 _ptr(rows), _ptr(cols)))
    return KernelSet(D.value, S.value, w, off, rows, cols, path=path)


def wedge_kernels(num_dirs, h):
    """Uniform wedge kernel set: num_dirs directions, window lengths h (one
    scale each); see lpa_wedge_kernels in include/lpa_ici.h."""
    hs = _i32(h)
    nw = ctypes.c_int()
    check(lib().lpa_wedge_kernels(int(num_dirs), hs.size, _ptr(hs), ctypes.byref(nw), None, None, None, None))
    K = int(num_dirs) * hs.size
    w = np.zeros(nw.value, np.float32)
    off, rows, cols = (np.zeros(K, np.int32) for _ in range(3))
    check(lib().lpa_wedge_kernels(int(num_dirs), hs.size, _ptr(hs), ctypes.byref(nw), _ptr(w), _ptr(off),
                                  _ptr(rows), _ptr(cols)))
    return KernelSet(num_dirs, hs.size, w, off, rows, cols)


WINDOWS = {"uniform": 0, "gaussian": 1, "square": 2}


def make_kernels(num_dirs, h, order=1, window="gaussian"):
    """Directional LPA kernel set: num_dirs directions, window lengths h (one
    scale each), polynomial order `order` along the direction, window
    "uniform", "gaussian" or "square"; see lpa_make_kernels in
    include/lpa_ici.h."""
    hs = _i32(h)
    nw = ctypes.c_int()
    wi = WINDOWS[window]
    check(lib().lpa_make_kernels(int(num_dirs), hs.size, _ptr(hs), int(order), wi, ctypes.byref(nw), None, None,
                                 None, None))
    K = int(num_dirs) * hs.size
    w = np.zeros(nw.value, np.float32)
    off, rows, cols = (np.zeros(K, np.int32) for _ in range(3))
    check(lib().lpa_make_kernels(int(num_dirs), hs.size, _ptr(hs), int(order), wi, ctypes.byref(nw), _ptr(w),
                                 _ptr(off), _ptr(rows), _ptr(cols)))
    return KernelSet(num_dirs, hs.size, w, off, rows, cols)


def write_kernels(path, ks):
    """Write a kernel set in the reference's text format (load_kernels reads it back)."""
    check(lib().lpa_write_kernels(str(path).encode(), ks.D, ks.S, _ptr(ks.weights), _ptr(ks.offsets),
                                  _ptr(ks.rows), _ptr(ks.cols)))


def kernel_variances(ks, sigma):
    out = np.zeros(ks.K, np.float32)
    lib().lpa_kernel_variances(ks.K, _ptr(ks.weights), _ptr(ks.offsets), _ptr(ks.rows), _ptr(ks.cols),
                               float(sigma), _ptr(out))
    return out


def make_noisy(clean, sigma, seed):
    clean = _f32(clean)
    out = np.empty_like(clean)
    lib().lpa_make_noisy(_ptr(clean), _ptr(out), clean.size, float(sigma), int(seed) & 0xFFFFFFFF)
    return out


def psnr(a, b):
    a, b = _f32(a), _f32(b)
    assert a.size == b.size
    return lib().lpa_psnr(_ptr(a), _ptr(b), a.size)


def synth_image(width, height, seed):
    out = np.empty((height, width), np.float32)
    lib().lpa_synth_image(_ptr(out), int(width), int(height), int(seed) & 0xFFFFFFFF)
    return out


def load_image(path, channels=1):
    """stbi_loadf semantics: [H, W] for one channel, [H, W, 3] for three."""
    W, H = ctypes.c_int(), ctypes.c_int()
    if channels == 1:
        check(lib().lpa_load_image(path.encode(), ctypes.byref(W), ctypes.byref(H), None, 0))
        out = np.empty((H.value, W.value), np.float32)
        check(lib().lpa_load_image(path.encode(), ctypes.byref(W), ctypes.byref(H), _ptr(out), out.size))
        return out
    check(lib().lpa_load_image_channels(path.encode(), channels, ctypes.byref(W), ctypes.byref(H), None, 0))
    out = np.empty((H.value, W.value, channels), np.float32)
    check(lib().lpa_load_image_channels(path.encode(), channels, ctypes.byref(W), ctypes.byref(H), _ptr(out),
                                        out.size))
    return out


def write_hdr(path, img):
    """[H, W] (replicated to RGB) or [H, W, 3]."""
    img = _f32(img)
    if img.ndim == 2:
        check(lib().lpa_write_hdr(path.encode(), img.shape[1], img.shape[0], _ptr(img)))
    else:
        check(lib().lpa_write_hdr_channels(path.encode(), img.shape[1], img.shape[0], img.shape[2], _ptr(img)))


def workspace_peak(reset=False):
    """Bytes of device memory the library's per-call workspaces held at once
    since the last reset (lpa_workspace_peak)."""
    v = ctypes.c_ulonglong()
    check(lib().lpa_workspace_peak(ctypes.byref(v), int(bool(reset))))
    return v.value


def tile_configs():
    out = []
    for i in range(lib().lpa_num_tile_configs()):
        c = _lib.TileConfig()
        check(lib().lpa_tile_config_get(i, ctypes.byref(c)))
        out.append({n: getattr(c, n) for n, _ in _lib.TileConfig._fields_})
    return out


# ---------------------------------------------------------------------------
# Device calls (torch tensors on the GPU)
# ---------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def _stream_ptr(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def same_layout(x):
    """An uninitialised device tensor with x's shape and row pitch (outputs
    y and dL/dx share the layout of the image x, whatever its pitch)."""
    torch = _torch()
    return torch.empty_strided(tuple(x.shape), x.stride(), dtype=x.dtype, device=x.device)


def image_struct(x, y, scale_map=None, sigma=20.0 / 255.0, gamma=2.0, variances=None, estimates=None):
    """An lpa_image over device tensors x, y (H x W float32, row stride =
    pitch), an optional uint8 [D, H, W] scale map and optional float32
    [D, H, W] per-direction estimates."""
    torch = _torch()
    assert x.is_cuda and y.is_cuda and x.dtype == torch.float32 and y.dtype == torch.float32
    assert x.dim() == 2 and x.shape == y.shape and x.stride(1) == 1 and y.stride(1) == 1
    assert x.stride(0) == y.stride(0)
    if scale_map is not None:
        assert scale_map.dtype == torch.uint8 and scale_map.is_contiguous()
        assert scale_map.shape[1:] == x.shape
    if estimates is not None:
        assert estimates.dtype == torch.float32 and estimates.is_contiguous() and estimates.is_cuda
        assert estimates.shape[1:] == x.shape
    img = LpaImage()
    img.x, img.y = x.data_ptr(), y.data_ptr()
    img.scale_map = scale_map.data_ptr() if scale_map is not None else None
    img.height, img.width = int(x.shape[0]), int(x.shape[1])
    img.pitch = int(x.stride(0))
    img.sigma, img.gamma = float(sigma), float(gamma)
    img.variances = variances.ctypes.data if variances is not None else None
    img.estimates = estimates.data_ptr() if estimates is not None else None
    return img


def denoise(ks, xs, sigmas=20.0 / 255.0, gammas=2.0, scale_maps=False, flags=0, stream=None, variances=None,
            keep_estimates=False):
    """Denoise one (H x W tensor) or several (list) device images.

    Returns y (and the uint8 [D,H,W] scale maps if scale_maps=True, and the
    float32 [D,H,W] per-direction estimates at the selected scales if
    keep_estimates=True, in that order), lists when a list went in.  flags:
    FWD_PLAIN for one launch per image, FWD_EXACT for separately rounded
    mul/add (the CPU path's bits)."""
    torch = _torch()
    single = not isinstance(xs, (list, tuple))
    xs = [xs] if single else list(xs)
    n = len(xs)
    sig = sigmas if isinstance(sigmas, (list, tuple)) else [sigmas] * n
    gam = gammas if isinstance(gammas, (list, tuple)) else [gammas] * n
    var = variances if isinstance(variances, (list, tuple)) else [variances] * n
    var = [None if v is None else _f32(v) for v in var]
    ys = [same_layout(x) for x in xs]
    maps = [torch.empty((ks.D,) + tuple(x.shape), dtype=torch.uint8, device=x.device) for x in xs] \
        if scale_maps else [None] * n
    ests = [torch.empty((ks.D,) + tuple(x.shape), dtype=torch.float32, device=x.device) for x in xs] \
        if keep_estimates else [None] * n
    arr = (LpaImage * n)(*[image_struct(x, y, m, s, g, v, e)
                           for x, y, m, s, g, v, e in zip(xs, ys, maps, sig, gam, var, ests)])
    check(lib().lpa_ici_forward(ks.handle, arr, n, int(flags), _stream_ptr(stream)))
    out = [ys] + ([maps] if scale_maps else []) + ([ests] if keep_estimates else [])
    if single:
        out = [o[0] for o in out]
    return out[0] if len(out) == 1 else tuple(out)


def estimates(ks, xs, maps, flags=0, stream=None):
    """Per-direction estimates at the scales `maps` (uint8 [D, H, W] device
    tensors from a forward) selected, with the forward's walk for the same
    flags: the values that forward aggregated.  xs: a device tensor or a list.
    Returns float32 [D, H, W] device tensors (one, or a list)."""
    torch = _torch()
    single = not isinstance(xs, (list, tuple))
    xs = [xs] if single else list(xs)
    maps = [maps] if single else list(maps)
    outs = [torch.empty((ks.D,) + tuple(x.shape), dtype=torch.float32, device=x.device) for x in xs]
    arr = (LpaImage * len(xs))(*[image_struct(x, x, m, 1.0, 1.0) for x, m in zip(xs, maps)])
    est_p = (ctypes.c_void_p * len(xs))(*[o.data_ptr() for o in outs])
    check(lib().lpa_ici_estimates(ks.handle, arr, est_p, len(xs), int(flags), _stream_ptr(stream)))
    return outs[0] if single else outs


def backward(ks, xs, ys, maps, dys, sigmas=20.0 / 255.0, gammas=2.0, dkernel=None, stream=None, estimates=None):
    """Adjoint of the forward with the scales frozen at `maps`.

    xs: the forward's inputs, ys: its outputs, maps: its uint8 scale maps,
    dys: upstream gradients (all lists of device tensors); estimates: the
    forward's per-direction estimates (keep_estimates=True), used instead of
    recomputing them.  Returns (dxs, dkernel) where dkernel (float32, one
    entry per run weight of the kernel set, see KernelSet.runs()) is
    accumulated into if given."""
    torch = _torch()
    n = len(xs)
    sig = sigmas if isinstance(sigmas, (list, tuple)) else [sigmas] * n
    gam = gammas if isinstance(gammas, (list, tuple)) else [gammas] * n
    dxs = [same_layout(x) for x in xs]
    for d, y in zip(dys, ys):
        assert d.shape == y.shape and d.stride() == y.stride() and d.dtype == torch.float32
    if dkernel is None:
        dkernel = torch.zeros(ks.info()["num_weights"], dtype=torch.float32, device=xs[0].device)
    ests = list(estimates) if estimates is not None else [None] * n
    arr = (LpaImage * n)(*[image_struct(x, y, m, s, g, None, e)
                           for x, y, m, s, g, e in zip(xs, ys, maps, sig, gam, ests)])
    dy_p = (ctypes.c_void_p * n)(*[d.data_ptr() for d in dys])
    dx_p = (ctypes.c_void_p * n)(*[d.data_ptr() for d in dxs])
    check(lib().lpa_ici_backward(ks.handle, arr, dy_p, dx_p, ctypes.c_void_p(dkernel.data_ptr()), n,
                                 _stream_ptr(stream)))
    return dxs, dkernel


class Plan:
    """A fixed batch with its device state uploaded once; run() enqueues the
    forward (a persistent launch, or one launch per image with FWD_PLAIN;
    replayed from a hipGraph with FWD_GRAPH).  Keeps references to the
    tensors it was built over."""

    def __init__(self, ks, xs, ys, maps=None, sigmas=20.0 / 255.0, gammas=2.0, flags=0, estimates=None):
        n = len(xs)
        maps = maps if maps is not None else [None] * n
        ests = estimates if estimates is not None else [None] * n
        sig = sigmas if isinstance(sigmas, (list, tuple)) else [sigmas] * n
        gam = gammas if isinstance(gammas, (list, tuple)) else [gammas] * n
        self._keep = (ks, list(xs), list(ys), list(maps), list(ests))
        arr = (LpaImage * n)(*[image_struct(x, y, m, s, g, None, e)
                               for x, y, m, s, g, e in zip(xs, ys, maps, sig, gam, ests)])
        h = ctypes.c_void_p()
        check(lib().lpa_plan_create(ks.handle, arr, n, int(flags), ctypes.byref(h)))
        self._h = h

    def run(self, stream=None):
        check(lib().lpa_plan_run(self._h, _stream_ptr(stream)))

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().lpa_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
