"""Benchmark of the MI355X LPA-ICI denoiser.

Headline (one JSON line on stdout): plain form, one 4096x4096 fp32 synthetic
image (sigma 20/255, gamma 2, 16x7 symmetric m=(1,0) kernel set), Mpixel/s
over K timed forward passes with the input resident in HBM (K steps between
two HIP events on the launch stream; each step also has its own event pair,
whose median is reported beside the mean).  The same line carries
  roofline      useful-tap FP32 FLOP/s of the forward kernel against the
                MI355X vector FP32 peak
  reference_window  the reference's timing window on the same image (from
                after H2D to after D2H: the forward plus the copy of y into
                pageable host memory), median of 5
  float_ring_walk, weighted_walk  the other two fast walks on the same image
  gaussian_m20  the same image with the 16x7 m=(2,0) gaussian-window set
  batched_uneven  256 images with sides U{128..2048}, per-image sigma/gamma:
                plain (one launch per image) vs one persistent launch vs its
                hipGraph replay, each form at its own autotuned tile
                configuration; the batched form's useful-tap rate
  fwd_bwd       64 x 512^2 forward + adjoint (dL/dx, dL/dkernel); the forward
                saves the uint8 scale maps and the adjoint recomputes the
                estimates from them; timed again with the forward keeping
                each direction's selected estimate for the adjoint to read
                ("kept_estimates"); the forward also one launch per image
  config5       32 images with sides U{1024..8192} (about 630 Mpx), fwd+bwd
                in the same two forms, peak device memory against 288 GB
  cpu_baseline  the CPU leg: the reference's own v0_0 and v1_0 code
                (oracle/_ref, one core each) on the 512^2 image and a 1024^2
                image, the 4096^2 rate extrapolated from them and so
                labelled, plus the oracle's v1_0-order walk on every host
                core; and the accuracy of every GPU leg above against the
                oracle on windows (scale agreement, |dy| where the scales
                agree, dPSNR, error against the FP64 walk, adjoint error)
Flags --skip-* drop the side legs.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time
from concurrent.futures import ThreadPoolExecutor

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "lpa-ici-cuda_amd"))

import numpy as np  # noqa: E402
import torch  # noqa: E402

import lpa_ici  # noqa: E402
from lpa_ici import workloads  # noqa: E402
from lpa_ici.metrics import MI355X_FP32_PEAK_TFLOPS, useful_taps  # noqa: E402

METRIC = "Mpixel/s at 4096^2 plain vs batched-uneven; fwd+bwd Mpixel/s; % FP32 VALU peak"
SIGMA, GAMMA = workloads.SIGMA, workloads.GAMMA


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WARMUP_MS = 100.0  # least wall time of untimed calls before a timed region


def timed(fn, steps, warmup, stream):
    """At least W untimed calls, and more until they have kept the device busy
    for WARMUP_MS (after a pause the GPU's clocks take some 20 ms of load to
    come back: with 3 warm-up calls the first timed 4096^2 passes run up to
    12 % slower than the steady 1.13 ms); then K calls between two events on
    `stream` (synchronized on both sides), each call also between its own
    pair.  Returns {"ms": mean per call over the K, "median_ms": median of the
    per-call times, "min_ms", "warmup_calls"}."""
    n, t0 = 0, time.perf_counter()
    while n < warmup or (time.perf_counter() - t0) * 1e3 < WARMUP_MS:
        fn()
        n += 1
        if n % 4 == 0:
            torch.cuda.synchronize()  # the host waits on the device, so the clock measures its work
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    ev[0].record(stream)
    for i in range(steps):
        fn()
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    per = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
    return {"ms": ev[0].elapsed_time(ev[-1]) / steps, "median_ms": statistics.median(per), "min_ms": min(per),
            "warmup_calls": n}


def rounded(d):
    """A record with its floats to four significant decimals, nested records too."""
    if isinstance(d, dict):
        return {k: rounded(v) for k, v in d.items()}
    if isinstance(d, list):
        return [rounded(v) for v in d]
    if isinstance(d, float):
        return float(f"{d:.5g}")
    return d


# VALU lane-operations per second without packed math: 256 CUs x 2.4 GHz x
# 4 SIMDs x 16 lanes (one 32-bit add or FMA per lane per clock).
MI355X_LANE_OPS = 256 * 2.4e9 * 64


def walk_of(ks, flags):
    if ks.info()["ring_form"] != 1 or flags & (lpa_ici.FWD_WEIGHTED | lpa_ici.FWD_EXACT):
        return "weighted FMA"
    return "float ring sums" if flags & lpa_ici.FWD_RING_FLOAT else "fixed-point ring sums"


def executed_ops(ks, maps, walk):
    """Arithmetic the hardware does per tap or run, counted from the scale
    maps (scales 0..min(s*+1, S-1) of every pixel and direction): one FMA per
    nonzero tap (weighted), one add per ring tap (float ring sums), two
    integer adds per ring run (fixed point)."""
    if walk == "weighted FMA":
        return useful_taps(maps, ks.runs()["nnz"], ks.S), "f32 fma"
    if walk == "float ring sums":
        return useful_taps(maps, ks.ring_taps_per_kernel(), ks.S), "f32 add"
    return 2 * useful_taps(maps, ks.ring_runs_per_kernel(), ks.S), "i32 add"


def windows(H, W, n, count, seed):
    """Corner, edge and random n x n windows of an H x W image."""
    rng = np.random.default_rng(seed)
    h, w = min(n, H), min(n, W)
    out = [(0, 0), (H - h, W - w), (0, W - w), ((H - h) // 2, 0)]
    out += [(int(rng.integers(0, H - h + 1)), int(rng.integers(0, W - w + 1))) for _ in range(max(0, count - 4))]
    return [(r, c, h, w) for r, c in out[:count]]


def window_data(x_host, y, maps, clean, wins, sigma=SIGMA, gamma=GAMMA):
    """What the CPU leg needs to check a forward on windows (host copies)."""
    y = y.cpu().numpy() if torch.is_tensor(y) else y
    maps = maps.cpu().numpy() if torch.is_tensor(maps) else maps
    parts = []
    for r0, c0, h, w in wins:
        parts.append({"win": (r0, c0, h, w), "y": y[r0:r0 + h, c0:c0 + w].copy(),
                      "maps": maps[:, r0:r0 + h, c0:c0 + w].copy(),
                      "clean": None if clean is None else clean[r0:r0 + h, c0:c0 + w].copy()})
    return {"x": x_host, "sigma": sigma, "gamma": gamma, "parts": parts}


def plain_leg(ks, x_host, steps, warmup, stream, label, flags=0):
    x = torch.from_numpy(x_host).cuda()
    y = torch.empty_like(x)
    maps = torch.empty((ks.D,) + tuple(x.shape), dtype=torch.uint8, device="cuda")
    plan = lpa_ici.Plan(ks, [x], [y], [maps], SIGMA, GAMMA, flags=lpa_ici.FWD_PLAIN | flags)
    with torch.cuda.stream(stream):
        t = timed(lambda: plan.run(stream), steps, warmup, stream)
    ms = t["ms"]
    taps = useful_taps(maps, ks.runs()["nnz"], ks.S)
    walk = walk_of(ks, flags)
    ex, op = executed_ops(ks, maps, walk)
    npx = x.numel()
    tflops = 2.0 * taps / (ms * 1e-3) / 1e12
    ex_rate = ex / (ms * 1e-3)
    log(f"[{label}] {x.shape[1]}x{x.shape[0]}: {ms:.3f} ms/pass (median {t['median_ms']:.3f}), "
        f"{npx / ms / 1e3:.1f} Mpx/s, {taps / npx:.0f} useful taps/px ({tflops:.2f} TFLOP/s algorithmic), "
        f"{walk}: {ex / npx:.0f} {op}/px executed ({ex_rate / MI355X_LANE_OPS:.3f} of the VALU lane-op rate)")
    plan.close()
    return {"ms": ms, "median_ms": t["median_ms"], "warmup_calls": t["warmup_calls"], "mpx_s": npx / ms / 1e3,
            "useful_taps": taps,
            "taps_per_px": taps / npx, "tflops": tflops, "executed_op": op, "executed_ops_per_px": ex / npx,
            "executed_tops": ex_rate / 1e12, "executed_frac_of_lane_ops": ex_rate / MI355X_LANE_OPS, "walk": walk,
            "y": y, "maps": maps, "x": x}


def reference_window(ks, x_host, stream, reps=5):
    """The reference's timing window (src/v4_1_lpa_ici_2D_sparse_shared.cu:
    1216-1240): the forward on a resident image, then y copied into pageable
    host memory; wall time, median of `reps` after one warm-up."""
    x = torch.from_numpy(x_host).cuda()
    y = torch.empty_like(x)
    host = torch.empty(tuple(x.shape), dtype=torch.float32)  # pageable, as the reference's std::vector
    plan = lpa_ici.Plan(ks, [x], [y], None, SIGMA, GAMMA, flags=lpa_ici.FWD_PLAIN)
    times = []
    with torch.cuda.stream(stream):
        for i in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan.run(stream)
            host.copy_(y)
            torch.cuda.synchronize()
            if i:
                times.append((time.perf_counter() - t0) * 1e3)
    plan.close()
    ms = statistics.median(times)
    return {"ms": ms, "mpx_s": x.numel() / ms / 1e3, "runs": reps}


def batched_leg(ks, steps, warmup, stream):
    ws = workloads.uneven(256)
    t0 = time.time()
    xs_host = ws.noisy_all()
    xs = [torch.from_numpy(x).cuda() for x in xs_host]
    log(f"[batched] generated {len(ws)} images, {ws.pixels / 1e6:.1f} Mpx in {time.time() - t0:.1f} s")
    sig, gam = ws.sigmas, ws.gammas
    npx = ws.pixels
    ys_plain = [torch.empty_like(x) for x in xs]
    ys_batch = [torch.empty_like(x) for x in xs]
    ys_graph = [torch.empty_like(x) for x in xs]
    out = {"images": len(ws), "mpx": npx / 1e6}
    # each form at the tile configuration the autotuner picks for it on this
    # workload (a plan keeps the configuration it was built with)
    from lpa_ici.autotune import autotune
    saved = ks.tile_config()
    plain_cfg, _ = autotune(ks, xs, flags=lpa_ici.FWD_PLAIN, sigmas=sig, gammas=gam, stream=stream)
    plans = {"plain": lpa_ici.Plan(ks, xs, ys_plain, None, sig, gam, flags=lpa_ici.FWD_PLAIN)}
    batch_cfg, _ = autotune(ks, xs, flags=0, sigmas=sig, gammas=gam, stream=stream)
    plans["batched"] = lpa_ici.Plan(ks, xs, ys_batch, None, sig, gam, flags=0)
    plans["graph"] = lpa_ici.Plan(ks, xs, ys_graph, None, sig, gam, flags=lpa_ici.FWD_GRAPH)
    out["tile_config"] = {"plain": plain_cfg, "batched": batch_cfg}
    log(f"[batched] tile configurations: plain {plain_cfg}, batched and graph {batch_cfg}")
    with torch.cuda.stream(stream):
        for name, plan in plans.items():
            t = timed(lambda: plan.run(stream), steps, warmup, stream)
            out[name] = {"ms": t["ms"], "median_ms": t["median_ms"], "mpx_s": npx / t["ms"] / 1e3}
            log(f"[batched] {name}: {t['ms']:.2f} ms (median {t['median_ms']:.2f}), {npx / t['ms'] / 1e3:.1f} Mpx/s")
    torch.cuda.synchronize()
    out["identical"] = all(torch.equal(a, b) and torch.equal(a, c) for a, b, c in zip(ys_plain, ys_batch, ys_graph))
    out["batched_over_plain"] = out["batched"]["mpx_s"] / out["plain"]["mpx_s"]
    out["graph_over_plain"] = out["graph"]["mpx_s"] / out["plain"]["mpx_s"]
    for p in plans.values():
        p.close()
    # useful taps of the batch, from the scale maps of one more (untimed)
    # batched run at the same configuration: the batched form's rate on them
    maps = [torch.empty((ks.D,) + tuple(x.shape), dtype=torch.uint8, device="cuda") for x in xs]
    ys_m = [torch.empty_like(x) for x in xs]
    pm = lpa_ici.Plan(ks, xs, ys_m, maps, sig, gam, flags=0)
    with torch.cuda.stream(stream):
        pm.run(stream)
    torch.cuda.synchronize()
    pm.close()
    ks.set_tile_config(saved)
    nnz = ks.runs()["nnz"]
    taps = sum(useful_taps(m, nnz, ks.S) for m in maps)
    for name in ("batched", "graph", "plain"):
        tf = 2.0 * taps / (out[name]["ms"] * 1e-3) / 1e12
        out[name]["tflops"] = tf
        out[name]["frac"] = tf / MI355X_FP32_PEAK_TFLOPS
    out["useful_taps_per_px"] = taps / npx
    out["y_identical_with_maps"] = all(torch.equal(a, b) for a, b in zip(ys_m, ys_batch))
    log(f"[batched] {taps / npx:.0f} useful taps/px: batched {out['batched']['tflops']:.1f} TFLOP/s "
        f"({out['batched']['frac']:.3f} of peak)")
    area = np.array([w * h for w, h in ws.sides])
    pick = [int(i) for i in list(np.argsort(area)[:2]) + list(np.argsort(area)[-2:])]
    check = [window_data(xs_host[i], ys_batch[i], maps[i], ws.clean(i), windows(*xs_host[i].shape, 96, 3, i),
                         ws.sigmas[i], ws.gammas[i]) for i in pick]
    return out, check


def adjoint_variant(ks, xs, ys, maps, sig, gam, dys, dxs, dk, steps, warmup, stream, variant):
    """Times forward and forward + adjoint over the images, in one of two
    forms: "recompute" as the workload states it (the forward saves only the
    uint8 scale maps and the adjoint recomputes each direction's estimate from
    them), or "kept_estimates" (the forward also keeps each direction's
    selected estimate, 16 float maps per image, for the adjoint to read).
    Returns (forward timing, forward + adjoint timing)."""
    import ctypes
    n = len(xs)
    sig = sig if isinstance(sig, list) else [sig] * n
    gam = gam if isinstance(gam, list) else [gam] * n
    dy_p = (ctypes.c_void_p * n)(*[d.data_ptr() for d in dys])
    dx_p = (ctypes.c_void_p * n)(*[d.data_ptr() for d in dxs])
    sp = ctypes.c_void_p(stream.cuda_stream)
    ests = ([torch.empty((ks.D,) + tuple(x.shape), dtype=torch.float32, device="cuda") for x in xs]
            if variant == "kept_estimates" else None)
    fwd = lpa_ici.Plan(ks, xs, ys, maps, sig, gam, flags=0, estimates=ests)
    arr = (lpa_ici.LpaImage * n)(*[lpa_ici.image_struct(x, y, m, s_, g_, None, ests[i] if ests else None)
                                   for i, (x, y, m, s_, g_) in enumerate(zip(xs, ys, maps, sig, gam))])

    def step():
        fwd.run(stream)
        lpa_ici.check(lpa_ici.lib().lpa_ici_backward(ks.handle, arr, dy_p, dx_p, ctypes.c_void_p(dk.data_ptr()),
                                                     n, sp))

    with torch.cuda.stream(stream):
        t_f = timed(lambda: fwd.run(stream), steps, warmup, stream)
        t_fb = timed(step, steps, warmup, stream)
    torch.cuda.synchronize()
    fwd.close()
    del ests
    return t_f, t_fb


def adjoint_record(ks, npx, sel_taps, t_f, t_fb):
    ms_f, ms_fb = t_f["ms"], t_fb["ms"]
    ms_b = ms_fb - ms_f
    return {"fwd_ms": ms_f, "fwd_bwd_ms": ms_fb, "fwd_bwd_median_ms": t_fb["median_ms"],
            "fwd_mpx_s": npx / ms_f / 1e3, "fwd_bwd_mpx_s": npx / ms_fb / 1e3, "bwd_ms": ms_b,
            "bwd_tflops_algorithmic": 8.0 * sel_taps / (ms_b * 1e-3) / 1e12}


def selected_taps(ks, maps):
    nnz = torch.from_numpy(ks.runs()["nnz"].astype(np.int64).reshape(ks.D, ks.S)).cuda()
    dsel = torch.arange(ks.D, device="cuda")[:, None, None]
    return int(sum(nnz[dsel, m.long()].sum().item() for m in maps))


def adjoint_check(x_host, g, maps, dx, sigma, wins):
    """What the CPU leg needs to check dL/dx on windows (host copies)."""
    return {"x": x_host, "g": g.cpu().numpy(), "maps": maps.cpu().numpy(), "dx": dx.cpu().numpy(), "sigma": sigma,
            "wins": wins}


def fwd_bwd_leg(ks, steps, warmup, stream):
    ws = workloads.batch_512(64)
    xs_host = ws.noisy_all()
    xs = [torch.from_numpy(x).cuda() for x in xs_host]
    ys = [torch.empty_like(x) for x in xs]
    maps = [torch.empty((ks.D,) + tuple(x.shape), dtype=torch.uint8, device="cuda") for x in xs]
    g = torch.Generator(device="cuda").manual_seed(5)
    dys = [torch.randn(x.shape, generator=g, device="cuda") for x in xs]
    dxs = [torch.empty_like(x) for x in xs]
    dk = torch.zeros(ks.info()["num_weights"], dtype=torch.float32, device="cuda")
    # the persistent forward at its own tile configuration for these images
    from lpa_ici.autotune import autotune
    saved = ks.tile_config()
    batch_cfg, _ = autotune(ks, xs, flags=0, stream=stream)
    times = {v: adjoint_variant(ks, xs, ys, maps, SIGMA, GAMMA, dys, dxs, dk, steps, warmup, stream, v)
             for v in ("recompute", "kept_estimates")}
    npx = ws.pixels
    # the same forward, one launch per image, at the plain form's own tile
    # configuration for these images
    plain_cfg, _ = autotune(ks, xs, flags=lpa_ici.FWD_PLAIN, stream=stream)
    plain = lpa_ici.Plan(ks, xs, ys, maps, SIGMA, GAMMA, flags=lpa_ici.FWD_PLAIN)
    ks.set_tile_config(saved)
    with torch.cuda.stream(stream):
        ms_plain = timed(lambda: plain.run(stream), steps, warmup, stream)["ms"]
    plain.close()
    log(f"[fwd+bwd] forward one launch per image: {ms_plain:.2f} ms (tile configuration {plain_cfg}), "
        f"persistent {times['recompute'][0]['ms']:.2f} ms (tile configuration {batch_cfg}): "
        f"{ms_plain / times['recompute'][0]['ms']:.2f}x")
    # Algorithmic adjoint work: per selected kernel tap of every pixel and
    # direction, 2 FLOP to recompute yhat, 2 for the dL/dx gather and 4 for
    # dL/dkernel (the reference's taps, whatever the walks execute).
    sel = selected_taps(ks, maps)
    rec = {"images": len(ws), "side": 512, "selected_taps_per_px": sel / npx,
           "fwd_plain_per_image_ms": ms_plain, "fwd_plain_tile_config": plain_cfg, "fwd_batched_tile_config": batch_cfg,
           "fwd_batched_over_plain": ms_plain / times["recompute"][0]["ms"]}
    for variant, (t_f, t_fb) in times.items():
        v = adjoint_record(ks, npx, sel, t_f, t_fb)
        log(f"[fwd+bwd {variant}] 64x512^2: fwd {v['fwd_ms']:.2f} ms, fwd+bwd {v['fwd_bwd_ms']:.2f} ms, "
            f"{v['fwd_bwd_mpx_s']:.1f} Mpx/s; adjoint {v['bwd_ms']:.2f} ms, {sel / npx:.0f} selected taps/px, "
            f"{v['bwd_tflops_algorithmic']:.1f} TFLOP/s algorithmic")
        if variant == "recompute":
            rec.update(v)
        else:
            rec[variant] = v
    # the last timed step left dL/dx of the last call; recompute image 0's
    # gradient alone for the check (dL/dx does not depend on dkernel)
    dx0, _ = lpa_ici.backward(ks, [xs[0]], [ys[0]], [maps[0]], [dys[0]], SIGMA, GAMMA)
    check = {"forward": window_data(xs_host[0], ys[0], maps[0], ws.clean(0), windows(512, 512, 128, 3, 1)),
             "adjoint": adjoint_check(xs_host[0], dys[0], maps[0], dx0[0], SIGMA, windows(512, 512, 48, 2, 2))}
    return rec, check


def config5_leg(ks, steps, warmup, stream):
    """32 images with sides U{1024..8192}, per-image sigma/gamma as the batched
    leg: forward (scale maps kept as uint8) + adjoint that recomputes the
    estimates from them, then the same with the estimates kept.  Peak device
    memory per form: the tensors' high-water mark plus the library
    workspaces' (an upper bound of the simultaneous peak)."""
    ws = workloads.large_uneven(32)
    torch.cuda.synchronize()
    total = torch.cuda.mem_get_info()[1]
    t0 = time.time()
    xs_host = ws.noisy_all()
    xs = [torch.from_numpy(x).cuda() for x in xs_host]
    npx = ws.pixels
    log(f"[config5] generated {len(ws)} images, {npx / 1e6:.1f} Mpx in {time.time() - t0:.1f} s")
    sig, gam = ws.sigmas, ws.gammas
    ys = [torch.empty_like(x) for x in xs]
    maps = [torch.empty((ks.D,) + tuple(x.shape), dtype=torch.uint8, device="cuda") for x in xs]
    g = torch.Generator(device="cuda").manual_seed(7)
    dys = [torch.randn(x.shape, generator=g, device="cuda") for x in xs]
    dxs = [torch.empty_like(x) for x in xs]
    dk = torch.zeros(ks.info()["num_weights"], dtype=torch.float32, device="cuda")
    rec = {"images": len(ws), "mpx": round(npx / 1e6, 1), "hbm_total_gib": round(total / 2**30, 1),
           "scale_maps_gib": round(npx * ks.D / 2**30, 2)}
    sel = None
    # the recompute form runs first, before any estimates exist, so each
    # form's peak is its own
    for variant in ("recompute", "kept_estimates"):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        lpa_ici.workspace_peak(reset=True)
        t_f, t_fb = adjoint_variant(ks, xs, ys, maps, sig, gam, dys, dxs, dk, steps, warmup, stream, variant)
        tensors, work = torch.cuda.max_memory_allocated(), lpa_ici.workspace_peak()
        if sel is None:
            sel = selected_taps(ks, maps)
        v = adjoint_record(ks, npx, sel, t_f, t_fb)
        v["peak_hbm_gib"] = (tensors + work) / 2**30
        v["peak_workspace_gib"] = work / 2**30
        log(f"[config5 {variant}] {len(ws)} images {npx / 1e6:.1f} Mpx: fwd {v['fwd_ms']:.1f} ms "
            f"({v['fwd_mpx_s']:.0f} Mpx/s), fwd+bwd {v['fwd_bwd_ms']:.1f} ms ({v['fwd_bwd_mpx_s']:.0f} Mpx/s); "
            f"peak device memory <= {v['peak_hbm_gib']:.1f} GiB (library workspaces {work / 2**30:.2f}) of "
            f"{total / 2**30:.0f} GiB")
        if variant == "recompute":
            rec.update(v)
        else:
            rec[variant] = v
    i = int(np.argmax([w * h for w, h in ws.sides]))
    H, W = xs_host[i].shape
    dx1, _ = lpa_ici.backward(ks, [xs[i]], [ys[i]], [maps[i]], [dys[i]], sig[i], gam[i])
    check = {"forward": window_data(xs_host[i], ys[i], maps[i], ws.clean(i), windows(H, W, 128, 3, 3), sig[i],
                                    gam[i]),
             "adjoint": adjoint_check(xs_host[i], dys[i], maps[i], dx1[0], sig[i], windows(H, W, 48, 2, 4))}
    return rec, check


def measured_traffic(walk):
    """HBM bytes per launch of the 4096^2 forward from the committed rocprofv3
    counter runs (profiles/forward_plain_4096_symmetric.md), or None."""
    key = {"fixed-point ring sums": "ring_fixed", "float ring sums": "ring_sums", "weighted FMA": "weighted_fma"}[walk]
    try:
        for line in open(os.path.join(REPO, "profiles", "forward_plain_4096_symmetric.traffic")):
            if line.startswith(key + ":"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


# ---------------------------------------------------------------------------
# The CPU leg: the only part of this file that loads oracle/ (the reference's
# CPU builds as the baseline, the oracle as the checker of the GPU outputs)
# ---------------------------------------------------------------------------

def cpu_leg(checks, ks_by_name, threads):
    sys.path.insert(0, REPO)
    import oracle

    def crop_of(H, W, r0, c0, h, w, m=31):
        ra, ca = max(0, r0 - m), max(0, c0 - m)
        rb, cb = min(H, r0 + h + m), min(W, c0 + w + m)
        return (slice(ra, rb), slice(ca, cb)), (slice(r0 - ra, r0 - ra + h), slice(c0 - ca, c0 - ca + w))

    def forward_accuracy(ks, data):
        """The north-star bar on windows: per-direction scale agreement with
        the reference-order FP32 walk, flips only at margins < 1e-6, |dy|
        where every direction agrees, dPSNR over the windows, and the error
        against the FP64 walk (its own selection, and frozen at the GPU's)."""
        x, sigma, gamma = data["x"], data["sigma"], data["gamma"]
        H, W = x.shape
        var = ks.variances(sigma)
        agree = np.zeros(ks.D)
        total = flips = 0
        max_dy = e64 = e64f = 0.0
        margin_ok = True
        ys, yos, cl = [], [], []
        for part in data["parts"]:
            r0, c0, h, w = part["win"]
            crop, inner = crop_of(H, W, r0, c0, h, w)
            xc = np.ascontiguousarray(x[crop])
            yo, mo, marg = oracle.walk_f32(xc, ks, var, gamma, threads=threads)
            sl = (slice(None),) + inner
            yo, mo, marg = yo[inner], mo[sl], marg[sl]
            ok = part["maps"] == mo
            agree += ok.reshape(ks.D, -1).sum(axis=1)
            total += h * w
            flips += int((~ok).sum())
            margin_ok &= bool(np.all(marg[~ok] < 1e-6))
            all_ok = ok.all(axis=0)
            max_dy = max(max_dy, float(np.abs(part["y"].astype(np.float64) - yo)[all_ok].max(initial=0.0)))
            y64 = oracle.walk_f64(xc, ks, sigma, gamma, threads=threads)[0][inner]
            frozen = np.zeros((ks.D,) + xc.shape, np.uint8)
            frozen[sl] = part["maps"]
            y64f = oracle.walk_f64(xc, ks, sigma, gamma, frozen=frozen, threads=threads)[0][inner]
            e64 = max(e64, float(np.abs(part["y"] - y64).max()))
            e64f = max(e64f, float(np.abs(part["y"] - y64f).max()))
            ys.append(part["y"].ravel())
            yos.append(yo.ravel())
            if part["clean"] is not None:
                cl.append(part["clean"].ravel())
        out = {"pixels": total, "scale_agreement_min": float((agree / total).min()), "flips": flips,
               "flips_at_margin_below_1e-6": margin_ok, "max_abs_dy_where_agree": max_dy,
               "max_abs_err_fp64": e64, "max_abs_err_fp64_frozen_scales": e64f}
        if cl:
            c = np.concatenate(cl)
            out["dpsnr_db"] = abs(lpa_ici.psnr(c, np.concatenate(ys)) - lpa_ici.psnr(c, np.concatenate(yos)))
        out["meets_bar"] = bool(out["scale_agreement_min"] >= 0.999 and margin_ok and max_dy <= 1e-5 and
                                out.get("dpsnr_db", 0.0) <= 0.005)
        return out

    def adjoint_accuracy(ks, data):
        """dL/dx against the FP64 adjoint on windows (a window's dL/dx depends
        on g and the maps within 31 px), relative to its largest value."""
        x, g, maps, dx = data["x"], data["g"], data["maps"], data["dx"]
        H, W = x.shape
        err = 0.0
        for r0, c0, h, w in data["wins"]:
            crop, inner = crop_of(H, W, r0, c0, h, w)
            dx64, _ = oracle.adjoint_f64(np.ascontiguousarray(x[crop]), np.ascontiguousarray(g[crop]), ks,
                                         data["sigma"], np.ascontiguousarray(maps[(slice(None),) + crop]))
            ref = dx64[inner]
            err = max(err, float(np.abs(dx[r0:r0 + h, c0:c0 + w] - ref).max() / np.abs(ref).max()))
        return {"dx_max_rel_err_fp64": err, "windows": len(data["wins"])}

    t0 = time.time()
    accuracy = {}
    for name, (set_name, kind, data) in checks.items():
        ks = ks_by_name[set_name]
        if kind == "forward":
            accuracy[name] = forward_accuracy(ks, data)
        elif kind == "forward_list":
            parts = [forward_accuracy(ks, d) for d in data]
            accuracy[name] = {"images": len(parts), "pixels": sum(p["pixels"] for p in parts),
                              "scale_agreement_min": min(p["scale_agreement_min"] for p in parts),
                              "flips": sum(p["flips"] for p in parts),
                              "max_abs_dy_where_agree": max(p["max_abs_dy_where_agree"] for p in parts),
                              "max_abs_err_fp64_frozen_scales": max(p["max_abs_err_fp64_frozen_scales"] for p in parts),
                              "dpsnr_db_max": max(p.get("dpsnr_db", 0.0) for p in parts),
                              "meets_bar": all(p["meets_bar"] for p in parts)}
        else:
            accuracy[name] = adjoint_accuracy(ks, data)
    log(f"[cpu] accuracy checks in {time.time() - t0:.1f} s: " +
        ", ".join(f"{k}: {'bar met' if v.get('meets_bar', True) else 'BAR MISSED'}" for k, v in accuracy.items()))

    # baseline: the reference's own v0_0 and v1_0 builds, one core each, run
    # side by side on four host cores, on the 512^2 image and on a 1024^2
    # image of the same generator
    ks = ks_by_name["symmetric"]
    var = ks.variances(SIGMA)
    x512 = workloads.single(512).noisy(0)
    x1024 = workloads.single(1024).noisy(0)
    kind = "reference" if oracle.ref_available("v0_0") and oracle.ref_available("v1_0") else "port"

    def run(version, x):
        t = time.perf_counter()
        if kind == "reference":
            oracle.ref_denoise(x, ks, var, GAMMA, version=version)
        elif version == "v1_0":
            oracle.v1_0(x, ks, var, GAMMA)
        else:
            oracle.v0_0(x, ks, var, GAMMA)
        sec = time.perf_counter() - t
        return {"seconds": sec, "mpx_s": x.size / sec / 1e6}

    jobs = {("v0_0", 512): x512, ("v1_0", 512): x512, ("v0_0", 1024): x1024, ("v1_0", 1024): x1024}
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)  # the reference's code prints progress lines: keep them off the JSON stream
    try:
        t0 = time.time()
        with ThreadPoolExecutor(max_workers=len(jobs)) as pool:
            futs = {k: pool.submit(run, k[0], x) for k, x in jobs.items()}
            res = {k: f.result() for k, f in futs.items()}
        wall = time.time() - t0
        # our v1_0-order walk (same bits as v0_0) split by rows over every host core
        t = time.perf_counter()
        oracle.walk_f32(x1024, ks, var, GAMMA, threads=threads)
        all_cores = {"seconds": time.perf_counter() - t}
        all_cores["mpx_s"] = x1024.size / all_cores["seconds"] / 1e6
    finally:
        os.dup2(saved, 1)
        os.close(saved)
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    v10 = res[("v1_0", 512)]
    baseline = {
        "value": v10["mpx_s"], "unit": "Mpixel/s", "cores": 1, "kind": kind,
        "sample": f"the reference's v1_0 CPU path (its faster one) on the 512^2 synthetic image, sigma 20/255, "
                  f"gamma 2, symmetric set: {v10['seconds']:.1f} s; v0_0 and v1_0 at 512^2 and 1024^2 ran side by "
                  f"side, one core each ({wall:.0f} s wall); host {cpu}, nproc {os.cpu_count()}",
        "runs": {f"{v}_{n}": {"seconds": r["seconds"], "mpx_s": r["mpx_s"]} for (v, n), r in res.items()},
        "extrapolated_4096_seconds": {v: 4096 * 4096 / (res[(v, 1024)]["mpx_s"] * 1e6) for v in ("v0_0", "v1_0")},
        "extrapolation": "4096^2 seconds from the 1024^2 rate (per-pixel work does not grow with the image); "
                         "not measured",
        "oracle_walk_all_cores_1024": dict(all_cores, cores=threads, kind="port"),
    }
    log(f"[cpu] baseline: " + ", ".join(f"{v} {n}^2 {r['seconds']:.1f} s ({r['mpx_s']:.4f} Mpx/s)"
                                         for (v, n), r in res.items()))
    return baseline, accuracy


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--skip-cpu", action="store_true", help="no CPU baseline and no accuracy checks")
    ap.add_argument("--skip-walks", action="store_true")
    ap.add_argument("--skip-gaussian", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-fwdbwd", action="store_true")
    ap.add_argument("--skip-config5", action="store_true")
    ap.add_argument("--tile-config", type=int, default=-1)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (the reference kernel for the HBM byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the denoiser has no multi-GPU path "
              "(images shard across GPUs with no communication)", file=sys.stderr)
        sys.exit(2)
    torch.cuda.init()
    stream = torch.cuda.Stream()
    ks = lpa_ici.load_kernels()
    t0 = time.time()
    head = workloads.single(args.size)
    clean = head.clean(0)
    x_host = lpa_ici.make_noisy(clean, head.sigmas[0], head.noise_seeds[0])
    log(f"[setup] {args.size}^2 image in {time.time() - t0:.1f} s; device {torch.cuda.get_device_name(0)}")
    tune = {}
    if args.tile_config >= 0:
        ks.set_tile_config(args.tile_config)
    else:
        from lpa_ici.autotune import autotune
        best, times = autotune(ks, [torch.from_numpy(x_host).cuda()], flags=lpa_ici.FWD_PLAIN, stream=stream,
                               cache=None)
        tune = {str(k): round(v, 4) for k, v in times.items()}
        log(f"[autotune] plain {args.size}^2: best config {best}; ms per config {tune}")

    main_leg = plain_leg(ks, x_host, args.steps, args.warmup, stream, "plain symmetric")
    wins = windows(args.size, args.size, 192, 8, 11)
    checks = {"headline": ("symmetric", "forward", window_data(x_host, main_leg["y"], main_leg["maps"], clean, wins))}
    npx = args.size * args.size
    rec = {
        "metric": METRIC,
        "value": round(main_leg["mpx_s"], 2),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_min_ms": WARMUP_MS,
        "warmup_calls": main_leg["warmup_calls"],
        "ms_per_step": round(main_leg["ms"], 4),
        "median_ms_per_step": round(main_leg["median_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain {args.size}x{args.size} fp32 synthetic (polygons, disks, ramps, stripes), "
                               "sigma 20/255, gamma 2, 16x7 symmetric m=(1,0) set, one launch per pass",
                   "pixels": npx, "tile_config": ks.tile_config(), "autotune_ms": tune,
                   "tile_configs": lpa_ici.tile_configs()},
        "roofline": {"bound": "valu", "achieved": round(main_leg["tflops"], 3), "peak": MI355X_FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(main_leg["tflops"] / MI355X_FP32_PEAK_TFLOPS, 4),
                     "traffic": measured_traffic(main_leg["walk"]) if args.size == 4096 else None,
                     "algorithmic_flop_per_launch": 2 * main_leg["useful_taps"],
                     "useful_taps_per_px": round(main_leg["taps_per_px"], 1),
                     "walk": main_leg["walk"],
                     "executed_op": main_leg["executed_op"],
                     "executed_ops_per_px": round(main_leg["executed_ops_per_px"], 1),
                     "executed_tops": round(main_leg["executed_tops"], 3),
                     "executed_frac_of_lane_ops": round(main_leg["executed_frac_of_lane_ops"], 4)},
    }
    headline_config = ks.tile_config()
    del main_leg
    rw = reference_window(ks, x_host, stream)
    rec["reference_window"] = rounded(dict(rw, what="forward + copy of y into pageable host memory, wall clock, "
                                                    "median of 5 (the reference's timer window)"))
    log(f"[reference window] {rw['ms']:.2f} ms, {rw['mpx_s']:.0f} Mpx/s")

    def tune_for(kset, flags, label):
        """Autotune the tile configuration for this walk (unless one was given)."""
        if args.tile_config >= 0:
            kset.set_tile_config(args.tile_config)
            return
        from lpa_ici.autotune import autotune
        best, _ = autotune(kset, [torch.from_numpy(x_host).cuda()], flags=lpa_ici.FWD_PLAIN | flags,
                           stream=stream, cache=None)
        log(f"[autotune] {label}: best config {best}")

    if not args.skip_walks and ks.info()["ring_form"] == 1:
        for key, flags in (("float_ring_walk", lpa_ici.FWD_RING_FLOAT), ("weighted_walk", lpa_ici.FWD_WEIGHTED)):
            tune_for(ks, flags, key)
            w = plain_leg(ks, x_host, args.steps, args.warmup, stream, f"plain symmetric {key}", flags=flags)
            rec[key] = rounded({"mpx_s": w["mpx_s"], "ms": w["ms"], "median_ms": w["median_ms"],
                                "tflops": w["tflops"], "frac": w["tflops"] / MI355X_FP32_PEAK_TFLOPS,
                                "executed_op": w["executed_op"], "executed_ops_per_px": w["executed_ops_per_px"],
                                "executed_frac_of_lane_ops": w["executed_frac_of_lane_ops"],
                                "tile_config": ks.tile_config()})
            checks[key] = ("symmetric", "forward", window_data(x_host, w["y"], w["maps"], clean, wins[:4]))
            del w
        ks.set_tile_config(headline_config)
    ksg = None
    if not args.skip_gaussian:
        ksg = lpa_ici.load_kernels(lpa_ici.GAUSSIAN_M20_KERNELS)
        tune_for(ksg, 0, "m20 gaussian")
        g = plain_leg(ksg, x_host, args.steps, args.warmup, stream, "plain m20 gaussian")
        rec["gaussian_m20"] = rounded({"mpx_s": g["mpx_s"], "ms": g["ms"], "median_ms": g["median_ms"],
                                       "tflops": g["tflops"], "frac": g["tflops"] / MI355X_FP32_PEAK_TFLOPS,
                                       "taps_per_px": g["taps_per_px"], "executed_op": g["executed_op"],
                                       "executed_frac_of_lane_ops": g["executed_frac_of_lane_ops"],
                                       "tile_config": ksg.tile_config()})
        checks["gaussian_m20"] = ("gaussian", "forward", window_data(x_host, g["y"], g["maps"], clean, wins))
        del g
    torch.cuda.empty_cache()
    if not args.skip_batched:
        b, chk = batched_leg(ks, max(5, args.steps // 4), 1, stream)
        rec["batched_uneven"] = rounded(b)
        checks["batched_uneven"] = ("symmetric", "forward_list", chk)
        torch.cuda.empty_cache()
    if not args.skip_fwdbwd:
        fb, chk = fwd_bwd_leg(ks, max(5, args.steps // 4), 1, stream)
        rec["fwd_bwd"] = rounded(fb)
        checks["fwd_bwd_forward"] = ("symmetric", "forward", chk["forward"])
        checks["fwd_bwd_adjoint"] = ("symmetric", "adjoint", chk["adjoint"])
        torch.cuda.empty_cache()
    if not args.skip_config5:
        c5, chk = config5_leg(ks, 5, 1, stream)
        rec["config5"] = rounded(c5)
        checks["config5_forward"] = ("symmetric", "forward", chk["forward"])
        checks["config5_adjoint"] = ("symmetric", "adjoint", chk["adjoint"])
        torch.cuda.empty_cache()
    if not args.skip_cpu:
        baseline, accuracy = cpu_leg(checks, {"symmetric": ks, "gaussian": ksg}, args.cpu_threads)
        rec["cpu_baseline"] = rounded(baseline)
        rec["accuracy"] = rounded(accuracy)
    rec["host"] = platform.node()
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
