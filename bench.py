"""Benchmark of the MI355X LPA-ICI denoiser.

Headline (one JSON line on stdout): plain form, one 4096x4096 fp32 synthetic
image (sigma 20/255, gamma 2, 16x7 symmetric m=(1,0) kernel set), Mpixel/s
over K timed forward passes with the input resident in HBM.  The same line
carries
  roofline      useful-tap FP32 FLOP/s of the forward kernel (HIP events on
                its stream) against the MI355X vector FP32 peak
  cpu_baseline  the reference's own v1_0 CPU code (oracle/_ref, 1 core) on a
                bounded crop of the same input, or our restatement if the
                reference build is absent
  gaussian_m20  the same image with the 16x7 m=(2,0) gaussian-window set
  batched_uneven  256 images with sides U{128..2048}, per-image sigma/gamma:
                plain (one launch per image) vs one persistent launch vs its
                hipGraph replay, each form at its own autotuned tile
                configuration
  fwd_bwd       64 x 512^2 forward + adjoint (dL/dx, dL/dkernel); the forward
                saves the uint8 scale maps and the adjoint recomputes the
                estimates from them; timed again with the forward keeping
                each direction's selected estimate for the adjoint to read
                ("kept_estimates"); the forward also one launch per image
  config5       32 images with sides U{1024..8192} (about 630 Mpx), fwd+bwd
                in the same two forms, and each form's peak device memory in
                use against the 288 GB of HBM
Flags --skip-* drop the side legs.
"""
import argparse
import json
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "lpa-ici-cuda_amd"))

import numpy as np  # noqa: E402
import torch  # noqa: E402

import lpa_ici  # noqa: E402
from lpa_ici.metrics import MI355X_FP32_PEAK_TFLOPS, useful_taps  # noqa: E402

METRIC = "Mpixel/s at 4096^2 plain vs batched-uneven; fwd+bwd Mpixel/s; % FP32 VALU peak"
SIGMA, GAMMA, SEED = 20.0 / 255.0, 2.0, 20250910


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, steps, warmup, stream):
    """Median-free total: W untimed calls, then K calls between two events on
    `stream`, synchronized on both sides.  Returns ms per call."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def rounded(d):
    """A leg's record with its floats to three decimals, nested records too."""
    return {k: rounded(v) if isinstance(v, dict) else round(v, 3) if isinstance(v, float) else v
            for k, v in d.items()}


def make_image(w, h, seed, sigma, noise_seed):
    return lpa_ici.make_noisy(lpa_ici.synth_image(w, h, seed), sigma, noise_seed)


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


def plain_leg(ks, x_host, steps, warmup, stream, label, flags=0):
    x = torch.from_numpy(x_host).cuda()
    y = torch.empty_like(x)
    maps = torch.empty((ks.D,) + tuple(x.shape), dtype=torch.uint8, device="cuda")
    plan = lpa_ici.Plan(ks, [x], [y], [maps], SIGMA, GAMMA, flags=lpa_ici.FWD_PLAIN | flags)
    with torch.cuda.stream(stream):
        ms = timed(lambda: plan.run(stream), steps, warmup, stream)
    taps = useful_taps(maps, ks.runs()["nnz"], ks.S)
    walk = walk_of(ks, flags)
    ex, op = executed_ops(ks, maps, walk)
    npx = x.numel()
    tflops = 2.0 * taps / (ms * 1e-3) / 1e12
    ex_rate = ex / (ms * 1e-3)
    log(f"[{label}] {x.shape[1]}x{x.shape[0]}: {ms:.3f} ms/pass, {npx / ms / 1e3:.1f} Mpx/s, "
        f"{taps / npx:.0f} useful taps/px ({tflops:.2f} TFLOP/s algorithmic), {walk}: {ex / npx:.0f} {op}/px "
        f"executed ({ex_rate / MI355X_LANE_OPS:.3f} of the VALU lane-op rate)")
    plan.close()
    return {"ms": ms, "mpx_s": npx / ms / 1e3, "useful_taps": taps, "taps_per_px": taps / npx, "tflops": tflops,
            "executed_op": op, "executed_ops_per_px": ex / npx, "executed_tops": ex_rate / 1e12,
            "executed_frac_of_lane_ops": ex_rate / MI355X_LANE_OPS, "walk": walk, "y": y, "x": x}


def batched_leg(ks, steps, warmup, stream, n_images=256, seed=3):
    rng = np.random.default_rng(seed)
    sides = rng.integers(128, 2049, size=(n_images, 2))
    sig = [float(rng.choice([10, 20, 30, 40])) / 255.0 for _ in range(n_images)]
    gam = [float(rng.choice([1.5, 2.0, 2.5])) for _ in range(n_images)]
    t0 = time.time()
    xs = [torch.from_numpy(make_image(int(w), int(h), 1000 + i, s, SEED + i)).cuda()
          for i, ((w, h), s) in enumerate(zip(sides, sig))]
    log(f"[batched] generated {n_images} images, {sum(x.numel() for x in xs) / 1e6:.1f} Mpx in {time.time() - t0:.1f} s")
    npx = sum(x.numel() for x in xs)
    ys_plain = [torch.empty_like(x) for x in xs]
    ys_batch = [torch.empty_like(x) for x in xs]
    ys_graph = [torch.empty_like(x) for x in xs]
    out = {"images": n_images, "mpx": npx / 1e6}
    # each form at the tile configuration the autotuner picks for it on this
    # workload (a plan keeps the configuration it was built with)
    from lpa_ici.autotune import autotune
    saved = ks.tile_config()
    plain_cfg, _ = autotune(ks, xs, flags=lpa_ici.FWD_PLAIN, sigmas=sig, gammas=gam, stream=stream)
    plans = {"plain": lpa_ici.Plan(ks, xs, ys_plain, None, sig, gam, flags=lpa_ici.FWD_PLAIN)}
    batch_cfg, _ = autotune(ks, xs, flags=0, sigmas=sig, gammas=gam, stream=stream)
    plans["batched"] = lpa_ici.Plan(ks, xs, ys_batch, None, sig, gam, flags=0)
    plans["graph"] = lpa_ici.Plan(ks, xs, ys_graph, None, sig, gam, flags=lpa_ici.FWD_GRAPH)
    ks.set_tile_config(saved)
    out["tile_config"] = {"plain": plain_cfg, "batched": batch_cfg}
    log(f"[batched] tile configurations: plain {plain_cfg}, batched and graph {batch_cfg}")
    with torch.cuda.stream(stream):
        for name, plan in plans.items():
            ms = timed(lambda: plan.run(stream), steps, warmup, stream)
            out[name] = {"ms": ms, "mpx_s": npx / ms / 1e3}
            log(f"[batched] {name}: {ms:.2f} ms, {npx / ms / 1e3:.1f} Mpx/s")
    torch.cuda.synchronize()
    out["identical"] = all(torch.equal(a, b) and torch.equal(a, c) for a, b, c in zip(ys_plain, ys_batch, ys_graph))
    out["batched_over_plain"] = out["batched"]["mpx_s"] / out["plain"]["mpx_s"]
    out["graph_over_plain"] = out["graph"]["mpx_s"] / out["plain"]["mpx_s"]
    for p in plans.values():
        p.close()
    return out


def adjoint_variant(ks, xs, ys, maps, sig, gam, dys, dxs, dk, steps, warmup, stream, variant, on_step=None):
    """Times forward and forward + adjoint over the images, in one of two
    forms: "recompute" as the workload states it (the forward saves only the
    uint8 scale maps and the adjoint recomputes each direction's estimate from
    them), or "kept_estimates" (the forward also keeps each direction's
    selected estimate, 16 float maps per image, for the adjoint to read).
    Returns (fwd_ms, fwd_bwd_ms)."""
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
        if on_step:
            on_step()

    with torch.cuda.stream(stream):
        ms_f = timed(lambda: fwd.run(stream), steps, warmup, stream)
        ms_fb = timed(step, steps, warmup, stream)
    if on_step:
        torch.cuda.synchronize()
        on_step()
    fwd.close()
    return ms_f, ms_fb


def fwd_bwd_leg(ks, steps, warmup, stream, n=64, side=512):
    xs = [torch.from_numpy(make_image(side, side, 2000 + i, SIGMA, SEED + i)).cuda() for i in range(n)]
    ys = [torch.empty_like(x) for x in xs]
    maps = [torch.empty((ks.D, side, side), dtype=torch.uint8, device="cuda") for _ in xs]
    g = torch.Generator(device="cuda").manual_seed(5)
    dys = [torch.randn(x.shape, generator=g, device="cuda") for x in xs]
    dxs = [torch.empty_like(x) for x in xs]
    dk = torch.zeros(ks.info()["num_weights"], dtype=torch.float32, device="cuda")
    times = {v: adjoint_variant(ks, xs, ys, maps, SIGMA, GAMMA, dys, dxs, dk, steps, warmup, stream, v)
             for v in ("recompute", "kept_estimates")}
    npx = n * side * side
    # the same forward, one launch per image, at the plain form's own tile
    # configuration for these images
    from lpa_ici.autotune import autotune
    saved = ks.tile_config()
    plain_cfg, _ = autotune(ks, xs, flags=lpa_ici.FWD_PLAIN, stream=stream)
    plain = lpa_ici.Plan(ks, xs, ys, maps, SIGMA, GAMMA, flags=lpa_ici.FWD_PLAIN)
    ks.set_tile_config(saved)
    with torch.cuda.stream(stream):
        ms_plain = timed(lambda: plain.run(stream), steps, warmup, stream)
    plain.close()
    log(f"[fwd+bwd] forward one launch per image: {ms_plain:.2f} ms (tile configuration {plain_cfg}), "
        f"persistent {times['recompute'][0]:.2f} ms: {ms_plain / times['recompute'][0]:.2f}x")
    # Algorithmic adjoint work: per selected kernel tap of every pixel and
    # direction, 2 FLOP to recompute yhat, 2 for the dL/dx gather and 4 for
    # dL/dkernel (the reference's taps, whatever the walks execute).
    nnz = torch.from_numpy(ks.runs()["nnz"].astype(np.int64).reshape(ks.D, ks.S)).cuda()
    dsel = torch.arange(ks.D, device="cuda")[:, None, None]
    sel_taps = int(sum(nnz[dsel, m.long()].sum().item() for m in maps))
    rec = {"images": n, "side": side, "selected_taps_per_px": sel_taps / npx,
           "fwd_plain_per_image_ms": ms_plain, "fwd_plain_tile_config": plain_cfg,
           "fwd_batched_over_plain": ms_plain / times["recompute"][0]}
    for variant, (ms_f, ms_fb) in times.items():
        ms_b = ms_fb - ms_f
        bwd_tflops = 8.0 * sel_taps / (ms_b * 1e-3) / 1e12
        log(f"[fwd+bwd {variant}] {n}x{side}^2: fwd {ms_f:.2f} ms, fwd+bwd {ms_fb:.2f} ms, "
            f"{npx / ms_fb / 1e3:.1f} Mpx/s; adjoint {ms_b:.2f} ms, {sel_taps / npx:.0f} selected taps/px, "
            f"{bwd_tflops:.1f} TFLOP/s algorithmic")
        v = {"fwd_ms": ms_f, "fwd_bwd_ms": ms_fb, "fwd_mpx_s": npx / ms_f / 1e3, "fwd_bwd_mpx_s": npx / ms_fb / 1e3,
             "bwd_ms": ms_b, "bwd_tflops_algorithmic": bwd_tflops}
        if variant == "recompute":
            rec.update(v)
        else:
            rec[variant] = v
    return rec


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


def config5_leg(ks, steps, warmup, stream, n=32, seed=5):
    """32 images with sides U{1024..8192}, per-image sigma/gamma as the batched
    leg: forward (scale maps kept as uint8) + adjoint that recomputes the
    estimates from them, then the same with the estimates kept; peak device
    memory in use reported."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(1024, 8193, size=(n, 2))
    sig = [float(rng.choice([10, 20, 30, 40])) / 255.0 for _ in range(n)]
    gam = [float(rng.choice([1.5, 2.0, 2.5])) for _ in range(n)]
    torch.cuda.synchronize()
    free0, total = torch.cuda.mem_get_info()
    torch.cuda.reset_peak_memory_stats()
    t0 = time.time()
    xs = [torch.from_numpy(make_image(int(w), int(h), 3000 + i, s, SEED + i)).cuda()
          for i, ((w, h), s) in enumerate(zip(sides, sig))]
    npx = sum(x.numel() for x in xs)
    log(f"[config5] generated {n} images, {npx / 1e6:.1f} Mpx in {time.time() - t0:.1f} s")
    ys = [torch.empty_like(x) for x in xs]
    maps = [torch.empty((ks.D,) + tuple(x.shape), dtype=torch.uint8, device="cuda") for x in xs]
    g = torch.Generator(device="cuda").manual_seed(7)
    dys = [torch.randn(x.shape, generator=g, device="cuda") for x in xs]
    dxs = [torch.empty_like(x) for x in xs]
    dk = torch.zeros(ks.info()["num_weights"], dtype=torch.float32, device="cuda")
    free_min = {"recompute": free0, "kept_estimates": free0}
    current = ["recompute"]

    def sample():
        free_min[current[0]] = min(free_min[current[0]], torch.cuda.mem_get_info()[0])

    # the recompute form runs first, before any estimates exist, so each
    # variant's peak is its own
    times = {}
    for variant in ("recompute", "kept_estimates"):
        current[0] = variant
        times[variant] = adjoint_variant(ks, xs, ys, maps, sig, gam, dys, dxs, dk, steps, warmup, stream, variant,
                                         sample)
    rec = {"images": n, "mpx": round(npx / 1e6, 1), "hbm_total_gib": round(total / 2**30, 1),
           "scale_maps_gib": round(npx * ks.D / 2**30, 2)}
    for variant, (ms_f, ms_fb) in times.items():
        peak = free0 - free_min[variant]
        log(f"[config5 {variant}] {n} images {npx / 1e6:.1f} Mpx: fwd {ms_f:.1f} ms ({npx / ms_f / 1e3:.0f} Mpx/s), "
            f"fwd+bwd {ms_fb:.1f} ms ({npx / ms_fb / 1e3:.0f} Mpx/s); peak device memory in use "
            f"{peak / 2**30:.1f} GiB of {total / 2**30:.0f} GiB")
        v = {"fwd_ms": ms_f, "fwd_mpx_s": npx / ms_f / 1e3, "fwd_bwd_ms": ms_fb, "fwd_bwd_mpx_s": npx / ms_fb / 1e3,
             "peak_hbm_gib": round(peak / 2**30, 2)}
        if variant == "recompute":
            rec.update(v)
        else:
            rec[variant] = v
    rec["torch_peak_allocated_gib"] = round(torch.cuda.max_memory_allocated() / 2**30, 2)
    return rec


def cpu_baseline(x_host, crop=640):
    """The reference's v1_0 CPU code (fastest of its two CPU versions) on a
    crop of the workload image, one core; our restatement if the reference
    build is not present."""
    sys.path.insert(0, REPO)
    import oracle
    ks = lpa_ici.load_kernels()
    var = ks.variances(SIGMA)
    h, w = x_host.shape
    sample = np.ascontiguousarray(x_host[(h - crop) // 2:(h + crop) // 2, (w - crop) // 2:(w + crop) // 2])
    kind = "reference" if oracle.ref_available("v1_0") else "port"
    # the reference's code prints progress lines on stdout: keep them off the JSON stream
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        t = time.time()
        if kind == "reference":
            oracle.ref_denoise(sample, ks, var, GAMMA, version="v1_0")
        else:
            oracle.v1_0(sample, ks, var, GAMMA)
        sec = time.time() - t
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
    return {"value": sample.size / sec / 1e6, "unit": "Mpixel/s", "cores": 1, "kind": kind,
            "sample": f"v1_0 CPU path on the central {crop}x{crop} crop of the 4096^2 noisy image, "
                      f"{sec:.1f} s; host {cpu}, nproc {os.cpu_count()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-gaussian", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-fwdbwd", action="store_true")
    ap.add_argument("--skip-config5", action="store_true")
    ap.add_argument("--tile-config", type=int, default=-1)
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
    x_host = make_image(args.size, args.size, 42, SIGMA, SEED)
    log(f"[setup] {args.size}^2 image in {time.time() - t0:.1f} s; device {torch.cuda.get_device_name(0)}")
    tune = {}
    if args.tile_config >= 0:
        ks.set_tile_config(args.tile_config)
    else:
        from lpa_ici.autotune import autotune
        best, times = autotune(ks, [torch.from_numpy(x_host).cuda()], flags=lpa_ici.FWD_PLAIN, stream=stream,
                               cache=None)
        tune = {str(k): round(v, 4) for k, v in times.items()}
        log(f"[autotune] plain 4096^2: best config {best}; ms per config {tune}")

    main_leg = plain_leg(ks, x_host, args.steps, args.warmup, stream, "plain symmetric")
    npx = args.size * args.size
    rec = {
        "metric": METRIC,
        "value": round(main_leg["mpx_s"], 2),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_leg["ms"], 4),
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

    def tune_for(kset, flags, label):
        """Autotune the tile configuration for this walk (unless one was given)."""
        if args.tile_config >= 0:
            kset.set_tile_config(args.tile_config)
            return
        from lpa_ici.autotune import autotune
        best, times = autotune(kset, [torch.from_numpy(x_host).cuda()], flags=lpa_ici.FWD_PLAIN | flags,
                               stream=stream, cache=None)
        log(f"[autotune] {label}: best config {best}")

    if main_leg["walk"] == "fixed-point ring sums":
        for key, flags in (("float_ring_walk", lpa_ici.FWD_RING_FLOAT), ("weighted_walk", lpa_ici.FWD_WEIGHTED)):
            tune_for(ks, flags, key)
            w = plain_leg(ks, x_host, args.steps, args.warmup, stream, f"plain symmetric {key}", flags=flags)
            rec[key] = {"mpx_s": round(w["mpx_s"], 2), "ms": round(w["ms"], 4), "tflops": round(w["tflops"], 3),
                        "frac": round(w["tflops"] / MI355X_FP32_PEAK_TFLOPS, 4),
                        "executed_op": w["executed_op"], "executed_ops_per_px": round(w["executed_ops_per_px"], 1),
                        "executed_frac_of_lane_ops": round(w["executed_frac_of_lane_ops"], 4),
                        "tile_config": ks.tile_config()}
        ks.set_tile_config(headline_config)
    if not args.skip_cpu:
        rec["cpu_baseline"] = cpu_baseline(x_host)
        log(f"[cpu] {rec['cpu_baseline']}")
    if not args.skip_gaussian:
        ksg = lpa_ici.load_kernels(lpa_ici.GAUSSIAN_M20_KERNELS)
        tune_for(ksg, 0, "m20 gaussian")
        g = plain_leg(ksg, x_host, args.steps, args.warmup, stream, "plain m20 gaussian")
        rec["gaussian_m20"] = {"mpx_s": round(g["mpx_s"], 2), "ms": round(g["ms"], 4),
                               "tflops": round(g["tflops"], 3), "taps_per_px": round(g["taps_per_px"], 1),
                               "tile_config": ksg.tile_config()}
    del main_leg
    if not args.skip_batched:
        b = batched_leg(ks, max(3, args.steps // 4), 1, stream)
        rec["batched_uneven"] = rounded(b)
    if not args.skip_fwdbwd:
        fb = fwd_bwd_leg(ks, max(3, args.steps // 4), 1, stream)
        rec["fwd_bwd"] = rounded(fb)
    if not args.skip_config5:
        c5 = config5_leg(ks, 2, 1, stream)
        rec["config5"] = rounded(c5)
    rec["host"] = platform.node()
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
