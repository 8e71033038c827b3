"""Autotuner for the forward's tile configuration.

For a kernel set and a workload (device images), every compiled tile
configuration — tile shape, pixels per lane, inner-loop body — is timed with
HIP events on the caller's stream, and the fastest is made the kernel set's
configuration.  Results are cached in a JSON file keyed by the kernel set's
SHA-256, the device name, the workload form (plain or batched) and the walk
flags, so a later process reuses the choice without timing again.
"""
import hashlib
import json
import os

import torch

from . import FWD_PLAIN, Plan, same_layout, tile_configs

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "lpa_ici", "autotune.json")


def kset_key(ks):
    h = hashlib.sha256()
    for a in (ks.weights, ks.offsets, ks.rows, ks.cols):
        h.update(a.tobytes())
    return h.hexdigest()[:16]


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, table):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def time_config(ks, index, xs, ys, flags, stream, reps=3, warmup=1, sigmas=20.0 / 255.0, gammas=2.0):
    """Milliseconds per forward of configuration `index` on this workload."""
    previous = ks.tile_config()
    ks.set_tile_config(index)
    try:
        plan = Plan(ks, xs, ys, None, sigmas, gammas, flags=flags)
        with torch.cuda.stream(stream):
            for _ in range(warmup):
                plan.run(stream)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                plan.run(stream)
            e1.record(stream)
        torch.cuda.synchronize()
        plan.close()
        return e0.elapsed_time(e1) / reps
    finally:
        ks.set_tile_config(previous)


def autotune(ks, xs, flags=FWD_PLAIN, sigmas=20.0 / 255.0, gammas=2.0, cache=DEFAULT_CACHE, force=False,
             stream=None, reps=3):
    """Pick the fastest tile configuration for kernel set `ks` on images `xs`
    (list of device tensors) and install it on `ks`.  Returns (index, table)
    where table maps each configuration index to its milliseconds (empty
    when the choice came from the cache)."""
    stream = stream or torch.cuda.current_stream()
    form = "plain" if flags & FWD_PLAIN else "batched"
    npx = sum(x.numel() for x in xs)
    walk = flags & ~FWD_PLAIN
    key = f"{kset_key(ks)}|{torch.cuda.get_device_name(0)}|{form}|walk{walk}|{len(xs)}x{npx}"
    table = _load(cache) if cache else {}
    if not force and key in table:
        ks.set_tile_config(int(table[key]["best"]))
        return int(table[key]["best"]), {}
    ys = [same_layout(x) for x in xs]
    times = {}
    for i in range(len(tile_configs())):
        times[i] = time_config(ks, i, xs, ys, flags, stream, reps=reps, sigmas=sigmas, gammas=gammas)
    best = min(times, key=times.get)
    ks.set_tile_config(best)
    if cache:
        table[key] = {"best": best, "ms": {str(k): round(v, 4) for k, v in times.items()}}
        _store(cache, table)
    return best, times
