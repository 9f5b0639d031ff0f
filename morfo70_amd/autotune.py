"""Autotuner for the line-solver schedule of the FUSED step.

Knobs (all results-neutral except the reduced-system tile, which changes rounding by a
few ulps, well inside the 1e-10 bar):
  * rows_block: threads per block of the y-sweep solver, 256 (ny <= 2048, 36 KB of LDS,
    four blocks per CU) or 512 (72 KB, two blocks per CU);
  * tile_cols: unknowns per thread of the reduced column systems, 8 or 16;
  * graph (batches only): replay the step from a hipGraph or launch eagerly.

Each candidate runs a few steps on the caller's own handle or batch, timed with HIP events
on the stream the work is enqueued on; the fastest is applied and cached per
(device, nx bucket, ny bucket) in a JSON file, so later handles of a similar size reuse it.
"""
from __future__ import annotations

import json
import os

from . import _lib

CACHE = os.environ.get("M70H_AUTOTUNE_CACHE",
                       os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.json"))


def _bucket(n: int) -> int:
    b = 1
    while b < n:
        b <<= 1
    return b


def _key(nx: int, ny: int) -> str:
    import torch
    name = torch.cuda.get_device_name(0).replace(" ", "_")
    return f"{name}:{_bucket(nx)}x{_bucket(ny)}"


def _load() -> dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(cache: dict):
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def _time(run_async, stream_ptr: int, steps: int) -> float:
    import torch
    s = torch.cuda.ExternalStream(stream_ptr)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    run_async(steps)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def _apply_handle(h, cfg: dict):
    h._chk(h._L.m70h_set_rows_block(h.handle, cfg["rows_block"]))
    h._chk(h._L.m70h_set_solver_tiles(h.handle, cfg["tile_cols"], 0))


def tune_handle(h, steps: int = 4, use_cache: bool = True) -> dict:
    """Picks rows_block and tile_cols for a FUSED handle (state advances by the trial steps).
    The handle must run on a stream set with set_stream (its own stream is used otherwise)."""
    cache = _load() if use_cache else {}
    key = _key(h.nx, h.ny)
    if key in cache:
        _apply_handle(h, cache[key])
        return cache[key]
    import ctypes
    stream = ctypes.c_void_p()
    cands = []
    for rb in ((256, 512) if h.ny <= 2048 else (512,)):
        for tc in (8, 16):
            cands.append({"rows_block": rb, "tile_cols": tc})
    best, best_ms = None, float("inf")
    sp = h._stream_ptr() if hasattr(h, "_stream_ptr") else None
    for cfg in cands:
        _apply_handle(h, cfg)
        h.run_async(1)                       # warm the configuration
        ms = _time(h.run_async, sp or stream.value or 0, steps)
        cfg = dict(cfg, ms_per_step=ms / steps)
        if ms < best_ms:
            best, best_ms = cfg, ms
    _apply_handle(h, best)
    cache[key] = best
    if use_cache:
        _save(cache)
    return best


def tune_batch(b, steps: int = 4) -> dict:
    """Picks rows_block and graph replay for a batch (its state advances by the trials)."""
    results = []
    for rb in (256, 512):
        try:
            b.set_rows_block(rb)
        except _lib.M70HError:
            continue
        for graph in (True, False):
            b.set_graph(graph)
            b.run(2)
            ms = _time(b.run_async, b.stream, steps)
            results.append(({"rows_block": rb, "graph": graph}, ms / steps))
    cfg, ms = min(results, key=lambda r: r[1])
    b.set_rows_block(cfg["rows_block"])
    b.set_graph(cfg["graph"])
    return dict(cfg, ms_per_step=ms, tried=[dict(c, ms_per_step=m) for c, m in results])
