"""Launch-shape autotuner for batched rendering.

Times a batch under a set of candidate tunings (rise_hip_batch_tuning:
paths in flight, traversal workgroup size, traversal stack, traversal
schedule) and returns the fastest.  The films do not depend on the tuning,
so any candidate may be used for the real work.  To keep tuning cheap the
candidates are timed on a proxy of the workload: the same requests at a
reduced sample count (``proxy_spp``), which keeps every scene, image size and
queue shape but scales the work down.
"""
from __future__ import annotations

import itertools
import time


def candidates(persistent=(1, 2), trace_block=(64, 128, 256), max_paths=(1 << 24, 1 << 25), stack=(0, 16, -1)):
    """The search space: every workgroup size, path-pool size and traversal
    stack (0: the smallest LDS stack that fits the trees, 16 LDS entries per
    lane, -1 private memory) for persistent traversal; one ray per thread
    (persistent 2) with the automatic stack only, since its stack choice is
    the same trade-off and it has never been the fastest."""
    for p, b, m, s in itertools.product(persistent, trace_block, max_paths, stack):
        if p == 2 and s != 0:
            continue
        yield {"persistent": p, "trace_block": b, "max_paths": m, "stack": s}


def time_batch(items, tuning: dict, repeats: int = 2) -> float:
    """Mean milliseconds per launch of a batch of `items` under `tuning`."""
    import torch

    from .render import RenderBatch

    b = RenderBatch(items, tuning=tuning)
    try:
        b.launch()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(repeats):
            b.launch()
        end.record()
        torch.cuda.synchronize()
        return start.elapsed_time(end) / repeats
    finally:
        b.release()


def tune(items, proxy_spp: int | None = 1, space=None, repeats: int = 2, budget_s: float = 60.0):
    """Returns (best tuning, [(tuning, ms), ...]) for a list of
    (rasterizer, RenderSettings) items.  proxy_spp: sample count of the proxy
    (None: the items as given)."""
    import dataclasses

    proxy = []
    for it in items:
        rast, st = it[0], it[1]
        if proxy_spp is not None:
            st = dataclasses.replace(st, spp=min(st.spp, proxy_spp))
        proxy.append((rast, st))
    results = []
    t0 = time.perf_counter()
    for tn in (space if space is not None else candidates()):
        ms = time_batch(proxy, tn, repeats)
        results.append((tn, ms))
        if time.perf_counter() - t0 > budget_s:
            break
    best = min(results, key=lambda r: r[1])[0]
    return best, results
