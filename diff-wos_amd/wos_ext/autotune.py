"""Launch autotuner for the persistent walk kernels.

The walk, shape-gradient and batch launches share four tuning knobs
(``wos_set_tuning``): walks a wave claims per queue visit (``chunk``),
persistent workgroups per CU (``wgs_per_cu``), BVH levels copied into LDS
(``top_levels``) and node visits per lane between refills (``window``).
Results never depend on them -- every walk re-derives its random stream from
(seed, point, walk) -- only speed does, and the best setting depends on the
scene (tree depth, query lengths) and the walk count.

``autotune(run, key)`` times a workload under candidate settings by
coordinate descent from the current setting, keeps the fastest, and caches
it per (device, key) in a JSON file, so later processes reuse it without
timing again.
"""
from __future__ import annotations

import json
import os
from typing import Callable, Dict, Sequence, Tuple

CANDIDATES: Dict[str, Sequence[int]] = {
    "chunk": (64, 128, 256),
    "wgs_per_cu": (4, 6, 8),
    "top_levels": (0, 4, 6),
    "window": (4, 8, 12, 16),
}
KNOBS = ("chunk", "wgs_per_cu", "top_levels", "window")


def cache_path() -> str:
    return os.environ.get("WOS_AMD_TUNING_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "wos_amd", "tuning.json"))


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
        pass   # a read-only home only costs re-tuning


def coordinate_descent(measure: Callable[[Tuple[int, ...]], float], start: Tuple[int, ...],
                       candidates: Dict[str, Sequence[int]] = CANDIDATES, passes: int = 2):
    """Minimise measure(setting) one knob at a time; returns (best, time, trials).
    Pure host logic: `measure` maps a (chunk, wgs_per_cu, top_levels, window)
    tuple to a time."""
    seen = {}

    def cost(t):
        if t not in seen:
            seen[t] = measure(t)
        return seen[t]

    best = tuple(start)
    best_t = cost(best)
    for _ in range(passes):
        improved = False
        for j, name in enumerate(KNOBS):
            for v in candidates[name]:
                t = best[:j] + (int(v),) + best[j + 1:]
                c = cost(t)
                if c < best_t:
                    best, best_t, improved = t, c, True
        if not improved:
            break
    return best, best_t, len(seen)


def autotune(run: Callable[[], object], key: str, reps: int = 2, use_cache: bool = True,
             candidates: Dict[str, Sequence[int]] = CANDIDATES):
    """Tunes the launch knobs for `run` (one launch of the workload on the
    current stream), applies the best setting and returns it as a dict."""
    import torch
    from . import get_tuning, set_tuning

    dev = torch.cuda.get_device_name(torch.cuda.current_device())
    full_key = f"{dev}|{key}"
    path = cache_path()
    table = _load(path) if use_cache else {}
    if full_key in table:
        t = table[full_key]["tuning"]
        set_tuning(*t)
        return dict(zip(KNOBS, t), cached=True)

    def measure(t):
        set_tuning(*t)
        run()   # warm-up at this setting
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    best, ms, trials = coordinate_descent(measure, tuple(get_tuning()), candidates)
    set_tuning(*best)
    if use_cache:
        table[full_key] = {"tuning": list(best), "ms": ms}
        _store(path, table)
    return dict(zip(KNOBS, best), ms=ms, trials=trials, cached=False)
