"""Autotuner for the launch shapes of the fused path.

The library's launch-shape knobs (``emsim_hip_tuning_set``) decide how the
work of one pass is split over thread blocks -- rows (row pairs) per
row-worker block, column tiles per column-kernel block, and the propagator
pass's column-transform tile (values per thread and columns per block) --
and never the arithmetic (each butterfly computes the same values whichever
thread runs it), so every setting gives bit-identical images.  ``tune`` times a
workload under candidate settings by coordinate descent (one knob at a time,
the others held at their best so far) and returns the fastest; ``cached``
keeps results per (device, workload) in a JSON file so a serving process
tunes once.

The FFT radix plan is fixed by the length (one leading radix-2/4/8 pass then
radix-8 passes); the column-transform tiles of each length are compiled in
(``col_shape`` picks one at run time); the batch width is the caller's
choice and can be tuned the same way with ``tune_batch``.
"""
import json
import os
import time
from pathlib import Path
from typing import Callable, Dict, Iterable, Optional, Sequence, Tuple

KNOBS: Dict[str, Tuple[int, ...]] = {
    "row_items_per_block": (0, 1, 2, 4, 8, 16),
    "col_tiles_per_block": (0, 1, 2, 4, 8),
    # column-transform tile of the propagator pass (256, 512, 1024 points):
    # 0 = 8 values per thread, 8 columns per block; 1 = half the columns per
    # block; 2 = 16 values per thread (two radix-8 butterflies per pass)
    "col_shape": (0, 1, 2),
}


def apply(cfg: Dict[str, int]) -> None:
    """Set the library's knobs (missing keys go back to the built-in rule)."""
    from .backend import hip_ext
    for key in KNOBS:
        hip_ext.call("emsim_hip_tuning_set", key.encode(), int(cfg.get(key, 0)))


def current() -> Dict[str, int]:
    from .backend import hip_ext
    return {key: int(hip_ext.raw("emsim_hip_tuning_get")(key.encode())) for key in KNOBS}


def _time(run: Callable[[], object], sync: Callable[[], None], reps: int, warmup: int) -> float:
    for _ in range(warmup):
        run()
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    sync()
    return (time.perf_counter() - t0) / reps


def tune(run: Callable[[], object], knobs: Dict[str, Sequence[int]] = None, reps: int = 3,
         warmup: int = 1, sync: Callable[[], None] = None,
         set_knobs: Callable[[Dict[str, int]], None] = apply,
         timer: Callable[..., float] = None):
    """Coordinate descent over the knobs.  Returns (best settings, table of
    (settings, seconds per run)).  ``set_knobs`` / ``timer`` are injectable
    for tests."""
    knobs = KNOBS if knobs is None else knobs
    if sync is None:
        import torch
        sync = torch.cuda.synchronize
    timer = timer or (lambda: _time(run, sync, reps, warmup))
    best = {k: 0 for k in knobs}
    table = []
    best_t = None
    for key, values in knobs.items():
        for v in values:
            cfg = dict(best, **{key: int(v)})
            set_knobs(cfg)
            t = timer()
            table.append((cfg, t))
            if best_t is None or t < best_t:
                best_t, best = t, cfg
    set_knobs(best)
    return best, table


def tune_batch(make_run: Callable[[int], Callable[[], object]], widths: Iterable[int],
               reps: int = 3, sync: Callable[[], None] = None) -> Tuple[int, list]:
    """Fastest batch width by images per second; make_run(B) returns a
    callable that simulates one batch of B images."""
    if sync is None:
        import torch
        sync = torch.cuda.synchronize
    table = []
    for b in widths:
        t = _time(make_run(b), sync, reps, 1)
        table.append((b, b / t))
    return max(table, key=lambda r: r[1])[0], table


def default_cache_path() -> Path:
    """EMSIM_TUNING_CACHE, else ``tuning_cache.json`` beside the built HIP
    library (inside the repository, so results travel with the build)."""
    env = os.environ.get("EMSIM_TUNING_CACHE")
    if env:
        return Path(env)
    return Path(__file__).resolve().parent / "backend" / "hip_ext" / "tuning_cache.json"


def cached(key: str, run: Callable[[], object], path: Optional[Path] = None, **kw) -> Dict[str, int]:
    """Settings for ``key`` (e.g. "gfx950/n512/b64/s100/mode0") from the
    cache file, tuning and storing them on a miss; applies them."""
    path = Path(path) if path is not None else default_cache_path()
    data = {}
    if path.exists():
        try:
            data = json.loads(path.read_text())
        except (OSError, ValueError):
            data = {}
    if key in data:
        cfg = {k: int(v) for k, v in data[key].items() if k in KNOBS}
        kw.get("set_knobs", apply)(cfg)
        return cfg
    cfg, _ = tune(run, **kw)
    data[key] = cfg
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    os.replace(tmp, path)
    return cfg
