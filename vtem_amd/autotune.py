"""In-project autotuner for the tile shape of the low-precision Gram (`vt_gram_lp`).

The kernel's workgroup owns 256 spatial-pair rows and runs 8 waves.  A wave takes
2, 4 or 8 row tiles of 16, so the workgroup has 8, 4 or 2 row groups and 1, 2 or 4 column
groups:

- 2 row tiles: each W operand is formed (and, for split-BF16, split into hi/lo) once per
  workgroup;
- 4 or 8 row tiles: each B fragment read from LDS feeds more row tiles, at the price of W
  formed by several column groups.

Every tiling issues the same MFMAs in the same order for each output element, so the Gram
is bit-identical whatever the choice; only the time differs.  The best one depends on the
form (split-BF16 / FP16), on the temporal order Dt (how many column tiles there are), on
how many workgroups the batch launches and on N (the reduction length).

`GramTuner.launch` runs one Gram launch with the tiling cached for its key
(precision, Dt, workgroup bucket, N bucket).  On a miss it times every tiling on that very
launch, reps times each (HIP events on the caller's stream, median), keeps the fastest and
records it in memory and in a JSON file.  The last timed run has already written the
result.  Tuning waits for the stream once per new key; later launches with the key do
not wait.

The cache file is `$VTEM_AMD_TUNE_CACHE`, else `~/.cache/vtem_amd/gram_tiles.json`.
`VTEM_AMD_TUNE=0` turns tuning off: every launch then uses the form's default tiling.
"""
from __future__ import annotations

import json
import math
import os
import threading
from typing import Callable, Dict, List, Optional

TILES = (2, 4, 8)                     # row tiles of 16 per wave
LP_BP = 16                            # spatial indices per workgroup block (256 pair rows)


def default_tile(precision: str) -> int:
    """The form's tiling when nothing is tuned: split-BF16 2, FP16 4."""
    return 2 if precision == "bf16x3" else 4


def _bucket(x: int) -> int:
    """Power-of-two bucket (ceil of log2, at least 0)."""
    return max(0, int(math.ceil(math.log2(max(1, int(x))))))


def tune_key(precision: str, Dt: int, max_Ds: int, batch: int, max_N: int) -> str:
    """Cache key: the form, Dt, the workgroup count's and N's power-of-two buckets."""
    nbp = (int(max_Ds) + LP_BP - 1) // LP_BP
    wg = nbp * (nbp + 1) // 2 * int(batch)
    return f"{precision}/dt{int(Dt)}/wg{_bucket(wg)}/n{_bucket(max_N)}"


def default_cache_path() -> str:
    p = os.environ.get("VTEM_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "vtem_amd", "gram_tiles.json")


class GramTuner:
    """Tile choice per key; `timer(run, tile, reps)` returns the median milliseconds of
    `reps` runs of `run(tile)`.  The default timer uses HIP events on the current stream."""

    def __init__(self, path: Optional[str] = None, enabled: Optional[bool] = None,
                 reps: int = 3, timer: Optional[Callable] = None):
        self.path = default_cache_path() if path is None else path
        self.enabled = (os.environ.get("VTEM_AMD_TUNE", "1") != "0") if enabled is None else enabled
        self.reps = int(reps)
        self.timer = timer or _event_timer
        self._lock = threading.Lock()
        self.choices: Dict[str, int] = {}
        self.timings: Dict[str, Dict[int, float]] = {}      # measured this process
        self._load()

    def _load(self):
        try:
            with open(self.path) as f:
                data = json.load(f)
            for k, v in data.get("choices", {}).items():
                if int(v) in TILES:
                    self.choices[k] = int(v)
        except (OSError, ValueError, AttributeError):
            pass

    def _save(self):
        try:
            os.makedirs(os.path.dirname(self.path) or ".", exist_ok=True)
            tmp = self.path + f".{os.getpid()}.tmp"
            with open(tmp, "w") as f:
                json.dump({"choices": self.choices}, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass                                  # the choice stays in memory

    def launch(self, run: Callable[[int], None], precision: str, Dt: int, max_Ds: int,
               batch: int, max_N: int) -> int:
        """Run the Gram once with the key's tiling (tuning it first on a miss); returns the
        tiling used.  run(tile) enqueues the launch on the current stream."""
        key = tune_key(precision, Dt, max_Ds, batch, max_N)
        with self._lock:
            tile = self.choices.get(key)
        if tile is not None or not self.enabled:
            tile = tile or default_tile(precision)
            run(tile)
            return tile
        times = {t: self.timer(run, t, self.reps) for t in TILES}
        # every tiling wrote the same Gram: the timed runs leave the result in place
        best = min(TILES, key=lambda t: (times[t], t != default_tile(precision)))
        with self._lock:
            self.choices[key] = best
            self.timings[key] = times
            self._save()
        return best


def _event_timer(run: Callable[[int], None], tile: int, reps: int) -> float:
    import torch
    ts: List[float] = []
    run(tile)                                     # first launch of the code object
    for _ in range(max(1, reps)):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        run(tile)
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    ts.sort()
    return ts[len(ts) // 2]


_TUNER: Optional[GramTuner] = None
_TUNER_LOCK = threading.Lock()


def gram_tuner() -> GramTuner:
    """The process-wide tuner (created on first use)."""
    global _TUNER
    with _TUNER_LOCK:
        if _TUNER is None:
            _TUNER = GramTuner()
        return _TUNER
