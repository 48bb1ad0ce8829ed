"""A small autotuner for launch choices, with an on-disk cache.

``choose(key, candidates)`` runs each candidate once (warm-up), times
``reps`` more runs with device events, and keeps the fastest name per key;
later calls (and later processes, through the JSON cache) reuse it without
timing.  Keys name the device, the operation, the topology (the pattern's
layout hash), the dtype and the form, so a choice is remade when any of
them changes.

The cache file is ``$TENSORMESH_AMD_TUNE_CACHE`` if set, else
``~/.cache/tensormesh_amd/tune.json``; if it cannot be written the tuner keeps
its choices in memory only.  ``TENSORMESH_AMD_TUNE=0`` disables timing: every
``choose`` returns the first candidate (the documented default).

What is tuned (callers decide; each choice is a launch variant of the same
computation):
* ``ElementAssembler.p1_path = "auto"``: fused rows / two-stage segments /
  two-stage rows / incidence for P1 scalar forms.  All four sum every entry in
  ascending element order; the fused kernel evaluates the element geometry
  from a rotated vertex order, so its values can differ from the two-stage
  paths' in the last bits (within 1e-15 relative) -- the choice is opt-in.
* ``NeoHookeanModel.nh_path = "auto"``: slot-buffer segment reduction or
  incidence-ordered block rows; same element values, same order.
* ``epw = "auto"`` on the element-matrix, Neo-Hookean and J2 assemblers:
  elements per wavefront of the wave kernels (1/2, 1/2/4, 1/4/8; counts a
  kernel does not admit for the case are skipped).  Each element keeps its
  own lane group and arithmetic, so the results are bit-identical.
* ``BatchedPCG.solve(chunk=None)``: iterations per graph replay between
  convergence polls, from the previous solve's iteration count on the same
  system (systems stop on the device, so the chunk changes only the number of
  polls and of idle iterations, not the result).
"""

from __future__ import annotations

import json
import os
import threading
from pathlib import Path
from typing import Callable, Dict, Optional

import torch

_LOCK = threading.Lock()


def _cache_path() -> Path:
    env = os.environ.get("TENSORMESH_AMD_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "tensormesh_amd" / "tune.json"


class Tuner:
    def __init__(self, path: Optional[Path] = None):
        self.path = path or _cache_path()
        self.enabled = os.environ.get("TENSORMESH_AMD_TUNE", "1") != "0"
        self._mem: Dict[str, object] = {}
        self._loaded = False

    # ---- cache ----
    def _load(self):
        if self._loaded:
            return
        self._loaded = True
        try:
            with open(self.path) as f:
                data = json.load(f)
            if isinstance(data, dict):
                self._mem.update(data)
        except (OSError, ValueError):
            pass

    def _save(self):
        try:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            tmp = self.path.with_suffix(f".{os.getpid()}.tmp")
            with open(tmp, "w") as f:
                json.dump(self._mem, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass   # read-only home: choices stay in memory

    def get(self, key: str):
        with _LOCK:
            self._load()
            return self._mem.get(key)

    def put(self, key: str, value) -> None:
        with _LOCK:
            self._load()
            self._mem[key] = value
            self._save()

    # ---- timing ----
    def choose(self, key: str, candidates: Dict[str, Callable[[], object]], reps: int = 3,
               device: Optional[torch.device] = None) -> str:
        """Name of the fastest candidate for ``key`` (timed once, then cached)."""
        names = list(candidates)
        if not self.enabled or len(names) == 1:
            return names[0]
        hit = self.get(key)
        if isinstance(hit, dict) and hit.get("choice") in candidates:
            return hit["choice"]
        from ._native import NativeError
        dev = device or torch.device("cuda", torch.cuda.current_device())
        times = {}
        for name in names:
            fn = candidates[name]
            try:
                fn()
            except NativeError:
                continue   # a launch shape the kernel does not admit for this case
            torch.cuda.synchronize(dev)
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(reps):
                fn()
            stop.record()
            torch.cuda.synchronize(dev)
            times[name] = start.elapsed_time(stop) / reps
        if not times:
            return names[0]
        best = min(times, key=times.get)
        self.put(key, {"choice": best, "ms": times})
        return best


_TUNER: Optional[Tuner] = None


def tuner() -> Tuner:
    global _TUNER
    if _TUNER is None:
        _TUNER = Tuner()
    return _TUNER


def device_tag(device) -> str:
    dev = torch.device(device)
    if dev.type != "cuda":
        return dev.type
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    return torch.cuda.get_device_name(idx).replace(" ", "_")


def pcg_chunk(previous_iterations: Optional[int]) -> int:
    """Iterations per replay for a solve expected to take about as many
    iterations as the previous solve of the same system: about 1/32 of them,
    as a power of two in [8, 128] (32 with no history).  Polls then cost
    about 32 host round trips per solve and the idle tail stays below 1/32 of
    the iterations."""
    if not previous_iterations or previous_iterations <= 0:
        return 32
    c = 8
    while c < 128 and c * 32 < previous_iterations:
        c *= 2
    return c
