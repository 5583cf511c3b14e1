"""Tile/blocking autotuner with a persistent JSON cache.

A tuning decision is keyed by the operation and a shape bucket (for example the batched fit
keys on the kernel kinds, the problem count and the largest size, rounded up to powers of
two). ``tune`` returns the cached choice for the key, or times every candidate with the
caller's ``measure`` function (seconds, lower is better), stores the fastest with all
timings, and returns it. The cache file is ``$QMLHIP_AUTOTUNE_CACHE`` or
``~/.cache/qml_amd/autotune.json``; reads and writes go through one process-wide mutex and the
file is replaced atomically, so concurrent tuners never see a torn file.
"""

from __future__ import annotations

import json
import math
import os
import tempfile
import threading
from typing import Callable, Iterable

_lock = threading.Lock()


def default_path() -> str:
    return os.environ.get("QMLHIP_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "qml_amd",
                                       "autotune.json"))


def pow2_bucket(x: int) -> int:
    return 1 if x <= 1 else 1 << math.ceil(math.log2(x))


def make_key(op: str, **shape) -> str:
    return op + "|" + ",".join(f"{k}={shape[k]}" for k in sorted(shape))


class Autotuner:
    def __init__(self, path: str | None = None):
        self.path = path or default_path()

    def _read(self) -> dict:
        try:
            with open(self.path) as f:
                data = json.load(f)
            return data if isinstance(data, dict) else {}
        except (OSError, ValueError):
            return {}

    def _write(self, data: dict) -> None:
        d = os.path.dirname(self.path) or "."
        os.makedirs(d, exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=d, prefix=".autotune.")
        with os.fdopen(fd, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def lookup(self, key: str):
        with _lock:
            ent = self._read().get(key)
        return None if ent is None else ent.get("choice")

    def store(self, key: str, choice, timings: dict) -> None:
        with _lock:
            data = self._read()
            data[key] = {"choice": choice, "timings": timings}
            self._write(data)

    def tune(self, key: str, candidates: Iterable, measure: Callable[[object], float],
             force: bool = False):
        candidates = list(candidates)
        if not force:
            hit = self.lookup(key)
            if hit is not None and hit in candidates:
                return hit
        timings = {}
        for c in candidates:
            timings[str(c)] = float(measure(c))
        best = min(candidates, key=lambda c: timings[str(c)])
        self.store(key, best, timings)
        return best
