"""Measured-autotuning cache shared by the single-problem forward
(``engines.DeviceProblem.autotune``) and the robust-scenario batch
(``robust.RobustBatch.autotune``): winners are kept in one JSON file,
``$PBDOSE_AUTOTUNE_CACHE`` (default ``~/.cache/pbdose/autotune.json``), keyed by
the problem's shape and the device name. No counterpart in the reference.
"""
from __future__ import annotations

import json
import os


def cache_path() -> str:
    p = os.environ.get("PBDOSE_AUTOTUNE_CACHE")
    if p:
        return p
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "pbdose", "autotune.json")


def load_cache() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(c: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass
