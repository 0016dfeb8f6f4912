"""Measured-autotuning cache shared by the single-problem forward
(``engines.DeviceProblem.autotune``) and the robust-scenario batch
(``robust.RobustBatch.autotune``): winners are kept in one JSON file,
``$PBDOSE_AUTOTUNE_CACHE`` (default ``~/.cache/pbdose/autotune.json``), keyed by
the problem's shape and the device name. No counterpart in the reference.
"""
from __future__ import annotations

import json
import os

import torch


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


def device_key(dev) -> str:
    """Identity of a GPU for the cache: architecture and compute-unit count (the
    marketing name is not reported the same way in every process)."""
    p = torch.cuda.get_device_properties(dev)
    return f"{getattr(p, 'gcnArchName', p.name)}/{p.multi_processor_count}"
