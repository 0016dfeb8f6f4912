"""Gamma index on the GPU (``pbh_gamma_index``), with the reference's interface.

``gamma_index`` keeps the signature and result keys of gamma.py:81-160;
``gamma_index_torch`` those of gamma.py:213-261; ``gamma_pass_rate_vs_criteria``
those of gamma.py:197-210. The arithmetic runs in FP64 in the HIP kernel for
every entry point (the reference's torch variant evaluates in FP32), so pass
decisions equal an FP64 evaluation of gamma.py on the same inputs. Inputs may
be NumPy arrays (copied to the device) or device tensors (used in place,
float32 or float64).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

_DT = {torch.float32: 0, torch.float64: 1}


def _device_dose(a, dev):
    if torch.is_tensor(a):
        t = a.detach()
        if t.dtype not in _DT:
            t = t.double()
        return t.to(dev).contiguous()
    return torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev).contiguous()


def _percentiles(x: torch.Tensor, qs):
    """np.percentile(x, q) with linear interpolation, on the device."""
    if x.numel() == 0:
        return [float("nan")] * len(qs)
    s, _ = torch.sort(x.reshape(-1))
    n = s.numel()
    out = []
    for q in qs:
        pos = q / 100.0 * (n - 1)
        lo = int(np.floor(pos))
        hi = min(lo + 1, n - 1)
        f = pos - lo
        a, b = float(s[lo]), float(s[hi])
        out.append(a + (b - a) * f)
    return out


def gamma_device(reference, evaluation, spacing=(0.2, 0.2, 0.2), dose_crit_pct: float = 3.0,
                 dta_mm: float = 3.0, threshold_pct: float = 10.0, normalization: str = "global",
                 ref_dose=None, want_map: bool = True, device="cuda"):
    """Runs the kernel; returns (counts, stats, gamma map or None, ref tensor)."""
    dev = _lib.require_cuda(device)
    ref = _device_dose(reference, dev)
    ev = _device_dose(evaluation, dev)
    if ref.shape != ev.shape or ref.dim() != 3:
        raise ValueError(f"shape mismatch: {tuple(ref.shape)} vs {tuple(ev.shape)}")
    if normalization not in ("global", "local"):
        raise ValueError(f"unknown normalization {normalization!r}")
    nz, ny, nx = ref.shape
    dz, dy, dx = (float(s) for s in spacing)
    g = _lib.PbhGrid(nx, ny, nz, 0, dx, dy, dz, 0.0, 0.0, 0.0)
    L = _lib.lib()
    ws = _lib.workspace(L.pbh_gamma_workspace_bytes(g), dev)
    counts = torch.zeros(3, dtype=torch.int64, device=dev)
    stats = torch.zeros(3, dtype=torch.float64, device=dev)
    gmap = torch.empty(ref.shape, dtype=torch.float64, device=dev) if want_map else None
    rmax = 0.0 if ref_dose is None else float(ref_dose)
    if ref_dose is not None and rmax <= 0:
        raise ValueError("reference dose max <= 0")
    _lib.check(L.pbh_gamma_index(g, _lib.ptr(ref), _DT[ref.dtype], _lib.ptr(ev), _DT[ev.dtype],
                                 float(dose_crit_pct), float(dta_mm), float(threshold_pct), rmax,
                                 int(normalization == "local"), _lib.ptr(gmap), _lib.ptr(counts),
                                 _lib.ptr(stats), _lib.ptr(ws), ws.numel(), _lib.stream_of(dev)),
               "pbh_gamma_index")
    c = counts.cpu().numpy()
    s = stats.cpu().numpy()
    if not s[2] > 0:
        raise ValueError("reference dose max <= 0")
    return c, s, gmap, ref


def gamma_index(reference, evaluation, spacing=(0.2, 0.2, 0.2), dose_crit_pct: float = 3.0,
                dta_mm: float = 3.0, threshold_pct: float = 10.0, normalization: str = "global",
                ref_dose=None, return_maps: bool = True, device="cuda") -> dict:
    """Gamma index (gamma.py:81-160). Maps, when returned, are NumPy arrays."""
    c, s, gmap, ref = gamma_device(reference, evaluation, spacing, dose_crit_pct, dta_mm, threshold_pct,
                                   normalization, ref_dose, want_map=True, device=device)
    ref_max = float(s[2])
    mask = ref >= (threshold_pct / 100.0) * ref_max
    if int(mask.sum()) == 0:
        mask = ref > 0
    p95, p99 = _percentiles(gmap[mask], (95, 99))
    n_eval, n_pass = int(c[0]), int(c[1])
    res = {
        "pass_rate": (100.0 * n_pass / n_eval) if n_eval else float("nan"),
        "gamma_mean": float(s[1]),
        "gamma_max": float(s[0]),
        "gamma_p95": p95,
        "gamma_p99": p99,
        "n_eval": n_eval,
        "n_pass": n_pass,
        "dose_crit_pct": dose_crit_pct,
        "dta_mm": dta_mm,
        "threshold_pct": threshold_pct,
        "normalization": normalization,
        "ref_max": ref_max,
        "n_dta_offsets": int(c[2]),
    }
    if return_maps:
        res["gamma"] = gmap.cpu().numpy()
        res["mask"] = mask.cpu().numpy()
        res["ref"] = ref.double().cpu().numpy()
        res["eval"] = _device_dose(evaluation, ref.device).double().cpu().numpy()
    return res


def gamma_index_torch(reference, evaluation, spacing=(0.2, 0.2, 0.2), dose_crit_pct: float = 3.0,
                      dta_mm: float = 3.0, threshold_pct: float = 10.0, ref_dose=None,
                      device: str = "cuda", threshold: float = 1e-6) -> dict:
    """Statistics only (gamma.py:213-261); ``threshold`` is accepted and unused there too."""
    r = gamma_index(reference, evaluation, spacing, dose_crit_pct, dta_mm, threshold_pct, "global",
                    ref_dose, return_maps=False, device=device)
    return r


def gamma_pass_rate_vs_criteria(reference, evaluation, spacing=(0.2, 0.2, 0.2),
                                criteria=((3.0, 3.0), (2.0, 2.0), (1.0, 1.0), (3.0, 2.0), (2.0, 3.0)),
                                threshold_pct: float = 10.0, device="cuda") -> list:
    """Pass rates over several (dose %, DTA mm) criteria (gamma.py:197-210)."""
    out = []
    for dp, dt in criteria:
        r = gamma_index(reference, evaluation, spacing, dp, dt, threshold_pct=threshold_pct,
                        return_maps=False, device=device)
        out.append({"dose_pct": dp, "dta_mm": dt, **r})
    return out


def pass_rate(reference, evaluation, spacing=(0.2, 0.2, 0.2), dose_crit_pct: float = 2.0,
              dta_mm: float = 2.0, threshold_pct: float = 10.0, device="cuda"):
    """(pass rate %, n_eval, gamma_max) without maps or percentiles: one kernel pass."""
    c, s, _, _ = gamma_device(reference, evaluation, spacing, dose_crit_pct, dta_mm, threshold_pct,
                              want_map=False, device=device)
    n = int(c[0])
    return (100.0 * int(c[1]) / n if n else float("nan")), n, float(s[0])
