"""Dose engines: the drop-in for the reference's per-layer-exact batched-matmul
engine, backed by the fused HIP kernels of libpbdose_hip.so.

    D(z,y,x) = sum_l sum_ij w[l,j,i] IDD_l(z) Gx_l(z,x,x_i) Gy_l(z,y,y_j)
    Gx_l(z,x,x_i) = exp(-(x - x_i)^2 / (2 s^2)) / (sqrt(2 pi) s),  s = max(sigma_l(z), 1e-9)

``torch_separable`` keeps the reference signature (engines.py:208-283): it
copies the problem's host arrays to the device on every call, runs the fused
kernel and synchronises, returning a (nz, ny, nx) float32 device tensor
(float64 from the FP64 kernel with dtype=torch.float64).
``DeviceProblem`` is the device-resident path with no per-call copies.
"""
from __future__ import annotations

import dataclasses
import time
import warnings
from typing import Optional

import numpy as np
import torch

from . import _lib
from .model import PBProblem

INV_SQRT_2PI = 1.0 / np.sqrt(2.0 * np.pi)
PRECISIONS = tuple(_lib.PREC)


_CHUNK_CACHE: dict = {}


class DeviceProblem:
    """A PBProblem resident on one GPU: weights (L, nsy, nsx) FP32, idd and
    sigma (L, nz) FP64, plus the C-ABI descriptors and cached workspaces."""

    def __init__(self, problem: PBProblem, device="cuda", weights=None, idd=None, sigma=None):
        self.dev = _lib.require_cuda(device)
        self.problem = problem
        g, lat = problem.grid, problem.lattice
        self.grid_s = _lib.grid_struct(g)
        self.lat_s = _lib.lattice_struct(lat)
        self.shape = g.shape
        w = lat.ensure_weights() if weights is None else weights
        self._w_src = w          # the weights as given, for the FP64 form
        self._w64 = None
        self.w = torch.as_tensor(w, dtype=torch.float32, device=self.dev).contiguous()
        self.idd = torch.as_tensor(problem.idd if idd is None else idd, dtype=torch.float64,
                                   device=self.dev).contiguous()
        self.sigma = torch.as_tensor(problem.sigma if sigma is None else sigma,
                                     dtype=torch.float64, device=self.dev).contiguous()
        L, nsy, nsx = lat.n_layers, lat.n_sy, lat.n_sx
        if tuple(self.w.shape) != (L, nsy, nsx):
            raise ValueError(f"weights shape {tuple(self.w.shape)} != {(L, nsy, nsx)}")
        if tuple(self.idd.shape) != (L, g.nz) or tuple(self.sigma.shape) != (L, g.nz):
            raise ValueError(f"idd/sigma must be (L, nz) = {(L, g.nz)}")
        _lib.check(_lib.lib().pbh_validate(self.grid_s, self.lat_s), "pbh_validate")
        self._ws = {}
        # active layers per depth slice, for the layer chunk: from the host tables when
        # the problem carries them, else one small copy from the device
        if idd is None and isinstance(problem.idd, np.ndarray):
            act = (problem.idd != 0).sum(axis=0)
        else:
            act = (self.idd != 0).sum(dim=0).cpu().numpy()
        self._active = np.ascontiguousarray(act, dtype=np.int32)
        self._chunk = {}
        self._tuned = {}      # form -> (layer chunk, tile rows) measured by autotune()

    def auto_chunk(self, precision: int) -> int:
        """Layer chunk of the automatic forward for this problem's active layers
        (pbh_forward_choose_chunk; results are cached per problem shape, so that
        per-call wrappers such as torch_separable pay for the model once)."""
        if precision not in self._chunk:
            g, lat = self.problem.grid, self.problem.lattice
            key = (g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ox, g.oy, g.oz, lat.n_sx, lat.n_sy,
                   lat.n_layers, lat.dx_spot, lat.dy_spot, lat.x0, lat.y0, self.dev.index,
                   precision, self._active.tobytes())
            c = _CHUNK_CACHE.get(key)
            if c is None:
                c = _lib.lib().pbh_forward_choose_chunk(self.grid_s, self.lat_s, precision,
                                                        self._active.ctypes.data, 0)
                _lib.check(min(c, 0), "pbh_forward_choose_chunk")
                if len(_CHUNK_CACHE) > 4096:
                    _CHUNK_CACHE.clear()
                _CHUNK_CACHE[key] = c
            self._chunk[precision] = int(c)
        return self._chunk[precision]

    def auto_shape(self, precision: int) -> tuple:
        """(layer chunk, tile rows) of the automatic forward for this problem's active
        layers (pbh_forward_choose_shape: 32- or 64-row tiles; cached like auto_chunk)."""
        key = ("shape", precision)
        if key not in self._chunk:
            g, lat = self.problem.grid, self.problem.lattice
            ckey = ("shape", g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ox, g.oy, g.oz, lat.n_sx, lat.n_sy,
                    lat.n_layers, lat.dx_spot, lat.dy_spot, lat.x0, lat.y0, self.dev.index,
                    precision, self._active.tobytes())
            c = _CHUNK_CACHE.get(ckey)
            if c is None:
                rows = _lib.C.c_int32(0)
                ch = _lib.lib().pbh_forward_choose_shape(self.grid_s, self.lat_s, precision,
                                                         self._active.ctypes.data, 0, _lib.C.byref(rows))
                _lib.check(min(ch, 0), "pbh_forward_choose_shape")
                if len(_CHUNK_CACHE) > 4096:
                    _CHUNK_CACHE.clear()
                c = (int(ch), int(rows.value))
                _CHUNK_CACHE[ckey] = c
            self._chunk[key] = c
        return self._chunk[key]

    def shape_candidates(self, precision: int) -> list:
        """(layer chunk, tile rows) pairs worth timing: for each tile height the
        model's chunk and the chunks that cut the busiest slice's active layers into
        1 to 6 equal parts (where the count of partial tiles changes)."""
        g, lat = self.problem.grid, self.problem.lattice
        heights = [32]
        # 64-row tiles: 16-bit forms with 2 : 5 voxel : spot pitches on both axes
        if (precision != _lib.PREC["fp32"] and abs(lat.dx_spot / g.dx - 2.5) < 1e-9
                and abs(lat.dy_spot / g.dy - 2.5) < 1e-9):
            heights.append(64)
        amax = max(1, int(self._active.max()) if self._active.size else 1)
        c_model, r_model = self.auto_shape(precision)
        cands = set()
        for rows in heights:
            cands |= {(max(1, -(-amax // k)), rows) for k in range(1, 7)}
            cands.add((self.auto_chunk(precision), 32))
        cands.add((c_model, r_model))
        return sorted(cands, key=lambda t: (t[1], -t[0]))

    def autotune(self, precision: str, reps: int = 20, use_cache: bool = True) -> tuple:
        """Measured (layer chunk, tile rows) of the forward for this problem: every
        pair of shape_candidates is timed over `reps` back-to-back calls, then the
        chunks within 2 of the fastest at its tile height, and the fastest of all is
        used by later forward() calls with layer_chunk = tile_rows = 0. Winners are
        cached per (problem shape, active layers, form, device). All candidates
        compute the same dose up to FP32 summation order."""
        from . import tuning
        p = _lib.PREC[precision]
        g, lat = self.problem.grid, self.problem.lattice
        key = "fwd|" + "|".join(str(v) for v in (
            g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, lat.n_sx, lat.n_sy, lat.n_layers, lat.dx_spot,
            lat.dy_spot, precision, tuning.device_key(self.dev), self._active.tobytes().hex()))
        cache = tuning.load_cache() if use_cache else {}
        if key in cache:
            best = tuple(cache[key]["shape"])
        else:
            out = torch.empty(self.shape, dtype=torch.float32, device=self.dev)
            results = []

            def time_shape(c, rows):
                self.forward(precision, out=out, layer_chunk=c, tile_rows=rows)
                torch.cuda.synchronize(self.dev)
                s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s.record()
                for _ in range(reps):
                    self.forward(precision, out=out, layer_chunk=c, tile_rows=rows)
                e.record()
                torch.cuda.synchronize(self.dev)
                results.append((s.elapsed_time(e) / reps, (c, rows)))

            for c, rows in self.shape_candidates(p):
                time_shape(c, rows)
            # refine: the neighbouring chunks of the fastest at its tile height
            _, (c0, r0) = min(results, key=lambda r: r[0])
            tried = {cr for _, cr in results}
            amax = max(1, int(self._active.max()) if self._active.size else 1)
            for c in range(max(1, c0 - 2), min(amax, c0 + 2) + 1):
                if (c, r0) not in tried:
                    time_shape(c, r0)
            best_ms, best = min(results, key=lambda r: r[0])
            self.tuning_results = results
            if use_cache:
                cache = tuning.load_cache()
                cache[key] = {"shape": list(best), "ms": best_ms,
                              "timings": [[ms, list(cr)] for ms, cr in results]}
                tuning.save_cache(cache)
        self._tuned[p] = (int(best[0]), int(best[1]))
        return self._tuned[p]

    def autotune_split(self, precision: str, reps: int = 20, use_cache: bool = True) -> int:
        """Measured layer split of the weight adjoint (``precision`` 'fp32', 'bf16',
        'bf16x3', 'fp16' or 'fp16x3'): the dispatch model's split and its neighbours (splits 2 to 16) are
        timed over `reps` objective + adjoint calls, and the fastest is used by later
        adjoint calls with split 0. Cached like ``autotune``. Every split gives the same
        dJ/dw bit for bit."""
        from . import tuning
        p = _lib.PREC[precision]
        g, lat = self.problem.grid, self.problem.lattice
        key = "adj|" + "|".join(str(v) for v in (
            g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, lat.n_sx, lat.n_sy, lat.n_layers, lat.dx_spot,
            lat.dy_spot, precision, tuning.device_key(self.dev), self._active.tobytes().hex()))
        cache = tuning.load_cache() if use_cache else {}
        if key in cache:
            best = int(cache[key]["split"])
        else:
            from . import planning
            zero = torch.zeros(self.shape, dtype=torch.float32, device=self.dev)
            model = self.auto_split(p)
            results = []
            for sp in sorted({max(1, min(16, model + d)) for d in (-3, -2, -1, 0, 1, 2, 3)} | {2, 4, 8}):
                planning.objective_gradient(self, zero, zero, precision=precision, split=sp)
                torch.cuda.synchronize(self.dev)
                s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s.record()
                for _ in range(reps):
                    planning.objective_gradient(self, zero, zero, precision=precision, split=sp)
                e.record()
                torch.cuda.synchronize(self.dev)
                results.append((s.elapsed_time(e) / reps, sp))
            best_ms, best = min(results)
            self.split_results = results
            if use_cache:
                cache = tuning.load_cache()
                cache[key] = {"split": best, "ms": best_ms, "timings": [[ms, sp] for ms, sp in results]}
                tuning.save_cache(cache)
        self._chunk[("split", p)] = int(best)
        return int(best)

    def auto_split(self, precision: int) -> int:
        """Layer split of the automatic adjoint: the measured one after autotune_split,
        else the dispatch model's for this problem's active layers
        (pbh_adjoint_choose_split; cached like auto_chunk)."""
        key = ("split", precision)
        if key not in self._chunk:
            g, lat = self.problem.grid, self.problem.lattice
            ckey = ("split", g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, lat.n_sx, lat.n_sy, lat.n_layers,
                    lat.dx_spot, lat.dy_spot, self.dev.index, precision, self._active.tobytes())
            c = _CHUNK_CACHE.get(ckey)
            if c is None:
                c = _lib.lib().pbh_adjoint_choose_split(self.grid_s, self.lat_s, precision,
                                                        self._active.ctypes.data, 0)
                _lib.check(min(c, 0), "pbh_adjoint_choose_split")
                if len(_CHUNK_CACHE) > 4096:
                    _CHUNK_CACHE.clear()
                _CHUNK_CACHE[ckey] = c
            self._chunk[key] = int(c)
        return self._chunk[key]

    def workspace(self, kind: str, precision: int, layer_chunk: int = 0, tile_rows: int = 0,
                  adj_precision: int = 0) -> torch.Tensor:
        key = (kind, precision, layer_chunk, tile_rows, adj_precision)
        if key not in self._ws:
            L = _lib.lib()
            if kind == "fwdobjadj":
                n = L.pbh_forward_objective_adjoint_workspace_bytes(self.grid_s, self.lat_s, precision, layer_chunk,
                                                                    tile_rows, adj_precision)
                if n == 0:
                    raise ValueError("no planning-step workspace for these forms and shape")
            elif kind == "fwd":
                n = L.pbh_forward_workspace_bytes_shaped(self.grid_s, self.lat_s, precision, layer_chunk,
                                                         tile_rows)
            elif kind == "obj":
                n = L.pbh_objective_workspace_bytes(self.grid_s)
            elif kind == "objadj":
                n = L.pbh_objective_adjoint_workspace_bytes(self.grid_s, self.lat_s, precision)
            else:
                n = L.pbh_adjoint_workspace_bytes(self.grid_s, self.lat_s, precision)
            self._ws[key] = _lib.workspace(n, self.dev)
        return self._ws[key]

    def shape_of(self, p: int, layer_chunk: int = 0, tile_rows: int = 0) -> tuple:
        """The (layer chunk, tile rows) a forward of form p runs with (as ``forward``)."""
        if int(layer_chunk) == 0 and int(tile_rows) == 0:
            return self._tuned.get(p) or self.auto_shape(p)
        if int(layer_chunk) == 0:
            return self.auto_chunk(p), int(tile_rows)
        return int(layer_chunk), int(tile_rows)

    def forward(self, precision: str = "fp32", out: Optional[torch.Tensor] = None,
                w: Optional[torch.Tensor] = None, layer_chunk: int = 0, kernel_events=None,
                tile_rows: int = 0) -> torch.Tensor:
        """Dose (nz, ny, nx) float32, asynchronous on the current stream.
        ``layer_chunk``: at most this many active layers of a depth slice per
        workgroup (>= L = one workgroup per dose tile). With layer_chunk and tile_rows
        both 0, both come from this problem's active layers per slice
        (pbh_forward_choose_shape); a chunk of 0 alone is pbh_forward_choose_chunk's.
        ``kernel_events``: a pair of recorded torch.cuda.Event(enable_timing=True)
        that pbh_forward_profiled re-records right before and after the fused
        contraction kernel (for timing that kernel alone).
        ``tile_rows``: dose-tile height, 32 or 64 (16-bit forms on the 2 : 5 pitch
        geometry); 0 = the default."""
        p = _lib.PREC[precision]
        if out is None:
            out = torch.empty(self.shape, dtype=torch.float32, device=self.dev)
        elif out.dtype != torch.float32 or tuple(out.shape) != self.shape or not out.is_contiguous():
            raise ValueError("out must be a contiguous float32 (nz, ny, nx) tensor")
        wt = self.w if w is None else w
        if wt.dtype != torch.float32 or tuple(wt.shape) != tuple(self.w.shape) or not wt.is_contiguous():
            raise ValueError("weights must be a contiguous float32 (L, nsy, nsx) tensor")
        layer_chunk, tile_rows = self.shape_of(p, layer_chunk, tile_rows)
        ws = self.workspace("fwd", p, int(layer_chunk), int(tile_rows))
        if kernel_events is not None:
            e0, e1 = kernel_events
            _lib.check(_lib.lib().pbh_forward_profiled(
                self.grid_s, self.lat_s, _lib.ptr(wt), _lib.ptr(self.idd), _lib.ptr(self.sigma),
                _lib.ptr(out), p, int(layer_chunk), int(tile_rows), _lib.ptr(ws), ws.numel(),
                _lib.C.c_void_p(e0.cuda_event), _lib.C.c_void_p(e1.cuda_event),
                _lib.stream_of(self.dev)), "pbh_forward_profiled")
            return out
        _lib.check(_lib.lib().pbh_forward_shaped(
            self.grid_s, self.lat_s, _lib.ptr(wt), _lib.ptr(self.idd), _lib.ptr(self.sigma),
            _lib.ptr(out), p, int(layer_chunk), int(tile_rows), _lib.ptr(ws), ws.numel(),
            _lib.stream_of(self.dev)), "pbh_forward_shaped")
        return out


    def forward_fp64(self, out: Optional[torch.Tensor] = None, w: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Dose (nz, ny, nx) float64 from FP64 weights and tables (pbh_forward_fp64),
        asynchronous on the current stream. ``w``: FP64 (L, nsy, nsx) weights on the
        device; default the problem's weights at full precision."""
        if out is None:
            out = torch.empty(self.shape, dtype=torch.float64, device=self.dev)
        elif out.dtype != torch.float64 or tuple(out.shape) != self.shape or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 (nz, ny, nx) tensor")
        if w is None:
            if self._w64 is None:
                self._w64 = torch.as_tensor(self._w_src, dtype=torch.float64, device=self.dev).contiguous()
            w = self._w64
        if w.dtype != torch.float64 or tuple(w.shape) != tuple(self.w.shape) or not w.is_contiguous():
            raise ValueError("weights must be a contiguous float64 (L, nsy, nsx) tensor")
        _lib.check(_lib.lib().pbh_forward_fp64(self.grid_s, self.lat_s, _lib.ptr(w), _lib.ptr(self.idd),
                                               _lib.ptr(self.sigma), _lib.ptr(out), _lib.stream_of(self.dev)),
                   "pbh_forward_fp64")
        return out


def torch_separable(problem: PBProblem, device: str = "cuda", dtype=None,
                    allow_tf32: bool = False, per_layer: bool = True,
                    layer_chunk: Optional[int] = None, return_timing: bool = False,
                    return_intermediates: bool = False, precision: str = "fp32"):
    """Per-layer-exact dose on the GPU, drop-in for engines.py:208-283.

    device        must be a CUDA (ROCm) device; there is no CPU path.
    dtype         None or torch.float32: the dose in FP32, computed in FP32 or in
                  the mixed form ``precision`` selects. torch.float64: FP64
                  weights, tables, arithmetic and dose.
    allow_tf32    accepted for compatibility; gfx950 has no TF32 mode, so it
                  has no effect (the reference measured none either) and True warns.
    per_layer     True (default): sigma exact per layer. False: the reference's
                  single-sigma approximation (engines.py:243-255), on the GPU:
                  one width per slice, the layer-weight average sigma_bar(z)
                  (engines.py:75-80), applied to A(z) = sum_l IDD_l(z) w_l.
    layer_chunk   accepted and ignored (with a warning when given): no
                  layer-sized intermediate exists.
    precision     'fp32' (default) or one of 'bf16', 'fp16', 'bf16x3',
                  'fp16x3', 'bf16x3_peak', 'fp16x3_peak'.
    Returns D (nz, ny, nx) on the device, or (D, wall_seconds).
    """
    if dtype not in (None, torch.float32, torch.float64):
        raise ValueError(f"dtype {dtype} not supported: float32 (default) or float64")
    if dtype == torch.float64 and precision != "fp32":
        raise ValueError(f"precision {precision!r} selects a float32 form; dtype=float64 computes in FP64")
    if return_intermediates:
        warnings.warn("return_intermediates is ignored: no intermediate is materialised", stacklevel=2)
    if layer_chunk is not None:
        warnings.warn("layer_chunk is ignored: no layer-sized intermediate exists (the kernels' own "
                      "layer chunks are chosen per problem)", stacklevel=2)
    if allow_tf32:
        warnings.warn("allow_tf32 has no effect: gfx950 has no TF32 mode; the dose is computed in the "
                      "form `precision` names", stacklevel=2)
    t0 = time.perf_counter()
    fp64 = dtype == torch.float64
    dp = DeviceProblem(problem, device) if per_layer else averaged_problem(problem, device, fp64=fp64)
    D = dp.forward_fp64() if fp64 else dp.forward(precision)
    torch.cuda.synchronize(dp.dev)
    dt = time.perf_counter() - t0
    return (D, dt) if return_timing else D


def averaged_problem(problem: PBProblem, device="cuda", fp64: bool = False) -> DeviceProblem:
    """The single-sigma approximation (per_layer=False, engines.py:184-191,
    :243-255) as a DeviceProblem of nz layers, written on the device by
    pbh_average_sigma: layer k carries the amplitude map A(z_k) = sum_l IDD_l(z_k)
    w_l, an IDD of 1 at slice k only, and sigma_bar on every slice, so that the
    forward gives slice z the dose Gy A(z) Gx^T with width sigma_bar(z). The
    sigma_bar(z) of the reference (engines.py:75-80) is ``.sigma[0]``. ``fp64``: the
    amplitude maps are kept in FP64 (pbh_average_sigma_fp64) for ``forward_fp64``."""
    dev = _lib.require_cuda(device)
    g, lat = problem.grid, problem.lattice
    w = torch.as_tensor(lat.ensure_weights(), dtype=torch.float64, device=dev).contiguous()
    idd = torch.as_tensor(problem.idd, dtype=torch.float64, device=dev).contiguous()
    sig = torch.as_tensor(problem.sigma, dtype=torch.float64, device=dev).contiguous()
    L, nz = lat.n_layers, g.nz
    if tuple(w.shape) != (L, lat.n_sy, lat.n_sx) or tuple(idd.shape) != (L, nz) or tuple(sig.shape) != (L, nz):
        raise ValueError("weights must be (L, nsy, nsx) and idd/sigma (L, nz)")
    wa = torch.empty((nz, lat.n_sy, lat.n_sx), dtype=torch.float64 if fp64 else torch.float32, device=dev)
    ia = torch.empty((nz, nz), dtype=torch.float64, device=dev)
    sa = torch.empty((nz, nz), dtype=torch.float64, device=dev)
    wl = torch.empty((L,), dtype=torch.float64, device=dev)
    name = "pbh_average_sigma_fp64" if fp64 else "pbh_average_sigma"
    _lib.check(getattr(_lib.lib(), name)(_lib.ptr(w), _lib.ptr(idd), _lib.ptr(sig), L, nz, lat.n_sy, lat.n_sx,
                                         _lib.ptr(wa), _lib.ptr(ia), _lib.ptr(sa), _lib.ptr(wl),
                                         _lib.stream_of(dev)), name)
    lat_a = dataclasses.replace(lat, n_layers=nz, energies=None, weights=None)
    pa = PBProblem(g, lat_a, problem.idd, problem.sigma, dict(problem.meta), problem.fluence_normalized)
    dp = DeviceProblem(pa, dev, weights=wa, idd=ia, sigma=sa)
    dp.layer_sums = wl
    return dp


def separable_hip(problem: PBProblem, device="cuda", precision: str = "fp32") -> torch.Tensor:
    """Alias of torch_separable(problem, device, precision=...)."""
    return torch_separable(problem, device, precision=precision)


def relative_error(a, b, mask=None) -> dict:
    """Global-normalised error of a against reference b (engines.py:364-378)."""
    a = np.asarray(a.detach().cpu() if torch.is_tensor(a) else a, dtype=np.float64)
    b = np.asarray(b.detach().cpu() if torch.is_tensor(b) else b, dtype=np.float64)
    if mask is None:
        mask = np.ones(b.shape, dtype=bool)
    d = (a - b)[mask]
    ref = float(np.max(np.abs(b))) or 1.0
    if d.size == 0:
        return {"max_abs_err": 0.0, "max_rel_err_global": 0.0, "rms_err": 0.0,
                "rms_rel_global": 0.0, "n_vox": int(mask.sum())}
    mx, rms = float(np.abs(d).max()), float(np.sqrt((d**2).mean()))
    return {"max_abs_err": mx, "max_rel_err_global": mx / ref, "rms_err": rms,
            "rms_rel_global": rms / ref, "n_vox": int(mask.sum())}


def normalize_dose(d, mode: str = "max", roi=None):
    """Dose divided by its max (or the ROI max) (engines.py:381-388)."""
    d = np.asarray(d, dtype=np.float64)
    ref = float(d[roi].max()) if (mode == "roi_max" and roi is not None) else float(d.max())
    return d / (ref if ref > 0 else 1.0), ref
