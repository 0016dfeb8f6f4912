"""Robust-scenario batches: fields x setup/range scenarios, forward dose,
weighted-quadratic objective and spot-weight adjoint for every (field,
scenario) item, run by the native batch entry points of libpbdose_hip.so
(``pbh_batch_create`` / ``pbh_batch_forward_adjoint``) and captured into one
HIP graph.

A scenario moves the patient by (dx, dy, dz) cm and scales every layer's CSDA
range by ``range_scale``; it is expressed as a modified problem, so every item
is an ordinary PBProblem:
  lateral shift  -> spot lattice origin (x0 + dx, y0 + dy)
  depth shift    -> dose grid depth origin (oz + dz)
  range scaling  -> layer energies E(R0 * s); per-layer IDD and sigma rebuilt by
                    the FP64 look-up-table builder on the device.
The prescription and voxel weights stay attached to the patient grid.
The standard set has 21 scenarios: {nominal, +-dx, +-dy, +-dz} x {1, 1 -+ r}.

``optimize`` runs inverse planning on a batch: projected-gradient descent of each
field's summed or worst-case objective over its scenarios, one graph replay per
iteration, with the weight update in the HIP kernel pbh_weight_step.

Launch parameters (forward layer chunks, adjoint layer split, one forward
launch per spot-chunk width or one at the widest, item order) are chosen by ``autotune``:
every candidate is built and timed on the device, and the winner is cached per
batch signature in a JSON file.
"""
from __future__ import annotations

import ctypes as C
import hashlib
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib, physics
from .model import DoseGrid, PBProblem, SpotLattice


@dataclass(frozen=True)
class Scenario:
    dx: float = 0.0
    dy: float = 0.0
    dz: float = 0.0
    range_scale: float = 1.0

    @property
    def name(self) -> str:
        return f"dx{self.dx:+.2f}_dy{self.dy:+.2f}_dz{self.dz:+.2f}_r{self.range_scale:.3f}"


def standard_scenarios(setup_cm: float = 0.3, range_frac: float = 0.035) -> List[Scenario]:
    """{nominal, +-x, +-y, +-z setup} x {R0, R0 (1 - r), R0 (1 + r)}: 21 scenarios."""
    setups = [(0, 0, 0), (setup_cm, 0, 0), (-setup_cm, 0, 0), (0, setup_cm, 0),
              (0, -setup_cm, 0), (0, 0, setup_cm), (0, 0, -setup_cm)]
    return [Scenario(float(dx), float(dy), float(dz), float(s))
            for s in (1.0, 1.0 - range_frac, 1.0 + range_frac) for dx, dy, dz in setups]


@dataclass
class Field:
    """A field on a water grid: lattice with energies (MeV) and weights (L, nsy, nsx)."""
    grid: DoseGrid
    lattice: SpotLattice


def robust_fields(n: int = 128, voxel: float = 0.2, spacing: float = 0.5) -> List[Field]:
    """The three fields of the robust-scenario workload on an n^3 water grid:
    31 x 31 spots x 25 layers, 41 x 41 x 40 and 27 x 35 x 45 (nsx x nsy x L), 5 mm
    pitch, energies linspace(70, 180, L) MeV, weights lognormal(0, 0.5) from
    seeds 0, 1, 2."""
    grid = DoseGrid.centered(n, voxel)
    out = []
    for seed, (nsx, nsy, L) in enumerate([(31, 31, 25), (41, 41, 40), (27, 35, 45)]):
        E = np.linspace(70.0, 180.0, L)
        w = np.random.default_rng(seed).lognormal(0.0, 0.5, (L, nsy, nsx))
        out.append(Field(grid, SpotLattice(nsx, nsy, L, spacing, spacing, energies=E, weights=w)))
    return out


def scenario_energies(energies, range_scale: float, device="cuda") -> np.ndarray:
    """Layer energies after scaling every CSDA range by range_scale (device tables)."""
    E0 = np.asarray(energies, dtype=np.float64)
    if range_scale == 1.0:
        return E0.copy()
    return physics.energy_from_range(physics.csda_range(E0, device) * range_scale, device)


def scenario_problem(field: Field, sc: Scenario, device="cuda", sigma_x0: float = 0.30,
                     idd_model=physics.DEFAULT_IDD):
    """(PBProblem, FP64 device idd, FP64 device sigma) of one field in one scenario."""
    g, lat = field.grid, field.lattice
    grid = DoseGrid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ox, g.oy, g.oz + sc.dz)
    E = scenario_energies(lat.energies, sc.range_scale, device)
    lat_s = SpotLattice(lat.n_sx, lat.n_sy, lat.n_layers, lat.dx_spot, lat.dy_spot,
                        lat.x0 + sc.dx, lat.y0 + sc.dy, energies=E, weights=lat.weights)
    idd, sig, _ = physics.build_luts(grid.nz, grid.oz, grid.dz, E, idd_model=idd_model,
                                     sigma_x0=sigma_x0, device=device)
    return PBProblem(grid, lat_s, idd.cpu().numpy(), sig.cpu().numpy()), idd, sig


def central_cube(grid: DoseGrid, half: float = 3.0, dose_value: float = 2.0, w_in: float = 1.0,
                 w_out: float = 0.1, device="cuda"):
    from .planning import central_cube as cc
    return cc(grid, half, dose_value, w_in, w_out, device)


# ---------------------------------------------------------------------------
# tuning
# ---------------------------------------------------------------------------
DEFAULT_TUNING = {"fwd_chunk": 0, "adj_split": 0, "group": 1, "order": 1, "tile_rows": 0, "persistent": 0}

# forward: automatic layer chunks per item, or one chunk (all layers of a tile
# in one workgroup); 32- or 64-row dose tiles; one workgroup per work unit, or a
# persistent launch claiming units from a counter; adjoint: workgroups per depth
# slice; items in the given order or largest first
CANDIDATES = [{"fwd_chunk": f, "adj_split": a, "group": g, "order": o, "tile_rows": r, "persistent": p}
              for f in (0, 4096) for a in (2, 4, 8) for g in (0, 1) for o in (0, 1)
              for r in (32, 64) for p in (0, 1)]


from .tuning import cache_path as _cache_path, load_cache as _load_cache, save_cache as _save_cache  # noqa: E402


def tuning_struct(t: dict) -> _lib.PbhBatchTuning:
    return _lib.PbhBatchTuning(int(t["fwd_chunk"]), int(t["adj_split"]), int(t["group"]), int(t["order"]),
                               int(t.get("tile_rows", 0)), int(t.get("persistent", 0)))


# ---------------------------------------------------------------------------
# the batch
# ---------------------------------------------------------------------------
class RobustBatch:
    """Fields x scenarios with per-item forward dose, objective and adjoint.

    presc / vox_w: per-field prescription and voxel weights (float32 device
    tensors of the grid shape); default: 2 Gy in the central 6 cm cube, weights
    1 inside / 0.1 outside. ``objective=False`` runs the forward doses only.
    ``tuning``: a dict like DEFAULT_TUNING, or "auto" (measured, cached).
    Outputs (device tensors, refreshed by every run): ``dose[k]``,
    ``grad_w[k]`` (dJ_k/dw of the item's field weights) and ``J[k]`` (FP64) for
    item k = field * n_scenarios + scenario; ``items[k] = (field, scenario)``.
    """

    def __init__(self, fields: Sequence[Field], scenarios: Sequence[Scenario], *,
                 precision: str = "bf16x3", presc=None, vox_w=None, objective: bool = True,
                 lut_grads: bool = False, tuning="auto", device="cuda", sigma_x0: float = 0.30):
        self.dev = _lib.require_cuda(device)
        if precision not in _lib.PREC:
            raise ValueError(f"unknown precision {precision!r}")
        self.precision = precision
        self.fields, self.scenarios = list(fields), list(scenarios)
        if not self.fields or not self.scenarios:
            raise ValueError("a batch needs at least one field and one scenario")
        self.w, self.Dp, self.wv = [], [], []
        for fi, f in enumerate(self.fields):
            w = torch.as_tensor(f.lattice.ensure_weights(), dtype=torch.float32, device=self.dev)
            self.w.append(w.contiguous())
            if not objective:
                self.Dp.append(None)
                self.wv.append(None)
            elif presc is None:
                Dp, wv = central_cube(f.grid, device=self.dev)
                self.Dp.append(Dp)
                self.wv.append(wv)
            else:
                self.Dp.append(presc[fi].to(self.dev, torch.float32).contiguous())
                self.wv.append(None if vox_w is None else vox_w[fi].to(self.dev, torch.float32).contiguous())
        self.objective = objective
        self.items, self.problems, self.idd, self.sigma = [], [], [], []
        for fi, f in enumerate(self.fields):
            for si, sc in enumerate(self.scenarios):
                prob, idd, sig = scenario_problem(f, sc, self.dev, sigma_x0)
                self.items.append((fi, si))
                self.problems.append(prob)
                self.idd.append(idd)
                self.sigma.append(sig)
        n = len(self.items)
        self.dose = [torch.empty(p.grid.shape, dtype=torch.float32, device=self.dev) for p in self.problems]
        self.grad_w = [torch.empty_like(self.w[fi]) for fi, _ in self.items] if objective else None
        self.grad_idd = [torch.empty_like(x) for x in self.idd] if (objective and lut_grads) else None
        self.grad_sigma = [torch.empty_like(x) for x in self.sigma] if (objective and lut_grads) else None
        self.J = torch.zeros(n, dtype=torch.float64, device=self.dev)
        self._desc = self._descriptors()
        if tuning == "auto":
            tuning = self.autotune()
        self.tuning = dict(DEFAULT_TUNING if tuning is None else tuning)
        self._handle, self._ws = self._create(self.tuning)
        self.graph = None

    # -- native plumbing --
    def _descriptors(self):
        arr = (_lib.PbhBatchItem * len(self.items))()
        Jp = self.J.data_ptr()
        for k, (fi, _) in enumerate(self.items):
            p = self.problems[k]
            it = arr[k]
            it.grid = _lib.grid_struct(p.grid)
            it.lattice = _lib.lattice_struct(p.lattice)
            it.w = self.w[fi].data_ptr()
            it.idd = self.idd[k].data_ptr()
            it.sigma = self.sigma[k].data_ptr()
            it.dose = self.dose[k].data_ptr()
            if self.objective:
                it.presc = self.Dp[fi].data_ptr()
                it.vox_w = self.wv[fi].data_ptr() if self.wv[fi] is not None else None
                it.j_out = Jp + 8 * k
                it.grad_w = self.grad_w[k].data_ptr()
                if self.grad_idd is not None:
                    it.grad_idd = self.grad_idd[k].data_ptr()
                    it.grad_sigma = self.grad_sigma[k].data_ptr()
        return arr

    def _create(self, tuning: dict):
        L = _lib.lib()
        prec = _lib.PREC[self.precision]
        t = tuning_struct(tuning)
        nbytes = L.pbh_batch_workspace_bytes(self._desc, len(self.items), prec, C.byref(t))
        if nbytes == 0:
            raise ValueError(f"invalid batch or tuning {tuning}")
        ws = _lib.workspace(nbytes, self.dev)
        h = C.c_void_p()
        _lib.check(L.pbh_batch_create(self._desc, len(self.items), prec, C.byref(t), _lib.ptr(ws), ws.numel(),
                                      _lib.stream_of(self.dev), C.byref(h)), "pbh_batch_create")
        return h, ws

    def _run(self, handle) -> None:
        _lib.check(_lib.lib().pbh_batch_forward_adjoint(handle, _lib.stream_of(self.dev)),
                   "pbh_batch_forward_adjoint")

    def close(self) -> None:
        if getattr(self, "_handle", None):
            _lib.lib().pbh_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def launch_count(self) -> int:
        return int(_lib.lib().pbh_batch_launch_count(self._handle))

    @property
    def workspace_bytes(self) -> int:
        return int(self._ws.numel())

    def signature(self) -> str:
        """Key of this batch's shape for the tuning cache."""
        from .tuning import device_key
        parts = [self.precision, str(self.objective), device_key(self.dev)]
        for p in self.problems:
            g, lat = p.grid, p.lattice
            parts.append(f"{g.nx}x{g.ny}x{g.nz}@{g.dx:g},{g.dy:g}:{lat.n_sx}x{lat.n_sy}x{lat.n_layers}"
                         f"@{lat.dx_spot:g},{lat.dy_spot:g}")
        return hashlib.sha1("|".join(parts).encode()).hexdigest()

    # -- running --
    def run_eager(self) -> torch.Tensor:
        """All items, launched directly on the current stream. Returns J."""
        self._run(self._handle)
        return self.J

    def capture(self):
        """Records the whole batch into one graph (after one eager warm-up run)."""
        self.run_eager()
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._run(self._handle)
        torch.cuda.synchronize(self.dev)
        self.graph = g
        return g

    def run(self) -> torch.Tensor:
        """Replays the captured batch (one graph launch). Returns J."""
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return self.J

    # -- tuning --
    def autotune(self, candidates=None, reps: int = 3, use_cache: bool = True) -> dict:
        """Measured best launch parameters for this batch (cached per signature)."""
        key = self.signature()
        cache = _load_cache() if use_cache else {}
        if key in cache:
            return dict(cache[key]["tuning"])
        results = []
        for cand in (candidates or CANDIDATES):
            h, ws = self._create(cand)
            try:
                self._run(h)
                torch.cuda.synchronize(self.dev)
                s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s.record()
                for _ in range(reps):
                    self._run(h)
                e.record()
                torch.cuda.synchronize(self.dev)
                results.append((s.elapsed_time(e) / reps, cand))
            finally:
                _lib.lib().pbh_batch_destroy(h)
                del ws
        best_ms, best = min(results, key=lambda r: r[0])
        if use_cache:
            cache = _load_cache()
            cache[key] = {"tuning": best, "ms": best_ms,
                          "timings": [[ms, c] for ms, c in results]}
            _save_cache(cache)
        self.tuning_results = results
        return dict(best)

    # -- results --
    def field_gradient(self, reduce: str = "sum") -> List[torch.Tensor]:
        """Per-field dJ/dw over its scenarios: 'sum' of the scenario gradients, or
        'worst' (the gradient of the scenario with the largest J)."""
        if not self.objective:
            raise ValueError("the batch was built without an objective")
        out = []
        Jh = self.J.cpu().numpy()
        for fi in range(len(self.fields)):
            ks = [k for k, (f, _) in enumerate(self.items) if f == fi]
            if reduce == "sum":
                g = self.grad_w[ks[0]].clone()
                for k in ks[1:]:
                    g += self.grad_w[k]
            elif reduce == "worst":
                g = self.grad_w[max(ks, key=lambda k: Jh[k])].clone()
            else:
                raise ValueError(f"unknown reduction {reduce!r}")
            out.append(g)
        return out

    def flops(self) -> float:
        """Algorithmic FLOP of the forward doses of all items (adjoint: the same again)."""
        tot = 0.0
        for p in self.problems:
            g, lat = p.grid, p.lattice
            a = lat.n_sy * lat.n_sx * g.ny + g.nx * lat.n_sx * g.ny
            b = lat.n_sy * lat.n_sx * g.nx + g.ny * lat.n_sy * g.nx
            tot += 2.0 * lat.n_layers * g.nz * min(a, b)
        return tot


# ---------------------------------------------------------------------------
# inverse planning on a robust batch
# ---------------------------------------------------------------------------
ROBUST_MODES = {"sum": 0, "worst": 1}


def weight_step(w_old: torch.Tensor, w_new: torch.Tensor, grads: Sequence[torch.Tensor],
                J: Optional[torch.Tensor], step: float, mode: str = "sum", w_min: float = 0.0,
                w_max: float = float("inf")) -> None:
    """w_new := clamp(w_old - step * g, w_min, w_max) on the device (pbh_weight_step);
    g = sum of ``grads`` (mode 'sum') or the gradient of the scenario with the
    largest J (mode 'worst')."""
    dev = w_old.device
    ptrs = torch.tensor([g.data_ptr() for g in grads], dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().pbh_weight_step(_lib.ptr(w_old), _lib.ptr(w_new), _lib.ptr(ptrs), _lib.ptr(J),
                                          len(grads), w_old.numel(), ROBUST_MODES[mode], float(step),
                                          float(w_min), float(w_max), _lib.stream_of(dev)),
               "pbh_weight_step")


def optimize(batch: "RobustBatch", iters: int = 20, mode: str = "sum", w_min: float = 0.0,
             w_max: float = float("inf"), step0: Optional[float] = None, grow: float = 1.5,
             shrink: float = 0.5, use_graph: bool = True) -> dict:
    """Projected-gradient descent of every field's robust objective over its scenarios:
    F_f(w) = sum_s J_fs(w) ('sum') or max_s J_fs(w) ('worst', min-max with the worst
    scenario's gradient as a subgradient), with w_min <= w <= w_max.

    Each iteration takes one trial step per field from the gradients of its last
    accepted point, evaluates all fields and scenarios with one batch run (graph
    replay) and keeps a field's step if its F decreased (step size x grow), else
    restores the field's weights (step size x shrink). F is non-increasing.
    Returns {'F': per-iteration list of per-field F, 'accepted': ..., 'steps': ...}."""
    if not batch.objective:
        raise ValueError("the batch was built without an objective")
    if mode not in ROBUST_MODES:
        raise ValueError(f"unknown robust mode {mode!r}")
    run = batch.run if use_graph else batch.run_eager
    nS, nF = len(batch.scenarios), len(batch.fields)

    def robust(Jh):
        per = Jh.reshape(nF, nS)
        return per.sum(axis=1) if mode == "sum" else per.max(axis=1)

    run()
    Jh = batch.J.cpu().numpy()
    F = robust(Jh)
    w_acc = [w.clone() for w in batch.w]
    g_acc = [g.clone() for g in batch.grad_w]
    J_acc = batch.J.clone()
    if step0 is None:
        steps = []
        for f in range(nF):
            gs = sum(g_acc[f * nS + s].abs().max().item() for s in range(nS))
            steps.append(0.1 * w_acc[f].abs().max().item() / gs if gs > 0 else 1.0)
    else:
        steps = [float(step0)] * nF
    hist = {"F": [F.tolist()], "accepted": [], "steps": []}
    for _ in range(iters):
        for f in range(nF):
            weight_step(w_acc[f], batch.w[f], g_acc[f * nS:(f + 1) * nS], J_acc[f * nS:(f + 1) * nS],
                        steps[f], mode, w_min, w_max)
        run()
        Jh = batch.J.cpu().numpy()
        Fn = robust(Jh)
        acc = []
        for f in range(nF):
            if Fn[f] < F[f]:
                F[f] = Fn[f]
                w_acc[f].copy_(batch.w[f])
                for s in range(nS):
                    g_acc[f * nS + s].copy_(batch.grad_w[f * nS + s])
                J_acc[f * nS:(f + 1) * nS].copy_(batch.J[f * nS:(f + 1) * nS])
                steps[f] *= grow
                acc.append(True)
            else:
                batch.w[f].copy_(w_acc[f])
                steps[f] *= shrink
                acc.append(False)
        hist["F"].append(F.tolist())
        hist["accepted"].append(acc)
        hist["steps"].append(list(steps))
    # leave the batch's weights at the best accepted point
    for f in range(nF):
        batch.w[f].copy_(w_acc[f])
    torch.cuda.synchronize(batch.dev)
    return hist
