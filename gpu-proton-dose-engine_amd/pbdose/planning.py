"""Inverse-planning layer: weighted-quadratic objective, hand-written adjoint and a
``torch.autograd.Function`` over spot weights and per-layer range shifts.

    J(w, delta) = sum_v wv_v (D_v(w, delta) - Dp_v)^2

D is the forward dose of ``engines.DeviceProblem`` (HIP kernels). A per-layer
range shift delta_l (cm) moves layer l's CSDA range R0_l -> R0_l + delta_l: the
layer is rebuilt at energy E(R0_l + delta_l) by the on-device FP64 LUT builder.
The gradients come from the adjoint kernels:

    dJ/dw      pbh_adjoint (deterministic sum over depth)
    dJ/ddelta  pbh_adjoint's dJ/dIDD and dJ/dsigma, contracted over depth by
               pbh_lut_chain with the exact FP64 derivatives of the look-up tables
               with respect to delta: pbh_build_lut_tangent differentiates every
               kernel of the LUT builder in forward mode along dE/dR at R0 + delta
               (pbh_interp_table_slope), in the same pass that builds the tables.

No counterpart exists in the reference (its engines are forward-only,
engines.py:208-283); the objective follows the workload definition used for
the dose benchmarks: D_presc inside a central cube, voxel weights 1 inside and
0.1 outside.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _lib, physics
from .engines import DeviceProblem
from .model import DoseGrid, PBProblem, SpotLattice


class FieldModel:
    """A field whose per-layer tables are rebuilt on the device for range shifts.

    grid, lattice (with energies), beam/IDD model. ``luts(delta)`` returns FP64
    (idd, sigma) device tensors for the energies E(R0 + delta);
    ``luts_and_derivatives(delta)`` adds their exact derivatives in delta."""

    def __init__(self, grid: DoseGrid, lattice: SpotLattice, *, idd_model=physics.DEFAULT_IDD,
                 sigma_x0: float = 0.30, sigma_theta0: float = 0.0, rho0: float = 0.0,
                 device="cuda", lut_step: float = 1e-3):
        if lattice.energies is None:
            raise ValueError("lattice.energies is required")
        self.dev = _lib.require_cuda(device)
        self.grid, self.lattice = grid, lattice
        self.kw = dict(idd_model=idd_model, sigma_x0=sigma_x0, sigma_theta0=sigma_theta0, rho0=rho0)
        self.lut_step = float(lut_step)
        self.E0 = torch.as_tensor(np.asarray(lattice.energies, dtype=np.float64), device=self.dev)
        E, R = physics.range_table(self.dev)
        self.R0 = self._interp(E, R, self.E0)
        idd, sig = self.luts(None)
        self.problem = PBProblem(grid, lattice, idd.cpu().numpy(), sig.cpu().numpy())
        self.dp = DeviceProblem(self.problem, self.dev, idd=idd, sigma=sig)

    def _interp(self, xp, fp, x):
        x = x.contiguous()
        out = torch.empty_like(x)
        _lib.check(_lib.lib().pbh_interp_table(_lib.ptr(xp), _lib.ptr(fp), xp.numel(), _lib.ptr(x),
                                               _lib.ptr(out), x.numel(), _lib.stream_of(self.dev)),
                   "pbh_interp_table")
        return out

    def _delta(self, delta):
        return torch.zeros_like(self.R0) if delta is None else delta.to(self.dev, torch.float64)

    def energies(self, delta: Optional[torch.Tensor]) -> torch.Tensor:
        if delta is None:
            return self.E0
        E, R = physics.range_table(self.dev)
        return self._interp(R, E, self.R0 + self._delta(delta))

    def luts(self, delta: Optional[torch.Tensor]):
        g = self.grid
        idd, sig, _ = physics.build_luts(g.nz, g.oz, g.dz, self.energies(delta), device=self.dev,
                                         **self.kw)
        return idd, sig

    def luts_and_derivatives(self, delta: Optional[torch.Tensor]):
        """(idd, sigma, d idd/d delta_l, d sigma/d delta_l), all (L, nz) FP64: one
        tangent build at E(R0 + delta) with direction dE/dR there."""
        g = self.grid
        E, dE = physics.energy_from_range_slope(self.R0 + self._delta(delta), device=self.dev)
        if delta is None:
            E = self.E0
        idd, sig, _, didd, dsig = physics.build_luts(g.nz, g.oz, g.dz, E, device=self.dev,
                                                     denergies=dE, **self.kw)
        return idd, sig, didd, dsig

    def lut_derivatives(self, delta: Optional[torch.Tensor], method: str = "analytic"):
        """d(idd)/d(delta_l), d(sigma)/d(delta_l) (L, nz) FP64. ``method``
        'analytic' (forward-mode derivative of the LUT builder) or 'fd' (central
        difference with step ``lut_step``, two more LUT builds)."""
        if method == "analytic":
            return self.luts_and_derivatives(delta)[2:]
        if method != "fd":
            raise ValueError(f"method {method!r}: 'analytic' or 'fd'")
        h = self.lut_step
        d0 = self._delta(delta)
        ip, sp = self.luts(d0 + h)
        im, sm = self.luts(d0 - h)
        return (ip - im) / (2.0 * h), (sp - sm) / (2.0 * h)


def objective(dp: DeviceProblem, dose: torch.Tensor, presc: torch.Tensor,
              vox_w: Optional[torch.Tensor] = None, want_residual: bool = True):
    """(J, R): J = sum wv (D - Dp)^2 (FP64 scalar tensor), R = dJ/dD = 2 wv (D - Dp)."""
    L = _lib.lib()
    R = torch.empty_like(dose) if want_residual else None
    j = torch.empty(1, dtype=torch.float64, device=dp.dev)
    ws = dp.workspace("obj", 0)
    _lib.check(L.pbh_objective(dp.grid_s, _lib.ptr(dose), _lib.ptr(presc), _lib.ptr(vox_w), _lib.ptr(R),
                               _lib.ptr(j), _lib.ptr(ws), ws.numel(), _lib.stream_of(dp.dev)),
               "pbh_objective")
    return j[0], R


# Arithmetic of the weight adjoint: FP32, or 16-bit MFMA operands (BF16 or FP16) with
# single or split (x3) products and FP32 accumulation
ADJOINT_FORMS = ("fp32", "bf16", "bf16x3", "fp16", "fp16x3")


def adjoint(dp: DeviceProblem, R: torch.Tensor, w: Optional[torch.Tensor] = None,
            lut_grads: bool = False, split: int = 0, precision: str = "fp32"):
    """dJ/dw (L, nsy, nsx) FP32 for the dose cotangent R; with ``lut_grads`` also
    dJ/dIDD and dJ/dsigma (L, nz) FP64 (FP32 arithmetic only). ``split``:
    workgroups per depth slice sharing its layers (0 = automatic).
    ``precision``: 'fp32', or 'bf16' / 'bf16x3' / 'fp16' / 'fp16x3' (16-bit MFMA, dJ/dw
    only; FP16 scales R by a power of two per tile group, exactly). Without table
    gradients, ``split`` 0 takes the split that pbh_adjoint_choose_split models fastest
    for this problem's active layers per slice."""
    if precision not in ADJOINT_FORMS:
        raise ValueError(f"adjoint precision {precision!r}: one of {', '.join(ADJOINT_FORMS)}")
    if lut_grads and precision != "fp32":
        raise ValueError("table gradients (lut_grads) are computed in FP32 only")
    L = _lib.lib()
    if R.dtype != torch.float32 or tuple(R.shape) != tuple(dp.shape):
        raise ValueError("R must be float32 with the dose shape")
    R = R.contiguous()
    wt = dp.w if w is None else w.contiguous()
    gw = torch.empty_like(dp.w)
    gi = torch.empty_like(dp.idd) if lut_grads else None
    gs = torch.empty_like(dp.sigma) if lut_grads else None
    ws = dp.workspace("adj", 0)
    if int(split) == 0 and not lut_grads:    # table gradients take every layer of a slice
        split = dp.auto_split(_lib.PREC[precision])
    _lib.check(L.pbh_adjoint_tuned(dp.grid_s, dp.lat_s, _lib.ptr(wt), _lib.ptr(dp.idd), _lib.ptr(dp.sigma),
                                   _lib.ptr(R), _lib.ptr(gw), _lib.ptr(gi), _lib.ptr(gs), _lib.PREC[precision],
                                   int(split),
                                   _lib.ptr(ws), ws.numel(), _lib.stream_of(dp.dev)), "pbh_adjoint_tuned")
    return (gw, gi, gs) if lut_grads else gw


def objective_gradient(dp: DeviceProblem, dose: torch.Tensor, presc: torch.Tensor,
                       vox_w: Optional[torch.Tensor] = None, w: Optional[torch.Tensor] = None,
                       split: int = 0, precision: str = "fp32"):
    """(J, dJ/dw) of J = sum wv (D - Dp)^2 in one call (pbh_objective_adjoint): the same
    values as ``objective`` followed by ``adjoint``, with one kernel launch fewer."""
    if precision not in ADJOINT_FORMS:
        raise ValueError(f"adjoint precision {precision!r}: one of {', '.join(ADJOINT_FORMS)}")
    for name, t in (("dose", dose), ("presc", presc), ("vox_w", vox_w)):
        if t is not None and (t.dtype != torch.float32 or tuple(t.shape) != tuple(dp.shape) or
                              not t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous float32 tensor with the dose shape")
    L = _lib.lib()
    p = _lib.PREC[precision]
    wt = dp.w if w is None else w.contiguous()
    gw = torch.empty_like(dp.w)
    j = torch.empty(1, dtype=torch.float64, device=dp.dev)
    if int(split) == 0:
        split = dp.auto_split(p)
    ws = dp.workspace("objadj", p)
    _lib.check(L.pbh_objective_adjoint(dp.grid_s, dp.lat_s, _lib.ptr(wt), _lib.ptr(dp.idd), _lib.ptr(dp.sigma),
                                       _lib.ptr(dose), _lib.ptr(presc), _lib.ptr(vox_w), _lib.ptr(gw),
                                       _lib.ptr(j), p, int(split), _lib.ptr(ws), ws.numel(),
                                       _lib.stream_of(dp.dev)), "pbh_objective_adjoint")
    return j[0], gw


def forward_objective_gradient(dp: DeviceProblem, presc: torch.Tensor, vox_w: Optional[torch.Tensor] = None,
                               w: Optional[torch.Tensor] = None, precision: str = "bf16x3",
                               forward_precision: Optional[str] = None, out: Optional[torch.Tensor] = None,
                               split: int = 0, layer_chunk: int = 0, tile_rows: int = 0):
    """One inverse-planning step in one call (pbh_forward_objective_adjoint): the dose
    of w, J = sum wv (D - Dp)^2 and dJ/dw. Returns (D, J, dJ/dw); D and dJ/dw equal
    ``dp.forward`` followed by ``objective_gradient`` bit for bit (J up to the grouping
    of its FP64 sum), with the objective formed inside the forward's combine.
    ``precision`` is the adjoint's (one of ADJOINT_FORMS); the forward's defaults
    to it ('fp32' forward for an 'fp32' adjoint)."""
    if precision not in ADJOINT_FORMS:
        raise ValueError(f"adjoint precision {precision!r}: one of {', '.join(ADJOINT_FORMS)}")
    fwd = forward_precision or precision
    for name, t in (("presc", presc), ("vox_w", vox_w), ("out", out)):
        if t is not None and (t.dtype != torch.float32 or tuple(t.shape) != tuple(dp.shape) or
                              not t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous float32 tensor with the dose shape")
    L = _lib.lib()
    pa, pf = _lib.PREC[precision], _lib.PREC[fwd]
    chunk, rows = dp.shape_of(pf, layer_chunk, tile_rows)
    wt = dp.w if w is None else w.contiguous()
    D = torch.empty(dp.shape, dtype=torch.float32, device=dp.dev) if out is None else out
    gw = torch.empty_like(dp.w)
    j = torch.empty(1, dtype=torch.float64, device=dp.dev)
    if int(split) == 0:
        split = dp.auto_split(pa)
    ws = dp.workspace("fwdobjadj", pf, int(chunk), int(rows), pa)
    _lib.check(L.pbh_forward_objective_adjoint(dp.grid_s, dp.lat_s, _lib.ptr(wt), _lib.ptr(dp.idd),
                                               _lib.ptr(dp.sigma), _lib.ptr(D), _lib.ptr(presc), _lib.ptr(vox_w),
                                               _lib.ptr(gw), _lib.ptr(j), pf, int(chunk), int(rows), pa,
                                               int(split), _lib.ptr(ws), ws.numel(), _lib.stream_of(dp.dev)),
               "pbh_forward_objective_adjoint")
    return D, j[0], gw


def lut_chain(grad_idd, grad_sigma, didd, dsigma) -> torch.Tensor:
    """Per-layer sum over depth of grad_idd * didd + grad_sigma * dsigma (FP64)."""
    Lr, nz = grad_idd.shape
    out = torch.empty(Lr, dtype=torch.float64, device=grad_idd.device)
    _lib.check(_lib.lib().pbh_lut_chain(_lib.ptr(grad_idd), _lib.ptr(grad_sigma), _lib.ptr(didd.contiguous()),
                                        _lib.ptr(dsigma.contiguous()), Lr, nz, _lib.ptr(out),
                                        _lib.stream_of(grad_idd.device)), "pbh_lut_chain")
    return out


class _Dose(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, delta, model: FieldModel, precision: str, adjoint_precision: str):
        dp = model.dp
        ctx.dtab = None
        if delta is not None:
            if ctx.needs_input_grad[1]:
                # the tables and their derivatives in one tangent build
                idd, sig, didd, dsig = model.luts_and_derivatives(delta.detach())
                ctx.dtab = (didd, dsig)
            else:
                idd, sig = model.luts(delta.detach())
            dp = DeviceProblem(model.problem, model.dev, weights=w.detach(), idd=idd, sigma=sig)
        D = dp.forward(precision, w=w.detach().contiguous())
        ctx.model, ctx.dp = model, dp
        ctx.save_for_backward(w, delta if delta is not None else torch.empty(0))
        ctx.has_delta = delta is not None
        ctx.adjoint_precision = adjoint_precision
        return D

    @staticmethod
    def backward(ctx, gD):
        w, delta = ctx.saved_tensors
        need_delta = ctx.has_delta and ctx.needs_input_grad[1]
        # table gradients (for the range shifts) are computed in FP32 only
        prec = "fp32" if need_delta else ctx.adjoint_precision
        out = adjoint(ctx.dp, gD.contiguous().float(), w=w.detach(), lut_grads=need_delta, precision=prec)
        if need_delta:
            gw, gi, gs = out
            didd, dsig = ctx.dtab if ctx.dtab is not None else ctx.model.lut_derivatives(delta.detach())
            gdelta = lut_chain(gi, gs, didd, dsig).to(delta.dtype)
        else:
            gw, gdelta = out, None
        return (gw if ctx.needs_input_grad[0] else None), gdelta, None, None, None


def dose(w: torch.Tensor, model: FieldModel, delta: Optional[torch.Tensor] = None,
         precision: str = "fp32", adjoint_precision: str = "fp32") -> torch.Tensor:
    """Differentiable forward dose D(w, delta) (nz, ny, nx) float32 on the device.
    ``precision``: the forward's arithmetic; ``adjoint_precision``: the backward's for
    dJ/dw (one of ADJOINT_FORMS; the range-shift gradient, which needs the table
    gradients, always runs the FP32 adjoint)."""
    if w.dtype != torch.float32:
        raise ValueError("spot weights must be float32")
    if adjoint_precision not in ADJOINT_FORMS:
        raise ValueError(f"adjoint precision {adjoint_precision!r}: one of {', '.join(ADJOINT_FORMS)}")
    return _Dose.apply(w, delta, model, precision, adjoint_precision)


def central_cube(grid: DoseGrid, half: float = 3.0, dose_value: float = 2.0, w_in: float = 1.0,
                 w_out: float = 0.1, device="cuda"):
    """(Dp, wv) device tensors: dose_value inside the cube |x|, |y| <= half,
    |z - z_mid| <= half, 0 outside; voxel weights w_in inside, w_out outside."""
    x, y, z = grid.x, grid.y, grid.z
    zc = 0.5 * (z[0] + z[-1])
    inside = ((np.abs(z - zc) <= half + 1e-9)[:, None, None]
              & (np.abs(y) <= half + 1e-9)[None, :, None]
              & (np.abs(x) <= half + 1e-9)[None, None, :])
    dev = _lib.require_cuda(device)
    Dp = torch.as_tensor(np.where(inside, dose_value, 0.0), dtype=torch.float32, device=dev)
    wv = torch.as_tensor(np.where(inside, w_in, w_out), dtype=torch.float32, device=dev)
    return Dp.contiguous(), wv.contiguous()
