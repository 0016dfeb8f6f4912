"""Proton physics look-up tables, computed in FP64 on the GPU.

Mirrors the names the reference's field builders and scripts use
(physics.py:42-65, :111-139, :292-387) with the arithmetic running in the
HIP kernels of libpbdose_hip.so (``pbh_range_table``, ``pbh_interp_table``,
``pbh_build_lut``). Host-side values returned as NumPy arrays are copies of
device results.

Depth grids passed to the LUT builders must be uniform (z_k = z0 + k dz), which
is what every ``DoseGrid`` produces (model.py:48-52).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

# physics.py:42-65 (model constants re-exported for callers that read them)
M_P_C2 = 938.27208816
M_E_C2 = 0.51099895000
K_BB = 0.307075
WATER_Z_OVER_A = 10.0 / 18.01528
WATER_I_EV = 75.0
WATER_X0 = 36.08
E_LO, E_HI = 0.5, 400.0
A_R, P_R = 0.0022, 1.77
BOHR_COEF = K_BB * WATER_Z_OVER_A * 1.0 * M_E_C2

_STRAGGLING = {"bohr": 0, "icru49": 1, "none": 2}


@dataclass
class IDDModel:
    """IDD construction parameters (physics.py:292-322)."""
    straggling_model: str = "bohr"
    straggling_coef: float = 0.012
    straggling_exp: float = 0.935
    energy_spread: float = 0.0033
    attenuation_length: float = 150.0
    n_fine: int = 6000
    e_cut: float = 0.05          # declared by the reference, unused there and here
    # False: one range spread sigma_R for the whole CSDA curve (physics.py:370-373).
    # True: the CSDA dose binned at depth z' is spread with hypot(sigma_B(z'), sigma_E),
    # where sigma_B(z') is the Bohr straggling accumulated up to z' -- the model fix the
    # reference lists first (README.md:180-184). Needs straggling_model "bohr" or "none".
    depth_straggling: bool = False


DEFAULT_IDD = IDDModel()


def beam_model(idd_model: IDDModel = DEFAULT_IDD, sigma_x0: float = 0.30,
               sigma_theta0: float = 0.0, rho0: float = 0.0,
               n_sub: int = 4, max_step: float = 0.02) -> _lib.PbhBeamModel:
    if idd_model.straggling_model not in _STRAGGLING:
        raise ValueError(f"unknown straggling model: {idd_model.straggling_model}")
    if idd_model.depth_straggling and idd_model.straggling_model == "icru49":
        raise ValueError("depth-dependent straggling is defined for the Bohr integral only")
    return _lib.PbhBeamModel(float(sigma_x0), float(sigma_theta0), float(rho0),
                             float(idd_model.energy_spread), float(idd_model.attenuation_length),
                             float(max_step), int(idd_model.n_fine), int(n_sub),
                             _STRAGGLING[idd_model.straggling_model],
                             int(bool(idd_model.depth_straggling)),
                             float(idd_model.straggling_coef), float(idd_model.straggling_exp))


_TABLES: dict = {}


def range_table(device="cuda"):
    """(E, R) CSDA range table as FP64 device tensors, built once per device.
    physics.py:111-127"""
    dev = _lib.require_cuda(device)
    key = (dev.type, dev.index)
    if key not in _TABLES:
        L = _lib.lib()
        n = L.pbh_range_table_size()
        E = torch.empty(n, dtype=torch.float64, device=dev)
        R = torch.empty(n, dtype=torch.float64, device=dev)
        _lib.check(L.pbh_range_table(_lib.ptr(E), _lib.ptr(R), _lib.stream_of(dev)),
                   "pbh_range_table")
        _TABLES[key] = (E, R)
    return _TABLES[key]


def _interp(x, xp, fp, dev):
    xt = torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev).reshape(-1).contiguous()
    out = torch.empty_like(xt)
    _lib.check(_lib.lib().pbh_interp_table(_lib.ptr(xp), _lib.ptr(fp), xp.numel(), _lib.ptr(xt),
                                           _lib.ptr(out), xt.numel(), _lib.stream_of(dev)),
               "pbh_interp_table")
    return out


def csda_range(E_kin, device="cuda"):
    """CSDA range [cm] of kinetic energy E [MeV]. physics.py:130-133"""
    dev = _lib.require_cuda(device)
    E, R = range_table(dev)
    shape = np.shape(E_kin)
    return _interp(E_kin, E, R, dev).cpu().numpy().reshape(shape)


def energy_from_range(R_cm, device="cuda"):
    """Energy [MeV] whose CSDA range is R (clipped to the table). physics.py:136-139"""
    dev = _lib.require_cuda(device)
    E, R = range_table(dev)
    shape = np.shape(R_cm)
    return _interp(R_cm, R, E, dev).cpu().numpy().reshape(shape)


def _uniform_axis(z):
    """(nz, z0, dz) of a uniform depth grid; a single depth has nz = 1 (dz unused)."""
    z = np.asarray(z, dtype=np.float64).reshape(-1)
    if z.size == 0:
        raise ValueError("depth grid needs at least one point")
    if z.size == 1:
        return 1, float(z[0]), 1.0
    dz = (z[-1] - z[0]) / (z.size - 1)
    if not np.allclose(z, z[0] + np.arange(z.size) * dz, rtol=0, atol=1e-12 * max(1.0, abs(z[-1]))):
        raise ValueError("depth grid must be uniform (z0 + k dz)")
    return z.size, float(z[0]), float(dz)


def _energies(energies, dev):
    if torch.is_tensor(energies):
        return energies.to(device=dev, dtype=torch.float64).reshape(-1).contiguous()
    return torch.as_tensor(np.asarray(energies, dtype=np.float64), device=dev).reshape(-1).contiguous()


def build_luts(nz: int, z0: float, dz: float, energies, *, idd_model: IDDModel = DEFAULT_IDD,
               sigma_x0: float = 0.30, sigma_theta0: float = 0.0, rho0: float = 0.0,
               device="cuda", denergies=None):
    """Per-layer (idd, sigma, info) on depths z0 + k dz as FP64 device tensors.

    idd (L, nz) is the fluence-normalised IDD (idd_pristine(normalize=False),
    physics.py:328-387); sigma (L, nz) is fermi_eyges_moments(...)['sigma_x']
    (physics.py:221-286), with a single depth (nz = 1) integrated over [0, z0]
    as physics.py:247-249 does; info (L, 4) = R0, sigma_R, depth and value of
    the IDD maximum. One device call for all layers (model.py:225-233 loops on
    the host), reading the range table built once per device.

    With ``denergies`` (one value per layer) the call also returns the tables'
    derivatives along that direction of the layer energies,
    (idd, sigma, info, didd, dsigma), from pbh_build_lut_tangent: the same
    tables bit for bit, with forward-mode FP64 derivatives of every kernel.
    """
    dev = _lib.require_cuda(device)
    L = _lib.lib()
    en = _energies(energies, dev)
    nl = en.numel()
    m = beam_model(idd_model, sigma_x0, sigma_theta0, rho0)
    tangent = denergies is not None
    query = L.pbh_lut_tangent_workspace_bytes if tangent else L.pbh_lut_workspace_bytes
    nbytes = query(m, int(nz), float(z0), float(dz), nl)
    if nbytes == 0:
        raise ValueError("invalid beam model or grid for the LUT builder")
    ws = _lib.workspace(nbytes, dev)
    tE, tR = range_table(dev)
    idd = torch.empty((nl, nz), dtype=torch.float64, device=dev)
    sig = torch.empty((nl, nz), dtype=torch.float64, device=dev)
    info = torch.empty((nl, 4), dtype=torch.float64, device=dev)
    if not tangent:
        _lib.check(L.pbh_build_lut(m, int(nz), float(z0), float(dz), nl, _lib.ptr(en), _lib.ptr(tE),
                                   _lib.ptr(tR), _lib.ptr(idd), _lib.ptr(sig), _lib.ptr(info),
                                   _lib.ptr(ws), ws.numel(), _lib.stream_of(dev)), "pbh_build_lut")
        return idd, sig, info
    den = _energies(denergies, dev)
    if den.numel() != nl:
        raise ValueError("denergies needs one value per layer")
    didd = torch.empty_like(idd)
    dsig = torch.empty_like(sig)
    _lib.check(L.pbh_build_lut_tangent(m, int(nz), float(z0), float(dz), nl, _lib.ptr(en), _lib.ptr(den),
                                       _lib.ptr(tE), _lib.ptr(tR), _lib.ptr(idd), _lib.ptr(sig),
                                       _lib.ptr(didd), _lib.ptr(dsig), _lib.ptr(info), _lib.ptr(ws),
                                       ws.numel(), _lib.stream_of(dev)), "pbh_build_lut_tangent")
    return idd, sig, info, didd, dsig


def energy_from_range_slope(R_cm, device="cuda"):
    """(E(R), dE/dR) as FP64 device tensors: energy_from_range (physics.py:136-139)
    and the slope of the table segment it interpolates (0 where R is clamped)."""
    dev = _lib.require_cuda(device)
    E, R = range_table(dev)
    x = _energies(R_cm, dev)
    out, dout = torch.empty_like(x), torch.empty_like(x)
    _lib.check(_lib.lib().pbh_interp_table_slope(_lib.ptr(R), _lib.ptr(E), R.numel(), _lib.ptr(x),
                                                 _lib.ptr(out), _lib.ptr(dout), x.numel(),
                                                 _lib.stream_of(dev)), "pbh_interp_table_slope")
    return out, dout


def idd_pristine(z, E0, model: IDDModel = DEFAULT_IDD, normalize: bool = True, device="cuda"):
    """(idd, sigma_R) of one layer on a uniform depth grid. physics.py:328-387"""
    nz, z0, dz = _uniform_axis(z)
    idd, _, info = build_luts(nz, z0, dz, [float(E0)], idd_model=model, device=device)
    out = idd[0].cpu().numpy()
    if normalize:
        pk = float(out.max())
        out = out / (pk if pk > 0 else 1.0)
    return out, float(info[0, 1])


def fermi_eyges_sigma(z, E0, *, sigma_x0=0.0, sigma_theta0=0.0, rho0=0.0, device="cuda"):
    """Lateral rms width sigma_x(z) [cm] of one layer on a uniform depth grid
    (fermi_eyges_moments(...)['sigma_x'], physics.py:221-286)."""
    nz, z0, dz = _uniform_axis(z)
    _, sig, _ = build_luts(nz, z0, dz, [float(E0)], sigma_x0=sigma_x0,
                           sigma_theta0=sigma_theta0, rho0=rho0, device=device)
    return sig[0].cpu().numpy()


def fermi_eyges_moments(z, E0, *, sigma_x0=0.0, sigma_theta0=0.0, rho0=0.0, device="cuda"):
    """{'sigma_x': ...} on a uniform depth grid (the field the dose path uses)."""
    return {"sigma_x": fermi_eyges_sigma(z, E0, sigma_x0=sigma_x0, sigma_theta0=sigma_theta0,
                                         rho0=rho0, device=device)}
