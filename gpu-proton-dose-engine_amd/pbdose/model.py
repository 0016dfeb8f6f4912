"""Problem model: voxel grid, spot lattice, pencil-beam problem, field builders.

Same types, fields and conventions as the reference (model.py:27-330): dose
layout (nz, ny, nx) with x fastest, voxel centres o + k d, centred spot lattice
x0 + (i - (n-1)/2) ds, weights[l, j, i]. The field builders compute every
layer's IDD and sigma in one FP64 device call (``physics.build_luts``) instead
of the reference's host loop over layers (model.py:225-233, :312-320).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import physics as ph


@dataclass
class DoseGrid:
    """Regular voxel grid (model.py:27-79)."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    ox: float = 0.0
    oy: float = 0.0
    oz: float = 0.0

    @property
    def shape(self) -> tuple:
        return (self.nz, self.ny, self.nx)

    @property
    def n_vox(self) -> int:
        return self.nx * self.ny * self.nz

    def axis(self, name: str) -> np.ndarray:
        n, d, o = {"x": (self.nx, self.dx, self.ox), "y": (self.ny, self.dy, self.oy),
                   "z": (self.nz, self.dz, self.oz)}[name]
        return o + np.arange(n) * d

    @property
    def x(self) -> np.ndarray:
        return self.axis("x")

    @property
    def y(self) -> np.ndarray:
        return self.axis("y")

    @property
    def z(self) -> np.ndarray:
        return self.axis("z")

    def voxel_volume(self) -> float:
        return self.dx * self.dy * self.dz

    def as_dict(self) -> dict:
        return dict(nx=self.nx, ny=self.ny, nz=self.nz, dx=self.dx, dy=self.dy, dz=self.dz,
                    ox=self.ox, oy=self.oy, oz=self.oz)

    @staticmethod
    def centered(n: int, d: float, axis: str = "z") -> "DoseGrid":
        """Cube centred laterally on 0; z starts at 0 unless axis != 'z'."""
        h = (n - 1) * d / 2.0
        return DoseGrid(n, n, n, d, d, d, -h, -h, 0.0 if axis == "z" else -h)


@dataclass
class SpotLattice:
    """IMPT spot lattice; weights[l, j, i] (model.py:85-131)."""
    n_sx: int
    n_sy: int
    n_layers: int
    dx_spot: float
    dy_spot: float
    x0: float = 0.0
    y0: float = 0.0
    energies: Optional[np.ndarray] = None
    weights: Optional[np.ndarray] = None

    @property
    def n_spots(self) -> int:
        return self.n_sx * self.n_sy * self.n_layers

    @property
    def spot_x(self) -> np.ndarray:
        return self.x0 + (np.arange(self.n_sx) - (self.n_sx - 1) / 2.0) * self.dx_spot

    @property
    def spot_y(self) -> np.ndarray:
        return self.y0 + (np.arange(self.n_sy) - (self.n_sy - 1) / 2.0) * self.dy_spot

    def ensure_weights(self, rng: np.random.Generator | None = None) -> np.ndarray:
        """Uniform [0, 1) weights from default_rng(0) if none were given."""
        if self.weights is None:
            self.weights = (rng or np.random.default_rng(0)).random(
                (self.n_layers, self.n_sy, self.n_sx))
        return self.weights

    def flat_spot_table(self) -> tuple:
        """(x, y, layer, weight) per spot, M = L * nsy * nsx entries."""
        w = self.ensure_weights()
        X, Y = np.meshgrid(self.spot_x, self.spot_y)
        shape = (self.n_layers, self.n_sy, self.n_sx)
        lay = np.broadcast_to(np.arange(self.n_layers)[:, None, None], shape)
        return (np.broadcast_to(X, shape).ravel().copy(), np.broadcast_to(Y, shape).ravel().copy(),
                lay.ravel().copy(), w.ravel().copy())


@dataclass
class PBProblem:
    """Grid + lattice + per-layer tables idd[l, k], sigma[l, k] (model.py:137-187)."""
    grid: DoseGrid
    lattice: SpotLattice
    idd: np.ndarray
    sigma: np.ndarray
    meta: dict = field(default_factory=dict)
    fluence_normalized: bool = True

    @property
    def n_layers(self) -> int:
        return self.lattice.n_layers

    @property
    def nz(self) -> int:
        return self.grid.nz

    def summary(self) -> str:
        g, L = self.grid, self.lattice
        lines = [f"grid {g.nx} x {g.ny} x {g.nz} ({g.n_vox / 1e6:.2f} M voxels, {g.dx * 10:.1f} mm)",
                 f"lattice {L.n_sx} x {L.n_sy} @ {L.dx_spot * 10:.1f} mm, {L.n_layers} layers, "
                 f"{L.n_spots:,} spots"]
        if L.energies is not None:
            lines.append(f"energies {np.min(L.energies):.1f} - {np.max(L.energies):.1f} MeV")
        lines += [f"{k}: {v}" for k, v in self.meta.items()]
        return "\n".join(lines)


def _tables(grid: DoseGrid, energies, idd_model, sigma_x0, sigma_theta0, device,
            fluence_normalized: bool = True):
    """(idd, sigma) host arrays of every layer, from one device call. With
    fluence_normalized=False each layer's IDD is divided by its maximum on the
    grid, as idd_pristine(normalize=True) does (physics.py:380-382)."""
    idd, sig, info = ph.build_luts(grid.nz, grid.oz, grid.dz, energies, idd_model=idd_model,
                                   sigma_x0=sigma_x0, sigma_theta0=sigma_theta0, device=device)
    idd = idd.cpu().numpy()
    if not fluence_normalized:
        pk = info[:, 3].cpu().numpy()
        idd = idd / np.where(pk > 0, pk, 1.0)[:, None]
    return idd, sig.cpu().numpy()


def gaussian_field_weights(n_layers: int, n_spot: int = 41, spot_spacing: float = 0.5,
                           spot_sigma: float = 0.28, seed: int = 0) -> np.ndarray:
    """Weights (L, n_spot, n_spot) of gaussian_impt_field: a Gaussian lateral
    envelope (peak 1) with 25 % per-spot jitter from default_rng(seed), clipped
    at 0. The draws are taken in the reference's order (model.py:207-220), so
    that the same seed gives the same weights bit for bit."""
    rng = np.random.default_rng(seed)
    s = (np.arange(n_spot) - (n_spot - 1) / 2.0) * spot_spacing
    X, Y = np.meshgrid(s, s)
    env = np.exp(-(X**2 + Y**2) / (2.0 * spot_sigma**2))
    env = env / env.max()
    return np.clip(env[None] * (1.0 + 0.25 * rng.standard_normal((n_layers, n_spot, n_spot))), 0.0, None)


def gaussian_impt_field(energies, grid: DoseGrid | None = None, *, n_spot: int = 41,
                        spot_spacing: float = 0.5, spot_sigma: float = 0.28, seed: int = 0,
                        idd_model: ph.IDDModel = ph.DEFAULT_IDD, beam_sigma_x0: float = 0.30,
                        beam_sigma_theta0: float = 0.0, fluence_normalized: bool = True,
                        device="cuda") -> PBProblem:
    """Gaussian lateral weight envelope with 25 % per-spot jitter (model.py:193-245)."""
    grid = grid or DoseGrid.centered(128, 0.2)
    energies = np.asarray(energies, dtype=np.float64)
    L = len(energies)
    w = gaussian_field_weights(L, n_spot, spot_spacing, spot_sigma, seed)
    lat = SpotLattice(n_spot, n_spot, L, spot_spacing, spot_spacing, energies=energies, weights=w)
    idd, sig = _tables(grid, energies, idd_model, beam_sigma_x0, beam_sigma_theta0, device,
                       fluence_normalized)
    meta = {"weights": f"Gaussian envelope sigma={spot_sigma * 10:.1f} mm, seed {seed}",
            "beam sigma_x0": f"{beam_sigma_x0 * 10:.1f} mm",
            "IDD": "fluence-normalised" if fluence_normalized else "peak-normalised per layer"}
    return PBProblem(grid, lat, idd, sig, meta=meta, fluence_normalized=fluence_normalized)


def clinical_layer_energies(n_layers: int = 40, r_min: float = 3.0, r_max: float = 25.5,
                            device="cuda") -> np.ndarray:
    """Energies of equally range-spaced layers (model.py:248-255)."""
    return ph.energy_from_range(np.linspace(r_min, r_max, n_layers), device=device)


def flat_field_weights(n_sx: int, n_sy: int, n_layers: int, *, field_half_width: float = 5.0,
                       spot_spacing: float = 0.5, edge: float = 0.8, radial_falloff: float = 0.06,
                       layer_jitter: float = 0.15, seed: int = 0) -> np.ndarray:
    """Flat-top lateral weights with a tanh edge and seeded per-spot jitter
    (model.py:258-287). As in the reference the profile is built on an
    n_sx x n_sx meshgrid, so the lattice must be square."""
    if n_sx != n_sy:
        raise ValueError("flat_field_weights builds a square lattice (n_sx == n_sy)")
    rng = np.random.default_rng(seed)
    s = (np.arange(n_sx) - (n_sx - 1) / 2.0) * spot_spacing
    X, Y = np.meshgrid(s, s)
    r = np.hypot(X, Y)
    prof = 0.5 * (1.0 - np.tanh((r - field_half_width) / max(edge, 1e-6)))
    prof = prof * np.exp(-np.clip(r - field_half_width, 0, None) / max(radial_falloff, 1e-6))
    w = np.broadcast_to(prof, (n_layers, n_sy, n_sx)).copy()
    if layer_jitter > 0:
        w *= 1.0 + layer_jitter * rng.standard_normal(w.shape)
    return np.clip(w, 0.0, None)


def flat_impt_field(energies, grid: DoseGrid | None = None, *, n_spot: int = 41,
                    spot_spacing: float = 0.5, field_half_width: float = 5.0, seed: int = 0,
                    idd_model: ph.IDDModel = ph.DEFAULT_IDD, beam_sigma_x0: float = 0.30,
                    beam_sigma_theta0: float = 0.0, fluence_normalized: bool = True,
                    device="cuda") -> PBProblem:
    """Flat-top IMPT field (model.py:290-330)."""
    grid = grid or DoseGrid.centered(128, 0.2)
    energies = np.asarray(energies, dtype=np.float64)
    L = len(energies)
    w = flat_field_weights(n_spot, n_spot, L, field_half_width=field_half_width,
                           spot_spacing=spot_spacing, seed=seed)
    lat = SpotLattice(n_spot, n_spot, L, spot_spacing, spot_spacing, energies=energies, weights=w)
    idd, sig = _tables(grid, energies, idd_model, beam_sigma_x0, beam_sigma_theta0, device,
                       fluence_normalized)
    meta = {"cross-section": f"flat {2 * field_half_width:.1f} x {2 * field_half_width:.1f} cm",
            "IDD": "fluence-normalised" if fluence_normalized else "peak-normalised per layer"}
    return PBProblem(grid, lat, idd, sig, meta=meta, fluence_normalized=fluence_normalized)


def water_field(grid: DoseGrid, n_sx: int, n_sy: int, spacing: float, energies, weights,
                *, x0: float = 0.0, y0: float = 0.0, sigma_x0: float = 0.30,
                idd_model: ph.IDDModel = ph.DEFAULT_IDD, device="cuda") -> PBProblem:
    """Field on a water grid with explicit weights (L, n_sy, n_sx)."""
    energies = np.asarray(energies, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)
    if w.shape != (len(energies), n_sy, n_sx):
        raise ValueError(f"weights shape {w.shape} != {(len(energies), n_sy, n_sx)}")
    lat = SpotLattice(n_sx, n_sy, len(energies), spacing, spacing, x0, y0, energies=energies,
                      weights=w)
    idd, sig = _tables(grid, energies, idd_model, sigma_x0, 0.0, device)
    return PBProblem(grid, lat, idd, sig)
