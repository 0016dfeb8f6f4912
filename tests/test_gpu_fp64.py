"""FP64 forward (pbh_forward_fp64, torch_separable(dtype=torch.float64)) vs the FP64 oracle.

Bar: max global relative error <= 1e-12 (max |D - D_ref| / max |D_ref|). Both sides
evaluate the same FP64 formula; they differ in summation order, in the library exp
against numpy's, and in how spot centres are rounded, all of order 1e-15.
"""
import numpy as np
import pytest
import torch

from oracle import pb_oracle as O
from test_gpu_forward import make_problem, oracle_dose

pytestmark = pytest.mark.gpu
TOL_FP64 = 1e-12


def check64(p, tol=TOL_FP64):
    from pbdose import engines
    D = engines.torch_separable(p, "cuda", dtype=torch.float64)
    assert D.dtype == torch.float64 and tuple(D.shape) == (p.grid.nz, p.grid.ny, p.grid.nx)
    D = D.cpu().numpy()
    ref = oracle_dose(p)
    r = O.relative_error(D, ref)
    print(f"  max_rel_err_global={r['max_rel_err_global']:.3e} rms={r['rms_rel_global']:.3e}",
          flush=True)
    assert r["max_rel_err_global"] <= tol
    return D, ref


def test_engine_check_problem_fp64():
    # engine_check.py:15-23 geometry (24x24x32 @ 2.5 mm, 3 layers of 3x3 spots, seed 7),
    # also against the textbook voxel x spot loop on the first 4 slices
    print("case fp64 engine_check 24x24x32", flush=True)
    from pbdose import model
    g = model.DoseGrid(24, 24, 32, 0.25, 0.25, 0.25, -3.0, -3.0, 0.0)
    E = np.array([80.0, 130.0, 180.0])
    w = np.random.default_rng(7).random((3, 3, 3))
    idd, sig = O.layer_luts(g.z, E)
    p = model.PBProblem(g, model.SpotLattice(3, 3, 3, 0.5, 0.5, energies=E, weights=w), idd, sig)
    D, _ = check64(p)
    B = O.brute_force_dose(g.x, g.y, 4, p.lattice.spot_x, p.lattice.spot_y, w, idd, sig)
    assert np.abs(D[:4] - B).max() <= TOL_FP64 * np.abs(B).max()


def test_small_64cube_fp64():
    print("case fp64 64^3 / 21x21 / 10 layers", flush=True)
    check64(make_problem(64, 64, 64, 0.2, 21, 21, 0.5, np.linspace(100, 150, 10), seed=4))


@pytest.mark.parametrize("shape", [(37, 70, 5), (130, 65, 3), (1, 1, 2), (200, 33, 4)])
def test_ragged_grids_fp64(shape):
    nx, ny, nz = shape
    print(f"case fp64 ragged grid {shape}", flush=True)
    check64(make_problem(nx, ny, nz, 0.2, 7, 5, 0.5, [90.0, 120.0], seed=3))


@pytest.mark.parametrize("spots", [(1, 1), (64, 64), (65, 3), (3, 65), (130, 70)])
def test_spot_chunks_fp64(spots):
    # the kernel takes spots in chunks of 64 per axis: full, one-past and several chunks
    nsx, nsy = spots
    print(f"case fp64 lattice {spots}", flush=True)
    check64(make_problem(64, 48, 4, 0.25, nsx, nsy, 0.15, [100.0, 160.0], seed=5, x0=0.37, y0=-0.21))


@pytest.mark.parametrize("pitch", [(0.2, 0.5), (0.3, 0.45), (0.2, 0.2), (0.25, 0.15), (0.25, 0.37), (0.2, 0.123)])
def test_pitch_ratios_fp64(pitch):
    # voxel / spot pitches that are small-integer multiples of one length (factor tables:
    # 2:5, 2:3, 1:1, 5:3) and ratios that are not (0.25 / 0.37, 0.2 / 0.123: every factor
    # evaluated where it is used)
    d, ds = pitch
    print(f"case fp64 pitches {pitch}", flush=True)
    check64(make_problem(72, 40, 6, d, 19, 13, ds, [95.0, 140.0, 175.0], seed=8, x0=0.11, y0=-0.07))


def test_zero_idd_layers_and_zero_weights_fp64():
    print("case fp64 zero IDD beyond range, zero weights", flush=True)
    w = np.random.default_rng(2).random((3, 9, 9))
    w[1] = 0.0
    w[2, 3:6, :] = 0.0
    p = make_problem(48, 48, 70, 0.2, 9, 9, 0.5, [60.0, 70.0, 120.0], w=w)
    assert (p.idd[0] == 0).any()
    D, _ = check64(p)
    # slices beyond every layer's range get exactly zero
    dead = np.flatnonzero((p.idd == 0).all(axis=0))
    assert len(dead) > 0 and not D[dead].any()


def test_weights_keep_full_precision():
    """Weights that FP32 cannot hold: the FP64 dose follows them, so it differs from the
    dose of the FP32-rounded weights by far more than the FP64 error bar."""
    print("case fp64 weights below FP32 resolution", flush=True)
    from pbdose import engines
    w = 1.0 + np.random.default_rng(8).random((2, 9, 9)) * 2.0 ** -30
    p = make_problem(40, 40, 8, 0.25, 9, 9, 0.5, [100.0, 140.0], w=w)
    D, ref = check64(p)
    dp = engines.DeviceProblem(p)
    w32 = torch.as_tensor(w.astype(np.float32).astype(np.float64), device=dp.dev)
    D32w = dp.forward_fp64(w=w32.contiguous()).cpu().numpy()
    gap = np.abs(D32w - ref).max() / np.abs(ref).max()
    print(f"  FP32-rounded weights: {gap:.3e}", flush=True)
    assert gap > 100 * TOL_FP64


def test_linearity_determinism_and_out_argument_fp64():
    from pbdose import engines
    p = make_problem(64, 40, 12, 0.2, 21, 15, 0.5, np.linspace(100, 150, 5), seed=11)
    dp = engines.DeviceProblem(p)
    a = dp.forward_fp64()
    out = torch.full(dp.shape, float("nan"), dtype=torch.float64, device=dp.dev)
    b = dp.forward_fp64(out=out)
    assert b.data_ptr() == out.data_ptr() and torch.equal(a, b)
    w2 = torch.rand(tuple(dp.w.shape), dtype=torch.float64, device=dp.dev)
    d2 = dp.forward_fp64(w=w2)
    d12 = dp.forward_fp64(w=(dp._w64 + 3.0 * w2).contiguous())
    ref = a + 3.0 * d2
    assert float((d12 - ref).abs().max()) <= 1e-13 * float(ref.abs().max())
    # the FP32 forward agrees to its own bar
    f32 = dp.forward().double()
    assert float((f32 - a).abs().max()) <= 1e-6 * float(a.abs().max())
    with pytest.raises(ValueError):
        dp.forward_fp64(out=torch.empty(dp.shape, dtype=torch.float32, device=dp.dev))
    with pytest.raises(ValueError):
        dp.forward_fp64(w=w2.float())
