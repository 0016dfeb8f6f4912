"""Forward dose on the GPU (through the C ABI) vs the FP64 oracle.

Bar: max global relative error <= 1e-6 for the FP32 form
(max |D - D_ref| / max |D_ref|, engines.py:364-378 convention).
"""
import numpy as np
import pytest
import torch

from oracle import pb_oracle as O

pytestmark = pytest.mark.gpu
TOL_FP32 = 1e-6


def make_problem(nx, ny, nz, d, nsx, nsy, ds, energies, w=None, x0=0.0, y0=0.0, oz=0.0,
                 seed=0, sigma_x0=0.3):
    from pbdose import model
    g = model.DoseGrid(nx, ny, nz, d, d, d, -(nx - 1) * d / 2, -(ny - 1) * d / 2, oz)
    energies = np.asarray(energies, dtype=np.float64)
    if w is None:
        w = np.random.default_rng(seed).lognormal(0.0, 0.5, (len(energies), nsy, nsx))
    idd, sig = O.layer_luts(g.z, energies, sigma_x0=sigma_x0)
    lat = model.SpotLattice(nsx, nsy, len(energies), ds, ds, x0, y0, energies=energies, weights=w)
    return model.PBProblem(g, lat, idd, sig)


def oracle_dose(p, z_idx=None):
    return O.separable_dose(p.grid.x, p.grid.y, p.lattice.spot_x, p.lattice.spot_y,
                            p.lattice.weights, p.idd, p.sigma, z_idx=z_idx)


def check(p, tol=TOL_FP32, z_idx=None, precision="fp32"):
    from pbdose import engines
    D = engines.torch_separable(p, "cuda", precision=precision).double().cpu().numpy()
    ref = oracle_dose(p, z_idx)
    got = D if z_idx is None else D[z_idx]
    r = O.relative_error(got, ref)
    print(f"  max_rel_err_global={r['max_rel_err_global']:.3e} rms={r['rms_rel_global']:.3e}",
          flush=True)
    assert r["max_rel_err_global"] <= tol
    return D, ref


def test_engine_check_problem():
    # engine_check.py:15-23 geometry: 24x24x32 @ 2.5 mm, 3 layers of 3x3 spots, seed 7
    print("case engine_check 24x24x32", flush=True)
    from pbdose import model
    g = model.DoseGrid(24, 24, 32, 0.25, 0.25, 0.25, -3.0, -3.0, 0.0)
    E = np.array([80.0, 130.0, 180.0])
    w = np.random.default_rng(7).random((3, 3, 3))
    idd, sig = O.layer_luts(g.z, E)
    p = model.PBProblem(g, model.SpotLattice(3, 3, 3, 0.5, 0.5, energies=E, weights=w), idd, sig)
    D, ref = check(p)
    B = O.brute_force_dose(g.x, g.y, 4, p.lattice.spot_x, p.lattice.spot_y, w, idd, sig)
    assert np.abs(D[:4] - B).max() <= TOL_FP32 * np.abs(B).max()


@pytest.mark.parametrize("precision,tol", [("fp32", TOL_FP32), ("bf16x3", 2e-5), ("fp16x3", 2e-6)])
def test_triton_check_problem(precision, tol):
    # triton_check2.py:14-22 geometry: 32x32x40 @ 2.5 mm, 3 layers (90, 140, 190 MeV) of
    # 5x5 spots, seed 3; per-layer exact against separable_numpy(per_layer=True) and the
    # single-sigma path against separable_numpy(per_layer=False), as that script checks
    print(f"case triton_check 32x32x40 {precision}", flush=True)
    from pbdose import engines, model
    g = model.DoseGrid(32, 32, 40, 0.25, 0.25, 0.25, -4.0, -4.0, 0.0)
    E = np.array([90.0, 140.0, 190.0])
    w = np.random.default_rng(3).random((3, 5, 5))
    idd, sig = O.layer_luts(g.z, E)
    lat = model.SpotLattice(5, 5, 3, 0.5, 0.5, energies=E, weights=w)
    p = model.PBProblem(g, lat, idd, sig)
    check(p, tol=tol, precision=precision)
    Da = engines.torch_separable(p, "cuda", per_layer=False, precision=precision).double().cpu().numpy()
    ref = O.separable_dose_averaged(g.x, g.y, lat.spot_x, lat.spot_y, w, idd, sig)
    r = O.relative_error(Da, ref)["max_rel_err_global"]
    print(f"  single sigma: {r:.3e}", flush=True)
    assert r <= tol


def test_config_small_64cube():
    # 64^3 @ 2 mm, 21x21 @ 5 mm, 10 layers 100-150 MeV, w = 1
    print("case 64^3 / 21x21 / 10 layers", flush=True)
    p = make_problem(64, 64, 64, 0.2, 21, 21, 0.5, np.linspace(100, 150, 10),
                     w=np.ones((10, 21, 21)))
    check(p)


def test_config_clinical_128cube_full_volume():
    """The clinical workload (128^3 @ 2 mm, 41 x 41 @ 5 mm, 40 layers 70-180 MeV,
    lognormal(0, 0.5) weights from seed 0): every voxel of the FP32 dose against
    the FP64 oracle."""
    print("case 128^3 / 41x41 / 40 layers (full volume vs the FP64 oracle)", flush=True)
    p = make_problem(128, 128, 128, 0.2, 41, 41, 0.5, np.linspace(70, 180, 40))
    check(p)


@pytest.mark.parametrize("shape", [(37, 70, 5), (130, 65, 3), (1, 1, 2), (200, 33, 4)])
def test_ragged_grids(shape):
    nx, ny, nz = shape
    print(f"case ragged grid {shape}", flush=True)
    p = make_problem(nx, ny, nz, 0.2, 7, 5, 0.5, [90.0, 120.0], seed=3)
    check(p)


@pytest.mark.parametrize("spots", [(1, 1), (60, 50), (97, 13), (49, 49)])
def test_lattice_chunking(spots):
    nsx, nsy = spots
    print(f"case lattice {spots}", flush=True)
    p = make_problem(64, 48, 6, 0.25, nsx, nsy, 0.3, [100.0, 160.0], seed=5, x0=0.37, y0=-0.21)
    check(p)


def test_zero_idd_layers_and_zero_weights():
    print("case zero IDD beyond range, zero weights", flush=True)
    w = np.random.default_rng(2).random((3, 9, 9))
    w[1] = 0.0
    w[2, 3:6, :] = 0.0
    p = make_problem(48, 48, 70, 0.2, 9, 9, 0.5, [60.0, 70.0, 120.0], w=w)
    assert (p.idd[0] == 0).any()
    check(p)


# Degenerate shapes at the 2.5 mm voxel / 5 mm spot pitch (the tabulated geometry of
# the mixed forms): one voxel, spot and layer; one depth slice; one-voxel-wide grids.
DEGENERATE = {
    "one_voxel_spot_layer": ((1, 1, 1), (1, 1), [120.0]),
    "one_slice": ((48, 40, 1), (9, 7), [90.0, 120.0, 150.0]),
    "one_column": ((1, 40, 6), (5, 9), [100.0, 130.0]),
    "one_row": ((40, 1, 6), (9, 5), [100.0, 130.0]),
}


@pytest.mark.parametrize("precision", ["fp32", "bf16", "bf16x3"])
@pytest.mark.parametrize("case", list(DEGENERATE))
def test_degenerate_shapes(case, precision):
    from pbdose import model
    (nx, ny, nz), (nsx, nsy), E = DEGENERATE[case]
    print(f"case degenerate {case} {precision}", flush=True)
    p = make_problem(nx, ny, max(nz, 2), 0.25, nsx, nsy, 0.5, E, seed=4)
    if nz == 1:
        # the oracle's moments need two depths: tables of the first slice of two
        g = p.grid
        g1 = model.DoseGrid(g.nx, g.ny, 1, g.dx, g.dy, g.dz, g.ox, g.oy, g.oz)
        p = model.PBProblem(g1, p.lattice, p.idd[:, :1].copy(), p.sigma[:, :1].copy())
    check(p, tol=TOL_FP32 if precision == "fp32" else MIXED_TOL[precision], precision=precision)


@pytest.mark.parametrize("precision", ["fp32", "bf16", "bf16x3"])
def test_all_zero_weights_give_zero_dose(precision):
    from pbdose import engines
    print(f"case all-zero weights {precision}", flush=True)
    p = make_problem(32, 32, 8, 0.25, 7, 7, 0.5, [100.0, 140.0], w=np.zeros((2, 7, 7)))
    D = engines.torch_separable(p, "cuda", precision=precision)
    assert int(torch.count_nonzero(D)) == 0

def test_linearity_and_determinism():
    from pbdose import engines
    p = make_problem(64, 64, 16, 0.2, 21, 21, 0.5, np.linspace(100, 150, 6), seed=11)
    dp = engines.DeviceProblem(p)
    a = dp.forward()
    b = dp.forward()
    assert torch.equal(a, b)
    w2 = torch.rand_like(dp.w)
    d2 = dp.forward(w=w2)
    d12 = dp.forward(w=(dp.w + 3.0 * w2).contiguous())
    ref = (a.double() + 3.0 * d2.double())
    assert (d12.double() - ref).abs().max() <= 1e-6 * ref.abs().max()


# Per-form bounds on the max global relative error vs the FP64 oracle. Single
# 16-bit products carry the factor rounding (2^-9 BF16, 2^-12 FP16); the split
# forms drop only lo*lo.
MIXED_TOL = {"bf16": 5e-3, "fp16": 1e-3, "bf16x3": 2e-5, "fp16x3": 2e-6,
             "bf16x3_peak": 5e-3, "fp16x3_peak": 1e-3}


@pytest.mark.parametrize("precision", list(MIXED_TOL))
def test_mixed_precision_forms(precision):
    print(f"case mixed {precision} on 64^3 / 21x21 / 10 layers", flush=True)
    p = make_problem(64, 64, 64, 0.2, 21, 21, 0.5, np.linspace(100, 150, 10), seed=4)
    check(p, tol=MIXED_TOL[precision], precision=precision)


@pytest.mark.parametrize("precision", ["fp16", "fp16x3", "fp16x3_peak"])
@pytest.mark.parametrize("scale_exp", [-26, -12, 17, 30])
def test_fp16_forms_at_any_weight_scale(precision, scale_exp):
    """FP16 has a narrow exponent range: weights of 2^17 (proton counts) would
    overflow it and weights of 2^-12 would lose the low half to subnormals. The
    weight split scales every (layer, spot chunk) by a power of two and the layer's
    U by its bound, so the relative error is that of O(1) weights."""
    print(f"case {precision} with weights x 2^{scale_exp}", flush=True)
    p = make_problem(64, 64, 24, 0.2, 21, 21, 0.5, np.linspace(100, 150, 6), seed=6)
    p.lattice.weights = p.lattice.weights * 2.0 ** scale_exp
    check(p, tol=MIXED_TOL[precision], precision=precision)


@pytest.mark.parametrize("precision", ["bf16x3_peak", "fp16x3_peak"])
def test_peak_forms_split_near_each_bragg_peak(precision):
    """The "_peak" forms: the weight-split launch writes each layer's Bragg-peak
    depth (np.argmax of its IDD on the grid) into the workspace, and one-layer
    fields are within the split bound on the slices within 1 cm of that depth and
    within the single-product bound elsewhere."""
    from pbdose import _lib, engines
    print(f"case peak depths {precision}", flush=True)
    for E in (95.0, 140.0):
        p = make_problem(64, 64, 96, 0.2, 15, 15, 0.5, [E], seed=8)
        dp = engines.DeviceProblem(p)
        D = dp.forward(precision).double().cpu().numpy()
        chunk, rows = dp.auto_shape(_lib.PREC[precision])
        ws = dp.workspace("fwd", _lib.PREC[precision], chunk, rows)
        r_peak = ws[:4].view(torch.float32).cpu().numpy()
        z_pk = p.grid.oz + np.argmax(p.idd[0]) * p.grid.dz
        assert abs(float(r_peak[0]) - z_pk) <= 1e-5
        ref = oracle_dose(p)
        # slices at exactly 1 cm are left out: the kernel compares in FP32
        near = np.abs(p.grid.z - z_pk) <= 0.99
        far = np.abs(p.grid.z - z_pk) >= 1.01
        assert near.any() and far.any()
        scale = np.abs(ref).max()
        e_near = float(np.abs(D[near] - ref[near]).max() / scale)
        e_far = float(np.abs(D[far] - ref[far]).max() / scale)
        split_tol = MIXED_TOL[precision.replace("_peak", "")]
        print(f"  E {E}: peak depth {float(r_peak[0]):.2f} cm; error within 1 cm {e_near:.2e} "
              f"(bar {split_tol:.0e}), elsewhere {e_far:.2e}", flush=True)
        assert e_near <= split_tol and e_far <= MIXED_TOL[precision]


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_mixed_ragged_and_chunked(precision):
    print(f"case mixed {precision} ragged 70x37x5, lattice 60x50", flush=True)
    p = make_problem(70, 37, 5, 0.25, 60, 50, 0.3, [90.0, 150.0], seed=9, x0=0.2)
    check(p, tol=MIXED_TOL[precision], precision=precision)


@pytest.mark.parametrize("precision", ["bf16", "fp16", "bf16x3", "fp16x3", "bf16x3_peak", "fp16x3_peak"])
@pytest.mark.parametrize("case", ["clinical_slab", "offset_ragged", "spot_chunks"])
def test_tile_heights_agree_bit_for_bit(precision, case):
    """64-row tiles (2 mm voxels, 5 mm spots) give the 32-row tiles' dose bit for
    bit: every voxel sums the same MFMA products in the same order (same layers,
    K steps, windows and FP16 scale), only the sharing of operands differs. Both
    are within the form's bound of the oracle, ragged tile edges included."""
    from pbdose import engines
    print(f"case tile heights {case} {precision}", flush=True)
    if case == "clinical_slab":
        p = make_problem(128, 128, 20, 0.2, 41, 41, 0.5, np.linspace(70, 180, 40), seed=2)
    elif case == "offset_ragged":
        p = make_problem(100, 90, 7, 0.2, 15, 13, 0.5, [85.0, 120.0, 160.0], x0=1.1, y0=-0.7, seed=3)
    else:
        p = make_problem(150, 70, 5, 0.2, 60, 55, 0.5, [90.0, 140.0], seed=5)
    dp = engines.DeviceProblem(p)
    ref = oracle_dose(p)
    for chunk in (3, 40):
        a = dp.forward(precision, layer_chunk=chunk, tile_rows=32).clone()
        b = dp.forward(precision, layer_chunk=chunk, tile_rows=64).clone()
        assert torch.equal(a, b), f"chunk {chunk}: {float((a - b).abs().max()):.3e}"
        r = O.relative_error(b.double().cpu().numpy(), ref)["max_rel_err_global"]
        assert r <= MIXED_TOL[precision]
    from pbdose import _lib
    c, rows = dp.auto_shape(_lib.PREC[precision])
    assert rows in (32, 64) and c >= 1


def test_tile_height_argument_checks():
    from pbdose import engines
    p = make_problem(64, 48, 4, 0.2, 11, 9, 0.5, [100.0, 130.0], seed=6)
    dp = engines.DeviceProblem(p)
    with pytest.raises(RuntimeError, match="unsupported"):
        dp.forward("fp32", layer_chunk=2, tile_rows=64)          # FP32 tiles are 32 rows
    with pytest.raises(RuntimeError, match="invalid dimension"):
        dp.forward("bf16x3", layer_chunk=2, tile_rows=48)
    q = make_problem(64, 48, 4, 0.2, 11, 9, 0.37, [100.0, 130.0], seed=6)
    with pytest.raises(RuntimeError, match="unsupported"):
        engines.DeviceProblem(q).forward("bf16x3", layer_chunk=2, tile_rows=64)   # no 2 : 5 table


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp16x3"])
def test_layer_chunk_parity_and_determinism(precision):
    """Active layers cut into chunks of 1, 3, 5, 7 and 16 (one chunk) per workgroup,
    and the automatic chunk: every result matches the oracle and repeats bit for bit."""
    from pbdose import engines
    print(f"case layer chunk sweep {precision}", flush=True)
    p = make_problem(128, 96, 12, 0.2, 41, 33, 0.5, np.linspace(80, 170, 16), seed=21)
    dp = engines.DeviceProblem(p)
    ref = oracle_dose(p)
    tol = TOL_FP32 if precision == "fp32" else MIXED_TOL[precision]
    for chunk in (0, 1, 3, 5, 7, 16):
        a = dp.forward(precision, layer_chunk=chunk).clone()
        b = dp.forward(precision, layer_chunk=chunk)
        assert torch.equal(a, b), f"layer chunk {chunk} not deterministic"
        r = O.relative_error(a.double().cpu().numpy(), ref)["max_rel_err_global"]
        print(f"  layer chunk {chunk}: {r:.3e}", flush=True)
        assert r <= tol


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3", "bf16"])
@pytest.mark.parametrize("rows", [32, 64])
def test_combine_of_two_to_six_chunks(precision, rows):
    """Layer chunks that leave 1, 2 or 3 partial tiles per dose tile (the combine's
    unrolled paths, every load before the first store) and 5 (its loop), at both tile
    heights, on a grid whose last band of tile rows is cut by the grid (ny = 100): each
    within the form's bound of the oracle and repeatable bit for bit."""
    from pbdose import engines
    print(f"case combine chunks {precision} {rows} rows", flush=True)
    p = make_problem(128, 100, 6, 0.2, 41, 33, 0.5, np.linspace(80, 170, 16), seed=31)
    dp = engines.DeviceProblem(p)
    ref = oracle_dose(p)
    for chunk in (8, 6, 4, 3):            # 16 active layers per slice: 2, 3, 4, 6 chunks
        a = dp.forward(precision, layer_chunk=chunk, tile_rows=rows).clone()
        b = dp.forward(precision, layer_chunk=chunk, tile_rows=rows)
        assert torch.equal(a, b), f"layer chunk {chunk} not deterministic"
        r = O.relative_error(a.double().cpu().numpy(), ref)["max_rel_err_global"]
        print(f"  layer chunk {chunk}: {r:.3e}", flush=True)
        assert r <= MIXED_TOL[precision]


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp16x3", "bf16"])
def test_direct_factor_path(precision):
    # spot pitch 0.37 cm against 0.2 cm voxels has no small integer ratio:
    # every Gaussian factor is evaluated directly instead of from a table
    print(f"case direct factors {precision}", flush=True)
    p = make_problem(100, 90, 8, 0.2, 23, 29, 0.37, [95.0, 125.0, 165.0], seed=13)
    check(p, tol=TOL_FP32 if precision == "fp32" else MIXED_TOL[precision], precision=precision)


def test_two_to_three_pitch_grid():
    # 2 mm voxels, 3 mm spots: parity-interleaved tables with P = 2, Q = 3 (run-time
    # slots), several spot chunks in x
    print("case 2 mm voxels / 3 mm spots, FP32", flush=True)
    p = make_problem(160, 72, 6, 0.2, 61, 21, 0.3, [85.0, 120.0, 150.0], seed=41)
    check(p)


def test_quarter_pitch_grid():
    # 2.5 mm voxels, 5 mm spots: table with P = 1, Q = 2
    print("case 2.5 mm voxels / 5 mm spots", flush=True)
    p = make_problem(96, 96, 10, 0.25, 31, 31, 0.5, np.linspace(90, 160, 8), seed=17)
    check(p)


@pytest.mark.parametrize("precision", ["bf16x3", "bf16", "fp16x3", "bf16x3_peak"])
@pytest.mark.parametrize("case", ["offset_field", "direct_factors", "spot_chunks"])
def test_windowed_contraction(precision, case):
    """The mixed kernels skip K steps whose spots lie beyond 6.66 sigma (Gaussian
    factor below 2^-32 of its peak) of every row or column of a wave. Small,
    off-centre fields and narrow beams on wide grids leave most waves with partial
    or empty windows; every form must stay within its bound, and voxels far from
    every spot must stay negligible."""
    print(f"case windowed {case} {precision}", flush=True)
    if case == "offset_field":
        p = make_problem(128, 96, 10, 0.2, 9, 7, 0.5, [70.0, 95.0], x0=-6.0, y0=4.0,
                         sigma_x0=0.12, seed=31)
    elif case == "direct_factors":
        p = make_problem(128, 64, 6, 0.2, 11, 9, 0.37, [80.0, 120.0], x0=5.0,
                         sigma_x0=0.15, seed=32)
    else:
        p = make_problem(160, 40, 4, 0.2, 97, 13, 0.3, [90.0, 140.0], sigma_x0=0.1, seed=33)
    D, ref = check(p, tol=MIXED_TOL[precision], precision=precision)
    far = ref < 1e-12 * ref.max()
    assert far.any()
    assert np.abs(D[far]).max() <= 1e-7 * ref.max()


def test_forward_autotune_picks_a_timed_candidate(tmp_path, monkeypatch):
    """DeviceProblem.autotune times (layer chunk, tile rows) candidates, 32- and
    64-row tiles among them, and later automatic forwards use the winner: the
    dose equals the explicit call with that shape bit for bit, and the cache
    returns the same winner without timing again."""
    from pbdose import _lib, engines
    monkeypatch.setenv("PBDOSE_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    print("case forward autotune bf16x3", flush=True)
    p = make_problem(128, 128, 24, 0.2, 41, 41, 0.5, np.linspace(70, 180, 40), seed=7)
    dp = engines.DeviceProblem(p)
    cands = dp.shape_candidates(_lib.PREC["bf16x3"])
    assert {r for _, r in cands} == {32, 64}
    best = dp.autotune("bf16x3", reps=3)
    timed = [cr for _, cr in dp.tuning_results]
    assert set(cands) <= set(timed) and len(timed) == len(set(timed)) and best in timed
    # the refinement times the neighbouring chunks of the fastest candidate
    fastest = min(dp.tuning_results[:len(cands)])[1]
    amax = int(dp._active.max())
    assert all((c, fastest[1]) in timed for c in range(max(1, fastest[0] - 2), min(amax, fastest[0] + 2) + 1))
    a = dp.forward("bf16x3").clone()
    b = dp.forward("bf16x3", layer_chunk=best[0], tile_rows=best[1])
    assert torch.equal(a, b)
    dp2 = engines.DeviceProblem(p)
    assert dp2.autotune("bf16x3") == best and not hasattr(dp2, "tuning_results")
    # FP32: 32-row tiles only
    assert {r for _, r in dp.shape_candidates(_lib.PREC["fp32"])} == {32}
