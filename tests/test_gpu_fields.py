"""The field builders and the reference's own benchmark problem through the
HIP path (GPU, C ABI), against the reference's values: fixtures written by
tests/golden/make_golden.py from the reference's functions, and the published
peak dose 635.2707 (02_performance_report.pdf section 1.2).

Bars: weights bit-equal (seeded RNG in the reference's order); tables 1e-12
relative; dose 1e-6 (FP32) and 1e-5 (bf16x3) relative to the peak.
"""
import numpy as np
import pytest
import torch

from oracle import pb_oracle as O

pytestmark = pytest.mark.gpu


def close(a, b, tol):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max()) <= tol * float(np.abs(b).max())


def test_field_builders_match_reference(ref_physics):
    """gaussian_impt_field, flat_impt_field (model.py:193-245, :290-330), with a
    divergent beam and peak-normalised IDDs, and clinical_layer_energies."""
    from pbdose import model
    print("case field builders on a 16^3 grid", flush=True)
    g = model.DoseGrid.centered(16, 0.4)
    p = model.gaussian_impt_field(np.array([100.0, 120.0, 140.0]), g, n_spot=9, spot_spacing=0.5, seed=3)
    np.testing.assert_array_equal(p.lattice.weights, ref_physics["gauss_w"])
    assert close(p.idd, ref_physics["gauss_idd"], 1e-12) and close(p.sigma, ref_physics["gauss_sigma"], 1e-12)
    p = model.flat_impt_field(np.array([95.0, 125.0, 155.0]), g, n_spot=9, spot_spacing=0.5,
                              field_half_width=1.5, seed=5)
    np.testing.assert_array_equal(p.lattice.weights, ref_physics["flatf_w"])
    assert close(p.idd, ref_physics["flatf_idd"], 1e-12) and close(p.sigma, ref_physics["flatf_sigma"], 1e-12)
    p = model.gaussian_impt_field(np.array([110.0, 170.0]), g, n_spot=7, spot_spacing=0.5, seed=4,
                                  beam_sigma_x0=0.25, beam_sigma_theta0=0.003, fluence_normalized=False)
    np.testing.assert_array_equal(p.lattice.weights, ref_physics["gaussd_w"])
    assert not p.fluence_normalized
    assert close(p.idd, ref_physics["gaussd_idd"], 1e-12) and close(p.sigma, ref_physics["gaussd_sigma"], 1e-12)
    assert np.abs(p.idd.max(axis=1) - 1.0).max() <= 1e-15
    np.testing.assert_allclose(model.clinical_layer_energies(40, 3.0, 25.5), ref_physics["clin40"], rtol=1e-13)


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-6), ("bf16x3", 1e-5)])
def test_reference_benchmark_problem(ref_physics, precision, tol):
    """run_benchmark.py:33-38 built by pbdose.model and computed by torch_separable:
    the peak is the published 635.2707, and the dose equals the reference's FP64
    separable dose at 4,352 sampled voxels (incl. the 256 hottest) and in every
    slice sum."""
    from pbdose import engines, model
    print(f"case reference benchmark problem 128^3, 67,240 spots, {precision}", flush=True)
    grid = model.DoseGrid.centered(128, 0.2)
    E = model.clinical_layer_energies(40, 3.0, 25.5)
    p = model.gaussian_impt_field(E, grid, n_spot=41, spot_spacing=0.5, seed=0)
    assert close(p.idd, ref_physics["bp_idd"], 1e-12) and close(p.sigma, ref_physics["bp_sigma"], 1e-12)
    D = engines.torch_separable(p, "cuda", precision=precision).double().cpu().numpy()
    peak = ref_physics["bp_max"][0]
    e_pub = abs(D.max() / 635.2707 - 1.0)
    e_peak = abs(D.max() / peak - 1.0)
    e_vox = float(np.abs(D.ravel()[ref_physics["bp_idx"]] - ref_physics["bp_val"]).max() / peak)
    e_sum = float(np.abs(D.sum(axis=(1, 2)) / ref_physics["bp_slice_sum"] - 1.0).max())
    print(f"  peak {D.max():.7f} (published 635.2707: {e_pub:.2e}); vs reference: peak {e_peak:.2e}, "
          f"voxels {e_vox:.2e}, slice sums {e_sum:.2e}", flush=True)
    assert e_pub <= tol + 1e-7     # the published value carries 7 digits
    assert e_peak <= tol and e_vox <= tol and e_sum <= tol


def test_single_depth_tables_and_dose(ref_physics):
    """nz = 1 (physics.py:247-249): sigma integrated over [0, z0] ([0, 1e-6] for
    z0 <= 0), the IDD read at z0; a one-slice field through the forward."""
    from pbdose import engines, model, physics
    print("case single-depth tables", flush=True)
    E = ref_physics["nz1_E"]
    for k, z0 in enumerate(ref_physics["nz1_z"]):
        idd, sig, _ = physics.build_luts(1, float(z0), 1.0, E)
        assert close(sig.cpu().numpy()[:, 0], ref_physics["nz1_sigma"][:, k], 1e-12)
        ri = ref_physics["nz1_idd"][:, k]
        assert float(np.abs(idd.cpu().numpy()[:, 0] - ri).max()) <= 1e-12 * max(1.0, float(np.abs(ri).max()))
    s = physics.fermi_eyges_moments([4.0], 150.0, sigma_x0=0.30)["sigma_x"]
    assert abs(s[0] / ref_physics["nz1_sigma"][1, 2] - 1.0) <= 1e-12
    g = model.DoseGrid(24, 20, 1, 0.2, 0.2, 0.2, -2.3, -1.9, 4.0)
    p = model.gaussian_impt_field(np.array([90.0, 150.0]), g, n_spot=7, spot_spacing=0.5, seed=1)
    D = engines.torch_separable(p, "cuda").double().cpu().numpy()
    ref = O.separable_dose(g.x, g.y, p.lattice.spot_x, p.lattice.spot_y, p.lattice.weights, p.idd, p.sigma)
    assert O.relative_error(D, ref)["max_rel_err_global"] <= 1e-6


@pytest.mark.parametrize("name", ["icru49", "none"])
def test_other_straggling_models(ref_physics, name):
    """IDDModel('icru49') / ('none'): sigma_R and the IDD vs the reference's values."""
    from pbdose import physics
    print(f"case straggling model {name}", flush=True)
    mdl = physics.IDDModel(straggling_model=name)
    z, E = ref_physics["small_z"], ref_physics["small_E"]
    idd, _, _ = physics.build_luts(len(z), float(z[0]), float(z[1] - z[0]), E, idd_model=mdl)
    assert close(idd.cpu().numpy(), ref_physics[f"{name}_idd"], 1e-12)
    _, _, info = physics.build_luts(len(z), float(z[0]), float(z[1] - z[0]), ref_physics[f"{name}_E"],
                                    idd_model=mdl)
    np.testing.assert_allclose(info[:, 1].cpu().numpy(), ref_physics[f"{name}_sigR"], rtol=1e-13)


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-6), ("bf16x3", 2e-5), ("fp64", 1e-12)])
def test_averaged_sigma_matches_reference(ref_physics, precision, tol):
    """per_layer=False (engines.py:184-191, :243-255): sigma_bar(z) and the dose
    against the reference's own values on its fixture field (16^3 at 4 mm, 9 x 9
    spots, 3 layers), then a 2 : 5 pitch field against the oracle's FP64
    single-sigma dose. fp64: dtype=torch.float64 (amplitude maps and dose in FP64)."""
    from pbdose import engines, model
    print(f"case averaged sigma, {precision}", flush=True)
    fp64 = precision == "fp64"

    def dose(p):
        if fp64:
            D = engines.torch_separable(p, "cuda", per_layer=False, dtype=torch.float64)
            assert D.dtype == torch.float64
            return D.cpu().numpy()
        return engines.torch_separable(p, "cuda", per_layer=False, precision=precision).double().cpu().numpy()

    g = model.DoseGrid.centered(16, 0.4)
    p = model.gaussian_impt_field(np.array([100.0, 120.0, 140.0]), g, n_spot=9, spot_spacing=0.5, seed=3)
    dp = engines.averaged_problem(p, "cuda", fp64=fp64)
    np.testing.assert_allclose(dp.sigma.cpu().numpy(), np.tile(ref_physics["avg_sigma_bar"], (16, 1)),
                               rtol=1e-12, atol=0)
    np.testing.assert_array_equal(dp.idd.cpu().numpy(), np.eye(16))
    e = O.relative_error(dose(p), ref_physics["avg_dose"])["max_rel_err_global"]
    print(f"  reference fixture field: {e:.2e}", flush=True)
    assert e <= tol

    g = model.DoseGrid(64, 48, 96, 0.2, 0.2, 0.2, -6.3, -4.7, 0.0)
    p = model.gaussian_impt_field(np.array([90.0, 115.0, 140.0, 160.0]), g, n_spot=21, spot_spacing=0.5, seed=7)
    D = dose(p)
    ref = O.separable_dose_averaged(g.x, g.y, p.lattice.spot_x, p.lattice.spot_y, p.lattice.weights,
                                    p.idd, p.sigma)
    err = O.relative_error(D, ref)["max_rel_err_global"]
    print(f"  64 x 48 x 96, 21 x 21 spots, 4 layers: {err:.2e}", flush=True)
    assert err <= tol


def test_averaged_sigma_equals_exact_when_widths_agree():
    """With every layer's sigma row equal, sigma_bar is that row and the
    single-sigma dose equals the per-layer-exact dose (FP32 both)."""
    from pbdose import engines, model
    print("case averaged sigma = exact for equal widths", flush=True)
    g = model.DoseGrid(40, 36, 30, 0.2, 0.2, 0.3, -3.9, -3.5, 0.0)
    p = model.gaussian_impt_field(np.array([80.0, 100.0, 120.0]), g, n_spot=11, spot_spacing=0.5, seed=2)
    p.sigma = np.tile(p.sigma[1], (3, 1))
    Da = engines.torch_separable(p, "cuda", per_layer=False).double().cpu().numpy()
    De = engines.torch_separable(p, "cuda").double().cpu().numpy()
    assert O.relative_error(Da, De)["max_rel_err_global"] <= 1e-6
