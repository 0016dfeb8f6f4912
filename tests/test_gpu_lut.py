"""On-device FP64 look-up tables vs the reference's own values (fixtures) and
the oracle. Tolerance: 1e-12 relative (transcendentals of the device library
may differ from the host's by an ulp; sums are reordered)."""
import numpy as np
import pytest
import torch

from oracle import pb_oracle as O

pytestmark = pytest.mark.gpu


def test_range_table_bits(ref_physics):
    from pbdose import physics
    E, R = physics.range_table("cuda")
    np.testing.assert_array_equal(E.cpu().numpy(), ref_physics["table_E"])
    np.testing.assert_allclose(R.cpu().numpy(), ref_physics["table_R"], rtol=1e-13, atol=0)


def test_range_interpolation(ref_physics):
    from pbdose import physics
    np.testing.assert_allclose(physics.csda_range(ref_physics["csda_E"]), ref_physics["csda_R"],
                               rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(physics.energy_from_range(ref_physics["efr_R"]),
                               ref_physics["efr_E"], rtol=1e-13)
    e = physics.energy_from_range(np.linspace(3.0, 25.5, 40))
    np.testing.assert_allclose(e, ref_physics["clin40"], rtol=1e-13)


@pytest.mark.parametrize("case", ["small", "c1", "c2"])
def test_layer_luts(ref_physics, case):
    from pbdose import physics
    z, E = ref_physics[f"{case}_z"], ref_physics[f"{case}_E"]
    print(f"LUT case {case}: {len(E)} layers x {len(z)} depths", flush=True)
    idd, sig, info = physics.build_luts(len(z), float(z[0]), float(z[1] - z[0]), E)
    ri, rs = ref_physics[f"{case}_idd"], ref_physics[f"{case}_sigma"]
    idd, sig, info = idd.cpu().numpy(), sig.cpu().numpy(), info.cpu().numpy()
    assert np.abs(idd - ri).max() <= 1e-12 * np.abs(ri).max()
    assert np.abs(sig - rs).max() <= 1e-12 * np.abs(rs).max()
    np.testing.assert_allclose(info[:, 1], ref_physics[f"{case}_sigR"], rtol=1e-13)
    np.testing.assert_allclose(info[:, 3], ri.max(axis=1), rtol=1e-12)


def test_lut_with_divergence_and_icru_model():
    from pbdose import physics
    z = np.arange(50) * 0.3 - 0.6          # starts below 0: IDD left of the grid is 0
    E = np.array([72.5, 133.0, 201.0])
    model = physics.IDDModel(straggling_model="icru49", energy_spread=0.005,
                             attenuation_length=120.0)
    idd, sig, _ = physics.build_luts(len(z), float(z[0]), 0.3, E, idd_model=model,
                                     sigma_x0=0.25, sigma_theta0=0.003, rho0=0.2)
    for l, e in enumerate(E):
        ri = O.idd_pristine(z, e, energy_spread=0.005, attenuation_length=120.0,
                            straggling="icru49")
        rs = O.fe_moments(z, e, 0.25, 0.003, 0.2)["sigma_x"]
        assert np.abs(idd[l].cpu().numpy() - ri).max() <= 1e-12 * ri.max()
        assert np.abs(sig[l].cpu().numpy() - rs).max() <= 1e-12 * rs.max()


@pytest.mark.parametrize("case", ["small", "c1", "c2"])
def test_depth_straggling_luts(ref_physics, case):
    """Depth-dependent range straggling on the device vs the oracle (1e-12);
    the lateral sigma and layer sigma_R are those of the reference model."""
    from pbdose import physics
    z, E = ref_physics[f"{case}_z"], ref_physics[f"{case}_E"]
    print(f"depth-straggling LUT case {case}: {len(E)} layers x {len(z)} depths", flush=True)
    model = physics.IDDModel(depth_straggling=True)
    idd, sig, info = physics.build_luts(len(z), float(z[0]), float(z[1] - z[0]), E,
                                        idd_model=model)
    idd, sig, info = idd.cpu().numpy(), sig.cpu().numpy(), info.cpu().numpy()
    ref = np.stack([O.idd_depth_straggling(z, e) for e in E])
    assert np.abs(idd - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(sig - ref_physics[f"{case}_sigma"]).max() <= 1e-12 * np.abs(sig).max()
    np.testing.assert_allclose(info[:, 1], ref_physics[f"{case}_sigR"], rtol=1e-13)


def test_depth_straggling_small_spreads_and_no_straggling():
    """No energy spread: the entrance bins have spreads below two fine-grid
    steps (discretely normalised kernels; zero spread at the surface). With
    straggling "none" the depth model equals the reference model."""
    from pbdose import physics
    z = np.arange(120) * 0.1 - 0.05
    E = np.array([E for E in (O.energy_from_range(10.0), 70.0, 180.0)])
    m0 = physics.IDDModel(depth_straggling=True, energy_spread=0.0)
    idd, _, _ = physics.build_luts(len(z), float(z[0]), 0.1, E, idd_model=m0)
    ref = np.stack([O.idd_depth_straggling(z, e, energy_spread=0.0) for e in E])
    assert np.abs(idd.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    m1 = physics.IDDModel(depth_straggling=True, straggling_model="none")
    m2 = physics.IDDModel(straggling_model="none")
    a, _, _ = physics.build_luts(len(z), float(z[0]), 0.1, E, idd_model=m1)
    b, _, _ = physics.build_luts(len(z), float(z[0]), 0.1, E, idd_model=m2)
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    with pytest.raises(ValueError):
        physics.beam_model(physics.IDDModel(depth_straggling=True, straggling_model="icru49"))


def _oracle_tables(z, E, shift, model="bohr", depth=False, sigma_x0=0.30):
    """The oracle's tables with every layer's range shifted by `shift` cm."""
    Es = O.shifted_energies(E, np.full(len(E), shift))
    if depth:
        idd = np.stack([O.idd_depth_straggling(z, e) for e in Es])
    else:
        idd = np.stack([O.idd_pristine(z, e, straggling=model) for e in Es])
    sig = np.stack([O.fe_moments(z, e, sigma_x0)["sigma_x"] for e in Es])
    return idd, sig


@pytest.mark.parametrize("case", ["c2", "icru49", "depth"])
def test_tangent_tables_vs_converged_differences(ref_physics, case):
    """pbh_build_lut_tangent: its tables equal pbh_build_lut's bit for bit, and
    its derivatives in a range shift (direction dE/dR from pbh_interp_table_slope)
    equal the limit of the oracle's finite differences.

    The tables are piecewise smooth in the shift with dense breaks: every
    range-table knot that crosses a CSDA bin edge, every fine-grid point that
    crosses R0 (the scattering power drops to 0 there) and every interpolation
    bracket that crosses a depth adds a kink or a step, some 1e-5 cm apart. So
    the derivative is checked where the oracle's differences at h = 1e-6 cm show
    no break within 2h: the central differences at h and 2h agree, and so do the
    one-sided second differences (a kink at the point itself shows only there),
    each to 1e-8 of the layer's largest derivative. There the derivative must
    equal the Richardson extrapolation of the central differences. Most entries
    qualify (asserted)."""
    from pbdose import physics
    z = ref_physics["c2_z"]
    if case == "c2":
        E, model, depth = ref_physics["c2_E"], physics.DEFAULT_IDD, False
    else:
        E = np.linspace(72.0, 195.0, 8)
        model = (physics.IDDModel(straggling_model="icru49") if case == "icru49"
                 else physics.IDDModel(depth_straggling=True))
        depth = case == "depth"
    print(f"case tangent tables {case}: {len(E)} layers x {len(z)} depths", flush=True)
    R0 = O.csda_range(E)
    Et, dE = physics.energy_from_range_slope(R0)
    np.testing.assert_array_equal(Et.cpu().numpy(), O.energy_from_range(R0))
    Eg, Rg = O.range_table()
    k = np.searchsorted(Rg, R0, side="right") - 1
    np.testing.assert_array_equal(dE.cpu().numpy(), (Eg[k + 1] - Eg[k]) / (Rg[k + 1] - Rg[k]))
    a = physics.build_luts(len(z), 0.0, 0.2, Et, idd_model=model)
    t = physics.build_luts(len(z), 0.0, 0.2, Et, idd_model=model, denergies=dE)
    for u, v in zip(a, t[:3]):
        assert torch.equal(u, v)
    ad = [t[3].cpu().numpy(), t[4].cpu().numpy()]
    mname = "icru49" if case == "icru49" else "bohr"
    h = 1e-6
    v = {s: _oracle_tables(z, E, s * h, mname, depth) for s in (-2, -1, 0, 1, 2)}
    for q, name in ((0, "idd"), (1, "sigma")):
        c1 = (v[1][q] - v[-1][q]) / (2 * h)
        c2 = (v[2][q] - v[-2][q]) / (4 * h)
        k1 = (v[1][q] - 2 * v[0][q] + v[-1][q]) / h          # h f'' when smooth
        k2 = (v[2][q] - 2 * v[0][q] + v[-2][q]) / (2 * h)    # 2 h f''
        sc = np.abs(c1).max(axis=1, keepdims=True)
        sc = np.where(sc > 0, sc, 1.0)
        ok = (np.abs(c2 - c1) <= 1e-8 * sc) & (np.abs(k2 - 2 * k1) <= 1e-8 * sc)
        rich = (4.0 * c1 - c2) / 3.0
        err = np.abs(ad[q] - rich) / sc
        frac = float(ok.mean())
        e_ok = float(err[ok].max())
        print(f"  d{name}/ddelta: {frac:.3f} of entries without a break within 2h; max error there "
              f"{e_ok:.2e}; median error everywhere {float(np.median(err)):.2e}", flush=True)
        assert frac >= (0.85 if q == 0 else 0.6)
        assert e_ok <= 1e-7
