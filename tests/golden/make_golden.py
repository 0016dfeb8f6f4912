"""Write the reference-physics fixtures in tests/golden/ref_physics.npz.

Run in the development container only (the reference tree is not present
where the tests run): ``python tests/golden/make_golden.py``.

Published digits pin the physics lookup tables to 4-7 significant figures
only, so the FP64 lookup-table functions of the reference
(physics.py:111-139 range table, :154-174 straggling, :221-286 Fermi-Eyges
moments, :328-387 pristine IDD; model.py:248-287 energy and weight builders)
are evaluated here, by importing the reference package read-only, on the
inputs the tests use. Only inputs and outputs are stored.
"""
from __future__ import annotations

import importlib
import importlib.util
import os
import sys

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_physics.npz")


def _load_reference():
    spec = importlib.util.spec_from_file_location(
        "pbdose_ref", os.path.join(REF, "__init__.py"), submodule_search_locations=[REF])
    pkg = importlib.util.module_from_spec(spec)
    sys.modules["pbdose_ref"] = pkg
    spec.loader.exec_module(pkg)
    ph = importlib.import_module("pbdose_ref.physics")
    md = importlib.import_module("pbdose_ref.model")
    return ph, md


def main():
    ph, md = _load_reference()
    out = {}
    # range table and its two interpolations
    E, R = ph._range_table()
    out["table_E"], out["table_R"] = E, R
    e_q = np.array([0.3, 0.5, 0.75, 1.0, 2.0, 5.0, 10.0, 33.3, 59.05787709884295, 70.0,
                    100.0, 137.5, 150.0, 180.0, 200.0, 250.0, 300.0, 399.9, 400.0, 450.0])
    out["sp_E"], out["sp_S"] = e_q, ph.stopping_power_water(e_q)
    out["csda_E"], out["csda_R"] = e_q, ph.csda_range(e_q)
    r_q = np.array([-1.0, 0.0, 1e-4, 0.3, 3.0, 4.08, 7.7, 12.6, 25.5, 26.0, 50.0, 80.0, 1e3])
    out["efr_R"], out["efr_E"] = r_q, ph.energy_from_range(r_q)
    out["clin40"] = md.clinical_layer_energies(40, 3.0, 25.5)

    # per-layer tables for the two single-field workloads and the engine check
    cases = {
        "c1": (np.arange(64) * 0.2, np.linspace(100.0, 150.0, 10)),
        "c2": (np.arange(128) * 0.2, np.linspace(70.0, 180.0, 40)),
        "small": (np.arange(32) * 0.25, np.array([80.0, 130.0, 180.0])),
    }
    for key, (z, energies) in cases.items():
        idd = np.stack([ph.idd_pristine(z, e, ph.DEFAULT_IDD, normalize=False)[0]
                        for e in energies])
        sig = np.stack([ph.fermi_eyges_moments(z, e, sigma_x0=0.30)["sigma_x"]
                        for e in energies])
        out[f"{key}_z"], out[f"{key}_E"] = z, energies
        out[f"{key}_idd"], out[f"{key}_sigma"] = idd, sig
        out[f"{key}_sigR"] = np.array([ph.DEFAULT_IDD.sigma_range(e)[0] for e in energies])
        out[f"{key}_strag"] = np.array([ph.range_straggling_bohr(e) for e in energies])

    # full moment set on one grid, with beam divergence and correlation on
    zm = np.linspace(0.0, 25.0, 51)
    mom = ph.fermi_eyges_moments(zm, 200.0, sigma_x0=0.3, sigma_theta0=0.004, rho0=0.5)
    out["mom_z"] = zm
    for k in ("A0", "A1", "A2", "B", "sigma_x"):
        out[f"mom_{k}"] = mom[k]

    # weight builders (seeded RNG use)
    out["flat_w"] = md.flat_field_weights(21, 21, 10, seed=0)
    p = md.gaussian_impt_field(np.array([100.0, 120.0, 140.0]), md.DoseGrid.centered(16, 0.4),
                               n_spot=9, spot_spacing=0.5, seed=3)
    out["gauss_w"], out["gauss_idd"], out["gauss_sigma"] = p.lattice.weights, p.idd, p.sigma
    # the other field builder, a divergent beam and peak-normalised IDDs
    p = md.flat_impt_field(np.array([95.0, 125.0, 155.0]), md.DoseGrid.centered(16, 0.4),
                           n_spot=9, spot_spacing=0.5, field_half_width=1.5, seed=5)
    out["flatf_w"], out["flatf_idd"], out["flatf_sigma"] = p.lattice.weights, p.idd, p.sigma
    p = md.gaussian_impt_field(np.array([110.0, 170.0]), md.DoseGrid.centered(16, 0.4),
                               n_spot=7, spot_spacing=0.5, seed=4, beam_sigma_x0=0.25,
                               beam_sigma_theta0=0.003, fluence_normalized=False)
    out["gaussd_w"], out["gaussd_idd"], out["gaussd_sigma"] = p.lattice.weights, p.idd, p.sigma

    # the two other range-straggling models of IDDModel (physics.py:310-322)
    e_s = np.array([70.0, 100.0, 150.0, 200.0, 230.0])
    for name in ("icru49", "none"):
        mdl = ph.IDDModel(straggling_model=name)
        sr = np.array([mdl.sigma_range(e) for e in e_s])
        out[f"{name}_E"], out[f"{name}_sigR"], out[f"{name}_strag"] = e_s, sr[:, 0], sr[:, 1]
        z, en = cases["small"]
        out[f"{name}_idd"] = np.stack([ph.idd_pristine(z, e, mdl, normalize=False)[0] for e in en])

    # single-depth grids: fermi_eyges_moments integrates over [0, z0] (physics.py:247-249)
    z1 = np.array([0.0, 0.2, 4.0, 12.3, 30.0])
    e1 = np.array([90.0, 150.0, 210.0])
    out["nz1_z"], out["nz1_E"] = z1, e1
    out["nz1_sigma"] = np.array([[ph.fermi_eyges_moments(np.array([zz]), e, sigma_x0=0.30)["sigma_x"][0]
                                  for zz in z1] for e in e1])
    out["nz1_idd"] = np.array([[ph.idd_pristine(np.array([zz]), e, ph.DEFAULT_IDD, normalize=False)[0][0]
                                for zz in z1] for e in e1])

    # the reference's own benchmark problem, run_benchmark.py:33-38: 128^3 at 2 mm, 41 x 41
    # spots at 5 mm, clinical_layer_energies(40, 3, 25.5), gaussian_impt_field seed 0;
    # its FP64 separable dose (engines.py:172-202) as the peak, per-slice sums and samples
    eng = importlib.import_module("pbdose_ref.engines")
    bp = md.gaussian_impt_field(md.clinical_layer_energies(40, 3.0, 25.5), md.DoseGrid.centered(128, 0.2),
                                n_spot=41, spot_spacing=0.5, seed=0)
    D = eng.separable_numpy(bp)
    rng = np.random.default_rng(11)
    flat = rng.choice(D.size, 4096, replace=False)
    hot = np.argsort(D.ravel())[-256:]
    idx = np.unique(np.concatenate([flat, hot]))
    out["bp_idd"], out["bp_sigma"] = bp.idd, bp.sigma
    out["bp_max"] = np.array([D.max()])
    out["bp_slice_sum"] = D.sum(axis=(1, 2))
    out["bp_idx"], out["bp_val"] = idx.astype(np.int64), D.ravel()[idx]
    # the averaged-sigma approximation (per_layer=False) on a small field
    sp = md.gaussian_impt_field(np.array([100.0, 120.0, 140.0]), md.DoseGrid.centered(16, 0.4),
                                n_spot=9, spot_spacing=0.5, seed=3)
    out["avg_dose"] = eng.separable_numpy(sp, per_layer=False)
    out["avg_sigma_bar"] = eng.layerwise_sigma_bar(sp)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {os.path.getsize(OUT)} bytes, {len(out)} arrays")


if __name__ == "__main__":
    main()
