"""FP64 CPU oracle for the per-layer-exact pencil-beam dose path.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``. The shipped package
never imports this module; it is the checker the HIP path is compared against.

Every function restates one piece of the reference's arithmetic and cites the
reference file:line it follows. Where NumPy's own numerics decide the last bits
(``linspace``, ``interp``, ``cumsum``, ``fft``) the oracle calls the same NumPy
routine, because the reference does.

Pinning (what fixes each function, beyond retyping the formula):

==============================  ==============================================
function                        pinned by
==============================  ==============================================
stopping_power                  NIST/reference table values (published 4-5
                                digits), Bethe-Bloch closed form at 3 energies
range_table / csda_range        reference golden ranges 200/250/300 MeV; the
energy_from_range               16-digit clinical_layer_energies endpoints;
                                round trip R(E(R)) = R; reference fixtures
straggling_bohr / sigma_range   published sigma_R at 100/150/200/230 MeV;
                                reference fixtures of all three models
                                (bohr, icru49, none)
fe_moments                      constant-T closed forms (A0=Tz, A2=Tz^3/6,
                                B=Tz^3/3); published sigma_x table at 200 MeV;
                                reference fixtures, incl. single depths
idd_pristine                    FFT == direct Gaussian convolution; energy
                                conservation of the CSDA binning; published
                                peak IDD values; reference fixtures
straggling_bohr_depth           published sigma_rho(z) table (5 depths,
                                03_validation_report section 7.6.2); an
                                independent depth-space quadrature; equals
                                straggling_bohr beyond R0
idd_depth_straggling            with no straggling it equals idd_pristine;
                                unit weight per source bin; the reference's
                                measured Monte Carlo / analytic entrance
                                ratios (+30 %, +7.8 %, +1.9 % at 2/4/6 mm)
separable_dose                  brute-force triple loop (tiny problems),
                                single-spot analytic Gaussian, linearity,
                                mirror symmetry, published peak dose 635.2707
                                of the reference's benchmark problem, and
                                reference fixtures of that problem's dose
separable_dose_averaged,        reference fixtures (dose, sigma_bar) of the
sigma_bar                       per_layer=False path; equals separable_dose
                                when every layer has the same sigma
gaussian_field_weights          reference fixtures (bit-equal weights)
relative_error, mean_rel_error  hand-computed cases
gamma_index                     hand-computed cases; gamma(ref, ref) == 0
objective_grad_w                dot-product test and central differences
==============================  ==============================================
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

# ---------------------------------------------------------------------------
# Physical constants (physics.py:42-65). These are CODATA/PDG values and the
# reference's model parameters; they must be the same numbers to reproduce it.
# ---------------------------------------------------------------------------
MP_MEV = 938.27208816            # proton rest energy
ME_MEV = 0.51099895000           # electron rest energy
K_BB = 0.307075                  # 4 pi N_A r_e^2 m_e c^2 [MeV cm^2/mol]
WATER_ZA = 10.0 / 18.01528       # Z/A of water
WATER_I = 75.0 * 1e-6            # mean excitation energy [MeV]
WATER_RHO = 1.0
WATER_X0 = 36.08                 # radiation length [cm]
E_HIGHLAND = 13.6
E_LO, E_HI = 0.5, 400.0          # CSDA table span [MeV]
P_R = 1.77                       # range-energy power used for the spread term
BOHR_COEF = K_BB * WATER_ZA * WATER_RHO * ME_MEV   # physics.py:151
INV_SQRT_2PI = 1.0 / np.sqrt(2.0 * np.pi)          # engines.py:43
SIGMA_FLOOR = 1e-9                                 # engines.py:57, :266


# ---------------------------------------------------------------------------
# Stopping power and the CSDA range table
# ---------------------------------------------------------------------------
def kinematics(E):
    """(beta, gamma, beta^2) of a proton of kinetic energy E. physics.py:72-77"""
    E = np.asarray(E, dtype=np.float64)
    g = 1.0 + E / MP_MEV
    b2 = np.clip(1.0 - 1.0 / g**2, 1e-14, 1.0)
    return np.sqrt(b2), g, b2


def stopping_power(E):
    """Bethe-Bloch linear stopping power of water [MeV/cm]. physics.py:80-105

    No shell or density-effect term; I = 75 eV. The evaluation order of the
    products follows the reference so the FP64 bits agree.
    """
    E = np.asarray(E, dtype=np.float64)
    _, g, b2 = kinematics(E)
    r = ME_MEV / MP_MEV
    tmax = 2.0 * ME_MEV * b2 * g**2 / (1.0 + 2.0 * g * r + r**2)
    arg = np.clip(2.0 * ME_MEV * b2 * g**2 * tmax / (WATER_I**2), 1.0 + 1e-12, None)
    return K_BB * WATER_ZA * WATER_RHO / b2 * (0.5 * np.log(arg) - b2)


def energy_grid():
    """The 3797-point energy grid of the range table. physics.py:118-123

    Four linspace segments whose shared end points (2, 20, 60 MeV) are
    removed by ``np.unique``.
    """
    parts = [np.linspace(E_LO, 2.0, 800), np.linspace(2.0, 20.0, 900),
             np.linspace(20.0, 60.0, 900), np.linspace(60.0, E_HI, 1200)]
    return np.unique(np.concatenate(parts))


@lru_cache(maxsize=1)
def range_table():
    """(E, R) with R(E) = cumulative trapezoid of 1/S from E_LO. physics.py:111-127"""
    E = energy_grid()
    inv = 1.0 / stopping_power(E)
    seg = 0.5 * (inv[1:] + inv[:-1]) * np.diff(E)
    R = np.concatenate([[0.0], np.cumsum(seg)])
    E.setflags(write=False)
    R.setflags(write=False)
    return E, R


def csda_range(E):
    """CSDA range [cm] by linear interpolation in the table. physics.py:130-133"""
    Eg, Rg = range_table()
    return np.interp(np.asarray(E, dtype=np.float64), Eg, Rg)


def energy_from_range(R):
    """Inverse of csda_range; R is clipped to [0, R_max]. physics.py:136-139"""
    Eg, Rg = range_table()
    return np.interp(np.clip(np.asarray(R, dtype=np.float64), 0.0, Rg[-1]), Rg, Eg)


def clinical_layer_energies(n_layers=40, r_min=3.0, r_max=25.5):
    """Energies of equally range-spaced layers. model.py:248-255"""
    return energy_from_range(np.linspace(r_min, r_max, n_layers))


# ---------------------------------------------------------------------------
# Range straggling and the IDD model
# ---------------------------------------------------------------------------
def straggling_bohr(E0, n=4000):
    """Bohr range straggling sigma [cm]. physics.py:154-174

    sigma^2 = int_{E_LO}^{E0} BOHR_COEF / (beta^2 S^3) dE, trapezoid on a
    uniform n-point grid.
    """
    E = np.linspace(E_LO, float(E0), n)
    if E[-1] <= E[0]:
        return 0.0
    S = stopping_power(E)
    _, _, b2 = kinematics(E)
    f = BOHR_COEF / (np.clip(b2, 1e-12, None) * S**3)
    return float(np.sqrt(np.trapezoid(f, E)))


def sigma_range(E0, energy_spread=0.0033, straggling="bohr",
                icru_coef=0.012, icru_exp=0.935):
    """Total range spread hypot(straggling, 1.77*spread*R0). physics.py:310-322"""
    R0 = float(csda_range(E0))
    if straggling == "bohr":
        s = straggling_bohr(E0)
    elif straggling == "icru49":
        s = icru_coef * R0 ** icru_exp
    elif straggling == "none":
        s = 0.0
    else:
        raise ValueError(straggling)
    e = P_R * energy_spread * R0
    return float(np.hypot(s, e)), float(s), float(e)


def idd_pristine(z, E0, energy_spread=0.0033, attenuation_length=150.0,
                 n_fine=6000, straggling="bohr", normalize=False, full=False):
    """Fluence-normalised pristine Bragg curve on z. physics.py:328-387

    1. energy-conserving CSDA binning on a 6000-point grid, with 6001 bin
       edges on [0, R0 + 8 sigma_R + 0.2];
    2. Gaussian range-spread convolution through a zero-padded real FFT;
    3. linear interpolation onto z with zero outside the fine grid.
    """
    z = np.asarray(z, dtype=np.float64)
    R0 = float(csda_range(E0))
    sR = sigma_range(E0, energy_spread, straggling)[0]
    top = R0 + 8.0 * sR + 0.2
    zf = np.linspace(0.0, top, n_fine)
    h = zf[1] - zf[0]
    edges = np.linspace(0.0, top, n_fine + 1)
    Ee = np.where(edges < R0, energy_from_range(np.clip(R0 - edges, 0.0, None)), 0.0)
    d_csda = np.exp(-zf / attenuation_length) * (Ee[:-1] - Ee[1:]) / h
    nfft = int(2 ** np.ceil(np.log2(2 * n_fine)))
    f = np.fft.rfftfreq(nfft, d=h)
    H = np.exp(-2.0 * (np.pi * f * sR) ** 2)
    d_fine = np.fft.irfft(np.fft.rfft(d_csda, nfft) * H, nfft)[:n_fine]
    out = np.interp(z, zf, d_fine, left=0.0, right=0.0)
    pk = float(out.max()) if out.max() > 0 else 1.0
    if normalize:
        out = out / pk
    if full:
        return dict(idd=out, sigma_R=sR, R0=R0, z_fine=zf, dose_fine=d_fine,
                    dose_csda=d_csda, h=h, peak=pk)
    return out


def idd_direct_convolution(E0, energy_spread=0.0033, attenuation_length=150.0,
                           n_fine=6000):
    """Closed-form check of the FFT step: the zero-padded FFT with the Gaussian
    multiplier exp(-2 (pi f sigma)^2) equals a direct sum against the sampled
    Gaussian h/(sqrt(2 pi) sigma) exp(-(k h)^2 / (2 sigma^2)) (Poisson summation;
    the 16384-point period is far beyond the kernel support)."""
    o = idd_pristine([0.0], E0, energy_spread, attenuation_length, n_fine, full=True)
    sR, h, x = o["sigma_R"], o["h"], o["dose_csda"]
    k = np.arange(-(n_fine - 1), n_fine)
    g = h * INV_SQRT_2PI / sR * np.exp(-(k * h) ** 2 / (2.0 * sR * sR))
    return np.convolve(x, g)[n_fine - 1: 2 * n_fine - 1], o["dose_fine"]


def straggling_bohr_depth(E0, z, n=4000):
    """Bohr range straggling accumulated up to depth z [cm].

    sigma_B(z)^2 = int_{E(R0-z)}^{E0} BOHR_COEF / (beta^2 S^3) dE
                 = int_0^z BOHR_COEF / (beta^2 S^2) dz'
    as written in the reference's own diagnosis of its constant-sigma_R IDD
    (README.md:180-184; 03_validation_report.pdf section 7.6.2). The integral is
    the cumulative trapezoid on straggling_bohr's grid, linspace(0.5, E0, n)
    (physics.py:168-174), read at the residual energy by np.interp; beyond R0
    it is straggling_bohr(E0).
    """
    z = np.asarray(z, dtype=np.float64)
    E = np.linspace(E_LO, float(E0), n)
    if E[-1] <= E[0]:
        return np.zeros_like(z)
    R0 = float(csda_range(E0))
    S = stopping_power(E)
    _, _, b2 = kinematics(E)
    f = BOHR_COEF / (np.clip(b2, 1e-12, None) * S**3)
    C = np.concatenate([[0.0], np.cumsum(np.diff(E) * (f[1:] + f[:-1]) / 2.0)])
    er = np.where(z < R0, energy_from_range(np.clip(R0 - z, 0.0, None)), E_LO)
    er = np.clip(er, E_LO, float(E0))
    return np.sqrt(np.maximum(C[-1] - np.interp(er, E, C), 0.0))


def idd_depth_straggling(z, E0, energy_spread=0.0033, attenuation_length=150.0,
                         n_fine=6000, straggling="bohr", full=False):
    """Pristine Bragg curve with a depth-dependent range spread.

    The reference convolves the whole CSDA curve with one sigma_R
    (physics.py:370-373) and documents that as its open model defect
    (README.md:180-184). Here the CSDA dose binned at depth z' (the same
    binning as idd_pristine) is spread with s(z') = hypot(sigma_B(z'),
    1.77 * energy_spread * R0), each bin keeping unit weight: the sampled
    Gaussian h/(sqrt(2 pi) s) exp(-(k h)^2 / (2 s^2)) for s >= 2h, the same
    Gaussian normalised by its discrete sum for 0 < s < 2h, and no spread for
    s = 0. With s constant this is idd_pristine's convolution (Poisson
    summation, see idd_direct_convolution).
    """
    if straggling not in ("bohr", "none"):
        raise ValueError(straggling)
    z = np.asarray(z, dtype=np.float64)
    o = idd_pristine([0.0], E0, energy_spread, attenuation_length, n_fine, straggling, full=True)
    h, dc, zf, R0 = o["h"], o["dose_csda"], o["z_fine"], o["R0"]
    se = P_R * energy_spread * R0
    sb = straggling_bohr_depth(E0, zf) if straggling == "bohr" else np.zeros_like(zf)
    s = np.hypot(sb, se)
    amp = np.ones_like(s)
    c = np.full_like(s, -1.0)
    pos = s > 0
    c[pos] = 1.0 / (2.0 * s[pos] ** 2)
    big = s >= 2.0 * h
    amp[big] = h / (np.sqrt(2.0 * np.pi) * s[big])
    for i in np.nonzero(pos & ~big)[0]:
        m = np.arange(1, int(math.ceil(12.0 * s[i] / h)) + 1)
        amp[i] = 1.0 / (1.0 + 2.0 * np.sum(np.exp(-(m * h) ** 2 * c[i])))
    src = np.arange(n_fine)

    def fine(js):
        t = (js[:, None] - src[None, :]) * h
        w = amp[None, :] * np.exp(-(t * t) * np.where(pos, c, 0.0)[None, :])
        w = np.where(pos[None, :], w, (t == 0).astype(np.float64))
        return w @ dc

    out = np.zeros_like(z)
    ins = (z >= 0.0) & (z <= zf[-1])
    j0 = np.clip(np.searchsorted(zf, z[ins], side="right") - 1, 0, n_fine - 1)
    j1 = np.minimum(j0 + 1, n_fine - 1)
    d0, d1 = fine(j0), fine(j1)
    frac = np.where(j1 > j0, (z[ins] - zf[j0]) / np.where(j1 > j0, zf[j1] - zf[j0], 1.0), 0.0)
    out[ins] = d0 + frac * (d1 - d0)
    if full:
        return dict(idd=out, spread=s, z_fine=zf, dose_csda=dc, h=h, R0=R0,
                    fine=fine)
    return out


# ---------------------------------------------------------------------------
# Fermi-Eyges moments
# ---------------------------------------------------------------------------
def scattering_power(E):
    """Local Highland scattering power T [rad^2/cm] with L = 1. physics.py:185-218

    With ``z_over_X0=None`` the reference uses ln L = 0, so T is the base
    term times (1 + 0) * (1.076 + 0).
    """
    beta, g, _ = kinematics(E)
    pc = np.sqrt(np.clip(g**2 - 1.0, 0.0, None)) * MP_MEV
    bp = np.clip(beta * pc, 1e-9, None)
    base = (E_HIGHLAND / bp) ** 2 / WATER_X0
    lnL = np.log(1.0)
    return base * (1.0 + 0.038 * lnL) * (1.076 + 0.038 * lnL)


def fine_count(z, n_sub=4, max_step=0.02):
    """Fine-grid size of the moment integrals. physics.py:256-258"""
    z = np.asarray(z, dtype=np.float64)
    span = float(z[-1] - z[0])
    need = int(np.ceil(span / max(max_step, 1e-6))) + 1 if span > 0 else 2
    return max((len(z) - 1) * n_sub + 1, need, 2)


def fe_moments(z, E0, sigma_x0=0.0, sigma_theta0=0.0, rho0=0.0, n_sub=4,
               max_step=0.02, T_fn=None):
    """Fermi-Eyges moments and lateral sigma on z. physics.py:221-286

    B_i = sum_{j : zf_i > zf_j} (zf_i - zf_j)^2 T_j h_j is the explicit double
    sum (not 2*A2). ``T_fn`` replaces the scattering power (closed-form tests).
    A single depth z0 is evaluated as the last point of the grid [0, z0], or
    [0, 1e-6] when z0 <= 0 (physics.py:247-249, :284-285).
    """
    z = np.asarray(z, dtype=np.float64).reshape(-1)
    if z.size == 0:
        raise ValueError("fe_moments needs at least one depth")
    if z.size == 1:
        z2 = np.array([0.0, z[0]]) if z[0] > 0 else np.array([0.0, 1e-6])
        res = fe_moments(z2, E0, sigma_x0, sigma_theta0, rho0, n_sub, max_step, T_fn)
        return {k: (np.asarray(v)[-1:] if k != "n_fine" else v) for k, v in res.items()}
    nf = fine_count(z, n_sub, max_step)
    zf = np.linspace(z[0], z[-1], nf)
    hf = np.gradient(zf)
    R0 = float(csda_range(E0))
    Ez = np.where(zf < R0, energy_from_range(np.clip(R0 - zf, 0.0, None)), 0.0)
    if T_fn is None:
        Tf = np.where(zf < R0, scattering_power(np.clip(Ez, 1e-3, None)), 0.0)
    else:
        Tf = np.asarray(T_fn(zf), dtype=np.float64) * np.ones_like(zf)
    A0f = np.cumsum(Tf * hf)
    A1f = np.cumsum(A0f * hf)
    A2f = np.cumsum(A1f * hf)
    dz = zf[:, None] - zf[None, :]
    Bf = np.where(dz > 0, dz**2 * Tf[None, :] * hf[None, :], 0.0).sum(axis=1)
    B = np.interp(z, zf, Bf)
    var = (sigma_x0**2 + 2.0 * rho0 * sigma_x0 * sigma_theta0 * z
           + sigma_theta0**2 * z**2 + B)
    return dict(A0=np.interp(z, zf, A0f), A1=np.interp(z, zf, A1f),
                A2=np.interp(z, zf, A2f), B=B, n_fine=nf,
                sigma_x=np.sqrt(np.clip(var, 0.0, None)))


def layer_luts(z, energies, sigma_x0=0.30, sigma_theta0=0.0, rho0=0.0,
               energy_spread=0.0033, attenuation_length=150.0):
    """(idd, sigma) tables of shape (L, nz), one row per layer, as the field
    builders make them (model.py:225-233, :312-320)."""
    z = np.asarray(z, dtype=np.float64)
    L = len(energies)
    idd = np.zeros((L, z.size))
    sig = np.zeros((L, z.size))
    for l, E in enumerate(np.asarray(energies, dtype=np.float64)):
        idd[l] = idd_pristine(z, E, energy_spread, attenuation_length)
        sig[l] = fe_moments(z, E, sigma_x0, sigma_theta0, rho0)["sigma_x"]
    return idd, sig


def shifted_energies(energies, range_shift):
    """Energies whose CSDA range is R0 + shift (cm), layer by layer."""
    R0 = csda_range(np.asarray(energies, dtype=np.float64))
    return energy_from_range(R0 + np.asarray(range_shift, dtype=np.float64))


# ---------------------------------------------------------------------------
# Problem model (model.py:27-131)
# ---------------------------------------------------------------------------
def axis(n, d, o):
    """Voxel centres o + k d. model.py:48-52"""
    return o + np.arange(n) * d


def centered_grid(n, d):
    """DoseGrid.centered(n, d): x/y centred on 0, z from 0. model.py:73-79"""
    h = (n - 1) * d / 2.0
    return dict(nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, ox=-h, oy=-h, oz=0.0)


def spot_axis(n, ds, x0=0.0):
    """Centred spot positions x0 + (i - (n-1)/2) ds. model.py:107-113"""
    return x0 + (np.arange(n) - (n - 1) / 2.0) * ds


def grid_axes(grid):
    return (axis(grid["nx"], grid["dx"], grid["ox"]),
            axis(grid["ny"], grid["dy"], grid["oy"]),
            axis(grid["nz"], grid["dz"], grid["oz"]))


# ---------------------------------------------------------------------------
# Dose engines
# ---------------------------------------------------------------------------
def lateral(vox, spots, sigma):
    """G[..., k, i] = exp(-(v_k - s_i)^2 / (2 s^2)) / (sqrt(2 pi) s), s >= 1e-9.
    engines.py:49-61"""
    d2 = (np.asarray(vox)[:, None] - np.asarray(spots)[None, :]) ** 2
    s = np.clip(np.asarray(sigma, dtype=np.float64), SIGMA_FLOOR, None)
    return np.exp(-d2 / (2.0 * s[..., None, None] ** 2)) * (INV_SQRT_2PI / s)[..., None, None]


def separable_dose(x, y, sx, sy, w, idd, sigma, z_idx=None):
    """FP64 per-layer-exact separable dose (nz, ny, nx). engines.py:172-202

    D[z] = sum_l IDD_l(z) Gy_l(z) W_l Gx_l(z)^T, evaluated slice-block by
    slice-block to bound memory; the contraction is the same two einsums.
    With ``z_idx`` only those slices are computed (returned in that order).
    """
    w = np.asarray(w, dtype=np.float64)
    idd = np.asarray(idd, dtype=np.float64)
    sigma = np.asarray(sigma, dtype=np.float64)
    L, nz = idd.shape
    zs = np.arange(nz) if z_idx is None else np.asarray(z_idx, dtype=np.int64)
    D = np.zeros((len(zs), len(y), len(x)))
    zb = max(1, int(4_000_000 // max(1, L * len(x) * len(sx))))
    for b0 in range(0, len(zs), zb):
        sel = zs[b0:b0 + zb]
        Gx = lateral(x, sx, sigma[:, sel])                 # (L, b, nx, nsx)
        Gy = lateral(y, sy, sigma[:, sel])                 # (L, b, ny, nsy)
        A = idd[:, sel, None, None] * w[:, None, :, :]     # (L, b, nsy, nsx)
        T = np.einsum("lbji,lbvj->lbiv", A, Gy)            # (L, b, nsx, ny)
        Dl = np.einsum("lbai,lbiy->lbay", Gx, T)           # (L, b, nx, ny)
        D[b0:b0 + len(sel)] = Dl.sum(axis=0).transpose(0, 2, 1)
    return D


def sigma_bar(w, sigma):
    """Layer-weight-averaged sigma(z) of the single-sigma approximation:
    sum_l (W_l / sum W) sigma_l(z), W_l = sum of layer l's weights. engines.py:75-80"""
    w = np.asarray(w, dtype=np.float64)
    wl = w.reshape(w.shape[0], -1).sum(axis=1)
    wl = wl / max(wl.sum(), 1e-30)
    return (np.asarray(sigma, dtype=np.float64) * wl[:, None]).sum(axis=0)


def separable_dose_averaged(x, y, sx, sy, w, idd, sigma):
    """FP64 single-sigma approximation (per_layer=False) as the PyTorch engine
    computes it, engines.py:243-254: A(z) = sum_l IDD_l(z) W_l (nz, nsy, nsx) and
    one Gaussian width sigma_bar(z) per slice, D[z] = Gy(z) A(z) Gx(z)^T, shape
    (nz, ny, nx).

    The NumPy engine's version (engines.py:184-191) pairs A's x-spot index with
    Gy's spot index in its first einsum and returns (nz, nx, ny) after its
    transpose; the two slips cancel when the x and y voxel axes and spot
    positions coincide (square centred fields, such as its fixture here), and
    the PyTorch engine is the drop-in target, so its order is the one kept."""
    w = np.asarray(w, dtype=np.float64)
    idd = np.asarray(idd, dtype=np.float64)
    L, nz = idd.shape
    A = (idd.T @ w.reshape(L, -1)).reshape(nz, w.shape[1], w.shape[2])
    sb = sigma_bar(w, sigma)
    Gx = lateral(x, sx, sb)                                  # (nz, nx, nsx)
    Gy = lateral(y, sy, sb)                                  # (nz, ny, nsy)
    T = np.einsum("zji,zvj->ziv", A, Gy)                     # (nz, nsx, ny)
    return np.einsum("zai,ziv->zva", Gx, T)                  # (nz, ny, nx)


def gaussian_field_weights(n_layers, n_spot=41, spot_spacing=0.5, spot_sigma=0.28, seed=0):
    """Weights of gaussian_impt_field: Gaussian envelope times (1 + 0.25 N(0,1))
    from default_rng(seed), clipped at 0. model.py:207-220"""
    rng = np.random.default_rng(seed)
    s = (np.arange(n_spot) - (n_spot - 1) / 2.0) * spot_spacing
    X, Y = np.meshgrid(s, s)
    env = np.exp(-(X**2 + Y**2) / (2.0 * spot_sigma**2))
    env = env / env.max()
    jit = 1.0 + 0.25 * rng.standard_normal((n_layers, n_spot, n_spot))
    return np.clip(env[None, :, :] * jit, 0.0, None)


def brute_force_dose(x, y, z_count, sx, sy, w, idd, sigma):
    """Textbook voxel x spot loop (tiny problems only). engines.py:102-136"""
    w = np.asarray(w, dtype=np.float64)
    L, nsy, nsx = w.shape
    D = np.zeros((z_count, len(y), len(x)))
    for kz in range(z_count):
        for l in range(L):
            s = max(float(sigma[l, kz]), SIGMA_FLOOR)
            c = float(idd[l, kz]) / (2.0 * math.pi * s * s)
            for j in range(nsy):
                gy = np.exp(-(y - sy[j]) ** 2 / (2.0 * s * s))
                for i in range(nsx):
                    if w[l, j, i] == 0.0:
                        continue
                    gx = np.exp(-(x - sx[i]) ** 2 / (2.0 * s * s))
                    D[kz] += w[l, j, i] * c * np.outer(gy, gx)
    return D


def relative_error(a, b, mask=None):
    """Global-normalised error metrics. engines.py:364-378"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is None:
        mask = np.ones(b.shape, dtype=bool)
    d = (a - b)[mask]
    ref = float(np.max(np.abs(b))) or 1.0
    if d.size == 0:
        return dict(max_abs_err=0.0, max_rel_err_global=0.0, rms_err=0.0,
                    rms_rel_global=0.0, n_vox=int(mask.sum()))
    mx = float(np.abs(d).max())
    rms = float(np.sqrt((d**2).mean()))
    return dict(max_abs_err=mx, max_rel_err_global=mx / ref, rms_err=rms,
                rms_rel_global=rms / ref, n_vox=int(mask.sum()))


def mean_rel_error(a, b, thr=0.10):
    """Mean and max pointwise relative error |a - b| / |b| on the voxels with
    |b| >= thr * max|b| (the dose-weighted region a 10 % gamma threshold keeps)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = np.abs(b) >= thr * np.abs(b).max()
    r = np.abs(a[m] - b[m]) / np.abs(b[m])
    return float(r.mean()), float(r.max())


# ---------------------------------------------------------------------------
# Gamma index (gamma.py:41-160)
# ---------------------------------------------------------------------------
def dta_offsets(spacing, dta_cm, max_vox=6):
    """Integer voxel offsets within the DTA sphere and their lengths. gamma.py:41-60"""
    dz, dy, dx = spacing
    n = [min(int(np.floor(dta_cm / s + 1e-9)), max_vox) for s in (dz, dy, dx)]
    offs, dist = [], []
    for a in range(-n[0], n[0] + 1):
        for b in range(-n[1], n[1] + 1):
            for c in range(-n[2], n[2] + 1):
                r = math.sqrt((a * dz) ** 2 + (b * dy) ** 2 + (c * dx) ** 2)
                if r <= dta_cm + 1e-12:
                    offs.append((a, b, c))
                    dist.append(r)
    return np.array(offs, dtype=np.int64), np.array(dist)


def _shifted(a, off):
    """a shifted by off with zero fill: out[p] = a[p + off]. gamma.py:63-75"""
    out = np.zeros_like(a)
    src, dst = [], []
    for o, n in zip(off, a.shape):
        src.append(slice(max(0, o), n + min(0, o)))
        dst.append(slice(max(0, -o), n + min(0, -o)))
    out[tuple(dst)] = a[tuple(src)]
    return out


def gamma_index(ref, ev, spacing=(0.2, 0.2, 0.2), dose_pct=3.0, dta_mm=3.0,
                threshold_pct=10.0, normalization="global", ref_dose=None, return_map=False):
    """Gamma index over integer-voxel DTA offsets. gamma.py:81-160"""
    ref = np.asarray(ref, dtype=np.float64)
    ev = np.asarray(ev, dtype=np.float64)
    dta = dta_mm / 10.0
    rmax = float(ref.max()) if ref_dose is None else float(ref_dose)
    if normalization == "global":
        crit = dose_pct / 100.0 * rmax
    else:
        crit = dose_pct / 100.0 * np.maximum(ref, 1e-12)
    mask = ref >= threshold_pct / 100.0 * rmax
    if mask.sum() == 0:
        mask = ref > 0
    offs, dist = dta_offsets(spacing, dta)
    g2 = np.full(ref.shape, np.inf)
    for off, r in zip(offs, dist):
        dd = (_shifted(ev, tuple(off)) - ref) / crit
        dr = r / dta
        np.minimum(g2, dr * dr + dd * dd, out=g2)
    gm = np.sqrt(g2)
    g = gm[mask]
    out = dict(pass_rate=float((g <= 1.0).mean() * 100.0), gamma_max=float(g.max()),
               gamma_mean=float(g.mean()), n_eval=int(g.size),
               n_pass=int((g <= 1.0).sum()), n_dta_offsets=len(offs), ref_max=rmax,
               gamma_p95=float(np.percentile(g, 95)), gamma_p99=float(np.percentile(g, 99)))
    if return_map:
        out["gamma"], out["mask"] = gm, mask
    return out


# ---------------------------------------------------------------------------
# Objective and its adjoint (weighted quadratic; FP64)
# ---------------------------------------------------------------------------
def presc_cube(grid, half=3.0, dose=2.0, w_in=1.0, w_out=0.1):
    """D_presc = dose inside the centred cube |x|,|y| <= half, z within
    +-half of the middle of the z range; voxel weights w_in/w_out."""
    x, y, z = grid_axes(grid)
    zc = 0.5 * (z[0] + z[-1])
    inside = ((np.abs(z - zc) <= half + 1e-9)[:, None, None]
              & (np.abs(y) <= half + 1e-9)[None, :, None]
              & (np.abs(x) <= half + 1e-9)[None, None, :])
    Dp = np.where(inside, dose, 0.0)
    wv = np.where(inside, w_in, w_out)
    return Dp, wv


def objective(D, Dp, wv):
    """J = sum_v wv (D_v - Dp_v)^2."""
    r = np.asarray(D, dtype=np.float64) - Dp
    return float(np.sum(wv * r * r))


def adjoint_w(x, y, sx, sy, idd, sigma, R):
    """dJ/dW[l,j,i] = sum_z IDD_l(z) (Gy_l(z)^T R_z Gx_l(z))[j,i] for a
    residual R = dJ/dD (nz, ny, nx). The transpose of separable_dose."""
    idd = np.asarray(idd, dtype=np.float64)
    L, nz = idd.shape
    out = np.zeros((L, len(sy), len(sx)))
    for kz in range(nz):
        Gx = lateral(x, sx, sigma[:, kz])   # (L, nx, nsx)
        Gy = lateral(y, sy, sigma[:, kz])   # (L, ny, nsy)
        Q = np.einsum("yx,lxi->lyi", R[kz], Gx)
        out += idd[:, kz, None, None] * np.einsum("lyj,lyi->lji", Gy, Q)
    return out


def adjoint_lut(x, y, sx, sy, w, idd, sigma, R):
    """dJ/dIDD_l(z) and dJ/dsigma_l(z) (L, nz) for a dose cotangent R, FP64.

    With P = Gy^T R_z Gx (unnormalised Gaussians g = exp(-d^2 / 2 s^2)) and
    n = 1 / (2 pi s^2):  dJ/dIDD = n <W, P>;
    dJ/ds = IDD n [ -(2/s) <W, P> + <W, Gy'^T R Gx + Gy^T R Gx'> ], g' = g d^2 / s^3,
    zero where sigma is at the 1e-9 floor."""
    w = np.asarray(w, dtype=np.float64)
    idd = np.asarray(idd, dtype=np.float64)
    sigma = np.asarray(sigma, dtype=np.float64)
    L, nz = idd.shape
    dx2 = (np.asarray(x)[:, None] - np.asarray(sx)[None, :]) ** 2
    dy2 = (np.asarray(y)[:, None] - np.asarray(sy)[None, :]) ** 2
    gi = np.zeros((L, nz))
    gs = np.zeros((L, nz))
    for l in range(L):
        for k in range(nz):
            s = max(float(sigma[l, k]), SIGMA_FLOOR)
            gx = np.exp(-dx2 / (2 * s * s))
            gy = np.exp(-dy2 / (2 * s * s))
            Q = R[k] @ gx
            P = gy.T @ Q
            P2 = (gy * dy2 / s**3).T @ Q + gy.T @ (R[k] @ (gx * dx2 / s**3))
            n = 1.0 / (2 * math.pi * s * s)
            wp = float(np.sum(w[l] * P))
            gi[l, k] = n * wp
            gs[l, k] = idd[l, k] * n * (-(2.0 / s) * wp + float(np.sum(w[l] * P2))) \
                if sigma[l, k] > SIGMA_FLOOR else 0.0
    return gi, gs


def spot_dose(x, y, sx_i, sy_j, idd_l, sigma_l):
    """Dose (nz, ny, nx) of one spot of unit weight in one layer."""
    s = np.clip(np.asarray(sigma_l, dtype=np.float64), SIGMA_FLOOR, None)
    gx = np.exp(-(np.asarray(x)[None, :] - sx_i) ** 2 / (2 * s[:, None] ** 2))
    gy = np.exp(-(np.asarray(y)[None, :] - sy_j) ** 2 / (2 * s[:, None] ** 2))
    c = np.asarray(idd_l) / (2 * np.pi * s * s)
    return c[:, None, None] * gy[:, :, None] * gx[:, None, :]
