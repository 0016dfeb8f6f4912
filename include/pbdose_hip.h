/*
 * pbdose_hip.h -- C ABI of the MI355X (gfx950) per-layer-exact Fermi-Eyges
 * pencil-beam dose library (libpbdose_hip.so).
 *
 * Model (one energy layer l, one depth slice z):
 *     dose[z] += IDD_l(z) * Gy_l(z) * W_l * Gx_l(z)^T
 *     G[k][i]  = exp(-(v_k - s_i)^2 / (2 s^2)) / (sqrt(2 pi) s),  s = max(sigma_l(z), 1e-9)
 * summed over layers, sigma exact per layer.
 *
 * Conventions shared by every entry point
 *   - all array arguments are DEVICE pointers, owned by the caller;
 *   - dose layout (nz, ny, nx), x fastest; weights (L, nsy, nsx), i fastest;
 *     lookup tables (L, nz); voxel centres o + k d; spot centres
 *     x0 + (i - (n-1)/2) ds;
 *   - `stream` is a hipStream_t (NULL = default stream); every call is
 *     asynchronous on it and never synchronises the device;
 *   - return 0 (PBH_OK) or a negative PBH_ERR_* code, readable through
 *     pbh_strerror(); no entry point exits the process;
 *   - no entry point allocates device memory: scratch space is a caller-owned
 *     workspace whose size the matching *_workspace_bytes() query returns.
 *
 * Reference interfaces replaced (penglinliu52/gpu-proton-dose-engine):
 *   engines.py:208-283  torch_separable(problem, per_layer=True)  -> pbh_forward
 *   engines.py:172-202  separable_numpy (the FP64 model)          -> pbh_forward
 *   physics.py:111-139  _range_table / csda_range / energy_from_range
 *                                                   -> pbh_range_table, pbh_interp_table
 *   physics.py:328-387  idd_pristine(normalize=False)
 *   physics.py:221-286  fermi_eyges_moments(...)["sigma_x"]      -> pbh_build_lut
 *   model.py:225-233    per-layer LUT loop of the field builders  -> pbh_build_lut
 *   gamma.py:41-160     dta_offsets / gamma_index                -> pbh_gamma_index
 *   gamma.py:213-261    gamma_index_torch                        -> pbh_gamma_index
 *   pencil_beam.cu:647-699 PB_ERR_* codes and pb_strerror         -> PBH_ERR_*, pbh_strerror
 *   pencil_beam.cu:717-730 pb_problem_validate                    -> pbh_validate
 *   pencil_beam.cu:907-919 pb_problem_bytes (sizing query)        -> *_workspace_bytes
 * The adjoint (pbh_adjoint, pbh_objective, pbh_lut_chain) and the
 * robust-scenario batch (pbh_batch_*) have no counterpart in the reference;
 * they serve inverse planning on top of the forward model above.
 */
#ifndef PBDOSE_HIP_H
#define PBDOSE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (pencil_beam.cu:647-653) ---- */
#define PBH_OK                 0
#define PBH_ERR_NULL_ARG      -1
#define PBH_ERR_BAD_DIM       -2
#define PBH_ERR_ZGRID_MISMATCH -3
#define PBH_ERR_HIP           -4
#define PBH_ERR_UNSUPPORTED   -5
#define PBH_ERR_NO_SMEM       -6
#define PBH_ERR_WORKSPACE     -7
#define PBH_ERR_SCRATCH       -8   /* problem too large for pbh_forward_noalloc's fixed scratch */

/* ---- arithmetic forms of the dose contraction ---- */
#define PBH_PREC_FP32        0  /* FP32 factors, FP32 MFMA, FP32 accumulation */
#define PBH_PREC_BF16        1  /* BF16 factors on MFMA, FP32 accumulation */
#define PBH_PREC_FP16        2  /* FP16 factors on MFMA, FP32 accumulation */
#define PBH_PREC_BF16X3      3  /* split hi+lo BF16, three products, everywhere */
#define PBH_PREC_FP16X3      4  /* split hi+lo FP16, three products, everywhere */
#define PBH_PREC_BF16X3_PEAK 5  /* BF16; split products only within 1 cm of each layer's Bragg peak */
#define PBH_PREC_FP16X3_PEAK 6  /* FP16; split products only within 1 cm of each layer's Bragg peak */

/* Voxel grid (model.py:27-79). */
typedef struct PbhGrid {
    int32_t nx, ny, nz, _pad;
    double dx, dy, dz;
    double ox, oy, oz;
} PbhGrid;

/* Spot lattice (model.py:85-131); weights live in a separate device array. */
typedef struct PbhLattice {
    int32_t n_sx, n_sy, n_layers, _pad;
    double dx_spot, dy_spot;
    double x0, y0;
} PbhLattice;

/* Beam / IDD model parameters (physics.py:292-325, model.py:193-205). */
typedef struct PbhBeamModel {
    double sigma_x0;            /* cm, initial beam size            (0.30) */
    double sigma_theta0;        /* rad, initial divergence          (0.0)  */
    double rho0;                /* position-angle correlation       (0.0)  */
    double energy_spread;       /* relative 1-sigma energy spread   (0.0033) */
    double attenuation_length;  /* cm, primary fluence attenuation  (150.0) */
    double fe_max_step;         /* cm, Fermi-Eyges fine-grid step    (0.02) */
    int32_t n_fine_idd;         /* fine binning grid points          (6000) */
    int32_t fe_n_sub;           /* Fermi-Eyges subdivision           (4)    */
    int32_t straggling;         /* 0 = Bohr integral, 1 = ICRU49 fit, 2 = none */
    int32_t straggling_depth;   /* 0: one range spread sigma_R for the whole CSDA curve
                                      (physics.py:370-373, the reference's model);
                                   1: depth-dependent spread, the fix the reference lists
                                      first (README.md:180-184): the CSDA dose binned at
                                      depth z' is spread with hypot(sigma_B(z'), sigma_E),
                                      sigma_B(z')^2 = int_{E(R0-z')}^{E0} BOHR/(beta^2 S^3) dE
                                      (the Bohr straggling accumulated up to z') and
                                      sigma_E = 1.77 energy_spread R0. Bohr or none only. */
    double icru_coef, icru_exp; /* ICRU49 fit sigma = coef * R0^exp (0.012, 0.935) */
} PbhBeamModel;

const char* pbh_strerror(int code);
const char* pbh_version(void);

/* Checks grid and lattice dimensions against the library limits
 * (pencil_beam.cu:717-730). */
int pbh_validate(const PbhGrid* g, const PbhLattice* lat);

/* ---------------------------------------------------------------------------
 * Physics look-up tables, FP64 on the device.
 * ------------------------------------------------------------------------- */

/* Number of points in the CSDA range table (3797). */
int pbh_range_table_size(void);

/* Builds the CSDA range table: E[n] (MeV) and R[n] (cm), n = pbh_range_table_size().
 * physics.py:111-127 */
int pbh_range_table(double* E, double* R, void* stream);

/* out[i] = linear interpolation of (xp, fp) at x[i] after clamping x to
 * [xp[0], xp[n-1]] (csda_range: xp = E, fp = R; energy_from_range: xp = R,
 * fp = E). physics.py:130-139 */
int pbh_interp_table(const double* xp, const double* fp, int n_table,
                     const double* x, double* out, int n, void* stream);

/* out[i] and dout[i] = d out[i] / d x[i]: pbh_interp_table and the slope of the
 * bracketing segment (0 where x is clamped to the table's ends). With
 * (xp, fp) = (R, E) this is dE/dR at R, the energy change of a range shift.
 * physics.py:136-139 and its derivative; no counterpart in the reference. */
int pbh_interp_table_slope(const double* xp, const double* fp, int n_table,
                           const double* x, double* out, double* dout, int n, void* stream);

/* Workspace for pbh_build_lut on the depths z0 + k dz, k < nz (nz >= 1), with
 * n_layers layers; 0 for invalid input. */
size_t pbh_lut_workspace_bytes(const PbhBeamModel* m, int nz, double z0, double dz, int n_layers);

/* Per-layer IDD (fluence-normalised, MeV/cm) and lateral sigma (cm) on the
 * depths z_k = z0 + k dz, k < nz, for layer energies energies[L] (device).
 * A single depth (nz = 1) follows physics.py:247-249: sigma is that of the
 * grid [0, z0] ([0, 1e-6] when z0 <= 0) at its end, the IDD is read at z0.
 * With m->straggling_depth = 1 the IDD uses the depth-dependent range spread
 * described at PbhBeamModel; sigma_R in layer_info is the spread at the end
 * of range, which both models share.
 * range_E, range_R: the CSDA range table from pbh_range_table (device), so
 * that it is built once; both NULL: the call builds it in its workspace.
 * Writes idd[L*nz], sigma[L*nz] (FP64) and, if layer_info != NULL,
 * layer_info[L][4] = {CSDA range R0, range spread sigma_R, depth of the IDD
 * maximum on the grid (first maximum), IDD maximum}.
 * physics.py:221-286, :328-387; model.py:225-233 */
int pbh_build_lut(const PbhBeamModel* m, int nz, double z0, double dz,
                  int n_layers, const double* energies,
                  const double* range_E, const double* range_R,
                  double* idd, double* sigma, double* layer_info,
                  void* workspace, size_t workspace_bytes, void* stream);

/* Workspace for pbh_build_lut_tangent; 0 for invalid input. */
size_t pbh_lut_tangent_workspace_bytes(const PbhBeamModel* m, int nz, double z0, double dz,
                                       int n_layers);

/* pbh_build_lut together with the tables' derivatives along a direction of the
 * layer energies: didd[l][k] = (d idd[l][k] / d E_l) denergies[l] (each
 * layer's tables depend on its own energy only), likewise dsigma. The tables equal
 * pbh_build_lut's bit for bit; the derivatives are those of the same discrete
 * algorithm (FP64 forward-mode differentiation of every kernel: range
 * straggling, CSDA binning on the energy-dependent fine grid, the Gaussian
 * convolution, the interpolation onto z, the scattering power inside the
 * Fermi-Eyges double sum), exact wherever the algorithm is differentiable
 * (not where a bin edge or fine-grid point crosses the range R0). For a range
 * shift d of layer l, E_l = E(R0_l + d) and denergies[l] = dE/dR there
 * (pbh_interp_table_slope). physics.py:111-139, :151-174, :221-286, :328-387
 * differentiated; no counterpart in the reference. */
int pbh_build_lut_tangent(const PbhBeamModel* m, int nz, double z0, double dz,
                          int n_layers, const double* energies, const double* denergies,
                          const double* range_E, const double* range_R,
                          double* idd, double* sigma, double* didd, double* dsigma,
                          double* layer_info, void* workspace, size_t workspace_bytes,
                          void* stream);

/* ---------------------------------------------------------------------------
 * Forward dose.
 * ------------------------------------------------------------------------- */
/* Workspace of pbh_forward (automatic layer chunk) and of pbh_forward_tuned
 * with a given layer_chunk (0 = automatic). */
size_t pbh_forward_workspace_bytes(const PbhGrid* g, const PbhLattice* lat, int precision);
size_t pbh_forward_workspace_bytes_tuned(const PbhGrid* g, const PbhLattice* lat, int precision,
                                         int layer_chunk);

/* dose (nz, ny, nx) FP32 := sum_l IDD_l Gy_l W_l Gx_l^T for weights w
 * (L, nsy, nsx) FP32 and FP64 tables idd, sigma (L, nz). The dose array is
 * overwritten, not accumulated. engines.py:208-283 (per_layer=True) */
int pbh_forward(const PbhGrid* g, const PbhLattice* lat,
                const float* w, const double* idd, const double* sigma,
                float* dose, int precision,
                void* workspace, size_t workspace_bytes, void* stream);

/* The forward dose in FP64: weights w (L, nsy, nsx), tables idd, sigma (L, nz)
 * and dose (nz, ny, nx) all FP64 (FP64 MFMA over per-tile factor tables where
 * the voxel and spot pitches are small-integer multiples of one length, FP64 FMA
 * on the vector ALUs otherwise), every voxel's
 * layers and spots summed in a fixed order (deterministic). No workspace; the
 * dose is overwritten. engines.py:208-283 torch_separable(dtype=torch.float64)
 * on a GPU (per_layer=True); model of engines.py:172-202. */
int pbh_forward_fp64(const PbhGrid* g, const PbhLattice* lat,
                     const double* w, const double* idd, const double* sigma,
                     double* dose, void* stream);

/* pbh_forward with an explicit layer chunk: the active layers (IDD != 0) of
 * each depth slice are cut into chunks of at most `layer_chunk` layers (0 =
 * automatic; >= L: one chunk), each computed by its own workgroup; a second
 * kernel adds the chunks' partial tiles to the dose in chunk order, so results
 * are deterministic and differ between chunk sizes only by FP32 summation
 * order. The workspace (pbh_forward_workspace_bytes_tuned) needs no
 * initialisation. Calls that share one workspace must not overlap. */
int pbh_forward_tuned(const PbhGrid* g, const PbhLattice* lat,
                      const float* w, const double* idd, const double* sigma,
                      float* dose, int precision, int layer_chunk,
                      void* workspace, size_t workspace_bytes, void* stream);

/* pbh_forward_tuned with a dose-tile height: tile_rows = 32, or 64 for the 16-bit
 * forms on grids whose voxel and spot pitches are 2 : 5 on both axes (other
 * geometries and the FP32 form return PBH_ERR_UNSUPPORTED); 0 = the default
 * (32). A 64 x 128 tile shares every gathered Gaussian operand between two
 * row blocks of its MFMA products and copies the weights and builds the x table
 * once for twice the rows, at fewer workgroups per CU. Results differ from
 * 32-row tiles only by FP32 summation order. The workspace is
 * pbh_forward_workspace_bytes_shaped with the same layer_chunk and tile_rows.
 * No counterpart in the reference (its tiles are cuBLAS's). */
int pbh_forward_shaped(const PbhGrid* g, const PbhLattice* lat,
                       const float* w, const double* idd, const double* sigma,
                       float* dose, int precision, int layer_chunk, int tile_rows,
                       void* workspace, size_t workspace_bytes, void* stream);
size_t pbh_forward_workspace_bytes_shaped(const PbhGrid* g, const PbhLattice* lat, int precision,
                                          int layer_chunk, int tile_rows);

/* pbh_forward_shaped that also records two caller-created HIP events
 * (hipEvent_t, passed as void*) on `stream`: ev_start right before the fused
 * contraction kernel and ev_end right after it, so that the caller can time
 * that kernel alone (the weight split before it and the layer-chunk combine
 * after it fall outside). Results equal pbh_forward_shaped's. No counterpart in
 * the reference (engines.py:208-283 times whole calls, benchmark.py:42-69). */
int pbh_forward_profiled(const PbhGrid* g, const PbhLattice* lat,
                         const float* w, const double* idd, const double* sigma,
                         float* dose, int precision, int layer_chunk, int tile_rows,
                         void* workspace, size_t workspace_bytes,
                         void* ev_start, void* ev_end, void* stream);

/* Layer chunk for pbh_forward_tuned from the number of active layers (IDD != 0)
 * of every depth slice, active_per_z[nz] (host memory): simulates the dispatch
 * of the forward's workgroups onto the resident slots (compute_units CUs, <= 0:
 * the current device's) and returns the chunk with the shortest makespan, or a
 * negative PBH_ERR_* code. Host only; no counterpart in the reference (its
 * layer_chunk, engines.py:208, bounds memory, not load balance). */
int pbh_forward_choose_chunk(const PbhGrid* g, const PbhLattice* lat, int precision,
                             const int32_t* active_per_z, int compute_units);

/* pbh_forward_tuned without a workspace argument and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The workspace is a fixed device array of
 * PBH_NOALLOC_SCRATCH_BYTES inside the library, one per GPU, held until the
 * process ends; a problem whose forward workspace
 * (pbh_forward_workspace_bytes_tuned with the same layer_chunk) is larger
 * returns PBH_ERR_SCRATCH (the 128^3, 41 x 41 x 40-layer field needs about
 * 17 MB at its automatic chunk). Calls share their GPU's array: calls that
 * overlap on one GPU race, so serialise them (one stream, or events). The first
 * call on a device resolves the array's address on that device (no allocation).
 * Results equal pbh_forward_tuned's bit for bit. */
#define PBH_NOALLOC_SCRATCH_BYTES (32u << 20)
int pbh_forward_noalloc(const PbhGrid* g, const PbhLattice* lat,
                        const float* w, const double* idd, const double* sigma,
                        float* dose, int precision, int layer_chunk, void* stream);

/* Tile height and layer chunk for pbh_forward_shaped from the active layers per
 * slice (as pbh_forward_choose_chunk): simulates the dispatch of 32- and 64-row
 * tiles (the latter for the 16-bit forms on the 2 : 5 pitch geometry) and returns
 * the chunk of the faster one, with its height in *tile_rows, or a negative
 * PBH_ERR_* code. Host only; no counterpart in the reference. */
int pbh_forward_choose_shape(const PbhGrid* g, const PbhLattice* lat, int precision,
                             const int32_t* active_per_z, int compute_units, int32_t* tile_rows);

/* The single-sigma approximation (per_layer=False, engines.py:184-191, :243-254)
 * as an ordinary problem for the forward entry points: from weights w (L, nsy,
 * nsx) FP64 and tables idd, sigma (L, nz) FP64 it writes a problem of nz layers,
 * w_avg (nz, nsy, nsx) FP32 = A(z) = sum_l IDD_l(z) w_l (engines.py:63-72),
 * idd_avg (nz, nz) FP64 = the identity, sigma_avg (nz, nz) FP64 = sigma_bar(z)
 * on every row, sigma_bar(z) = sum_l (W_l / sum W) sigma_l(z) with W_l the sum of
 * layer l's weights (engines.py:75-80; layer_sums[L] receives the W_l). A
 * forward of that problem on the same grid gives slice z the dose
 * Gy A(z) Gx^T with both factors of width sigma_bar(z). FP64 sums in layer
 * order. */
int pbh_average_sigma(const double* w, const double* idd, const double* sigma, int n_layers, int nz,
                      int nsy, int nsx, float* w_avg, double* idd_avg, double* sigma_avg,
                      double* layer_sums, void* stream);

/* pbh_average_sigma with the amplitude maps w_avg (nz, nsy, nsx) kept in FP64, for
 * pbh_forward_fp64: torch_separable(dtype=torch.float64, per_layer=False)
 * (engines.py:243-255 in FP64). */
int pbh_average_sigma_fp64(const double* w, const double* idd, const double* sigma, int n_layers, int nz,
                           int nsy, int nsx, double* w_avg, double* idd_avg, double* sigma_avg,
                           double* layer_sums, void* stream);

/* ---------------------------------------------------------------------------
 * Adjoint of the forward map (inverse planning).
 * ------------------------------------------------------------------------- */
size_t pbh_adjoint_workspace_bytes(const PbhGrid* g, const PbhLattice* lat, int precision);

/* grad_w (L, nsy, nsx) FP32 := sum_z IDD_l(z) Gy_l(z)^T R_z Gx_l(z) for a dose
 * cotangent R (nz, ny, nx). Deterministic (no atomics). precision:
 * PBH_PREC_FP32, or PBH_PREC_BF16 / PBH_PREC_BF16X3 / PBH_PREC_FP16 /
 * PBH_PREC_FP16X3 (16-bit MFMA, plain or split products, FP32 accumulation; FP16
 * with a power-of-two scale of R per tile group, so any scale of R is exact)
 * when no table gradient is requested and the
 * voxel and spot pitches have a small integer ratio (table geometry); other
 * combinations return PBH_ERR_UNSUPPORTED.
 * If grad_sigma != NULL it also writes grad_sigma (L, nz) FP64:
 *     d<R, dose>/d sigma_l(z)  and, if grad_idd != NULL, grad_idd (L, nz) FP64:
 *     d<R, dose>/d IDD_l(z). */
int pbh_adjoint(const PbhGrid* g, const PbhLattice* lat,
                const float* w, const double* idd, const double* sigma,
                const float* R, float* grad_w, double* grad_idd, double* grad_sigma,
                int precision, void* workspace, size_t workspace_bytes, void* stream);

/* pbh_adjoint with an explicit layer split: `split` workgroups share the layers
 * of each depth slice (1 to 16; 0 = automatic from the geometry). Results do not
 * depend on the split. */
int pbh_adjoint_tuned(const PbhGrid* g, const PbhLattice* lat,
                      const float* w, const double* idd, const double* sigma,
                      const float* R, float* grad_w, double* grad_idd, double* grad_sigma,
                      int precision, int split, void* workspace, size_t workspace_bytes, void* stream);

/* Layer split for pbh_adjoint_tuned from the number of active layers (IDD != 0)
 * of every depth slice, active_per_z[nz] (host memory): simulates the dispatch
 * of the adjoint's workgroups onto the resident slots (compute_units CUs, <= 0:
 * the current device's) and returns the split (1 to 16) with the shortest
 * makespan, or a negative PBH_ERR_* code. precision: PBH_PREC_FP32, PBH_PREC_BF16,
 * PBH_PREC_BF16X3, PBH_PREC_FP16 or PBH_PREC_FP16X3. Host only; no counterpart in the reference (its engines are
 * forward-only). */
int pbh_adjoint_choose_split(const PbhGrid* g, const PbhLattice* lat, int precision,
                             const int32_t* active_per_z, int compute_units);

/* R = 2 wv (dose - presc) and J = sum wv (dose - presc)^2 (J written to
 * j_out[0] as FP64). Fused residual of the weighted-quadratic objective. */
size_t pbh_objective_workspace_bytes(const PbhGrid* g);
int pbh_objective(const PbhGrid* g, const float* dose, const float* presc,
                  const float* vox_w, float* R, double* j_out,
                  void* workspace, size_t workspace_bytes, void* stream);

/* pbh_objective followed by pbh_adjoint_tuned without table gradients, in one
 * call: J = sum wv (dose - presc)^2 into j_out[0] (FP64) and grad_w = dJ/dw.
 * The residual R stays in the workspace, and J is summed by one more workgroup
 * of the adjoint's depth-reduction launch, one kernel launch fewer than the two
 * calls. Results equal the two calls' bit for bit. No counterpart in the
 * reference (its engines are forward-only). */
size_t pbh_objective_adjoint_workspace_bytes(const PbhGrid* g, const PbhLattice* lat, int precision);
int pbh_objective_adjoint(const PbhGrid* g, const PbhLattice* lat,
                          const float* w, const double* idd, const double* sigma,
                          const float* dose, const float* presc, const float* vox_w,
                          float* grad_w, double* j_out, int precision, int split,
                          void* workspace, size_t workspace_bytes, void* stream);

/* out[l] = sum_z (grad_idd[l,z] didd[l,z] + grad_sigma[l,z] dsigma[l,z]) for
 * (L, nz) FP64 arrays: the chain rule from the look-up tables to a per-layer
 * parameter (e.g. a range shift, with didd/dsigma the tables' derivatives). */
/* One step of inverse planning in one call: the forward dose of w
 * (pbh_forward_shaped with fwd_precision, layer_chunk, tile_rows) into dose, J =
 * sum vox_w (dose - presc)^2 (FP64) into *j_out and dJ/dw (adj_precision FP32, BF16
 * or BF16X3, layer split as pbh_adjoint_tuned) into grad_w. The layer-chunk combine
 * of the forward runs over every dose tile and also forms R = dJ/dD and J's partial
 * sums, so no launch re-reads the dose: the objective kernel of
 * pbh_objective_adjoint is not needed. dose and grad_w equal pbh_forward_shaped +
 * pbh_objective_adjoint's bit for bit; J differs by the grouping of its FP64 sum.
 * Workspace: pbh_forward_objective_adjoint_workspace_bytes. No counterpart in the
 * reference (its engines are forward-only). */
size_t pbh_forward_objective_adjoint_workspace_bytes(const PbhGrid* g, const PbhLattice* lat,
                                                     int fwd_precision, int layer_chunk, int tile_rows,
                                                     int adj_precision);
int pbh_forward_objective_adjoint(const PbhGrid* g, const PbhLattice* lat, const float* w,
                                  const double* idd, const double* sigma, float* dose,
                                  const float* presc, const float* vox_w, float* grad_w,
                                  double* j_out, int fwd_precision, int layer_chunk, int tile_rows,
                                  int adj_precision, int split, void* workspace,
                                  size_t workspace_bytes, void* stream);

int pbh_lut_chain(const double* grad_idd, const double* grad_sigma, const double* didd,
                  const double* dsigma, int n_layers, int nz, double* out, void* stream);

/* ---------------------------------------------------------------------------
 * Robust-scenario batches: many independent problems (fields x setup / range
 * scenarios; grids and lattices may differ per item) computed by one launch
 * per kernel stage: [peak depths] -> forward -> objective -> adjoint -> depth
 * reduction. A setup shift is expressed by the caller as a shifted lattice
 * origin (x0, y0) or depth origin (oz), a range scaling as tables built for the
 * scaled energies, so every item is an ordinary problem.
 * No counterpart in the reference (its engines run one field at a time,
 * engines.py:208-283).
 * ------------------------------------------------------------------------- */
typedef struct PbhBatchItem {
    PbhGrid grid;
    PbhLattice lattice;
    const float* w;          /* (L, nsy, nsx) spot weights */
    const double* idd;       /* (L, nz) */
    const double* sigma;     /* (L, nz) */
    float* dose;             /* (nz, ny, nx) out */
    const float* presc;      /* (nz, ny, nx) prescription, or NULL: forward only */
    const float* vox_w;      /* (nz, ny, nx) voxel weights, or NULL: all 1 */
    double* j_out;           /* J = sum vox_w (dose - presc)^2 (one FP64), or NULL */
    float* grad_w;           /* (L, nsy, nsx) dJ/dw out; required when presc != NULL */
    double* grad_idd;        /* (L, nz) dJ/dIDD out, or NULL */
    double* grad_sigma;      /* (L, nz) dJ/dsigma out, or NULL */
} PbhBatchItem;

typedef struct PbhBatchTuning {
    int32_t fwd_chunk;       /* forward layer chunk: at most this many active layers per
                                workgroup (0 = automatic per item) */
    int32_t adj_split;       /* adjoint layer split 1 to 16 (0 = per-item default) */
    int32_t group;           /* 0: one forward launch at the widest spot chunk;
                                1: one forward launch per chunk width */
    int32_t order;           /* 0: items in the given order; 1: largest work first */
    int32_t tile_rows;       /* forward dose-tile rows: 0 or 32, or 64 for the 16-bit forms
                                (items without 2 : 5 pitch tables keep 32) */
    int32_t persistent;      /* 0: one workgroup per (item, tile, layer chunk) of a forward
                                launch; 1: a persistent forward launch of as many workgroups
                                as fit on the GPU at once, each claiming the next
                                (item, tile, chunk) of the group's work list from a global
                                atomic counter (the last workgroup out resets it, so the
                                launch can be replayed from a graph) */
} PbhBatchTuning;

typedef struct PbhBatch PbhBatch;   /* opaque host-side plan (host memory only) */

/* Workspace of a batch (descriptor tables, residuals, per-slice partials).
 * tuning may be NULL (all fields 0). Returns 0 for invalid input. */
size_t pbh_batch_workspace_bytes(const PbhBatchItem* items, int n_items, int precision,
                                 const PbhBatchTuning* tuning);

/* Validates the items, plans the launches and copies the device descriptor
 * tables into the workspace (zero-filled by the caller before this call); it
 * synchronises `stream` once. The items' device arrays must stay valid until
 * pbh_batch_destroy. */
int pbh_batch_create(const PbhBatchItem* items, int n_items, int precision,
                     const PbhBatchTuning* tuning, void* workspace, size_t workspace_bytes,
                     void* stream, PbhBatch** out);

/* The adjoint of a batch follows `precision`: FP32 for PBH_PREC_FP32, otherwise
 * the forward's 16-bit type (BF16 or FP16) with single products for the plain
 * forms and split products for the x3 and near-peak forms; items that request
 * table gradients or have no table geometry use the FP32 adjoint. */

/* Runs the whole batch on `stream`: kernel launches only (no copies, memsets
 * or allocations), so it can be captured into a HIP graph and replayed.
 * Deterministic: results equal those of the single-problem entry points with
 * the same layer splits. Calls on one batch must not overlap. */
int pbh_batch_forward_adjoint(const PbhBatch* batch, void* stream);

/* Kernel launches one pbh_batch_forward_adjoint call makes. */
int pbh_batch_launch_count(const PbhBatch* batch);

void pbh_batch_destroy(PbhBatch* batch);

/* ---------------------------------------------------------------------------
 * Inverse planning: projected-gradient step of one field's spot weights from
 * the dJ/dw of its scenarios (e.g. the grad_w outputs of a robust batch).
 * No counterpart in the reference.
 * ------------------------------------------------------------------------- */
#define PBH_ROBUST_SUM    0   /* g = sum over scenarios, in scenario order */
#define PBH_ROBUST_WORST  1   /* g = gradient of the first scenario with the largest J */

/* w_new := min(max(w_old - step * g, w_min), w_max) elementwise over n weights
 * (FP32, no fused multiply-add). grads: device array of n_scen device pointers;
 * J: device FP64 array of n_scen objectives (read for PBH_ROBUST_WORST).
 * w_new may equal w_old. */
int pbh_weight_step(const float* w_old, float* w_new, const float* const* grads, const double* J,
                    int n_scen, size_t n, int mode, float step, float w_min, float w_max, void* stream);

/* ---------------------------------------------------------------------------
 * Gamma index with integer-voxel DTA offsets. gamma.py:41-160 (gamma_index);
 * also replaces gamma.py:213-261 (gamma_index_torch, which evaluates in FP32:
 * here the arithmetic is FP64 for both).
 * ------------------------------------------------------------------------- */
#define PBH_DTYPE_F32 0
#define PBH_DTYPE_F64 1

size_t pbh_gamma_workspace_bytes(const PbhGrid* g);

/* Number of DTA offsets for the grid spacing (dz, dy, dx) and dta_mm
 * (gamma.py:41-60: per axis at most min(floor(DTA/d + 1e-9), 6) voxels, kept
 * if |offset| <= DTA + 1e-12). Host-only. */
int pbh_gamma_count_offsets(const PbhGrid* g, double dta_mm);

/* ref, eval: (nz, ny, nx) doses of dtype PBH_DTYPE_F32 / _F64 on the grid g
 * (only its sizes and spacings are used). dose_pct, dta_mm, threshold_pct as
 * in gamma.py:81-86; ref_max <= 0 means max(ref) (computed on the device);
 * local != 0 selects local normalisation (crit = dose% * max(ref, 1e-12)).
 * Writes (device memory) counts[3] = {n_eval, n_pass, n_dta_offsets} and
 * stats[3] = {gamma_max, gamma_mean, ref_max used}; gamma_out (nz, ny, nx)
 * FP64 may be NULL. Evaluated voxels: ref >= threshold% * ref_max, or ref > 0
 * if there are none (gamma.py:139-141). Deterministic. */
int pbh_gamma_index(const PbhGrid* g, const void* ref, int ref_dtype, const void* eval, int eval_dtype,
                    double dose_pct, double dta_mm, double threshold_pct, double ref_max, int local,
                    double* gamma_out, unsigned long long* counts, double* stats,
                    void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PBDOSE_HIP_H */
