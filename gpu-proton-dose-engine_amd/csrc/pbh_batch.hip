// pbh_batch.hip -- robust-scenario batches: host planning of one launch per
// kernel stage over many independent problems.
//
// Every item is an ordinary problem (grid, lattice, weights, tables). The
// forward, objective, adjoint and depth-reduction kernels of the single-problem
// entry points are instantiated as batch kernels that find their item through a
// prefix table of workgroup counts (batch_item), so a batch of 63 field
// scenarios costs 6-8 launches instead of 6 per item, every launch has tens of
// thousands of workgroups, and the run is a pure sequence of kernel launches
// that a HIP graph can capture. Items whose kernels differ in template
// parameters (spot-chunk width, table or direct factors) go to separate
// launches ("groups").
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "pbh_batch.h"

using namespace pbh;

float* pbh_forward_ws_peaks(const FwdArgs& a, void* ws);
extern "C" int pbh_validate(const PbhGrid* g, const PbhLattice* lat);

namespace {


size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Launch {
    int first;      // first descriptor of the group
    int count;      // descriptors in the group
    int wg0_off;    // offset of the group's prefix table (count + 1 entries)
    int total;      // workgroups
    int key;        // FP32: padded chunk width; mixed: stage B K steps
    int mode;       // bit 0: table factors; bit 1: compile-time table slots (mixed)
};

bool is_mixed(int precision) { return precision != PBH_PREC_FP32; }

bool is_peak(int precision)
{
    return precision == PBH_PREC_BF16X3_PEAK || precision == PBH_PREC_FP16X3_PEAK;
}

PbhBatchTuning tuning_or_default(const PbhBatchTuning* t)
{
    PbhBatchTuning r{0, 0, 0, 0, 0, 0};
    if (t) r = *t;
    return r;
}

// Template key of an item's forward: FP32 -> padded chunk width rounded to
// {16, 24, 32, 40, 48}; mixed -> stage B K steps (1..3).
int fwd_key(const PbhBatchItem& it, int precision)
{
    if (is_mixed(precision)) return (chunk_pad(it.lattice.n_sx, 1) + 15) / 16;
    const int p = chunk_pad(it.lattice.n_sx, it.lattice.n_sy);
    return std::max(16, (p + 7) & ~7);
}

double fwd_work(const PbhBatchItem& it)
{
    const PbhGrid& g = it.grid;
    const PbhLattice& l = it.lattice;
    const double a = (double)l.n_sy * l.n_sx * g.ny + (double)g.nx * l.n_sx * g.ny;
    const double b = (double)l.n_sy * l.n_sx * g.nx + (double)g.ny * l.n_sy * g.nx;
    return 2.0 * l.n_layers * g.nz * std::min(a, b);
}

struct Layout {
    // offsets in the workspace
    size_t fwd_items, fwd_wg0, cmb_wg0, peak_items, obj_items, adj_items, adj_wg0, red_items, red_wg0, counters;
    size_t table_end;
    std::vector<size_t> fwd_ws, R, adj_part, obj_part;
    std::vector<size_t> fwd_ws_bytes;
    size_t total;
};

int check_items(const PbhBatchItem* items, int n, int precision)
{
    if (!items) return PBH_ERR_NULL_ARG;
    if (n <= 0 || n > (1 << 20)) return PBH_ERR_BAD_DIM;
    if (precision < PBH_PREC_FP32 || precision > PBH_PREC_FP16X3_PEAK) return PBH_ERR_UNSUPPORTED;
    for (int k = 0; k < n; ++k) {
        const PbhBatchItem& it = items[k];
        int rc = pbh_validate(&it.grid, &it.lattice);
        if (rc != PBH_OK) return rc;
        if (!it.w || !it.idd || !it.sigma || !it.dose) return PBH_ERR_NULL_ARG;
        if (it.presc && !it.grad_w) return PBH_ERR_NULL_ARG;
    }
    return PBH_OK;
}

int check_tuning(const PbhBatchTuning& t)
{
    if (t.fwd_chunk < 0) return PBH_ERR_UNSUPPORTED;
    if (t.adj_split < 0 || t.adj_split > 16)
        return PBH_ERR_UNSUPPORTED;
    if (t.group != 0 && t.group != 1) return PBH_ERR_UNSUPPORTED;
    if (t.order != 0 && t.order != 1) return PBH_ERR_UNSUPPORTED;
    if (t.tile_rows != 0 && t.tile_rows != 32 && t.tile_rows != 64) return PBH_ERR_UNSUPPORTED;
    if (t.persistent != 0 && t.persistent != 1) return PBH_ERR_UNSUPPORTED;
    return PBH_OK;
}

// Dose-tile rows of an item's forward: 64 when requested for a 16-bit form on the
// P = 2, Q = 5 table geometry, else 32.
int item_rows(const PbhBatchItem& it, int precision, const PbhBatchTuning& t)
{
    if (t.tile_rows != 64 || !is_mixed(precision)) return FT_TY;
    const FwdArgs a = make_fwd_args(&it.grid, &it.lattice, nullptr, nullptr, nullptr, nullptr, FT_TY);
    FwdPlan pl;
    if (pbh_plan_forward(a, 48, 64, nullptr, 0, a.L, &pl) != PBH_OK) return FT_TY;
    return (pl.table && pl.tx.P == 2 && pl.tx.Q == 5 && pl.ty.P == 2 && pl.ty.Q == 5) ? 64 : FT_TY;
}

Layout make_layout(const PbhBatchItem* items, int n, int precision, const PbhBatchTuning& t)
{
    Layout L;
    int n_obj = 0;
    for (int k = 0; k < n; ++k) n_obj += items[k].presc != nullptr;
    size_t off = 0;
    // descriptor tables; each group's prefix table needs count + 1 entries, and
    // there are at most n groups per stage
    L.fwd_items = off; off = align256(off + (size_t)n * sizeof(BatchFwdItem));
    L.fwd_wg0 = off; off = align256(off + (size_t)2 * n * sizeof(int) + sizeof(int));
    L.cmb_wg0 = off; off = align256(off + (size_t)(n + 1) * sizeof(int));
    L.peak_items = off; off = align256(off + (size_t)n * sizeof(BatchPeakItem));
    L.obj_items = off; off = align256(off + (size_t)std::max(n_obj, 1) * sizeof(BatchObjItem));
    L.adj_items = off; off = align256(off + (size_t)std::max(n_obj, 1) * sizeof(AdjArgs));
    L.adj_wg0 = off; off = align256(off + (size_t)2 * std::max(n_obj, 1) * sizeof(int) + sizeof(int));
    L.red_items = off; off = align256(off + (size_t)std::max(n_obj, 1) * sizeof(BatchRedItem));
    L.red_wg0 = off; off = align256(off + (size_t)(std::max(n_obj, 1) + 1) * sizeof(int));
    // claim counters of persistent forward launches, two per group (zero-filled by the
    // caller with the workspace, and left at zero by every launch)
    L.counters = off; off = align256(off + (size_t)2 * n * sizeof(unsigned));
    L.table_end = off;
    L.fwd_ws.resize(n);
    L.fwd_ws_bytes.resize(n);
    L.R.assign(n, 0);
    L.adj_part.assign(n, 0);
    L.obj_part.assign(n, 0);
    for (int k = 0; k < n; ++k) {
        const PbhBatchItem& it = items[k];
        const FwdArgs a = make_fwd_args(&it.grid, &it.lattice, nullptr, nullptr, nullptr, nullptr,
                                        item_rows(it, precision, t));
        const size_t fb = pbh_forward_ws_bytes(a, t.fwd_chunk);
        L.fwd_ws[k] = off;
        L.fwd_ws_bytes[k] = fb;
        off = align256(off + fb);
        if (it.presc) {
            const size_t nv = (size_t)it.grid.nx * it.grid.ny * it.grid.nz;
            const size_t ne = (size_t)it.lattice.n_layers * it.lattice.n_sy * it.lattice.n_sx;
            L.R[k] = off; off = align256(off + nv * sizeof(float));
            L.adj_part[k] = off; off = align256(off + (size_t)it.grid.nz * ne * sizeof(float));
            L.obj_part[k] = off; off = align256(off + (size_t)(OBJ_BATCH_BLOCKS + 1) * sizeof(double));
        }
    }
    L.total = off;
    return L;
}

}  // namespace

struct PbhBatch {
    int n = 0, precision = 0, n_obj = 0;
    bool peaks = false, sig = false;
    int max_wchunks = 0;     // mixed forms: largest (layer, spot chunk) count of an item
    int adj_prec = PBH_PREC_FP32;   // arithmetic of the adjoint kernels
    BatchFwdItem* d_fwd = nullptr;
    int* d_fwd_wg0 = nullptr;
    int* d_cmb_wg0 = nullptr;    // combine: prefix of the tiles of items with several layer chunks
    int cmb_total = 0;
    BatchPeakItem* d_peak = nullptr;
    BatchObjItem* d_obj = nullptr;
    AdjArgs* d_adj = nullptr;
    int* d_adj_wg0 = nullptr;
    BatchRedItem* d_red = nullptr;
    int* d_red_wg0 = nullptr;
    int red_total = 0;
    std::vector<Launch> fwd, adj;
    bool persistent = false;
    unsigned* d_counters = nullptr;    // two per forward group (persistent launches)
};

// Splits items (already in launch order) into groups of equal key and mode.
static std::vector<Launch> make_groups(const std::vector<int>& keys, const std::vector<char>& modes)
{
    std::vector<Launch> out;
    const int n = (int)keys.size();
    for (int k = 0; k < n; ++k) {
        if (out.empty() || out.back().key != keys[k] || out.back().mode != (int)modes[k])
            out.push_back(Launch{k, 0, 0, 0, keys[k], (int)modes[k]});
        out.back().count++;
    }
    int off = 0;
    for (auto& g : out) {
        g.wg0_off = off;
        off += g.count + 1;
    }
    return out;
}

extern "C" size_t pbh_batch_workspace_bytes(const PbhBatchItem* items, int n_items, int precision,
                                            const PbhBatchTuning* tuning)
{
    if (check_items(items, n_items, precision) != PBH_OK) return 0;
    const PbhBatchTuning t = tuning_or_default(tuning);
    if (check_tuning(t) != PBH_OK) return 0;
    return make_layout(items, n_items, precision, t).total;
}

extern "C" int pbh_batch_create(const PbhBatchItem* items, int n, int precision, const PbhBatchTuning* tuning,
                                void* workspace, size_t workspace_bytes, void* stream, PbhBatch** out)
{
    if (!out) return PBH_ERR_NULL_ARG;
    *out = nullptr;
    int rc = check_items(items, n, precision);
    if (rc != PBH_OK) return rc;
    const PbhBatchTuning t = tuning_or_default(tuning);
    if ((rc = check_tuning(t)) != PBH_OK) return rc;
    if (!workspace) return PBH_ERR_NULL_ARG;
    const Layout L = make_layout(items, n, precision, t);
    if (workspace_bytes < L.total) return PBH_ERR_WORKSPACE;
    char* ws = (char*)workspace;

    // ---- launch order: grouped by template key (widest first), then by work ----
    std::vector<int> key(n);
    for (int k = 0; k < n; ++k) key[k] = fwd_key(items[k], precision);
    const int kmax = *std::max_element(key.begin(), key.end());
    if (t.group == 0) std::fill(key.begin(), key.end(), kmax);
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k) order[k] = k;
    std::vector<int> rows(n);
    for (int k = 0; k < n; ++k) rows[k] = item_rows(items[k], precision, t);
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) {
        if (key[p] != key[q]) return key[p] > key[q];
        if (rows[p] != rows[q]) return rows[p] > rows[q];
        return t.order == 1 && fwd_work(items[p]) > fwd_work(items[q]);
    });

    // ---- forward descriptors ----
    std::vector<BatchFwdItem> hf(n);
    std::vector<BatchPeakItem> hp(n);
    std::vector<int> keys(n);
    std::vector<char> tabs(n);
    std::vector<int> wgs(n);
    const bool mixed = is_mixed(precision);
    for (int p = 0; p < n; ++p) {
        const int k = order[p];
        const PbhBatchItem& it = items[k];
        BatchFwdItem& f = hf[p];
        f.a = make_fwd_args(&it.grid, &it.lattice, it.w, it.idd, it.sigma, it.dose, item_rows(it, precision, t));
        const int kx = mixed ? 16 * key[k] : key[k];
        const int ky = mixed ? 64 : key[k];
        rc = pbh_plan_forward(f.a, kx, ky, ws + L.fwd_ws[k], L.fwd_ws_bytes[k], t.fwd_chunk, &f.pl);
        if (rc != PBH_OK) return rc;
        // batch items scan their layers in the forward kernel: measured faster here than the
        // precomputed lists the single-problem mixed forward uses, also with the lists path
        // compiled into the batched kernel without spills (168 VGPRs)
        f.pl.layers = nullptr;
        f.pl.nlayers = nullptr;
        f.oz = it.grid.oz;
        f.dz = it.grid.dz;
        f.r_peak = is_peak(precision) ? pbh_forward_ws_peaks(f.a, ws + L.fwd_ws[k]) : nullptr;
        hp[p] = BatchPeakItem{it.idd, it.lattice.n_layers, it.grid.nz, it.grid.oz, it.grid.dz, (float*)f.r_peak};
        keys[p] = key[k];
        // forward mode: bit 0 table factors, bit 1 compile-time slots, bit 2 64-row tiles
        tabs[p] = (char)((mixed ? pbh_mixed_slot_mode(f.pl) : ((f.pl.table ? 1 : 0) | (fp32_q5(f.pl) ? 2 : 0))) |
                         (f.a.tile_rows == 64 ? 4 : 0));
        const long long w = (long long)f.a.tiles_x * f.a.tiles_y * f.a.nz * f.pl.nchunk;
        if (w > 0x3fffffffLL) return PBH_ERR_UNSUPPORTED;
        wgs[p] = (int)w;
    }
    PbhBatch* b = new (std::nothrow) PbhBatch;
    if (!b) return PBH_ERR_UNSUPPORTED;
    b->n = n;
    b->precision = precision;
    b->peaks = is_peak(precision);
    // the adjoint follows the forward's arithmetic: FP32, or the forward's 16-bit type with
    // single products (plain forms) or split products (x3 and near-peak forms)
    b->adj_prec = precision == PBH_PREC_FP32 ? PBH_PREC_FP32
                  : precision == PBH_PREC_BF16 ? PBH_PREC_BF16
                  : precision == PBH_PREC_FP16 ? PBH_PREC_FP16
                  : (precision == PBH_PREC_FP16X3 || precision == PBH_PREC_FP16X3_PEAK) ? PBH_PREC_FP16X3
                                                                                          : PBH_PREC_BF16X3;
    b->fwd = make_groups(keys, tabs);
    for (int k = 0; k < n; ++k)
        b->max_wchunks = std::max(b->max_wchunks, items[k].lattice.n_layers * chunk_count(items[k].lattice.n_sy) *
                                                      chunk_count(items[k].lattice.n_sx));
    std::vector<int> cmb_wg0(n + 1, 0);
    {
        long long acc = 0;
        for (int p = 0; p < n; ++p) {
            cmb_wg0[p] = (int)acc;
            if (hf[p].pl.nchunk > 1) acc += (long long)hf[p].a.tiles_x * hf[p].a.tiles_y * hf[p].a.nz;
        }
        cmb_wg0[n] = (int)acc;
        b->cmb_total = (int)acc;
    }
    std::vector<int> fwd_wg0(2 * n + 1, 0);
    for (auto& g : b->fwd) {
        long long acc = 0;
        for (int q = 0; q < g.count; ++q) {
            fwd_wg0[g.wg0_off + q] = (int)acc;
            acc += wgs[g.first + q];
        }
        fwd_wg0[g.wg0_off + g.count] = (int)acc;
        if (acc > 0x7fffffffLL) {
            delete b;
            return PBH_ERR_UNSUPPORTED;
        }
        g.total = (int)acc;
    }

    // ---- objective / adjoint / reduction descriptors (items with a prescription) ----
    std::vector<BatchObjItem> ho;
    std::vector<AdjArgs> ha;
    std::vector<BatchRedItem> hr;
    std::vector<int> akeys, aw;
    std::vector<char> atabs;
    for (int p = 0; p < n; ++p) {
        const int k = order[p];
        const PbhBatchItem& it = items[k];
        if (!it.presc) continue;
        float* R = (float*)(ws + L.R[k]);
        float* part = (float*)(ws + L.adj_part[k]);
        const size_t nv = (size_t)it.grid.nx * it.grid.ny * it.grid.nz;
        ho.push_back(BatchObjItem{it.dose, it.presc, it.vox_w, R, (double*)(ws + L.obj_part[k]),
                                  it.j_out ? it.j_out : (double*)(ws + L.obj_part[k]) + OBJ_BATCH_BLOCKS, nv});
        AdjArgs A;
        rc = pbh_make_adj_args(&it.grid, &it.lattice, it.w, it.idd, it.sigma, R, part, it.grad_idd,
                               it.grad_sigma, t.adj_split, &A);
        if (rc != PBH_OK) {
            delete b;
            return rc;
        }
        ha.push_back(A);
        b->sig = b->sig || it.grad_sigma != nullptr;
        akeys.push_back(0);
        // adjoint mode: bit 0 table factors, bit 1 16-bit kernel, bit 2 P = 2, Q = 5 slots
        const bool mix = b->adj_prec != PBH_PREC_FP32 && A.tab && !it.grad_idd && !it.grad_sigma &&
                         pbh_adj_mixed_fits(A);
        atabs.push_back((char)((mix ? 3 : (A.tab ? 1 : 0)) | (pbh_adj_mixed_q5(A) ? 4 : 0)));
        aw.push_back(xcd_grid(it.grid.nz, A.split));     // a multiple of 8 (XCD-aware slice order)
        const size_t ne = (size_t)it.lattice.n_layers * it.lattice.n_sy * it.lattice.n_sx;
        hr.push_back(BatchRedItem{part, it.idd, it.grad_w, ne, it.grid.nz,
                                  it.lattice.n_sy * it.lattice.n_sx});
    }
    b->n_obj = (int)ho.size();
    std::vector<int> adj_wg0, red_wg0;
    if (b->n_obj) {
        // adjoint groups by table mode (stable: tables first keeps the launch order)
        std::vector<int> ao(b->n_obj);
        for (int q = 0; q < b->n_obj; ++q) ao[q] = q;
        std::stable_sort(ao.begin(), ao.end(), [&](int p, int q) { return atabs[p] > atabs[q]; });
        std::vector<AdjArgs> ha2(b->n_obj);
        std::vector<char> atabs2(b->n_obj);
        std::vector<int> aw2(b->n_obj);
        for (int q = 0; q < b->n_obj; ++q) {
            ha2[q] = ha[ao[q]];
            atabs2[q] = atabs[ao[q]];
            aw2[q] = aw[ao[q]];
        }
        ha.swap(ha2);
        b->adj = make_groups(akeys, atabs2);
        adj_wg0.assign(2 * b->n_obj + 1, 0);
        for (auto& g : b->adj) {
            long long acc = 0;
            for (int q = 0; q < g.count; ++q) {
                adj_wg0[g.wg0_off + q] = (int)acc;
                acc += aw2[g.first + q];
            }
            if (acc > 0x7fffffffLL) {
                delete b;
                return PBH_ERR_UNSUPPORTED;
            }
            adj_wg0[g.wg0_off + g.count] = (int)acc;
            g.total = (int)acc;
        }
        red_wg0.assign(b->n_obj + 1, 0);
        long long acc = 0;
        for (int q = 0; q < b->n_obj; ++q) {
            red_wg0[q] = (int)acc;
            acc += (long long)(hr[q].n / (size_t)hr[q].per_layer) *
                   ((hr[q].per_layer + RED_ELEMS - 1) / RED_ELEMS);
        }
        red_wg0[b->n_obj] = (int)acc;
        if (acc > 0x7fffffffLL) {
            delete b;
            return PBH_ERR_UNSUPPORTED;
        }
        b->red_total = (int)acc;
    }

    // ---- copy the tables to the device ----
    b->d_fwd = (BatchFwdItem*)(ws + L.fwd_items);
    b->d_fwd_wg0 = (int*)(ws + L.fwd_wg0);
    b->d_cmb_wg0 = (int*)(ws + L.cmb_wg0);
    b->d_peak = (BatchPeakItem*)(ws + L.peak_items);
    b->d_obj = (BatchObjItem*)(ws + L.obj_items);
    b->d_adj = (AdjArgs*)(ws + L.adj_items);
    b->d_adj_wg0 = (int*)(ws + L.adj_wg0);
    b->d_red = (BatchRedItem*)(ws + L.red_items);
    b->d_red_wg0 = (int*)(ws + L.red_wg0);
    b->d_counters = (unsigned*)(ws + L.counters);
    b->persistent = t.persistent != 0 && is_mixed(precision);
    std::vector<char> host(L.table_end, 0);
    std::memcpy(host.data() + L.fwd_items, hf.data(), hf.size() * sizeof(BatchFwdItem));
    std::memcpy(host.data() + L.fwd_wg0, fwd_wg0.data(), fwd_wg0.size() * sizeof(int));
    std::memcpy(host.data() + L.cmb_wg0, cmb_wg0.data(), cmb_wg0.size() * sizeof(int));
    std::memcpy(host.data() + L.peak_items, hp.data(), hp.size() * sizeof(BatchPeakItem));
    if (b->n_obj) {
        std::memcpy(host.data() + L.obj_items, ho.data(), ho.size() * sizeof(BatchObjItem));
        std::memcpy(host.data() + L.adj_items, ha.data(), ha.size() * sizeof(AdjArgs));
        std::memcpy(host.data() + L.adj_wg0, adj_wg0.data(), adj_wg0.size() * sizeof(int));
        std::memcpy(host.data() + L.red_items, hr.data(), hr.size() * sizeof(BatchRedItem));
        std::memcpy(host.data() + L.red_wg0, red_wg0.data(), red_wg0.size() * sizeof(int));
    }
    hipStream_t s = (hipStream_t)stream;
    if (hipMemcpyAsync(ws, host.data(), L.table_end, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        delete b;
        return PBH_ERR_HIP;
    }
    *out = b;
    return PBH_OK;
}

extern "C" int pbh_batch_forward_adjoint(const PbhBatch* b, void* stream)
{
    if (!b) return PBH_ERR_NULL_ARG;
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if (b->peaks && (rc = pbh_launch_peak_batch(b->d_peak, b->n, s)) != PBH_OK) return rc;
    if (is_mixed(b->precision) && (rc = pbh_launch_wsplit_batch(b->d_fwd, b->n, b->max_wchunks, b->precision, s)) != PBH_OK)
        return rc;
    for (const Launch& g : b->fwd) {
        if (b->precision == PBH_PREC_FP32)
            rc = pbh_launch_fwd_fp32_batch(b->d_fwd + g.first, b->d_fwd_wg0 + g.wg0_off, g.count, g.total, g.key,
                                           g.mode & 1, (g.mode & 2) != 0, s);
        else
            rc = pbh_launch_fwd_mixed_batch(b->d_fwd + g.first, b->d_fwd_wg0 + g.wg0_off, g.count, g.total,
                                            b->precision, g.key, g.mode & 1, (g.mode & 2) != 0,
                                            (g.mode & 4) ? 64 : 32,
                                            b->persistent ? b->d_counters + 2 * (&g - b->fwd.data()) : nullptr, s);
        if (rc != PBH_OK) return rc;
    }
    if (b->cmb_total > 0 && (rc = pbh_launch_combine_batch(b->d_fwd, b->d_cmb_wg0, b->n, b->cmb_total, s)) != PBH_OK)
        return rc;
    if (b->n_obj == 0) return PBH_OK;
    if ((rc = pbh_launch_obj_batch(b->d_obj, b->n_obj, s)) != PBH_OK) return rc;
    for (const Launch& g : b->adj) {
        if (g.mode & 2)
            rc = pbh_launch_adj_mixed_batch(b->d_adj + g.first, b->d_adj_wg0 + g.wg0_off, g.count, g.total,
                                            b->adj_prec, (g.mode & 4) != 0, s);
        else
            rc = pbh_launch_adj_batch(b->d_adj + g.first, b->d_adj_wg0 + g.wg0_off, g.count, g.total, g.mode & 1,
                                      b->sig, (g.mode & 4) != 0, s);
        if (rc != PBH_OK) return rc;
    }
    return pbh_launch_reduce_batch(b->d_red, b->d_red_wg0, b->n_obj, b->red_total, s);
}

extern "C" int pbh_batch_launch_count(const PbhBatch* b)
{
    if (!b) return PBH_ERR_NULL_ARG;
    int n = (b->peaks ? 1 : 0) + (is_mixed(b->precision) ? 1 : 0) + (int)b->fwd.size() + (b->cmb_total > 0 ? 1 : 0);
    if (b->n_obj) n += 2 + (int)b->adj.size() + 1;
    return n;
}

extern "C" void pbh_batch_destroy(PbhBatch* b)
{
    delete b;
}
