// C ABI of the QML engine: chunk registration (the Chunk constructor's
// bookkeeping), batch planning, and the per-iteration launch sequence.
#include "lyspeq_qml.h"
#include "qml_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

using namespace qml;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return fail(QML_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));   \
    } while (0)

constexpr double MY_PI = 3.14159265358979323846;
constexpr double DOUBLE_EPSILON = 1e-15;

struct TableHost {
    int R;
    double dv;
    int nv, nz;
    double v_length;
    double *d_qt = nullptr, *d_q2 = nullptr, *d_s = nullptr;
};

struct ChunkHost {
    qml_chunk_info info{};
    std::vector<double> wave, delta, ivar, U;  // wave in units of lambda_Lya; U is nu x N
    int N = 0, Np = 0, nu = 0;
    int n = 0, m = 0, kbu = 0, nzq = 0, zm0 = 0, nq = 0, fis = 0;
    int nvu = 0;
    long long in_off = 0;
    // resolution matrix: rows of R (N x rnb) and the fine grid (rM values)
    int rm = 0, rnb = 0, ros = 1, rcol0 = 0, rM = 0;
    int rdia = 0;   // the matrix came in diagonal storage (its grid is the pixels')
    std::vector<double> rband, fine;
    long long rin_off = 0;
};

struct Batch {
    std::vector<int> ids;
    ChunkDev *d_chunks = nullptr;
    Work *d_wl = nullptr, *d_wr = nullptr, *d_wg = nullptr;
    int nwl = 0, nwr = 0, nwg = 0;
    // stack slabs: the left and right work lists of slab s are
    // [sl_off[s], sl_off[s + 1]) of d_wl and [sr_off[s], sr_off[s + 1]) of d_wr
    int nslab = 1;
    std::vector<int> sl_off, sr_off;
    std::vector<hipEvent_t> evS;   // 3 per slab after the first (xgen, left, right), timing only
    int np_max = 0, nvu_max = 0, nu_max = 0;
    int dp_rb = 1, dp_cb = 1;   // d-pass operand blocks of the batch
    int nq_max = 0;             // most N_Q_MATRICES of a chunk in the batch
    int rm_kbu = -1;   // resolution-matrix chunks: largest KBIN_UPP
    RmLayout rm_lay;   // and the LDS layout of their sandwich launch
    hipEvent_t evA[7]{};   // stage A (assembly .. solves) on stream st
    hipEvent_t evB[6]{};   // stage B (stack generation .. Gram) on stream st2
    hipEvent_t evR{};      // after the resolution-matrix sandwiches (stage A)
};

double median_sorted(const std::vector<double> &a) {
    const size_t n = a.size();
    if (n % 2 == 0) return (a[n / 2 - 1] + a[n / 2]) / 2;
    return a[n / 2];
}

// One-sided Jacobi SVD of the N x nt column set V (column-major, columns
// contiguous).  Returns singular values (descending) and overwrites V with
// the matching left singular vectors.
void jacobi_svd(std::vector<double> &V, int N, int nt, std::vector<double> &sv) {
    auto col = [&](int j) { return V.data() + (size_t)j * N; };
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < nt; ++p)
            for (int q = p + 1; q < nt; ++q) {
                double a = 0, b = 0, g = 0;
                const double *vp = col(p), *vq = col(q);
                for (int i = 0; i < N; ++i) { a += vp[i] * vp[i]; b += vq[i] * vq[i]; g += vp[i] * vq[i]; }
                if (std::fabs(g) <= 1e-17 * std::sqrt(a * b) || g == 0.0) continue;
                rotated = true;
                const double zeta = (b - a) / (2 * g);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                const double cs = 1 / std::sqrt(1 + t * t), sn = cs * t;
                double *wp = col(p), *wq = col(q);
                for (int i = 0; i < N; ++i) {
                    const double x = wp[i], y = wq[i];
                    wp[i] = cs * x - sn * y;
                    wq[i] = sn * x + cs * y;
                }
            }
        if (!rotated) break;
    }
    std::vector<double> s(nt);
    for (int j = 0; j < nt; ++j) {
        double a = 0;
        for (int i = 0; i < N; ++i) a += col(j)[i] * col(j)[i];
        s[j] = std::sqrt(a);
    }
    std::vector<int> order(nt);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return s[x] > s[y]; });
    std::vector<double> out((size_t)N * nt);
    sv.assign(nt, 0.0);
    for (int j = 0; j < nt; ++j) {
        const int o = order[j];
        sv[j] = s[o];
        for (int i = 0; i < N; ++i) out[(size_t)j * N + i] = (s[o] > 0) ? col(o)[i] / s[o] : 0.0;
    }
    V.swap(out);
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

struct qml_ctx {
    int device = 0, prec = 0;
    double scratch_gb = 32.0;
    int max_batch = 256;
    int np_buckets = 0;   // 1: a batch also ends where the padded size changes
    int pipeline = 1;
    // staging of the split-BF16 products' operands (qml_set_kernel_stages):
    // direct-to-LDS loads with two stages for all three products gave the
    // fastest overlapped iteration on the 4096-chunk DESI batch
    int nst_left = 3, nst_right = 3, nst_gram = 3;
    int slab = 0;                     // stack members per slab (0: the whole stack at once)
    int poison = 0;   // LYSPEQ_QML_POISON=1 at commit: NaN-filled scratch and LDS
    hipStream_t st = nullptr, st2 = nullptr;
    bool bins_set = false, committed = false;
    bool plan_ok = false;   // a batch plan is in place (cleared by release_plan)
    std::vector<double> kedges, kcenters, zcenters;
    qml_bins bins{};
    BinsDev binsdev{};
    int T = 0;
    std::vector<TableHost> tables;
    TableDev *d_tables = nullptr;
    std::vector<ChunkHost> chunks;

    double *d_in = nullptr;
    std::vector<ChunkOut> outs;
    ChunkOut *d_outs = nullptr;
    double *d_fc = nullptr, *d_dbt = nullptr;
    int *d_status = nullptr;
    int *d_skip = nullptr;            // chunks left out of the accumulation (qml_set_chunk_active)
    // stage-A arenas are double-buffered so that batch b+1 is factorised
    // while batch b runs its dense products
    double *d_mat[2] = {}, *d_vec[2] = {}, *d_pj[2] = {}, *d_sj[2] = {};
    void *d_lic[2] = {};
    double *d_gram = nullptr;
    void *d_X = nullptr, *d_Y = nullptr;
    void *d_G = nullptr;   // the G store when the stack is slabbed (else G is written over X)
    double max_chunk_bytes = 0;   // the largest chunk's scratch in the current plan
    double *d_theta = nullptr, *d_F = nullptr, *d_dbtsum = nullptr;
    double *d_rin = nullptr;   // resolution-matrix rows and fine grids
    std::vector<Batch> batches;
    hipEvent_t ev_begin{}, ev_end{}, ev_acc{};
    hipEvent_t ev_fork{}, ev_join{};   // order stage B's stream inside one iteration
    qml_timings tm{};
    // pinned host staging of theta and of the sums: the iteration's copies
    // read and write fixed addresses, so a captured sequence can replay them
    double *h_theta = nullptr, *h_F = nullptr, *h_dbt = nullptr;
    // qml_set_graph: the iteration's launch sequence captured once and
    // replayed; dropped when the stage overlap setting changes
    int use_graph = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_pipeline = -1;
};

namespace {
void drop_graph(qml_ctx *c) {
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->graph_pipeline = -1;
}
void release_plan(qml_ctx *c);
int plan_batches(qml_ctx *c);
}  // namespace

extern "C" {

const char *qml_last_error(void) { return g_err.c_str(); }

int qml_create(qml_ctx **out, const qml_opts *opts) {
    if (!out) return fail(QML_ERR_ARG, "qml_create: null output pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(QML_ERR_HIP, "qml_create: no HIP device available");
    auto *c = new qml_ctx();
    if (opts) {
        c->device = opts->device;
        c->prec = opts->precision;
        if (opts->scratch_gb > 0) c->scratch_gb = opts->scratch_gb;
        if (opts->max_batch_chunks > 0) c->max_batch = opts->max_batch_chunks;
    }
    if (c->prec < 0 || c->prec > 3) {
        delete c;
        return fail(QML_ERR_ARG, "qml_create: unknown precision");
    }
    if (c->device < 0 || c->device >= ndev) {
        delete c;
        return fail(QML_ERR_ARG, "qml_create: device ordinal out of range");
    }
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess ||
        hipStreamCreate(&c->st2) != hipSuccess ||
        hipEventCreate(&c->ev_begin) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess ||
        hipEventCreate(&c->ev_acc) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        delete c;
        return fail(QML_ERR_HIP, "qml_create: stream/event creation failed");
    }
    // per-device address of the allocation-free median bootstrap's scratch
    if (const int e = boot_noalloc_setup(c->device); e != (int)hipSuccess) {
        qml_destroy(c);
        return fail(QML_ERR_HIP, std::string("qml_create: scratch symbol lookup: ") +
                                     hipGetErrorString((hipError_t)e));
    }
    *out = c;
    return QML_OK;
}

int qml_destroy(qml_ctx *c) {
    if (!c) return QML_OK;
    hipSetDevice(c->device);
    if (c->st) hipStreamSynchronize(c->st);
    if (c->st2) hipStreamSynchronize(c->st2);
    for (auto &t : c->tables) { hipFree(t.d_qt); hipFree(t.d_q2); hipFree(t.d_s); }
    release_plan(c);
    void *ptrs[] = {c->d_tables, c->d_in, c->d_outs, c->d_fc, c->d_dbt, c->d_status, c->d_skip,
                    c->d_theta, c->d_F, c->d_dbtsum, c->d_rin};
    for (void *p : ptrs) if (p) hipFree(p);
    if (c->ev_begin) hipEventDestroy(c->ev_begin);
    if (c->ev_end) hipEventDestroy(c->ev_end);
    if (c->ev_acc) hipEventDestroy(c->ev_acc);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    double *pinned[] = {c->h_theta, c->h_F, c->h_dbt};
    for (double *p : pinned) if (p) hipHostFree(p);
    if (c->st) hipStreamDestroy(c->st);
    if (c->st2) hipStreamDestroy(c->st2);
    delete c;
    return QML_OK;
}

int qml_set_bins(qml_ctx *c, const qml_bins *b) {
    if (!c || !b) return fail(QML_ERR_ARG, "qml_set_bins: null argument");
    if (c->committed || !c->chunks.empty() || !c->tables.empty())
        return fail(QML_ERR_ARG, "qml_set_bins: bins must be set before tables and chunks");
    if (b->nk <= 0 || b->nk > MAX_NK - 1 || b->nz <= 0 || b->nz > MAX_ZB || !b->kband_edges || !b->zbin_centers ||
        !(b->z_bin_width > 0))
        return fail(QML_ERR_ARG, "qml_set_bins: invalid bins");
    if (b->redshift_growth_on > 0 && !std::isfinite(b->fiducial_redshift_power))
        return fail(QML_ERR_ARG, "qml_set_bins: redshift growth needs a finite fiducial_redshift_power");
    c->bins = *b;
    c->kedges.assign(b->kband_edges, b->kband_edges + b->nk + 1);
    c->zcenters.assign(b->zbin_centers, b->zbin_centers + b->nz);
    c->kcenters.resize(b->nk);
    for (int k = 0; k < b->nk; ++k) c->kcenters[k] = (c->kedges[k] + c->kedges[k + 1]) / 2.;
    c->bins.kband_edges = c->kedges.data();
    c->bins.zbin_centers = c->zcenters.data();
    std::memset(&c->binsdev, 0, sizeof(c->binsdev));
    for (int z = 0; z < b->nz; ++z) c->binsdev.zc[z] = c->zcenters[z];
    for (int k = 0; k <= b->nk; ++k) c->binsdev.ke[k] = c->kedges[k];
    c->binsdev.dz = b->z_bin_width;
    c->binsdev.nz = b->nz;
    c->binsdev.nk = b->nk;
    c->binsdev.method = b->z_binning_method;
    c->binsdev.turn_off_sfid = b->turn_off_sfid;
    c->binsdev.growth = b->redshift_growth_on > 0;
    c->binsdev.growth_b = b->fiducial_redshift_power;
    c->T = b->nk * b->nz;
    int nvec = 0;
    if (b->cont_logl_order >= 0 || b->cont_lam_order >= 0) {
        nvec = 1 + std::max(0, b->cont_logl_order) + std::max(0, b->cont_lam_order);
    }
    if (nvec > MAX_NU) return fail(QML_ERR_LIMIT, "qml_set_bins: more than 4 continuum templates");
    c->bins_set = true;
    return QML_OK;
}

int qml_add_table(qml_ctx *c, int R_fwhm, double dv_kms, int nv, double v_length, int nz,
                  const double *deriv, const double *signal, int *r_index) {
    if (!c || !c->bins_set) return fail(QML_ERR_ARG, "qml_add_table: set bins first");
    if (c->committed) return fail(QML_ERR_ARG, "qml_add_table: context already committed");
    if (nv < 2 || !deriv || (!signal && !c->bins.turn_off_sfid) || (!c->bins.turn_off_sfid && nz < 2))
        return fail(QML_ERR_ARG, "qml_add_table: invalid table");
    hipSetDevice(c->device);
    const int nk = c->bins.nk;
    TableHost t;
    t.R = R_fwhm;
    t.dv = dv_kms;
    t.nv = nv;
    t.nz = std::max(nz, 2);
    t.v_length = v_length;
    std::vector<double> q((size_t)nv * nk);
    for (int kn = 0; kn < nk; ++kn)
        for (int j = 0; j < nv; ++j) q[(size_t)j * nk + kn] = deriv[(size_t)kn * nv + j];
    std::vector<double> s((size_t)nv * t.nz, 0.0);
    if (signal && !c->bins.turn_off_sfid) std::copy(signal, signal + (size_t)nv * nz, s.begin());
    HIPCHK(hipMalloc(&t.d_qt, q.size() * sizeof(double)));
    HIPCHK(hipMemcpy(t.d_qt, deriv, q.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&t.d_s, s.size() * sizeof(double)));
    // interpolation pairs (q[j][kn], q[j+1][kn]); the last row repeats its node
    std::vector<double> q2(2 * q.size());
    for (int j = 0; j < nv; ++j)
        for (int kn = 0; kn < nk; ++kn) {
            q2[2 * ((size_t)j * nk + kn)] = q[(size_t)j * nk + kn];
            q2[2 * ((size_t)j * nk + kn) + 1] = q[(size_t)std::min(j + 1, nv - 1) * nk + kn];
        }
    HIPCHK(hipMalloc(&t.d_q2, q2.size() * sizeof(double)));
    HIPCHK(hipMemcpy(t.d_q2, q2.data(), q2.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t.d_s, s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice));
    c->tables.push_back(t);
    if (r_index) *r_index = (int)c->tables.size() - 1;
    return QML_OK;
}

}  // extern "C"

namespace {
// Per-chunk scratch (elements of each arena) at padded size Np with m stack
// members.  With the stack slabbed (slab > 0 members) the X and Y arenas hold
// one slab and the G store (packed upper tiles of every member) the rest.
struct Need { size_t mat, vec, pj, sj, stk, lic, gram, g; };
Need chunk_need(int prec, const ChunkHost &ch, size_t Np, size_t m, int slab) {
    Need n{};
    const size_t nt = Np / TILE, ntu = nt * (nt + 1) / 2;
    const size_t mt = (m + TILE - 1) / TILE;
    // A, L, Li, S, then for a resolution-matrix chunk its kbu Q_k and S
    n.mat = (4 + (ch.rm ? (size_t)ch.kbu + 1 : 0)) * Np * Np;
    n.vec = (size_t)VEC_SLOTS * Np;
    n.pj = m * Np * (1 + MAX_NU) + m * MAX_NU * Np + (size_t)ch.nzq * ch.nvu;
    n.sj = m * MAX_NU * MAX_NU + ch.n;
    n.stk = (slab > 0 ? std::min(m, (size_t)slab) : m) * Np * Np;
    n.g = (slab > 0) ? m * ntu * TILE * TILE : 0;
    n.lic = (prec == 0) ? 0 : Np * Np;
    n.gram = ntu * (mt * (mt + 1) / 2) * TILE * TILE;
    return n;
}
size_t need_bytes(int prec, const Need &n) {
    const size_t es = (prec == 0) ? 8 : 4;
    // stage-A arenas (mat, vec, aux, s, Li copy) exist twice for the two-stream pipeline
    return 8 * (2 * (n.mat + n.vec + n.pj + n.sj) + n.gram) + es * (2 * n.stk + 2 * n.lic + n.g);
}
}  // namespace

extern "C" {

static int add_chunk_impl(qml_ctx *c, const qml_chunk_in *in, const qml_rmat *rm, qml_chunk_info *info,
                          int *chunk_id) {
    if (!c || !in || !c->bins_set) return fail(QML_ERR_ARG, "qml_add_chunk: set bins first");
    if (c->committed) return fail(QML_ERR_ARG, "qml_add_chunk: context already committed");
    if (in->size <= 0 || !in->wave || !in->delta || !in->ivar)
        return fail(QML_ERR_ARG, "qml_add_chunk: empty chunk");
    const qml_bins &b = c->bins;
    ChunkHost ch;
    // QSOFile(qmaster, i1, i2) + _cutMaskedBoundary (io/qso_file.cpp:61-89,139-157)
    int i1 = 0, i2 = in->size;
    while (i1 < in->size && in->ivar[i1] == 0) ++i1;
    while (i2 > 0 && in->ivar[i2 - 1] == 0) --i2;
    if (i2 - i1 <= 0) return fail(QML_ERR_EMPTY, "Empty spectrum when masked pixels at boundary are removed!");
    const int N = i2 - i1;
    ch.wave.assign(in->wave + i1, in->wave + i2);
    ch.delta.assign(in->delta + i1, in->delta + i2);
    ch.ivar.assign(in->ivar + i1, in->ivar + i2);
    int real = 0;
    for (double v : ch.ivar) real += (v != 0);
    // recalcDvDLam (io/qso_file.cpp:325-329)
    double dv = -1;
    if (N > 1) {
        std::vector<double> d(N - 1);
        for (int i = 0; i < N - 1; ++i) d[i] = std::log(ch.wave[i + 1] / ch.wave[i]) * SPEED_OF_LIGHT;
        std::sort(d.begin(), d.end());
        dv = median_sorted(d);
    }
    // short chunk (core/chunk_estimate.cpp:172-187)
    const double lenv = SPEED_OF_LIGHT * std::log(ch.wave[N - 1] / ch.wave[0]);
    const int minnopix = (int)(lenv / b.max_pixel_length_v);
    const double min_len = 0.90 / 2 * b.max_forest_length_v;
    if (real < minnopix || lenv < min_len) return fail(QML_ERR_SHORT_CHUNK, "Short chunk");
    // findSpecResIndex (core/sq_table.cpp:388-397); with a resolution matrix
    // the index is 0 (core/chunk_estimate.cpp:189-193)
    const double rounded = std::round(dv * 10) / 10;
    int rindex = -1;
    if (rm) {
        if (!c->tables.empty()) rindex = 0;
    } else {
        for (size_t r = 0; r < c->tables.size(); ++r)
            if (c->tables[r].R == in->R_fwhm && std::fabs(rounded - c->tables[r].dv) < 0.01) {
                rindex = (int)r;
                break;
            }
    }
    if (rindex < 0) {
        char msg[128];
        std::snprintf(msg, sizeof msg, "SpecRes R: %d, dv: %g not found in tables!", in->R_fwhm, dv);
        return fail(QML_ERR_NO_TABLE, msg);
    }
    // readMinMaxMedRedshift (io/qso_file.cpp:280-300)
    const double zlo = ch.wave[0] / LYA_REST - 1, zhi = ch.wave[N - 1] / LYA_REST - 1;
    const double zmed = median_sorted(ch.wave) / LYA_REST - 1;
    const int kbu = (int)(std::lower_bound(c->kcenters.begin(), c->kcenters.end(), 0.85 * (MY_PI / dv)) -
                          c->kcenters.begin());
    auto findz = [&](double z) {   // bins::findRedshiftBin
        const double zle = c->zcenters[0] - b.z_bin_width / 2.;
        if (z < zle) return -1;
        int r = (int)((z - zle) / b.z_bin_width);
        if (r >= b.nz) r = b.nz;
        return r;
    };
    const int zb = findz(zmed), zl = findz(zlo), zu = findz(zhi);
    // _setNQandFisherIndex (core/chunk_estimate.cpp:240-257)
    int nq = zu - zl + 1, fis = b.nk * zl;
    if (b.z_binning_method == 1) {
        if (zl >= 0 && zlo < c->zcenters[zl] && zl != 0) { ++nq; fis -= b.nk; }
        if (zu >= 0 && zu < b.nz && c->zcenters[zu] < zhi && zu < b.nz - 1) ++nq;
    }
    nq *= b.nk;
    if (zl < 0 || zu >= b.nz || fis < 0 || fis + nq > c->T)
        return fail(QML_ERR_LIMIT, "qml_add_chunk: chunk lies outside the redshift bins");
    if (rm) {
        // R rows for the cut pixels [i1, i2) (Resolution(rmaster, i1, i2),
        // mathtools/matrix_helper.cpp:936-970) and the matrix's grid
        const int nel = rm->nelem;
        ch.rm = 1;
        ch.rnb = nel;
        ch.rband.assign((size_t)N * nel, 0.0);
        ch.rdia = rm->kind == QML_RMAT_DIA;
        if (rm->kind == QML_RMAT_DIA) {
            // row i, slot t <-> column i - nel/2 + t on diagonal nel-1-t; the
            // grid gets nel/2 points beyond each end so that every slot lies
            // on it (slots off the matrix keep weight 0)
            ch.ros = 1;
            ch.rcol0 = 0;
            ch.rM = N + nel - 1;
            for (int i = 0; i < N; ++i)
                for (int t = 0; t < nel; ++t) {
                    const int col = i - nel / 2 + t;
                    if (col >= 0 && col < N) ch.rband[(size_t)i * nel + t] = rm->values[(size_t)(nel - 1 - t) * in->size + i1 + col];
                }
        } else {
            ch.ros = rm->oversampling;
            ch.rcol0 = 0;
            ch.rM = N * ch.ros + nel - 1;
            for (int i = 0; i < N; ++i)
                for (int t = 0; t < nel; ++t) ch.rband[(size_t)i * nel + t] = rm->values[(size_t)(i1 + i) * nel + t];
        }
        ch.nvu = 0;   // the derivative table is not used
    } else {
        const TableHost &tb = c->tables[rindex];
        const double tdv = tb.v_length / (tb.nv - 1.0);
        ch.nvu = (int)std::min<double>(tb.nv, std::floor(lenv / tdv) + 4.0);
    }
    if (rm) {
        // the grid of the matrix (core/chunk_estimate.cpp:936-944), units of lambda_Lya
        ch.fine.resize(ch.rM);
        if (rm->kind == QML_RMAT_DIA) {
            const int h = ch.rnb / 2;
            const double w0 = ch.wave[0], w1 = ch.wave[N - 1];
            const double d0 = (N > 1) ? ch.wave[1] - w0 : 1.0, d1 = (N > 1) ? w1 - ch.wave[N - 2] : 1.0;
            for (int a = 0; a < ch.rM; ++a) {
                const int p = a - h;
                const double w = (p < 0) ? w0 + p * d0 : (p >= N) ? w1 + (p - N + 1) * d1 : ch.wave[p];
                ch.fine[a] = w / LYA_REST;
            }
        } else {
            std::vector<double> dl(N - 1);
            for (int i = 0; i < N - 1; ++i) dl[i] = ch.wave[i + 1] - ch.wave[i];
            std::sort(dl.begin(), dl.end());
            const double fine_dl = median_sorted(dl) / LYA_REST / ch.ros;
            const int disp = ch.rnb / 2;
            for (int a = 0; a < ch.rM; ++a) ch.fine[a] = ch.wave[0] / LYA_REST + (a - disp) * fine_dl;
        }
    }
    for (double &w : ch.wave) w /= LYA_REST;
    ch.N = N;
    ch.nq = nq;
    ch.fis = fis;
    ch.kbu = kbu;
    ch.nzq = nq / b.nk;
    ch.zm0 = fis / b.nk;
    ch.n = ch.nzq * kbu;
    if (ch.nzq > 8) return fail(QML_ERR_LIMIT, "qml_add_chunk: a chunk spans more than 8 redshift bins");
    // continuum templates + SVD cut (core/chunk_estimate.cpp:576-665)
    int nt = 0;
    if (b.cont_logl_order >= 0 || b.cont_lam_order >= 0)
        nt = 1 + std::max(0, b.cont_logl_order) + std::max(0, b.cont_lam_order);
    ch.nu = 0;
    if (nt > 0) {
        std::vector<double> V((size_t)nt * N);
        for (int i = 0; i < N; ++i) V[i] = 1. / std::sqrt((double)N);
        int row = 1;
        for (int cmo = 1; cmo <= b.cont_logl_order; ++cmo, ++row) {
            double nrm = 0;
            for (int i = 0; i < N; ++i) {
                const double v = std::pow(std::log(ch.wave[i]), cmo);
                V[(size_t)row * N + i] = v;
                nrm += v * v;
            }
            const double sc = 1.0 / std::sqrt(nrm);
            for (int i = 0; i < N; ++i) V[(size_t)row * N + i] *= sc;
        }
        for (int cmo = 1; cmo <= b.cont_lam_order; ++cmo, ++row) {
            double nrm = 0;
            for (int i = 0; i < N; ++i) {
                const double v = std::pow(ch.wave[i], cmo);
                V[(size_t)row * N + i] = v;
                nrm += v * v;
            }
            const double sc = 1.0 / std::sqrt(nrm);
            for (int i = 0; i < N; ++i) V[(size_t)row * N + i] *= sc;
        }
        std::vector<double> sv;
        jacobi_svd(V, N, nt, sv);
        int use = nt;
        while (use > 0) {
            if (sv[use - 1] / sv[0] > DOUBLE_EPSILON) break;
            --use;
        }
        ch.nu = use;
        ch.U.assign(V.begin(), V.begin() + (size_t)use * N);
    }
    qml_chunk_info &inf = ch.info;
    inf.size = N;
    inf.real_size = real;
    inf.res_index = rindex;
    inf.kbin_upp = kbu;
    inf.zbin = zb;
    inf.zbin_low = zl;
    inf.zbin_upp = zu;
    inf.n_q_matrices = nq;
    inf.fisher_index_start = fis;
    inf.n_stored = ch.n;
    inf.n_templates = ch.nu;
    inf.dv_kms = dv;
    inf.lower_z = zlo;
    inf.upper_z = zhi;
    inf.median_z = zmed;
    if (info) *info = inf;
    {   // the matrices must fit: Chunk::_setStoredMatrices (core/chunk_estimate.cpp:258-282)
        // throws "Not all matrices are stored.", here against the device scratch budget
        const size_t bt = need_bytes(c->prec, chunk_need(c->prec, ch, align_up(N, TILE), (size_t)ch.n + 2, c->slab));
        if (bt > (size_t)(c->scratch_gb * 1073741824.0)) {
            char msg[200];
            std::snprintf(msg, sizeof msg,
                          "Not all matrices are stored: the chunk needs %.2f GB of device scratch, above the "
                          "%.2f GB budget", bt / 1073741824.0, c->scratch_gb);
            return fail(QML_ERR_MEMORY, msg);
        }
    }
    c->chunks.push_back(std::move(ch));
    if (chunk_id) *chunk_id = (int)c->chunks.size() - 1;
    return QML_OK;
}

int qml_add_chunk(qml_ctx *c, const qml_chunk_in *in, qml_chunk_info *info, int *chunk_id) {
    return add_chunk_impl(c, in, nullptr, info, chunk_id);
}

int qml_add_chunk_rmat(qml_ctx *c, const qml_chunk_in *in, const qml_rmat *rm, qml_chunk_info *info, int *chunk_id) {
    if (!rm || !rm->values || rm->nelem < 1 || rm->nelem % 2 == 0)
        return fail(QML_ERR_ARG, "qml_add_chunk_rmat: resolution matrix needs an odd number of elements per row");
    if (rm->kind == QML_RMAT_DIA) {
        if (15 + rm->nelem > RM_ROWS) return fail(QML_ERR_LIMIT, "qml_add_chunk_rmat: more than 81 diagonals");
    } else if (rm->kind == QML_RMAT_OVERSAMPLED) {
        if (rm->oversampling < 1 || 15 * rm->oversampling + rm->nelem > RM_ROWS)
            return fail(QML_ERR_LIMIT, "qml_add_chunk_rmat: 15 x oversampling + nelem exceeds 96");
    } else {
        return fail(QML_ERR_ARG, "qml_add_chunk_rmat: unknown matrix kind");
    }
    if (in && in->size < 2) return fail(QML_ERR_ARG, "qml_add_chunk_rmat: chunk too small");
    return add_chunk_impl(c, in, rm, info, chunk_id);
}

}  // extern "C"

namespace {
// Frees the batch plan: work lists, events, scratch arenas and the graph
// that referenced them.
void release_plan(qml_ctx *c) {
    drop_graph(c);
    for (auto &b : c->batches) {
        hipFree(b.d_chunks); hipFree(b.d_wl); hipFree(b.d_wr); hipFree(b.d_wg);
        for (auto &e : b.evA) if (e) hipEventDestroy(e);
        for (auto &e : b.evB) if (e) hipEventDestroy(e);
        for (auto &e : b.evS) if (e) hipEventDestroy(e);
        if (b.evR) hipEventDestroy(b.evR);
    }
    c->batches.clear();
    c->plan_ok = false;
    for (int k = 0; k < 2; ++k) {
        void *set[] = {c->d_mat[k], c->d_vec[k], c->d_pj[k], c->d_sj[k], c->d_lic[k]};
        for (void *p : set) if (p) hipFree(p);
        c->d_mat[k] = c->d_vec[k] = c->d_pj[k] = c->d_sj[k] = nullptr;
        c->d_lic[k] = nullptr;
    }
    void *rest[] = {c->d_gram, c->d_X, c->d_Y, c->d_G};
    for (void *p : rest) if (p) hipFree(p);
    c->d_gram = nullptr;
    c->d_X = c->d_Y = c->d_G = nullptr;
}

// Size-bucketed batch plan: chunks sorted by padded size (then stored Q
// count), descending, are cut into batches of at most c->max_batch chunks
// whose scratch fits c->scratch_gb; allocates the arenas of the largest
// batch and uploads every batch's chunk descriptors and work lists.
int plan_batches(qml_ctx *c) {
    const int nch = (int)c->chunks.size();
    c->max_chunk_bytes = 0;
    const size_t es = (c->prec == 0) ? 8 : 4;
    auto need = [&](const ChunkHost &ch) { return chunk_need(c->prec, ch, ch.Np, ch.m, c->slab); };
    auto bytes = [&](const Need &n) { return need_bytes(c->prec, n); };
    std::vector<int> order(nch);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (c->chunks[a].Np != c->chunks[b].Np) return c->chunks[a].Np > c->chunks[b].Np;
        return c->chunks[a].n > c->chunks[b].n;
    });
    const size_t budget = (size_t)(c->scratch_gb * 1073741824.0);
    std::vector<std::vector<int>> groups;
    {
        std::vector<int> cur;
        size_t used = 0;
        for (int id : order) {
            const size_t bt = bytes(need(c->chunks[id]));
            if (bt > budget) {
                char msg[160];
                std::snprintf(msg, sizeof msg, "qml_commit: chunk %d needs %.2f GB, above the %.2f GB scratch budget",
                              id, bt / 1073741824.0, c->scratch_gb);
                return fail(QML_ERR_MEMORY, msg);
            }
            const bool new_np = c->np_buckets && !cur.empty() && c->chunks[cur.front()].Np != c->chunks[id].Np;
            if (!cur.empty() && (used + bt > budget || (int)cur.size() >= c->max_batch || new_np)) {
                groups.push_back(cur);
                cur.clear();
                used = 0;
            }
            cur.push_back(id);
            used += bt;
        }
        if (!cur.empty()) groups.push_back(cur);
    }
    Need mx{};
    std::vector<std::vector<ChunkDev>> devs;
    for (auto &g : groups) {
        Need acc{};
        std::vector<ChunkDev> cd;
        for (int id : g) {
            const auto &ch = c->chunks[id];
            const Need n = need(ch);
            ChunkDev d{};
            d.N = ch.N;
            d.Np = ch.Np;
            d.nt = ch.Np / TILE;
            d.n = ch.n;
            d.m = ch.m;
            d.kbu = ch.kbu;
            d.nzq = ch.nzq;
            d.zm0 = ch.zm0;
            d.table = ch.info.res_index;
            d.nu = ch.nu;
            d.gidx = id;
            d.ntu = d.nt * (d.nt + 1) / 2;
            d.mt = (ch.m + TILE - 1) / TILE;
            d.in_off = ch.in_off;
            d.mat_off = (long long)acc.mat;
            d.vec_off = (long long)acc.vec;
            d.pj_off = (long long)acc.pj;
            d.at_off = (long long)(acc.pj + (size_t)ch.m * ch.Np * (1 + MAX_NU));
            d.th_off = d.at_off + (long long)ch.m * MAX_NU * ch.Np;
            d.nvu = ch.nvu;
            d.sj_off = (long long)acc.sj;
            // the split form stores two bf16 planes per element: offsets in bf16 units
            const long long su = (c->prec >= 2) ? 2 : 1;
            d.stk_off = (long long)acc.stk * su;
            d.g_off = (c->slab > 0) ? (long long)acc.g * su : d.stk_off;
            d.gram_off = (long long)acc.gram;
            d.lic_off = (c->prec == 0) ? (long long)(acc.mat + 2 * (size_t)ch.Np * ch.Np) : (long long)acc.lic * su;
            d.rm = ch.rm;
            d.rnb = ch.rnb;
            d.ros = ch.ros;
            d.rcol0 = ch.rcol0;
            d.rM = ch.rM;
            d.rin_off = ch.rin_off;
            d.qb_off = (long long)(acc.mat + 4 * (size_t)ch.Np * ch.Np);
            acc.mat += n.mat; acc.vec += n.vec; acc.pj += n.pj; acc.sj += n.sj;
            acc.stk += n.stk; acc.lic += n.lic; acc.gram += n.gram; acc.g += n.g;
            c->max_chunk_bytes = std::max(c->max_chunk_bytes, (double)bytes(n));
            cd.push_back(d);
        }
        mx.mat = std::max(mx.mat, acc.mat); mx.vec = std::max(mx.vec, acc.vec);
        mx.pj = std::max(mx.pj, acc.pj); mx.sj = std::max(mx.sj, acc.sj);
        mx.stk = std::max(mx.stk, acc.stk); mx.lic = std::max(mx.lic, acc.lic);
        mx.gram = std::max(mx.gram, acc.gram); mx.g = std::max(mx.g, acc.g);
        devs.push_back(cd);
    }
    for (int k = 0; k < 2; ++k) {
        HIPCHK(hipMalloc(&c->d_mat[k], std::max<size_t>(mx.mat, 1) * 8));
        HIPCHK(hipMalloc(&c->d_vec[k], std::max<size_t>(mx.vec, 1) * 8));
        HIPCHK(hipMalloc(&c->d_pj[k], std::max<size_t>(mx.pj, 1) * 8));
        HIPCHK(hipMalloc(&c->d_sj[k], std::max<size_t>(mx.sj, 1) * 8));
        if (c->prec != 0) HIPCHK(hipMalloc(&c->d_lic[k], std::max<size_t>(mx.lic, 1) * es));
        HIPCHK(hipMemset(c->d_vec[k], 0, std::max<size_t>(mx.vec, 1) * 8));
    }
    HIPCHK(hipMalloc(&c->d_gram, std::max<size_t>(mx.gram, 1) * 8));
    HIPCHK(hipMalloc(&c->d_X, std::max<size_t>(mx.stk, 1) * es));
    HIPCHK(hipMalloc(&c->d_Y, std::max<size_t>(mx.stk, 1) * es));
    if (c->slab > 0) HIPCHK(hipMalloc(&c->d_G, std::max<size_t>(mx.g, 1) * es));
    // LYSPEQ_QML_POISON=1 fills every scratch arena with all-ones bytes (NaN
    // in FP64, FP32 and bf16), so that a read of memory the iteration did
    // not write shows up in the results.  The vector arena keeps its zeros.
    // LDS is also filled with NaN bytes before each stage of every batch.
    if (const char *pz = std::getenv("LYSPEQ_QML_POISON"); pz && pz[0] == '1') {
        c->poison = 1;
        for (int k = 0; k < 2; ++k) {
            HIPCHK(hipMemset(c->d_mat[k], 0xff, std::max<size_t>(mx.mat, 1) * 8));
            HIPCHK(hipMemset(c->d_pj[k], 0xff, std::max<size_t>(mx.pj, 1) * 8));
            HIPCHK(hipMemset(c->d_sj[k], 0xff, std::max<size_t>(mx.sj, 1) * 8));
            if (c->d_lic[k]) HIPCHK(hipMemset(c->d_lic[k], 0xff, std::max<size_t>(mx.lic, 1) * es));
        }
        HIPCHK(hipMemset(c->d_gram, 0xff, std::max<size_t>(mx.gram, 1) * 8));
        HIPCHK(hipMemset(c->d_X, 0xff, std::max<size_t>(mx.stk, 1) * es));
        HIPCHK(hipMemset(c->d_Y, 0xff, std::max<size_t>(mx.stk, 1) * es));
        if (c->d_G) HIPCHK(hipMemset(c->d_G, 0xff, std::max<size_t>(mx.g, 1) * es));
    }

    for (size_t bi = 0; bi < groups.size(); ++bi) {
        Batch B;
        B.ids = groups[bi];
        const auto &cd = devs[bi];
        std::vector<Work> wl, wr, wg;
        int ra_max = 0, rb_max = 0, nb_max = 0;
        for (int li = 0; li < (int)cd.size(); ++li) {
            const ChunkDev &d = cd[li];
            B.np_max = std::max(B.np_max, d.Np);
            B.nvu_max = std::max(B.nvu_max, d.nvu);
            B.nu_max = std::max(B.nu_max, d.nu);
            B.nq_max = std::max(B.nq_max, c->outs[d.gidx].nq);
            B.dp_rb = std::max(B.dp_rb, (d.kbu + 1 + 15) / 16);
            B.dp_cb = std::max(B.dp_cb, ((d.nzq + 1) * (1 + d.nu) + 15) / 16);
            if (d.rm) {
                B.rm_kbu = std::max(B.rm_kbu, d.kbu);
                ra_max = std::max(ra_max, 15 * d.ros + d.rnb);
                rb_max = std::max(rb_max, 31 * d.ros + d.rnb);
                nb_max = std::max(nb_max, d.rnb);
            }
            const int g = (c->prec >= 2) ? 2 : 1;
            const int mtg = (d.mt + g - 1) / g;
            for (int s = 0; s < d.ntu; ++s)
                for (int I = 0; I < mtg; ++I)
                    for (int J = I; J < mtg; ++J) wg.push_back({li, s, I, J});
        }
        if (B.rm_kbu >= 0) {
            // 16 x 32 output tiles where their LDS fits in 64 KB, their a range
            // in 80 rows and their b range in 4 passes; else 16 x 16 (which
            // always fits under the registration limit on 15 ros + rnb)
            B.rm_lay = rmat_layout(32, ra_max, rb_max, nb_max);
            if (B.rm_lay.bytes > 65536 || B.rm_lay.em != 10 || B.rm_lay.npass > 4)
                B.rm_lay = rmat_layout(16, ra_max, ra_max, nb_max);
            if (!rmat_layout_ok(B.rm_lay)) return fail(QML_ERR_LIMIT, "qml_commit: no sandwich kernel for the resolution matrices");
        }
        // Output tiles of the whitening products: 64x64 for FP64 and FP32,
        // 128x128 for split BF16, whose left product P = Li M is lower
        // triangular (J <= I only).  The lists run slab by slab (one slab of
        // every member without slabbing), chunk by chunk within a slab.
        {
            int mmax = 0;
            for (const ChunkDev &d : cd) mmax = std::max(mmax, d.m);
            const int S = (c->slab > 0) ? c->slab : std::max(mmax, 1);
            B.nslab = (mmax + S - 1) / S;
            B.sl_off.assign(1, 0);
            B.sr_off.assign(1, 0);
            const int g = (c->prec >= 2) ? 2 : 1;
            for (int sb = 0; sb < B.nslab; ++sb) {
                for (int li = 0; li < (int)cd.size(); ++li) {
                    const ChunkDev &d = cd[li];
                    const int nt = (d.nt + g - 1) / g;
                    const int ja = sb * S, jb = std::min(d.m, (sb + 1) * S);
                    // the left product's tiles go four members at a time, tile by
                    // tile, so that the workgroups in flight on an XCD share Li
                    // rows across members (5 % faster than member by member on
                    // DESI-like batches); every tile is computed alone, so the
                    // order does not change any result
                    for (int j0 = ja; j0 < jb; j0 += 4)
                        for (int I = 0; I < nt; ++I)
                            for (int J = 0; J < ((c->prec >= 2) ? I + 1 : nt); ++J)
                                for (int j = j0; j < std::min(jb, j0 + 4); ++j) wl.push_back({li, j, I, J});
                    for (int j = ja; j < jb; ++j)
                        for (int I = 0; I < nt; ++I)
                            for (int J = I; J < nt; ++J) wr.push_back({li, j, I, J});
                }
                B.sl_off.push_back((int)wl.size());
                B.sr_off.push_back((int)wr.size());
            }
        }
        B.nwl = (int)wl.size();
        B.nwr = (int)wr.size();
        B.nwg = (int)wg.size();
        HIPCHK(hipMalloc(&B.d_chunks, cd.size() * sizeof(ChunkDev)));
        HIPCHK(hipMemcpy(B.d_chunks, cd.data(), cd.size() * sizeof(ChunkDev), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&B.d_wl, std::max<size_t>(wl.size(), 1) * sizeof(Work)));
        HIPCHK(hipMalloc(&B.d_wr, std::max<size_t>(wr.size(), 1) * sizeof(Work)));
        HIPCHK(hipMalloc(&B.d_wg, std::max<size_t>(wg.size(), 1) * sizeof(Work)));
        if (!wl.empty()) HIPCHK(hipMemcpy(B.d_wl, wl.data(), wl.size() * sizeof(Work), hipMemcpyHostToDevice));
        if (!wr.empty()) HIPCHK(hipMemcpy(B.d_wr, wr.data(), wr.size() * sizeof(Work), hipMemcpyHostToDevice));
        if (!wg.empty()) HIPCHK(hipMemcpy(B.d_wg, wg.data(), wg.size() * sizeof(Work), hipMemcpyHostToDevice));
        for (auto &e : B.evA) HIPCHK(hipEventCreate(&e));
        for (auto &e : B.evB) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventCreate(&B.evR));
        B.evS.assign(3 * (B.nslab - 1), nullptr);
        for (auto &e : B.evS) HIPCHK(hipEventCreate(&e));
        c->batches.push_back(B);
    }
    c->plan_ok = true;
    return QML_OK;
}
}  // namespace

extern "C" {

int qml_commit(qml_ctx *c) {
    if (!c || !c->bins_set) return fail(QML_ERR_ARG, "qml_commit: set bins first");
    if (c->committed) return fail(QML_ERR_ARG, "qml_commit: already committed");
    HIPCHK(hipSetDevice(c->device));
    const int nch = (int)c->chunks.size();
    const size_t es = (c->prec == 0) ? 8 : 4;
    // tables
    std::vector<TableDev> td(c->tables.size());
    for (size_t r = 0; r < c->tables.size(); ++r) {
        const auto &t = c->tables[r];
        td[r].qt = t.d_qt;
        td[r].q2 = t.d_q2;
        td[r].s = t.d_s;
        td[r].nv = t.nv;
        td[r].nz = t.nz;
        td[r].v1 = 0.0;
        td[r].dv = t.v_length / (t.nv - 1.0);
        td[r].z1 = c->zcenters[0] - c->bins.z_bin_width;
        td[r].dz = c->bins.z_bin_width * (c->bins.nz + 1) / (t.nz - 1.0);
    }
    if (!td.empty()) {
        HIPCHK(hipMalloc(&c->d_tables, td.size() * sizeof(TableDev)));
        HIPCHK(hipMemcpy(c->d_tables, td.data(), td.size() * sizeof(TableDev), hipMemcpyHostToDevice));
    }
    // inputs and persistent outputs
    size_t in_total = 0, fc_total = 0, dbt_total = 0;
    c->outs.resize(nch);
    for (int i = 0; i < nch; ++i) {
        auto &ch = c->chunks[i];
        ch.Np = (int)align_up(ch.N, TILE);
        ch.m = ch.n + 2;
        ch.in_off = (long long)in_total;
        in_total += (size_t)IN_SLOTS * ch.Np;
        c->outs[i].fc_off = (long long)fc_total;
        c->outs[i].dbt_off = (long long)dbt_total;
        c->outs[i].nq = ch.nq;
        c->outs[i].fis = ch.fis;
        fc_total += (size_t)ch.nq * ch.nq;
        dbt_total += (size_t)3 * ch.nq;
    }
    std::vector<double> hin(std::max<size_t>(in_total, 1), 0.0);
    for (int i = 0; i < nch; ++i) {
        const auto &ch = c->chunks[i];
        double *base = hin.data() + ch.in_off;
        const size_t Np = ch.Np;
        for (int p = 0; p < ch.Np; ++p) {
            const bool real = p < ch.N;
            base[IN_WAVE * Np + p] = real ? ch.wave[p] : ch.wave[ch.N - 1];
            base[IN_DELTA * Np + p] = real ? ch.delta[p] : 0.0;
            const double iv = real ? ch.ivar[p] : 0.0;
            base[IN_IVAR * Np + p] = iv;
            base[IN_D * Np + p] = std::sqrt(iv);
            base[IN_NPRIME * Np + p] = (iv != 0) ? 1.0 / iv : 0.0;
            for (int t = 0; t < ch.nu; ++t) base[(IN_U0 + t) * Np + p] = real ? ch.U[(size_t)t * ch.N + p] : 0.0;
        }
    }
    HIPCHK(hipMalloc(&c->d_in, hin.size() * sizeof(double)));
    HIPCHK(hipMemcpy(c->d_in, hin.data(), hin.size() * sizeof(double), hipMemcpyHostToDevice));
    {   // resolution-matrix rows and grids
        size_t rin_total = 0;
        for (auto &ch : c->chunks)
            if (ch.rm) {
                ch.rin_off = (long long)rin_total;
                rin_total += ch.rband.size() + ch.fine.size();
            }
        if (rin_total) {
            std::vector<double> hr(rin_total);
            for (const auto &ch : c->chunks)
                if (ch.rm) {
                    std::copy(ch.rband.begin(), ch.rband.end(), hr.begin() + ch.rin_off);
                    std::copy(ch.fine.begin(), ch.fine.end(), hr.begin() + ch.rin_off + ch.rband.size());
                }
            HIPCHK(hipMalloc(&c->d_rin, rin_total * sizeof(double)));
            HIPCHK(hipMemcpy(c->d_rin, hr.data(), rin_total * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    HIPCHK(hipMalloc(&c->d_outs, std::max(nch, 1) * sizeof(ChunkOut)));
    if (nch) HIPCHK(hipMemcpy(c->d_outs, c->outs.data(), nch * sizeof(ChunkOut), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&c->d_fc, std::max<size_t>(fc_total, 1) * sizeof(double)));
    HIPCHK(hipMalloc(&c->d_dbt, std::max<size_t>(dbt_total, 1) * sizeof(double)));
    HIPCHK(hipMalloc(&c->d_status, std::max(nch, 1) * sizeof(int)));
    HIPCHK(hipMemset(c->d_status, 0, std::max(nch, 1) * sizeof(int)));
    HIPCHK(hipMalloc(&c->d_skip, std::max(nch, 1) * sizeof(int)));
    HIPCHK(hipMemset(c->d_skip, 0, std::max(nch, 1) * sizeof(int)));
    HIPCHK(hipMalloc(&c->d_theta, std::max(c->T, 1) * sizeof(double)));
    HIPCHK(hipMalloc(&c->d_F, (size_t)c->T * c->T * sizeof(double)));
    HIPCHK(hipMalloc(&c->d_dbtsum, (size_t)3 * c->T * sizeof(double)));

    {
        const int rc = plan_batches(c);
        if (rc != QML_OK) return rc;
    }
    // the accumulation writes the upper triangle of the Fisher sum only
    HIPCHK(hipMemset(c->d_F, 0, (size_t)c->T * c->T * sizeof(double)));
    HIPCHK(hipHostMalloc((void **)&c->h_theta, std::max(c->T, 1) * sizeof(double)));
    HIPCHK(hipHostMalloc((void **)&c->h_F, std::max<size_t>((size_t)c->T * c->T, 1) * sizeof(double)));
    HIPCHK(hipHostMalloc((void **)&c->h_dbt, std::max(3 * c->T, 1) * sizeof(double)));
    c->committed = true;
    return QML_OK;
}

}  // extern "C"

namespace {
// Enqueues one iteration: theta from its pinned staging buffer, both stages of
// every batch, the accumulation and the copy of the sums into pinned host
// memory.  Stage B's stream is forked from c->st after the theta copy and
// joined back into it at the end, so that the sequence begins and ends on
// c->st (what a stream capture needs).  With timing every stage boundary
// records its event; otherwise only the events that order the two streams are
// recorded.
// The launch parameters of batch bi on stream st, with its arena set bi % 2.
LaunchCtx batch_launch(const qml_ctx *c, size_t bi) {
    const Batch &B = c->batches[bi];
    const int set = (int)(bi % 2);
    LaunchCtx L;
    L.stream = c->st;
    L.chunks = B.d_chunks;
    L.nchunks = (int)B.ids.size();
    L.np_max = B.np_max;
    L.nvu_max = B.nvu_max;
    L.nu_max = B.nu_max;
    L.dp_rb = B.dp_rb;
    L.nq_max = B.nq_max;
    L.dp_cb = B.dp_cb;
    L.rin = c->d_rin;
    L.rm_kbu = B.rm_kbu;
    L.rm_lay = B.rm_lay;
    L.in = c->d_in;
    L.tables = c->d_tables;
    L.bins = c->binsdev;
    L.mat = c->d_mat[set];
    L.vec = c->d_vec[set];
    L.pj = c->d_pj[set];
    L.sj = c->d_sj[set];
    return L;
}

int enqueue_iteration(qml_ctx *c, bool timing) {
    const int T = c->T;
    hipStream_t sB = c->pipeline ? c->st2 : c->st;   // stream of the dense-product stage
    auto mark = [&](hipEvent_t e, hipStream_t s) { return timing ? hipEventRecord(e, s) : hipSuccess; };
    HIPCHK(hipMemcpyAsync(c->d_theta, c->h_theta, T * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (sB != c->st) {
        HIPCHK(hipEventRecord(c->ev_fork, c->st));
        HIPCHK(hipStreamWaitEvent(sB, c->ev_fork, 0));
    }
    for (size_t bi = 0; bi < c->batches.size(); ++bi) {
        Batch &B = c->batches[bi];
        const int set = (int)(bi % 2);
        const LaunchCtx L = batch_launch(c, bi);
        // stage A on st: this arena set was last read by stage B of batch bi - 2
        if (bi >= 2) HIPCHK(hipStreamWaitEvent(c->st, c->batches[bi - 2].evB[5], 0));
        if (c->poison) launch_poison_lds(c->st);
        HIPCHK(mark(B.evA[0], c->st));
        if (!launch_rmat(L)) return fail(QML_ERR_LIMIT, "qml_iterate: no sandwich kernel for the resolution matrices");
        HIPCHK(mark(B.evR, c->st));
        launch_assemble(L, c->d_theta);
        HIPCHK(mark(B.evA[1], c->st));
        launch_cholesky(L, c->d_status);
        HIPCHK(mark(B.evA[2], c->st));
        launch_trinv(L);
        HIPCHK(mark(B.evA[3], c->st));
        launch_solve(L);
        HIPCHK(mark(B.evA[4], c->st));
        launch_dpass(L);
        HIPCHK(mark(B.evA[5], c->st));
        launch_small(L);
        if (c->prec != 0) launch_convert_li(L, c->d_lic[set], c->prec);
        HIPCHK(hipEventRecord(B.evA[6], c->st));
        HIPCHK(hipGetLastError());
        // stage B on sB, after stage A of this batch
        LaunchCtx L2 = L;
        L2.stream = sB;
        const void *li = (c->prec == 0) ? (const void *)c->d_mat[set] : c->d_lic[set];
        HIPCHK(hipStreamWaitEvent(sB, B.evA[6], 0));
        if (c->poison) launch_poison_lds(sB);
        HIPCHK(mark(B.evB[0], sB));
        // the stack slab by slab: X and Y hold members [jlo, jlo + S) of every
        // chunk, and each slab's G goes to the G store (over X without slabs)
        void *G = (c->slab > 0) ? c->d_G : c->d_X;
        const int S = (c->slab > 0) ? c->slab : (1 << 30);
        for (int sb = 0; sb < B.nslab; ++sb) {
            L2.jlo = sb * (c->slab > 0 ? S : 0);
            L2.jhi = (c->slab > 0) ? L2.jlo + S : (1 << 30);
            const bool last = sb + 1 == B.nslab;
            launch_xgen(L2, c->d_X, c->prec);
            HIPCHK(mark(last ? B.evB[1] : B.evS[3 * sb], sB));
            launch_gemm_whiten_left(sB, c->prec, c->nst_left, B.d_chunks, B.d_wl + B.sl_off[sb],
                                    B.sl_off[sb + 1] - B.sl_off[sb], li, c->d_X, c->d_Y, L2.jlo);
            HIPCHK(mark(last ? B.evB[2] : B.evS[3 * sb + 1], sB));
            launch_gemm_whiten_right(sB, c->prec, c->nst_right, B.d_chunks, B.d_wr + B.sr_off[sb],
                                     B.sr_off[sb + 1] - B.sr_off[sb], li, c->d_Y, G, L2.jlo);
            HIPCHK(mark(last ? B.evB[3] : B.evS[3 * sb + 2], sB));
        }
        launch_gemm_gram(sB, c->prec, c->nst_gram, B.d_chunks, B.d_wg, B.nwg, G, c->d_gram);
        HIPCHK(mark(B.evB[4], sB));
        launch_gram_finalize(L2, c->d_outs, c->d_gram, c->d_status, c->d_fc, c->d_dbt);
        HIPCHK(hipEventRecord(B.evB[5], sB));
        HIPCHK(hipGetLastError());
    }
    HIPCHK(mark(c->ev_acc, sB));
    launch_accumulate(sB, c->d_outs, c->d_status, c->d_skip, (int)c->chunks.size(), T, c->d_fc, c->d_dbt, c->d_F,
                      c->d_dbtsum);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_F, c->d_F, (size_t)T * T * sizeof(double), hipMemcpyDeviceToHost, sB));
    HIPCHK(hipMemcpyAsync(c->h_dbt, c->d_dbtsum, (size_t)3 * T * sizeof(double), hipMemcpyDeviceToHost, sB));
    if (sB != c->st) {
        HIPCHK(hipEventRecord(c->ev_join, sB));
        HIPCHK(hipStreamWaitEvent(c->st, c->ev_join, 0));
    }
    return QML_OK;
}

// Captures enqueue_iteration into c->graph and instantiates it.
int capture_iteration(qml_ctx *c) {
    drop_graph(c);
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_iteration(c, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->st, &g);
    if (rc != QML_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(QML_ERR_HIP, std::string("qml_iterate: stream capture: ") + hipGetErrorString(e));
    c->graph = g;
    const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        drop_graph(c);
        return fail(QML_ERR_HIP, std::string("qml_iterate: graph instantiation: ") + hipGetErrorString(ei));
    }
    c->graph_pipeline = c->pipeline;
    return QML_OK;
}
}  // namespace

extern "C" {

int qml_iterate(qml_ctx *c, const double *theta, double *fisher_sum, double *dbt_sum) {
    if (!c || !c->committed) return fail(QML_ERR_ARG, "qml_iterate: commit first");
    if (!c->plan_ok)
        return fail(QML_ERR_MEMORY, "qml_iterate: no batch plan (a re-plan failed; call qml_autotune with a cap "
                                    "that fits, or create a new context)");
    if (!theta || !fisher_sum || !dbt_sum) return fail(QML_ERR_ARG, "qml_iterate: null argument");
    HIPCHK(hipSetDevice(c->device));
    const int T = c->T;
    // the previous iteration has completed (both streams were synchronised)
    std::memcpy(c->h_theta, theta, (size_t)T * sizeof(double));
    if (c->use_graph) {
        if (!c->gexec || c->graph_pipeline != c->pipeline) {
            const int rc = capture_iteration(c);
            if (rc != QML_OK) return rc;
        }
        HIPCHK(hipEventRecord(c->ev_begin, c->st));
        HIPCHK(hipGraphLaunch(c->gexec, c->st));
        HIPCHK(hipEventRecord(c->ev_end, c->st));
    } else {
        HIPCHK(hipEventRecord(c->ev_begin, c->st));
        const int rc = enqueue_iteration(c, true);
        if (rc != QML_OK) return rc;
        HIPCHK(hipEventRecord(c->ev_end, c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipStreamSynchronize(c->st2));
    for (int i = 0; i < T; ++i)
        for (int j = i; j < T; ++j) fisher_sum[(size_t)i * T + j] += c->h_F[(size_t)i * T + j];
    for (int k = 0; k < 3 * T; ++k) dbt_sum[k] += c->h_dbt[k];
    qml_timings tm{};
    if (!c->use_graph) {   // a replayed graph records no per-stage events
        for (auto &B : c->batches) {
            for (int k = 0; k < 6; ++k) {
                float ms = 0;
                hipEventElapsedTime(&ms, k == 0 ? B.evR : B.evA[k], B.evA[k + 1]);
                tm.kernel_ms[k] += ms;
            }
            {
                float ms = 0;
                hipEventElapsedTime(&ms, B.evA[0], B.evR);
                tm.kernel_ms[QML_K_RMAT] += ms;
            }
            // stage B: slab by slab, xgen / left / right from the event before
            hipEvent_t prev = B.evB[0];
            for (int sb = 0; sb < B.nslab; ++sb) {
                const bool last = sb + 1 == B.nslab;
                hipEvent_t e3[3] = {last ? B.evB[1] : B.evS[3 * sb], last ? B.evB[2] : B.evS[3 * sb + 1],
                                    last ? B.evB[3] : B.evS[3 * sb + 2]};
                for (int k = 0; k < 3; ++k) {
                    float ms = 0;
                    hipEventElapsedTime(&ms, prev, e3[k]);
                    tm.kernel_ms[6 + k] += ms;
                    prev = e3[k];
                }
            }
            for (int k = 3; k < 5; ++k) {
                float ms = 0;
                hipEventElapsedTime(&ms, B.evB[k], B.evB[k + 1]);
                tm.kernel_ms[6 + k] += ms;
            }
        }
        float ms = 0;
        hipEventElapsedTime(&ms, c->ev_acc, c->ev_end);
        tm.kernel_ms[QML_K_ACCUMULATE] = ms;
        tm.accumulate_ms = ms;
    }
    tm.assemble_ms = tm.kernel_ms[QML_K_ASSEMBLE] + tm.kernel_ms[QML_K_RMAT];
    tm.invert_ms = tm.kernel_ms[QML_K_CHOLESKY] + tm.kernel_ms[QML_K_TRINV] + tm.kernel_ms[QML_K_SOLVE];
    tm.weight_ms = tm.kernel_ms[QML_K_DPASS] + tm.kernel_ms[QML_K_SMALL] + tm.kernel_ms[QML_K_XGEN] +
                   tm.kernel_ms[QML_K_WHITEN_LEFT] + tm.kernel_ms[QML_K_WHITEN_RIGHT];
    tm.fisher_ms = tm.kernel_ms[QML_K_GRAM] + tm.kernel_ms[QML_K_FINALIZE];
    float tot = 0;
    hipEventElapsedTime(&tot, c->ev_begin, c->ev_end);
    tm.total_ms = tot;
    tm.n_batches = (long long)c->batches.size();
    c->tm = tm;
    return QML_OK;
}

int qml_get_chunk_result(qml_ctx *c, int id, double *fisher, double *dbt, int *status) {
    if (!c || !c->committed || id < 0 || id >= (int)c->chunks.size())
        return fail(QML_ERR_ARG, "qml_get_chunk_result: bad chunk id or not committed");
    HIPCHK(hipSetDevice(c->device));
    const ChunkOut &o = c->outs[id];
    if (fisher)
        HIPCHK(hipMemcpy(fisher, c->d_fc + o.fc_off, (size_t)o.nq * o.nq * sizeof(double), hipMemcpyDeviceToHost));
    if (dbt) HIPCHK(hipMemcpy(dbt, c->d_dbt + o.dbt_off, (size_t)3 * o.nq * sizeof(double), hipMemcpyDeviceToHost));
    if (status) HIPCHK(hipMemcpy(status, c->d_status + id, sizeof(int), hipMemcpyDeviceToHost));
    return QML_OK;
}

int qml_chunk_matrix(qml_ctx *c, int id, int which, int arg, double *out, int *n) {
    if (!c || !c->committed || id < 0 || id >= (int)c->chunks.size())
        return fail(QML_ERR_ARG, "qml_chunk_matrix: bad chunk id or not committed");
    const ChunkHost &ch = c->chunks[id];
    const int nk = c->binsdev.nk;
    // v and z live on the matrix grid: the oversampled one of an oversampled
    // resolution matrix (a diagonal matrix keeps the pixels' grid)
    const bool fine = (which == QML_CMAT_V || which == QML_CMAT_Z) && ch.rm && !ch.rdia;
    int kn = 0, zm = 0;
    switch (which) {
        case QML_CMAT_V:
        case QML_CMAT_Z: break;
        case QML_CMAT_SFID:
            if (c->bins.turn_off_sfid || (!ch.rm && !c->tables[ch.info.res_index].d_s))
                return fail(QML_ERR_ARG, "qml_chunk_matrix: the fiducial signal is off or the table has none");
            break;
        case QML_CMAT_Q:
            if (arg < 0 || arg >= ch.nq) return fail(QML_ERR_ARG, "qml_chunk_matrix: i_kz outside the chunk");
            kn = (arg + ch.fis) % nk;
            if (kn >= ch.kbu) return fail(QML_ERR_ARG, "qml_chunk_matrix: the band of i_kz is not stored (KBIN_UPP)");
            break;
        case QML_CMAT_ZWEIGHT:
        case QML_CMAT_GROWTH:
            if (arg < 0 || arg >= c->binsdev.nz) return fail(QML_ERR_ARG, "qml_chunk_matrix: z bin outside the bins");
            zm = arg;
            break;
        default: return fail(QML_ERR_ARG, "qml_chunk_matrix: unknown matrix");
    }
    const int nn = fine ? ch.rM : ch.N;
    if (n) *n = nn;
    if (!out) return QML_OK;
    if (!c->plan_ok) return fail(QML_ERR_MEMORY, "qml_chunk_matrix: no batch plan");
    size_t bi = 0;
    int li = -1;
    for (size_t b = 0; b < c->batches.size() && li < 0; ++b) {
        const std::vector<int> &ids = c->batches[b].ids;
        const auto it = std::find(ids.begin(), ids.end(), id);
        if (it != ids.end()) {
            bi = b;
            li = (int)(it - ids.begin());
        }
    }
    if (li < 0) return fail(QML_ERR_ARG, "qml_chunk_matrix: the chunk is in no batch");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->st2) HIPCHK(hipStreamSynchronize(c->st2));
    const LaunchCtx L = batch_launch(c, bi);
    // R Shat R^T and the R Qhat_k R^T of the batch's resolution-matrix chunks
    // into their arena matrices, as the iteration's first step makes them
    if (ch.rm && (which == QML_CMAT_SFID || which == QML_CMAT_Q) && !launch_rmat(L))
        return fail(QML_ERR_LIMIT, "qml_chunk_matrix: no sandwich kernel for the resolution matrices");
    double *d_out = nullptr;
    const size_t bytes = (size_t)nn * nn * sizeof(double);
    HIPCHK(hipMalloc(&d_out, bytes));
    launch_chunk_matrix(c->st, c->batches[bi].d_chunks + li, c->d_in, c->d_rin, c->d_tables, c->binsdev, L.mat, which,
                        fine ? 1 : 0, kn, zm, nn, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    hipFree(d_out);
    if (e != hipSuccess) return fail(QML_ERR_HIP, std::string("qml_chunk_matrix: ") + hipGetErrorString(e));
    return QML_OK;
}

int qml_chunk_marginalize(qml_ctx *c, int id, double *cinv) {
    if (!c || !c->committed || id < 0 || id >= (int)c->chunks.size() || !cinv)
        return fail(QML_ERR_ARG, "qml_chunk_marginalize: bad chunk id, null matrix or not committed");
    const ChunkHost &ch = c->chunks[id];
    if (ch.nu == 0) return QML_OK;
    const int N = ch.N;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->st2) HIPCHK(hipStreamSynchronize(c->st2));
    double *d = nullptr;
    const size_t bytes = ((size_t)N * N + N) * sizeof(double);
    HIPCHK(hipMalloc(&d, bytes));
    hipError_t e = hipMemcpyAsync(d, cinv, (size_t)N * N * sizeof(double), hipMemcpyHostToDevice, c->st);
    for (int t = 0; t < ch.nu && e == hipSuccess; ++t) {
        launch_marginalize(c->st, d, c->d_in + ch.in_off + (size_t)(IN_U0 + t) * ch.Np, N, d + (size_t)N * N);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(cinv, d, (size_t)N * N * sizeof(double), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    hipFree(d);
    if (e != hipSuccess) return fail(QML_ERR_HIP, std::string("qml_chunk_marginalize: ") + hipGetErrorString(e));
    return QML_OK;
}

int qml_get_timings(qml_ctx *c, qml_timings *t) {
    if (!c || !t) return fail(QML_ERR_ARG, "qml_get_timings: null argument");
    *t = c->tm;
    return QML_OK;
}

int qml_set_pipeline(qml_ctx *c, int on) {
    if (!c) return fail(QML_ERR_ARG, "qml_set_pipeline: null context");
    c->pipeline = on ? 1 : 0;
    return QML_OK;
}

int qml_autotune(qml_ctx *c, const double *theta, const int *caps, int ncaps, int reps, int *best_cap,
                 double *best_ms) {
    if (!c || !c->committed) return fail(QML_ERR_ARG, "qml_autotune: commit first");
    if (!theta || !caps || ncaps < 1) return fail(QML_ERR_ARG, "qml_autotune: no theta or no candidates");
    for (int i = 0; i < ncaps; ++i)
        if (caps[i] < 1) return fail(QML_ERR_ARG, "qml_autotune: a candidate cap below 1");
    HIPCHK(hipSetDevice(c->device));
    reps = std::max(reps, 1);
    const int T = c->T;
    std::vector<double> F((size_t)T * T), dbt((size_t)3 * T);
    int best = -1;
    double best_t = 0.0;
    const int prev_cap = c->max_batch, prev_buckets = c->np_buckets;
    int best_buckets = prev_buckets;
    // a plan that cannot be made (e.g. a cap whose arenas do not fit) ends the
    // search; the previous cap's plan is restored so the context stays usable
    auto replan = [&](int cap, int buckets) {
        release_plan(c);
        c->max_batch = cap;
        c->np_buckets = buckets;
        const int rc = plan_batches(c);
        if (rc != QML_OK) {
            const std::string why = g_err;
            release_plan(c);
            c->max_batch = prev_cap;
            c->np_buckets = prev_buckets;
            if (plan_batches(c) != QML_OK) return fail(rc, why + " (and the previous plan could not be restored)");
            return fail(rc, why);
        }
        return QML_OK;
    };
    // each cap with and without a batch break at every change of padded size
    for (int i = 0; i < 2 * ncaps; ++i) {
        const int cap = caps[i / 2], buckets = i % 2;
        {
            const int rc = replan(cap, buckets);
            if (rc != QML_OK) return rc;
        }
        double t = 0.0;
        for (int r = 0; r <= reps; ++r) {   // r = 0 warms the new plan up
            const int rc = qml_iterate(c, theta, F.data(), dbt.data());
            if (rc != QML_OK) return rc;
            if (r > 0 && (r == 1 || c->tm.total_ms < t)) t = c->tm.total_ms;
        }
        if (best < 0 || t < best_t) {
            best = cap;
            best_buckets = buckets;
            best_t = t;
        }
    }
    if (c->max_batch != best || c->np_buckets != best_buckets) {
        const int rc = replan(best, best_buckets);
        if (rc != QML_OK) return rc;
    }
    if (c->prec >= 2) {
        // staging of the split-BF16 products' operands, timed on the chosen
        // plan one product at a time (left, right, Gram), each keeping the
        // best choice found for the products before it
        int st[3] = {c->nst_left, c->nst_right, c->nst_gram};
        double bt = -1.0;
        for (int p = 0; p < 3; ++p) {
            int bp = st[p];
            for (int v = 1; v <= 4; ++v) {
                if (p == 1 && v == 2) continue;   // the right product has no two-register-set form
                int trial[3] = {st[0], st[1], st[2]};
                trial[p] = v;
                const int rs = qml_set_kernel_stages(c, trial[0], trial[1], trial[2]);
                if (rs != QML_OK) return rs;
                double t = 0.0;
                for (int r = 0; r <= reps; ++r) {   // r = 0 warms the variant up
                    const int rc = qml_iterate(c, theta, F.data(), dbt.data());
                    if (rc != QML_OK) return rc;
                    if (r > 0 && (r == 1 || c->tm.total_ms < t)) t = c->tm.total_ms;
                }
                if (bt < 0 || t < bt) {
                    bt = t;
                    bp = v;
                }
            }
            st[p] = bp;
        }
        const int rs = qml_set_kernel_stages(c, st[0], st[1], st[2]);
        if (rs != QML_OK) return rs;
        best_t = bt;
    }
    if (best_cap) *best_cap = best;
    if (best_ms) *best_ms = best_t;
    return QML_OK;
}

int qml_set_kernel_stages(qml_ctx *c, int left, int right, int gram) {
    if (!c) return fail(QML_ERR_ARG, "qml_set_kernel_stages: null context");
    if (left < 1 || left > 4 || right < 1 || right > 4 || right == 2 || gram < 1 || gram > 4)
        return fail(QML_ERR_ARG, "qml_set_kernel_stages: staging must be 1 to 4 (right: 1, 3 or 4)");
    if (left != c->nst_left || right != c->nst_right || gram != c->nst_gram) {
        HIPCHK(hipSetDevice(c->device));
        drop_graph(c);   // a captured sequence holds the launches of the old kernels
    }
    c->nst_left = left;
    c->nst_right = right;
    c->nst_gram = gram;
    return QML_OK;
}

int qml_set_np_buckets(qml_ctx *c, int on) {
    if (!c) return fail(QML_ERR_ARG, "qml_set_np_buckets: null context");
    on = on ? 1 : 0;
    if (on == c->np_buckets) return QML_OK;
    const int prev = c->np_buckets;
    c->np_buckets = on;
    if (c->committed) {   // plan again
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->st));
        HIPCHK(hipStreamSynchronize(c->st2));
        release_plan(c);
        const int rc = plan_batches(c);
        if (rc != QML_OK) {
            const std::string why = g_err;
            release_plan(c);
            c->np_buckets = prev;
            plan_batches(c);
            return fail(rc, why);
        }
    }
    return QML_OK;
}

int qml_get_batch(qml_ctx *c, int batch, qml_batch_info *info) {
    if (!c || !info) return fail(QML_ERR_ARG, "qml_get_batch: null argument");
    if (!c->committed) return fail(QML_ERR_ARG, "qml_get_batch: commit first");
    if (batch < 0 || batch >= (int)c->batches.size()) return fail(QML_ERR_ARG, "qml_get_batch: no such batch");
    const Batch &B = c->batches[batch];
    info->n_chunks = (int)B.ids.size();
    info->np_min = 1 << 30;
    info->np_max = 0;
    info->members_max = 0;
    for (int id : B.ids) {
        const ChunkHost &ch = c->chunks[id];
        info->np_min = std::min(info->np_min, ch.Np);
        info->np_max = std::max(info->np_max, ch.Np);
        info->members_max = std::max(info->members_max, ch.m);
    }
    return QML_OK;
}

int qml_set_stack_slab(qml_ctx *c, int members) {
    if (!c) return fail(QML_ERR_ARG, "qml_set_stack_slab: null context");
    if (members < 0) return fail(QML_ERR_ARG, "qml_set_stack_slab: members must be >= 0");
    if (members == c->slab) return QML_OK;
    const int prev = c->slab;
    c->slab = members;
    if (c->committed) {   // plan again under the new arenas
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->st));
        HIPCHK(hipStreamSynchronize(c->st2));
        release_plan(c);
        const int rc = plan_batches(c);
        if (rc != QML_OK) {
            const std::string why = g_err;
            release_plan(c);
            c->slab = prev;
            plan_batches(c);
            return fail(rc, why);
        }
    }
    return QML_OK;
}

int qml_get_schedule(qml_ctx *c, qml_schedule *s) {
    if (!c || !s) return fail(QML_ERR_ARG, "qml_get_schedule: null argument");
    s->stack_slab = c->slab;
    s->max_chunk_gb = c->max_chunk_bytes / 1073741824.0;
    s->max_batch_chunks = c->max_batch;
    s->n_batches = (int)c->batches.size();
    s->stages_left = c->nst_left;
    s->stages_right = c->nst_right;
    s->np_buckets = c->np_buckets;
    s->stages_gram = c->nst_gram;
    s->pipeline = c->pipeline;
    s->graph = c->use_graph;
    return QML_OK;
}

int qml_set_graph(qml_ctx *c, int on) {
    if (!c) return fail(QML_ERR_ARG, "qml_set_graph: null context");
    c->use_graph = on ? 1 : 0;
    if (!c->use_graph) {
        HIPCHK(hipSetDevice(c->device));
        drop_graph(c);
    }
    return QML_OK;
}

int qml_set_chunk_active(qml_ctx *c, int id, int active) {
    if (!c || !c->committed || id < 0 || id >= (int)c->chunks.size())
        return fail(QML_ERR_ARG, "qml_set_chunk_active: bad chunk id or not committed");
    HIPCHK(hipSetDevice(c->device));
    const int skip = active ? 0 : 1;
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipStreamSynchronize(c->st2));
    HIPCHK(hipMemcpy(c->d_skip + id, &skip, sizeof(int), hipMemcpyHostToDevice));
    return QML_OK;
}

int qml_poisson_draws(unsigned long long seed, long long n, int *out) {
    if (n < 0 || (n > 0 && !out)) return fail(QML_ERR_ARG, "qml_poisson_draws: bad count or null output");
    // PoissonRNG of core/bootstrapper.hpp:19-37: a 64-bit Mersenne twister
    // seeded once, unit-mean Poisson draws from the standard library
    std::mt19937_64 rng(seed);
    std::poisson_distribution<int> pd;
    for (long long i = 0; i < n; ++i) out[i] = pd(rng);
    return QML_OK;
}

}  // extern "C"

namespace {
// The chunks of each spectrum as a CSR list in chunk order (qptr: nqso + 1
// offsets into qchunk); every spectrum must hold a chunk.
int spectra_lists(const qml_ctx *c, int nqso, const int *qso_of_chunk, std::vector<int> &qptr,
                  std::vector<int> &qchunk, const char *who) {
    const int nch = (int)c->chunks.size();
    qptr.assign(nqso + 1, 0);
    for (int i = 0; i < nch; ++i) {
        const int q = qso_of_chunk[i];
        if (q < -1 || q >= nqso) return fail(QML_ERR_ARG, std::string(who) + ": spectrum index out of range");
        if (q >= 0) ++qptr[q + 1];
    }
    for (int q = 0; q < nqso; ++q) {
        if (qptr[q + 1] == 0) return fail(QML_ERR_ARG, std::string(who) + ": a spectrum without chunks");
        qptr[q + 1] += qptr[q];
    }
    qchunk.resize(qptr[nqso]);
    std::vector<int> fill(qptr.begin(), qptr.end() - 1);
    for (int i = 0; i < nch; ++i)
        if (qso_of_chunk[i] >= 0) qchunk[fill[qso_of_chunk[i]]++] = i;   // chunk order within a spectrum
    return QML_OK;
}
}  // namespace

extern "C" {

int qml_bootstrap(qml_ctx *c, int nboots, unsigned long long seed, int nqso, const int *qso_of_chunk,
                  double *realizations, double *mean, double *cov) {
    if (!c || !c->committed) return fail(QML_ERR_ARG, "qml_bootstrap: commit and iterate first");
    if (nboots < 2 || nqso < 1 || !qso_of_chunk || !mean || !cov)
        return fail(QML_ERR_ARG, "qml_bootstrap: nboots >= 2, nqso >= 1 and the outputs are required");
    const int T = c->T;
    std::vector<int> qptr, qchunk;
    {
        const int rc = spectra_lists(c, nqso, qso_of_chunk, qptr, qchunk, "qml_bootstrap");
        if (rc != QML_OK) return rc;
    }
    // the weights of spectrum q are draws q * nboots .. (q + 1) * nboots - 1
    std::vector<int> P((size_t)nqso * nboots);
    {
        const int rc = qml_poisson_draws(seed, (long long)P.size(), P.data());
        if (rc != QML_OK) return rc;
    }
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipStreamSynchronize(c->st2));
    int *d_qptr = nullptr, *d_qchunk = nullptr, *d_P = nullptr;
    double *d_theta = nullptr, *d_R = nullptr, *d_mean = nullptr, *d_cov = nullptr;
    auto release = [&]() {
        void *ps[] = {d_qptr, d_qchunk, d_P, d_theta, d_R, d_mean, d_cov};
        for (void *p : ps) if (p) hipFree(p);
    };
    auto run = [&]() -> int {
        HIPCHK(hipMalloc(&d_qptr, qptr.size() * sizeof(int)));
        HIPCHK(hipMalloc(&d_qchunk, std::max<size_t>(qchunk.size(), 1) * sizeof(int)));
        HIPCHK(hipMalloc(&d_P, P.size() * sizeof(int)));
        HIPCHK(hipMalloc(&d_theta, (size_t)nqso * T * sizeof(double)));
        HIPCHK(hipMalloc(&d_R, (size_t)nboots * T * sizeof(double)));
        HIPCHK(hipMalloc(&d_mean, (size_t)T * sizeof(double)));
        HIPCHK(hipMalloc(&d_cov, (size_t)T * T * sizeof(double)));
        HIPCHK(hipMemcpy(d_qptr, qptr.data(), qptr.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!qchunk.empty())
            HIPCHK(hipMemcpy(d_qchunk, qchunk.data(), qchunk.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_P, P.data(), P.size() * sizeof(int), hipMemcpyHostToDevice));
        launch_bootstrap(c->st, c->d_outs, c->d_dbt, d_qptr, d_qchunk, nqso, T, d_P, nboots, d_theta, d_R, d_mean,
                         d_cov);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
        if (realizations)
            HIPCHK(hipMemcpy(realizations, d_R, (size_t)nboots * T * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(mean, d_mean, (size_t)T * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(cov, d_cov, (size_t)T * T * sizeof(double), hipMemcpyDeviceToHost));
        return QML_OK;
    };
    const int rc = run();
    release();
    return rc;
}

int qml_bootstrap_complete(qml_ctx *c, int nboots, unsigned long long seed, int nqso, const int *qso_of_chunk,
                           const int *weights, double *realizations, int *valid, int *nvalid, double *mean,
                           double *cov) {
    if (!c || !c->committed) return fail(QML_ERR_ARG, "qml_bootstrap_complete: commit and iterate first");
    if (nboots < 2 || nqso < 1 || !qso_of_chunk || !mean || !cov || !nvalid)
        return fail(QML_ERR_ARG, "qml_bootstrap_complete: nboots >= 2, nqso >= 1 and the outputs are required");
    const int T = c->T, Np = (int)align_up(T, 64);
    std::vector<int> qptr, qchunk;
    {
        const int rc = spectra_lists(c, nqso, qso_of_chunk, qptr, qchunk, "qml_bootstrap_complete");
        if (rc != QML_OK) return rc;
    }
    // istart and ndim of each spectrum (collapseBootstrap), its Fisher block
    // offsets, and the spectra covering each band, in queue order
    std::vector<int> qstart(nqso), qdim(nqso);
    std::vector<long long> fq_off(nqso + 1, 0);
    for (int q = 0; q < nqso; ++q) {
        int lo = T, hi = 0;
        for (int k = qptr[q]; k < qptr[q + 1]; ++k) {
            const ChunkOut &o = c->outs[qchunk[k]];
            lo = std::min(lo, o.fis);
            hi = std::max(hi, o.fis + o.nq);
        }
        qstart[q] = lo;
        qdim[q] = hi - lo;
        fq_off[q + 1] = fq_off[q] + (long long)qdim[q] * qdim[q];
    }
    std::vector<int> band_ptr(T + 1, 0), band_q;
    for (int q = 0; q < nqso; ++q)
        for (int i = qstart[q]; i < qstart[q] + qdim[q]; ++i) ++band_ptr[i + 1];
    for (int i = 0; i < T; ++i) band_ptr[i + 1] += band_ptr[i];
    band_q.resize(band_ptr[T]);
    {
        std::vector<int> fill(band_ptr.begin(), band_ptr.end() - 1);
        for (int q = 0; q < nqso; ++q)
            for (int i = qstart[q]; i < qstart[q] + qdim[q]; ++i) band_q[fill[i]++] = q;
    }
    // weights: realisation-major, one draw per spectrum (_oneBoot draws one
    // per spectrum in queue order for every realisation)
    std::vector<int> W((size_t)nboots * nqso);
    if (weights) {
        std::copy(weights, weights + W.size(), W.begin());
    } else {
        const int rc = qml_poisson_draws(seed, (long long)W.size(), W.data());
        if (rc != QML_OK) return rc;
    }
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipStreamSynchronize(c->st2));
    const int BATCH = std::min(nboots, 128);
    std::vector<void *> allocs;
    auto dalloc = [&](void **p, size_t bytes) -> hipError_t {
        const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 8));
        if (e == hipSuccess) allocs.push_back(*p);
        return e;
    };
    std::vector<double> R((size_t)nboots * T);
    std::vector<int> status(nboots, 0);
    auto run = [&]() -> int {
        int *d_qptr, *d_qchunk, *d_qstart, *d_qdim, *d_bptr, *d_bq, *d_W, *d_status;
        long long *d_fqoff;
        double *d_fq, *d_theta, *d_mat, *d_rhs, *d_R, *d_mean, *d_cov;
        ChunkDev *d_cd;
        HIPCHK(dalloc((void **)&d_qptr, qptr.size() * sizeof(int)));
        HIPCHK(dalloc((void **)&d_qchunk, qchunk.size() * sizeof(int)));
        HIPCHK(dalloc((void **)&d_qstart, nqso * sizeof(int)));
        HIPCHK(dalloc((void **)&d_qdim, nqso * sizeof(int)));
        HIPCHK(dalloc((void **)&d_fqoff, (nqso + 1) * sizeof(long long)));
        HIPCHK(dalloc((void **)&d_bptr, band_ptr.size() * sizeof(int)));
        HIPCHK(dalloc((void **)&d_bq, band_q.size() * sizeof(int)));
        HIPCHK(dalloc((void **)&d_W, W.size() * sizeof(int)));
        HIPCHK(dalloc((void **)&d_fq, (size_t)fq_off[nqso] * sizeof(double)));
        HIPCHK(dalloc((void **)&d_theta, (size_t)nqso * T * sizeof(double)));
        HIPCHK(dalloc((void **)&d_mat, (size_t)BATCH * 3 * Np * Np * sizeof(double)));
        HIPCHK(dalloc((void **)&d_rhs, (size_t)BATCH * Np * sizeof(double)));
        HIPCHK(dalloc((void **)&d_R, R.size() * sizeof(double)));
        HIPCHK(dalloc((void **)&d_status, BATCH * sizeof(int)));
        HIPCHK(dalloc((void **)&d_mean, (size_t)T * sizeof(double)));
        HIPCHK(dalloc((void **)&d_cov, (size_t)T * T * sizeof(double)));
        HIPCHK(dalloc((void **)&d_cd, BATCH * sizeof(ChunkDev)));
        HIPCHK(hipMemcpy(d_qptr, qptr.data(), qptr.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_qchunk, qchunk.data(), qchunk.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_qstart, qstart.data(), nqso * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_qdim, qdim.data(), nqso * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_fqoff, fq_off.data(), (nqso + 1) * sizeof(long long), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_bptr, band_ptr.data(), band_ptr.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!band_q.empty())
            HIPCHK(hipMemcpy(d_bq, band_q.data(), band_q.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_W, W.data(), W.size() * sizeof(int), hipMemcpyHostToDevice));
        // realisation b of a batch is factorised as a "chunk" of size Np at
        // b * 3 Np^2 (A, L, Li), its status at b
        std::vector<ChunkDev> cd(BATCH);
        for (int b = 0; b < BATCH; ++b) {
            cd[b] = ChunkDev{};
            cd[b].N = T;
            cd[b].Np = Np;
            cd[b].nt = Np / TILE;
            cd[b].gidx = b;
            cd[b].mat_off = (long long)b * 3 * Np * Np;
        }
        HIPCHK(hipMemcpy(d_cd, cd.data(), BATCH * sizeof(ChunkDev), hipMemcpyHostToDevice));
        launch_boot_collapse(c->st, c->d_outs, c->d_dbt, d_qptr, d_qchunk, nqso, T, d_theta);
        launch_boot_collapse_fisher(c->st, c->d_outs, c->d_fc, d_qptr, d_qchunk, d_qstart, d_qdim, d_fqoff, nqso,
                                    d_fq);
        HIPCHK(hipGetLastError());
        LaunchCtx L{};
        L.stream = c->st;
        L.chunks = d_cd;
        L.np_max = Np;
        L.mat = d_mat;
        for (int b0 = 0; b0 < nboots; b0 += BATCH) {
            const int nb = std::min(BATCH, nboots - b0);
            L.nchunks = nb;
            launch_boot_build(c->st, d_W, nqso, b0, nb, T, Np, d_bptr, d_bq, d_qstart, d_qdim, d_fqoff, d_fq,
                              d_theta, d_mat, d_rhs);
            launch_cholesky(L, d_status);
            launch_boot_solve(c->st, d_mat, d_rhs, b0, nb, T, Np, d_R);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(status.data() + b0, d_status, nb * sizeof(int), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
        }
        HIPCHK(hipMemcpy(R.data(), d_R, R.size() * sizeof(double), hipMemcpyDeviceToHost));
        // the valid realisations, in order (allpowers + ii T, ii counting the valid ones)
        int nv = 0;
        for (int b = 0; b < nboots; ++b) {
            if (valid) valid[b] = status[b] == 0;
            if (status[b] != 0) continue;
            if (nv != b) std::copy(R.begin() + (size_t)b * T, R.begin() + (size_t)(b + 1) * T, R.begin() + (size_t)nv * T);
            ++nv;
        }
        *nvalid = nv;
        if (realizations) std::copy(R.begin(), R.begin() + (size_t)nv * T, realizations);
        if (nv < 2) return fail(QML_ERR_ARG, "qml_bootstrap_complete: fewer than two valid realisations");
        HIPCHK(hipMemcpy(d_R, R.data(), (size_t)nv * T * sizeof(double), hipMemcpyHostToDevice));
        launch_boot_meancov(c->st, d_R, nv, T, d_mean, d_cov);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
        HIPCHK(hipMemcpy(mean, d_mean, (size_t)T * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(cov, d_cov, (size_t)T * T * sizeof(double), hipMemcpyDeviceToHost));
        return QML_OK;
    };
    const int rc = run();
    for (void *p : allocs) hipFree(p);
    return rc;
}

int qml_bootstrap_median(qml_ctx *c, int nboots, int T, const double *realizations, double *median,
                         double *mad_cov) {
    if (!c || nboots < 1 || T < 1 || !realizations || !median || !mad_cov)
        return fail(QML_ERR_ARG, "qml_bootstrap_median: a context, nboots >= 1, T >= 1 and every array are required");
    if (T > 65535) return fail(QML_ERR_LIMIT, "qml_bootstrap_median: more than 65535 bands");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipStreamSynchronize(c->st2));
    const size_t nrt = (size_t)nboots * T;
    double *d_R = nullptr, *d_X = nullptr, *d_med = nullptr, *d_mad = nullptr;
    auto run = [&]() -> int {
        HIPCHK(hipMalloc(&d_R, nrt * sizeof(double)));
        HIPCHK(hipMalloc(&d_X, nrt * sizeof(double)));
        HIPCHK(hipMalloc(&d_med, (size_t)T * sizeof(double)));
        HIPCHK(hipMalloc(&d_mad, (size_t)T * T * sizeof(double)));
        HIPCHK(hipMemcpy(d_R, realizations, nrt * sizeof(double), hipMemcpyHostToDevice));
        launch_boot_median(c->st, d_R, nboots, T, d_X, d_med, d_mad);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
        HIPCHK(hipMemcpy(median, d_med, (size_t)T * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(mad_cov, d_mad, (size_t)T * T * sizeof(double), hipMemcpyDeviceToHost));
        return QML_OK;
    };
    const int rc = run();
    void *ps[] = {d_R, d_X, d_med, d_mad};
    for (void *p : ps) if (p) hipFree(p);
    return rc;
}

int qml_bootstrap_median_noalloc(qml_ctx *c, int nboots, int T, const double *d_realizations, double *d_median,
                                 double *d_mad_cov) {
    if (!c || nboots < 1 || T < 1 || !d_realizations || !d_median || !d_mad_cov)
        return fail(QML_ERR_ARG,
                    "qml_bootstrap_median_noalloc: a context, nboots >= 1, T >= 1 and every array are required");
    if (T > 65535) return fail(QML_ERR_LIMIT, "qml_bootstrap_median_noalloc: more than 65535 bands");
    if ((long long)nboots * T > BOOT_NOALLOC_MAX) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "qml_bootstrap_median_noalloc: nboots x T = %lld exceeds the %lld-element scratch "
                      "(use qml_bootstrap_median)", (long long)nboots * T, (long long)BOOT_NOALLOC_MAX);
        return fail(QML_ERR_LIMIT, msg);
    }
    HIPCHK(hipSetDevice(c->device));
    // the arrays must be device memory of this context's device: a host
    // pointer would fault in the kernels
    const void *arrays[] = {d_realizations, d_median, d_mad_cov};
    for (const void *a : arrays) {
        hipPointerAttribute_t at{};
        if (hipPointerGetAttributes(&at, a) != hipSuccess || at.device != c->device ||
            (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)) {
            (void)hipGetLastError();
            return fail(QML_ERR_ARG, "qml_bootstrap_median_noalloc: arrays must be device memory of the "
                                     "context's device");
        }
    }
    double *X = boot_noalloc_scratch(c->device);
    if (!X) return fail(QML_ERR_HIP, "qml_bootstrap_median_noalloc: scratch not set up on this device");
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipStreamSynchronize(c->st2));
    // the transpose writes every scratch element the selections read, so
    // nothing needs clearing between calls
    launch_boot_median(c->st, d_realizations, nboots, T, X, d_median, d_mad_cov);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    return QML_OK;
}

int qml_num_chunks(qml_ctx *c) { return c ? (int)c->chunks.size() : 0; }
int qml_total_kz_bins(qml_ctx *c) { return c ? c->T : 0; }

int qml_make_tables(qml_ctx *c, const qml_table_spec *s, int npairs, const int *R_fwhm, const double *dv_kms,
                    double *deriv, double *signal) {
    if (!c || !s || npairs < 0 || (npairs > 0 && (!R_fwhm || !dv_kms)))
        return fail(QML_ERR_ARG, "qml_make_tables: null argument");
    if (c->kedges.empty()) return fail(QML_ERR_ARG, "qml_make_tables: set the bins first");
    if (s->nv < 2 || s->nz < 2 || !(s->v_length > 0) || s->filon_panels < 0 || s->filon_panels > 128)
        return fail(QML_ERR_ARG, "qml_make_tables: invalid table grid");
    for (int p = 0; p < npairs; ++p)
        if (R_fwhm[p] <= 0 || !(dv_kms[p] > 0)) return fail(QML_ERR_ARG, "qml_make_tables: invalid (R, dv) pair");
    if (npairs == 0 || (!deriv && !signal)) return QML_OK;
    HIPCHK(hipSetDevice(c->device));
    const int nk = (int)c->kedges.size() - 1, nv = s->nv, nz = s->nz;
    const int M = s->filon_panels > 0 ? s->filon_panels : 64;
    const double dxv = s->v_length / (nv - 1.0);
    // the signal grid (core/sq_table.cpp:221-227)
    const long long nrft = 1LL << ((int)std::ceil(std::log2(4.0 * nv)) + 1);
    const int nhalf = (int)(nrft / 2), nmode = nhalf + 1;
    const double length = dxv / 2.0 * (double)nrft, dk = 2.0 * 3.14159265358979323846 / length;
    const double z1 = c->zcenters[0] - c->bins.z_bin_width;
    const double dz = c->bins.z_bin_width * (c->bins.nz + 1) / (nz - 1.0);
    const PD13 pd{s->pd13[0], s->pd13[1], s->pd13[2], s->pd13[3], s->pd13[4], s->pd13[5]};
    // pairs per pass: outputs of up to 2 GiB on the device
    const size_t per = (size_t)8 * ((deriv ? (size_t)nk * nv : 0) + (signal ? (size_t)nv * nz : 0));
    const int nb = (int)std::max<size_t>(1, std::min<size_t>(npairs, ((size_t)2 << 30) / per));
    struct Bufs {
        std::vector<void *> p;
        ~Bufs() { for (void *x : p) hipFree(x); }
    } bufs;
    auto dalloc = [&](void **ptr, size_t bytes) {
        const hipError_t e = hipMalloc(ptr, std::max<size_t>(bytes, 8));
        if (e == hipSuccess) bufs.p.push_back(*ptr);
        return e;
    };
    double *d_pairs = nullptr, *d_ke = nullptr, *d_pk = nullptr, *d_q = nullptr, *d_s = nullptr;
    int *d_mend = nullptr;
    HIPCHK(dalloc((void **)&d_pairs, (size_t)2 * nb * sizeof(double)));
    HIPCHK(dalloc((void **)&d_mend, (size_t)nb * sizeof(int)));
    HIPCHK(dalloc((void **)&d_ke, (size_t)(nk + 1) * sizeof(double)));
    if (deriv) HIPCHK(dalloc((void **)&d_q, (size_t)nb * nk * nv * sizeof(double)));
    if (signal) {
        HIPCHK(dalloc((void **)&d_pk, (size_t)nz * nmode * sizeof(double)));
        HIPCHK(dalloc((void **)&d_s, (size_t)nb * nv * nz * sizeof(double)));
    }
    HIPCHK(hipMemcpy(d_ke, c->kedges.data(), (nk + 1) * sizeof(double), hipMemcpyHostToDevice));
    if (signal) launch_pfid_grid(c->st, dk, nmode, z1, dz, nz, pd, d_pk);
    std::vector<double> hp(2 * (size_t)nb);
    std::vector<int> hm(nb);
    for (int p0 = 0; p0 < npairs; p0 += nb) {
        const int n = std::min(nb, npairs - p0);
        for (int i = 0; i < n; ++i) {
            const double R = SPEED_OF_LIGHT / R_fwhm[p0 + i] / 2.35482004503;
            hp[2 * i] = R;
            hp[2 * i + 1] = dv_kms[p0 + i];
            // modes with W^2 < exp(-92) ~ 1e-40 of W^2(0) are left out
            hm[i] = (int)std::min<double>(nhalf, std::ceil(9.6 / (R * dk)));
        }
        HIPCHK(hipMemcpyAsync(d_pairs, hp.data(), 2 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(d_mend, hm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->st));
        if (deriv) launch_deriv_tables(c->st, d_pairs, n, d_ke, nk, nv, dxv, M, d_q);
        if (signal)
            launch_signal_tables(c->st, d_pairs, d_mend, n, d_pk, nmode, nhalf, dk, dxv, 1.0 / length, nv, nz, d_s);
        HIPCHK(hipGetLastError());
        if (deriv)
            HIPCHK(hipMemcpyAsync(deriv + (size_t)p0 * nk * nv, d_q, (size_t)n * nk * nv * sizeof(double),
                                  hipMemcpyDeviceToHost, c->st));
        if (signal)
            HIPCHK(hipMemcpyAsync(signal + (size_t)p0 * nv * nz, d_s, (size_t)n * nv * nz * sizeof(double),
                                  hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    }
    return QML_OK;
}

}  // extern "C"
