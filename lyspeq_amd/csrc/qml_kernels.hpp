// Launch interface between the host orchestration and the kernels.
#pragma once

#include "qml_common.hpp"

namespace qml {

// LDS carve-up of k_rm_sandwich (doubles): Qh qr x 33, T 16 x tw, R rows
// 16 x nbm and tj x nbm, grid-point u then (sin, cos) for npts points; em
// Qh entries per thread (qr = 8 em), npass passes of 32 b columns.
struct RmLayout {
    int qr = 0, tw = 0, nbm = 0, npts = 0, tj = 16, em = 12, npass = 0;
    size_t bytes = 0;
};
// Layout for output tiles of 16 x tj and the batch's largest a range
// (15 ros + rnb), b range ((tj - 1) ros + rnb) and rnb.
RmLayout rmat_layout(int tj, int ra_max, int rb_max, int nb_max);
// Whether launch_rmat has a kernel for this layout.
bool rmat_layout_ok(const RmLayout &lay);

struct LaunchCtx {
    hipStream_t stream;
    const ChunkDev *chunks;   // device, nchunks entries
    int nchunks;
    int np_max;               // largest Np in the batch
    int nvu_max;              // largest theta-table length in the batch
    int nu_max;               // most continuum templates of a chunk in the batch
    int dp_rb, dp_cb;         // d-pass operand blocks: most (KBIN_UPP + 1)/16 and (z bins + 1)(1 + nu)/16, rounded up
    int jlo = 0, jhi = 1 << 30;   // members [jlo, jhi) of the stack slab (all members: 0, large)
    int nq_max;               // most N_Q_MATRICES of a chunk in the batch
    const double *in;         // per-chunk input vectors
    const TableDev *tables;   // device table descriptors
    BinsDev bins;
    double *mat;              // FP64 matrix arena
    double *vec;              // FP64 vector arena
    double *pj;               // aux arena: P_j, At_j, theta tables
    double *sj;               // s_j and d arena
    // resolution-matrix mode
    const double *rin;        // R rows and fine grids
    int rm_kbu;               // largest KBIN_UPP of the batch's resolution-matrix chunks; -1 if none
    RmLayout rm_lay;          // LDS layout of the batch's sandwich launch
};

// kernels_chunk.hip
void launch_assemble(const LaunchCtx &L, const double *theta);
// kernels_rmat.hip: Q_k = R Qhat_k R^T (k < KBIN_UPP) and S = R Shat R^T for
// the batch's resolution-matrix chunks, into their qb_off matrices; false if
// no kernel matches the batch's layout (nothing is launched).
bool launch_rmat(const LaunchCtx &L);
// kernels_factor.hip: L = chol(A) and Li = L^-1 (both formed by
// launch_cholesky; launch_trinv is kept for the timing slot and is empty)
void launch_cholesky(const LaunchCtx &L, int *status);
void launch_trinv(const LaunchCtx &L);
void launch_solve(const LaunchCtx &L);
void launch_dpass(const LaunchCtx &L);
void launch_small(const LaunchCtx &L);
void launch_xgen(const LaunchCtx &L, void *X, int prec);
void launch_convert_li(const LaunchCtx &L, void *lic, int prec);
void launch_gram_finalize(const LaunchCtx &L, const ChunkOut *outs, const double *gram, const int *status,
                          double *fc, double *dbt);
// Debug aid: fills the LDS of every CU with all-ones bytes (NaN in FP64,
// FP32 and bf16), so that a kernel reading LDS it did not write shows it.
void launch_poison_lds(hipStream_t st);
// One matrix of Chunk's protected builders for the chunk at *chunk (n x n,
// row-major into out; which as QML_CMAT_*, fine: v/z on the oversampled grid).
void launch_chunk_matrix(hipStream_t st, const ChunkDev *chunk, const double *in, const double *rin,
                         const TableDev *tabs, const BinsDev &bins, const double *mat, int which, int fine, int kn,
                         int zm, int n, double *out);
// cinv (n x n, row-major, upper triangle read) minus its Sherman-Morrison
// projection on template u; y is n doubles of scratch.
void launch_marginalize(hipStream_t st, double *cinv, const double *u, int n, double *y);
// Adds the results of chunks with status 0 and skip 0, in chunk-id order.
void launch_accumulate(hipStream_t st, const ChunkOut *outs, const int *status, const int *skip, int nchunks,
                       int T, const double *fc, const double *dbt, double *Fsum, double *dbt_sum);

// kernels_gemm.hip.  prec: 0 FP64, 1 FP32, 2 BF16x3, 3 BF16x4.
// nst: staging of the split-BF16 operands (1, 2 register-staged with one or
// two LDS stages, 2 not for the right product; 3, 4 direct-to-LDS loads with
// two or three stages); the FP64/FP32 products ignore it.
// The stack arenas X and Y hold the members [jlo, jlo + slab) of each chunk
// (member j at slot j - jlo); the work lists name global members.
// Y_j = Li X_j for every member of the work list (lower-triangular Li).
void launch_gemm_whiten_left(hipStream_t st, int prec, int nst, const ChunkDev *chunks, const Work *work,
                             int nwork, const void *li, const void *X, void *Y, int jlo);
// G_j = Y_j Li^T on upper tiles, written tile-packed into the G store.
void launch_gemm_whiten_right(hipStream_t st, int prec, int nst, const ChunkDev *chunks, const Work *work,
                              int nwork, const void *li, const void *Y, void *G, int jlo);
// Gram partials E_s = sum over G tile s of G_j G_j'.
void launch_gemm_gram(hipStream_t st, int prec, int nst, const ChunkDev *chunks, const Work *work, int nwork,
                      const void *G, double *gram);

// kernels_boot.hip: FastBootstrap realisations, their mean and covariance.
// qptr/qchunk: CSR list of each spectrum's chunk ids (nqso + 1 offsets); P is
// nqso x nboots Poisson weights; theta nqso x T and R nboots x T scratch.
void launch_bootstrap(hipStream_t st, const ChunkOut *outs, const double *dbt, const int *qptr, const int *qchunk,
                      int nqso, int T, const int *P, int nboots, double *theta, double *R, double *mean,
                      double *cov);

void launch_boot_collapse(hipStream_t st, const ChunkOut *outs, const double *dbt, const int *qptr,
                          const int *qchunk, int nqso, int T, double *theta);
void launch_boot_meancov(hipStream_t st, const double *R, int nboots, int T, double *mean, double *cov);
// Median bootstrap: R nboots x T (device) -> median (T) and the median
// covariance (T x T); X is T x nboots scratch (deviations from the median).
void launch_boot_median(hipStream_t st, const double *R, int nboots, int T, double *X, double *median,
                        double *mad);
// Allocation-free median bootstrap: a static device array of
// BOOT_NOALLOC_MAX doubles holds the transposed realisations.
// boot_noalloc_setup looks its address up on the current device (once per
// device; returns a hipError_t code); boot_noalloc_scratch returns it, or
// null before the setup.
constexpr long long BOOT_NOALLOC_MAX = 1LL << 23;
constexpr int BOOT_NOALLOC_DEVICES = 64;
int boot_noalloc_setup(int device);
double *boot_noalloc_scratch(int device);
// The complete form: per-spectrum collapsed Fisher upper blocks (ndim x ndim
// at fq_off); per batch of nb realisations from b0, F_b and its right-hand
// side in the matrix layout of the batched Cholesky (A, L, Li per
// realisation, Np x Np each), then x = Li^T Li rhs into R.
void launch_boot_collapse_fisher(hipStream_t st, const ChunkOut *outs, const double *fc, const int *qptr,
                                 const int *qchunk, const int *qstart, const int *qdim, const long long *fq_off,
                                 int nqso, double *fq);
void launch_boot_build(hipStream_t st, const int *W, int nqso, int b0, int nb, int T, int Np, const int *band_ptr,
                       const int *band_q, const int *qstart, const int *qdim, const long long *fq_off,
                       const double *fq, const double *theta, double *mat, double *rhs);
void launch_boot_solve(hipStream_t st, const double *mat, const double *rhs, int b0, int nb, int T, int Np,
                       double *R);

// S/Q lookup tables (kernels_tables.hip)
struct PD13 {
    double A, n, alpha, B, beta, lambda;
};
void launch_deriv_tables(hipStream_t st, const double *pairs, int npairs, const double *kedges, int nk, int nv,
                         double dxv, int M, double *out);
void launch_pfid_grid(hipStream_t st, double dk, int nmode, double z1, double dz, int nz, PD13 pd, double *pk);
void launch_signal_tables(hipStream_t st, const double *pairs, const int *mend, int npairs, const double *pk,
                          int nmode, int nhalf, double dk, double dxv, double inv_len, int nv, int nz, double *out);

}  // namespace qml
