// Dense MFMA products of the QML engine (gfx950).
//
//   whiten_left :  Y_j = Li X_j              (Li lower triangular: K-range of
//                                             tile row I is [0, 64 (I+1)); X_j is
//                                             symmetric and read row-wise as B^T)
//   whiten_right:  G_j = Y_j Li^T            (upper tiles only, K-range
//                                             [0, 64 (J+1)), tile-packed out)
// The split-BF16x3 products instead write X_j = M + M^T (M the lower
// triangle of X_j with its diagonal halved, stored as T = M^T):
//   whiten_left :  P_j = Li M                (lower 128-tiles, K-range of tile
//                                             (I, J) is [128 J, 128 (I+1)))
//   whiten_right:  G_j = P_j Li^T + Li P_j^T (upper tiles, K-range [0, 128 (I+1)))
//   gram        :  E_s = sum_{k in G tile s} G_j[k] G_j'[k]   (split-K partials,
//                                             x2 for off-diagonal G tiles)
//
// FP64 and FP32 (k_gemm): one 256-thread workgroup computes a 64x64 output
// tile; each of its four waves owns a 32x32 quarter made of 2x2 16x16 MFMA
// blocks:
//   FP64   v_mfma_f64_16x16x4f64, FP64 operands and accumulation;
//   FP32   v_mfma_f32_16x16x4f32, FP32 operands and accumulation.
// Split BF16 (k_gemm_bf3): 128x128 output tiles on operands stored as two
// bf16 planes (hi, lo), v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
//   BF16x3 a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi;
//   BF16x4 the same with a_lo b_lo added first (four MFMAs per block).
// The work list holds (chunk, member or split, I, J) per workgroup; work is
// assigned so that consecutive items stay on one XCD (blockIdx % 8 selects
// the XCD), keeping a chunk's operands in that XCD's L2.
#include "qml_common.hpp"
#include "qml_kernels.hpp"

#include <climits>

namespace qml {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

enum { MODE_LEFT = 0, MODE_RIGHT = 1, MODE_GRAM = 2 };

__device__ __forceinline__ int upper_tile_index(int I, int J, int nt) {
    return I * nt - I * (I - 1) / 2 + (J - I);
}

__device__ __forceinline__ bool packed_tile_is_diag(int s, int nt) {
    int I = 0;
    while (s >= nt - I) { s -= nt - I; ++I; }
    return s == 0;
}

__device__ __forceinline__ int xcd_logical(int nwork) {
    const int per = (nwork + 7) / 8;
    return (blockIdx.x % 8) * per + blockIdx.x / 8;
}

// Problem description of one work item.
template <class T>
struct Prob {
    const T *A;  // row-major, rows of the output tile, K contiguous (lda)
    const T *B;  // NN: row-major K x N (ldb);  NT: row-major N x K (ldb)
    int lda, ldb;
    int kbeg, kend;
    int arows, brows;   // valid rows of A / B^T (members for the Gram); others read as 0
    bool nt;
};

// ------------------------------------------------------ FP64 / FP32 path --
template <class T> struct MF;
template <> struct MF<double> {
    using acc_t = d4;
    static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct MF<float> {
    using acc_t = f4;
    static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <class T>
__device__ __forceinline__ void mainloop_k4(const Prob<T> &p, int I0, int J0,
                                            typename MF<T>::acc_t (&acc)[2][2]) {
    constexpr int KB = 16, PAD = 4;
    __shared__ T As[KB][64 + PAD];
    __shared__ T Bs[KB][64 + PAD];
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    const int wm = w / 2, wn = w % 2;
    // A pattern: row tid/4, four consecutive k at (tid%4)*4
    const int ar = tid / 4, ak = (tid % 4) * 4;
    // NN pattern for B: k row tid/16, four consecutive n at (tid%16)*4
    const int bk = tid / 16, bn = (tid % 16) * 4;
    T ra[4], rb[4];
    auto gload = [&](int k0) {
        const bool aok = (I0 + ar) < p.arows;
        const T *pa = p.A + (size_t)(I0 + ar) * p.lda + k0 + ak;
#pragma unroll
        for (int u = 0; u < 4; ++u) ra[u] = aok ? pa[u] : (T)0;
        if (p.nt) {
            const bool bok = (J0 + ar) < p.brows;
            const T *pb = p.B + (size_t)(J0 + ar) * p.ldb + k0 + ak;
#pragma unroll
            for (int u = 0; u < 4; ++u) rb[u] = bok ? pb[u] : (T)0;
        } else {
            const T *pb = p.B + (size_t)(k0 + bk) * p.ldb + J0 + bn;
#pragma unroll
            for (int u = 0; u < 4; ++u) rb[u] = pb[u];
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int u = 0; u < 4; ++u) As[ak + u][ar] = ra[u];
        if (p.nt) {
#pragma unroll
            for (int u = 0; u < 4; ++u) Bs[ak + u][ar] = rb[u];
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) Bs[bk][bn + u] = rb[u];
        }
    };
    if (p.kbeg >= p.kend) return;
    gload(p.kbeg);
    sstore();
    __syncthreads();
    for (int k0 = p.kbeg; k0 < p.kend; k0 += KB) {
        const bool more = k0 + KB < p.kend;
        if (more) gload(k0 + KB);
#pragma unroll
        for (int kk = 0; kk < KB; kk += 4) {
            T a[2], b[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                a[q] = As[kk + lane / 16][wm * 32 + q * 16 + lane % 16];
                b[q] = Bs[kk + lane / 16][wn * 32 + q * 16 + lane % 16];
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) acc[x][y] = MF<T>::mfma(a[x], b[y], acc[x][y]);
        }
        __syncthreads();
        if (more) sstore();
        __syncthreads();
    }
}

// ------------------------------------------------------------------ kernel --
template <int PREC, int MODE>
__global__ __launch_bounds__(256) void k_gemm(const ChunkDev *cs, const Work *work, int nwork,
                                              const void *Abase, const void *Bbase, void *Cbase, int jlo) {
    using T = typename std::conditional<PREC == 0, double, float>::type;
    using acc_t = typename std::conditional<PREC == 0, d4, f4>::type;
    const int li = xcd_logical(nwork);
    if (li >= nwork) return;
    const Work wk = work[li];
    const ChunkDev c = cs[wk.c];
    const int Np = c.Np;
    const size_t msz = (size_t)Np * Np;
    Prob<T> p;
    int I0, J0;
    if (MODE == MODE_LEFT) {
        p.A = (const T *)Abase + c.lic_off;
        // X_j is symmetric, so its rows serve as the K-contiguous rows of B^T
        p.B = (const T *)Bbase + c.stk_off + (wk.j - jlo) * msz;
        p.lda = p.ldb = Np;
        p.kbeg = 0;
        p.kend = (wk.I + 1) * TILE;
        p.arows = p.brows = Np;
        p.nt = (PREC != 0);
        I0 = wk.I * TILE;
        J0 = wk.J * TILE;
    } else if (MODE == MODE_RIGHT) {
        p.A = (const T *)Abase + c.stk_off + (wk.j - jlo) * msz;
        p.B = (const T *)Bbase + c.lic_off;
        p.lda = p.ldb = Np;
        p.kbeg = 0;
        p.kend = (wk.J + 1) * TILE;
        p.arows = p.brows = Np;
        p.nt = true;
        I0 = wk.I * TILE;
        J0 = wk.J * TILE;
    } else {
        const int ktot = c.ntu * TILE * TILE;
        p.A = (const T *)Abase + c.g_off;
        p.B = p.A;
        p.lda = p.ldb = ktot;
        p.kbeg = wk.j * TILE * TILE;
        p.kend = p.kbeg + TILE * TILE;
        p.arows = p.brows = c.m;
        p.nt = true;
        I0 = wk.I * TILE;
        J0 = wk.J * TILE;
    }
    acc_t acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[x][y][u] = 0;
    mainloop_k4<T>(p, I0, J0, acc);

    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int wm = w / 2, wn = w % 2;
    // row of accumulator element u inside a 16x16 block: the FP64 MFMA
    // holds rows u*4 + lane/16, the FP32-output MFMAs rows 4*(lane/16) + u
    auto brow = [&](int u) { return (PREC == 0) ? (4 * u + lane / 16) : (4 * (lane / 16) + u); };
    if (MODE == MODE_LEFT) {
        T *C = (T *)Cbase + c.stk_off + (wk.j - jlo) * msz;
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = I0 + wm * 32 + x * 16 + brow(u);
                    const int q = J0 + wn * 32 + y * 16 + lane % 16;
                    C[(size_t)r * Np + q] = (T)acc[x][y][u];
                }
    } else if (MODE == MODE_RIGHT) {
        T *C = (T *)Cbase + c.g_off + (size_t)wk.j * c.ntu * TILE * TILE
             + (size_t)upper_tile_index(wk.I, wk.J, c.nt) * TILE * TILE;
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = wm * 32 + x * 16 + brow(u);
                    const int q = wn * 32 + y * 16 + lane % 16;
                    C[r * TILE + q] = (T)acc[x][y][u];
                }
    } else {
        const int ntile = c.mt * (c.mt + 1) / 2;
        const double wgt = packed_tile_is_diag(wk.j, c.nt) ? 1.0 : 2.0;
        double *C = (double *)Cbase + c.gram_off
                  + ((size_t)wk.j * ntile + upper_tile_index(wk.I, wk.J, c.mt)) * TILE * TILE;
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = wm * 32 + x * 16 + brow(u);
                    const int q = wn * 32 + y * 16 + lane % 16;
                    C[r * TILE + q] = wgt * (double)acc[x][y][u];
                }
    }
}

void launch_bf3(hipStream_t st, int mode, int nprod, int nst, const ChunkDev *chunks, const Work *work, int nwork,
                const void *A, const void *B, void *C, int jlo);

template <int MODE>
static void launch_mode(hipStream_t st, int prec, int nst, const ChunkDev *chunks, const Work *work, int nwork,
                        const void *A, const void *B, void *C, int jlo) {
    if (nwork <= 0) return;
    if (prec >= 2) {   // split BF16: three products, or four with lo * lo
        launch_bf3(st, MODE, prec == 3 ? 4 : 3, nst, chunks, work, nwork, A, B, C, jlo);
        return;
    }
    const int grid = 8 * ((nwork + 7) / 8);
    if (prec == 0)
        k_gemm<0, MODE><<<grid, 256, 0, st>>>(chunks, work, nwork, A, B, C, jlo);
    else
        k_gemm<1, MODE><<<grid, 256, 0, st>>>(chunks, work, nwork, A, B, C, jlo);
}

void launch_gemm_whiten_left(hipStream_t st, int prec, int nst, const ChunkDev *chunks, const Work *work,
                             int nwork, const void *li, const void *X, void *Y, int jlo) {
    launch_mode<MODE_LEFT>(st, prec, nst, chunks, work, nwork, li, X, Y, jlo);
}
void launch_gemm_whiten_right(hipStream_t st, int prec, int nst, const ChunkDev *chunks, const Work *work,
                              int nwork, const void *li, const void *Y, void *G, int jlo) {
    launch_mode<MODE_RIGHT>(st, prec, nst, chunks, work, nwork, Y, li, G, jlo);
}
void launch_gemm_gram(hipStream_t st, int prec, int nst, const ChunkDev *chunks, const Work *work, int nwork,
                      const void *G, double *gram) {
    launch_mode<MODE_GRAM>(st, prec, nst, chunks, work, nwork, G, G, gram, 0);
}

}  // namespace qml

namespace qml {

// ------------------------------------------------- split-BF16x3 GEMMs --
// Operands are stored pre-split: element (r, k) of an operand is hi[o] +
// hi[lo_off + o] (two bf16 planes), so the main loop moves bytes and issues
// MFMAs only.  The whitening products' operands (Li, the stack, P) are in
// the blocked layout of bf3_tiled, o = bf3_tiled(r, k, Np), so that a
// staged 16 x 32 block is one contiguous kilobyte; the Gram's (the G store,
// one row per member) are row-major, o = r ld + k.  Every product is A B^T with both operands K-major
// (X_j is symmetric; Li^T and the Gram are NT by construction).
//
// A workgroup computes a 128x128 tile with one wave per 64x64 quarter
// (4x4 blocks of 16x16, three v_mfma_f32_16x16x32_bf16 per block and K step:
// lo*hi, hi*lo, hi*hi; NPROD = 4 adds lo*lo first).  The lo*lo term is as
// large as what the two-plane split leaves out of each operand (both about
// 2^-16 of the product); adding it about halves the error of the products.  Waves whose quarter lies outside the matrix, or
// strictly below the diagonal for the upper-only products, skip their MFMAs
// (a separate copy of the loop) but still stage their share of the operands.
// NST LDS stages: the whitening products use one (168 VGPRs and 36 KB of
// LDS, three workgroups per CU; the next stage waits in registers), the Gram
// two (two workgroups per CU, a second register set two stages ahead).
//
// LDS: per stage four planes (A hi, A lo, B hi, B lo) of rows x 32 bf16
// (64 bytes, unpadded).  The four 16-byte chunks of row r sit at slot
// c ^ (((r >> 1) ^ ((r >> 3) & 1)) & 3): any 16 consecutive rows (from a multiple
// of 16) read at one chunk, and the staging writes (four rows' four chunks
// per 16 lanes), then fall on 16 distinct 16-byte slots of a 256-byte span
// of the 64 banks, and any aligned 8 of them on 8 distinct slots of 128
// bytes.
//
// The MFMA is issued as B-fragment x A-fragment, so that each lane ends with
// four consecutive columns of one output row (8-byte bf16 / 32-byte FP64
// stores) instead of four rows of one column.
constexpr int BK = 32;

// the swizzle repeats every 16 rows, so fragment block x sits at a fixed
// offset x * 16 * BK from block 0
__device__ __forceinline__ int bf3_slot(int r, int c) { return r * BK + ((c ^ bf3_swz(r & 15)) << 3); }

// 128x128 output tiles of four waves (2 x 2), each a 64x64 quarter
constexpr int BF3_BM = 128, BF3_BN = 128;
constexpr int BF3_PL = BF3_BM * BK;   // elements per LDS plane
constexpr int BF3_STG = 4 * BF3_PL;   // elements per stage: A hi, A lo, B hi, B lo
constexpr int BF3_EP = 72;            // row pitch of the staged epilogue block, in bf16

// Operands and K range of one work item (tile I, J of member or split j).
struct Bf3Item {
    const __bf16 *Ahi, *Bhi;   // hi planes; the lo planes follow at alo / blo elements
    long long lda, ldb, alo, blo;
    int arows, brows, kbeg, kend, Mv, Nv;
};

template <int MODE>
__device__ __forceinline__ Bf3Item bf3_item(const ChunkDev &c, const Work &wk, int jlo, const __bf16 *Abase,
                                            const __bf16 *Bbase) {
    Bf3Item t;
    const int Np = c.Np;
    const long long msz = (long long)Np * Np;
    const int I0 = wk.I * BF3_BM, J0 = wk.J * BF3_BN;
    t.kbeg = 0;
    if (MODE == MODE_LEFT) {
        // P = Li M, lower tiles (J <= I): Li[i,k] needs k <= i and the stored
        // T = M^T (upper triangle of X_j, diagonal halved) needs k >= n
        t.Ahi = Abase + c.lic_off;
        t.Bhi = Bbase + c.stk_off + 2 * (wk.j - jlo) * msz;
        t.lda = t.ldb = Np;
        t.alo = t.blo = msz;
        t.arows = t.brows = Np;
        t.kbeg = J0;
        t.kend = min(Np, I0 + BF3_BM);
        t.Mv = t.Nv = Np;
    } else if (MODE == MODE_RIGHT) {
        // G = P Li^T + Li P^T, upper tiles (I <= J): both terms need
        // k <= min(i, n) <= i, so k < I0 + BM; the second term swaps the
        // operand bases (same row offsets)
        t.Ahi = Abase + c.stk_off + 2 * (wk.j - jlo) * msz;
        t.Bhi = Bbase + c.lic_off;
        t.lda = t.ldb = Np;
        t.alo = t.blo = msz;
        t.arows = t.brows = Np;
        t.kend = min(Np, I0 + BF3_BM);
        t.Mv = t.Nv = Np;
    } else {
        const long long ktot = (long long)c.ntu * TILE * TILE;
        t.Ahi = t.Bhi = Abase + c.g_off;
        t.lda = t.ldb = 2 * ktot;
        t.alo = t.blo = ktot;
        t.arows = t.brows = c.m;
        t.kbeg = wk.j * TILE * TILE;
        t.kend = t.kbeg + TILE * TILE;
        t.Mv = t.Nv = c.m;
    }
    if (MODE != MODE_GRAM) {
        // past the chunk's N pixels every product is an exact zero (D = 0
        // there), so the sweep ends at N rounded up to a K step
        const int kn32 = (c.N + BK - 1) / BK * BK;
        t.kend = max(t.kbeg, min(t.kend, kn32));
    }
    return t;
}

// One BK step of a wave's 64x64 quarter out of LDS stage s0: fragment row
// offsets aoff / boff (rows 16x + lane%16 of the wave's slab, chunk lane/16).
template <int NPROD>
__device__ __forceinline__ void bf3_mfma(const __bf16 *s0, int aoff, int boff, f4 (&acc)[4][4]) {
    bf8 bh[4], bl[4];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        bh[y] = *(const bf8 *)(s0 + 2 * BF3_PL + boff + y * 16 * BK);
        bl[y] = *(const bf8 *)(s0 + 3 * BF3_PL + boff + y * 16 * BK);
    }
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        const bf8 ah = *(const bf8 *)(s0 + aoff + x * 16 * BK);
        const bf8 al = *(const bf8 *)(s0 + BF3_PL + aoff + x * 16 * BK);
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            if constexpr (NPROD == 4)
                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[y], al, acc[x][y], 0, 0, 0);
            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[y], al, acc[x][y], 0, 0, 0);
            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[y], ah, acc[x][y], 0, 0, 0);
            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[y], ah, acc[x][y], 0, 0, 0);
        }
    }
}

// Output of a wave's 64x64 quarter at (r0, c0); ep is the wave's own LDS
// region of 64 rows x BF3_EP (the stage buffers are free: the caller passed
// a barrier after the last step).
//   Gram: FP64 partials, x2 for off-diagonal G tiles, into the split's block.
//   LEFT / RIGHT: the FP32 block split into bf16 hi and lo planes, written
//   through LDS: the wave stages one plane (64 rows of 128 bytes at a row
//   pitch of 144 bytes, against bank conflicts) in its region, then each
//   store instruction writes eight full 128-byte rows, 16 bytes per lane.
//   One wave's LDS operations complete in order, so no barrier is needed.
template <int MODE>
__device__ __forceinline__ void bf3_epilogue(f4 (&acc)[4][4], const ChunkDev &c, const Work &wk, int jlo, int r0,
                                             int c0, void *Cbase, __bf16 *ep, int lane) {
    // lane holds C[16x + lane%16][16y + 4(lane/16) + u], u = 0..3
    const int orow = lane % 16, ocol = 4 * (lane / 16);
    if (MODE == MODE_GRAM) {
        const int ntile = c.mt * (c.mt + 1) / 2;
        const double wgt = packed_tile_is_diag(wk.j, c.nt) ? 1.0 : 2.0;
        double *C = (double *)Cbase + c.gram_off
                  + ((size_t)wk.j * ntile + upper_tile_index(r0 / TILE, c0 / TILE, c.mt)) * TILE * TILE;
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                d4 v;
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = wgt * (double)acc[x][y][u];
                double *o = C + (16 * x + orow) * TILE + 16 * y + ocol;
                *(double2 *)o = double2{v[0], v[1]};
                *(double2 *)(o + 2) = double2{v[2], v[3]};
            }
        return;
    }
    const int Np = c.Np;
    const long long msz = (long long)Np * Np;
    __bf16 *C = (__bf16 *)Cbase;
    long long base, ldc, lo;
    if (MODE == MODE_LEFT) {
        // P in the blocked layout: the quarter's eight 1 KB blocks
        base = c.stk_off + 2 * (wk.j - jlo) * msz;
        ldc = Np;
        lo = msz;
    } else {
        const long long ktot = (long long)c.ntu * TILE * TILE;
        base = c.g_off + 2 * wk.j * ktot + (long long)upper_tile_index(r0 / TILE, c0 / TILE, c.nt) * TILE * TILE;
        ldc = TILE;
        lo = ktot;
    }
    typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
    // hi plane first; acc then keeps the residual v - hi for the lo plane
#pragma unroll 1
    for (int plane = 0; plane < 2; ++plane) {
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                bf4 h;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    h[u] = (__bf16)acc[x][y][u];
                    acc[x][y][u] -= (float)h[u];
                }
                *(bf4 *)(ep + (16 * x + orow) * BF3_EP + 16 * y + ocol) = h;
            }
        __bf16 *Cp = C + base + plane * lo;
        typedef unsigned int u4 __attribute__((ext_vector_type(4)));
        if (MODE == MODE_LEFT) {
            // store s8 writes block (r0/16 + s8/2, c0/32 + s8%2) whole: lane i
            // fills slot i, row i/4 of the block, chunk (i%4) ^ bf3_swz(i/4)
            const int rr = lane / 4, ch = (lane % 4) ^ bf3_swz(rr);
#pragma unroll
            for (int s8 = 0; s8 < 8; ++s8) {
                const int rb = s8 / 2, kb = s8 % 2;
                __builtin_nontemporal_store(*(const u4 *)(ep + (16 * rb + rr) * BF3_EP + 32 * kb + 8 * ch),
                                            (u4 *)(Cp + bf3_tiled(r0 + 16 * rb, c0 + 32 * kb, Np) + 8 * lane));
            }
        } else {
#pragma unroll
            for (int s8 = 0; s8 < 8; ++s8) {
                const int r = 8 * s8 + lane / 8, ch = lane % 8;
                __builtin_nontemporal_store(*(const u4 *)(ep + r * BF3_EP + 8 * ch),
                                            (u4 *)(Cp + (long long)r * ldc + 8 * ch));
            }
        }
    }
}

template <int MODE, int NST, int NPROD>
__global__ __launch_bounds__(256, NST == 1 ? 3 : 2)
void k_gemm_bf3(const ChunkDev *cs, const Work *work, int nwork, const __bf16 *Abase, const __bf16 *Bbase,
                void *Cbase, int jlo) {
    constexpr int BM = BF3_BM, BN = BF3_BN;
    constexpr int RP = 256 / 4;                 // rows staged per pass
    constexpr int PA = BM / RP, PB = BN / RP;   // staging passes per plane
    constexpr int PLA = BF3_PL, PLB = BF3_PL;   // elements per LDS plane
    constexpr int STG = BF3_STG;                // elements per stage
    const int li = xcd_logical(nwork);
    if (li >= nwork) return;
    const Work wk = work[li];
    const ChunkDev c = cs[wk.c];
    const int I0 = wk.I * BM, J0 = wk.J * BN;
    const Bf3Item it = bf3_item<MODE>(c, wk, jlo, Abase, Bbase);
    const __bf16 *Ahi = it.Ahi, *Bhi = it.Bhi;
    const long long lda = it.lda, ldb = it.ldb, alo = it.alo, blo = it.blo;
    const int arows = it.arows, brows = it.brows, kbeg = it.kbeg, kend = it.kend, Mv = it.Mv, Nv = it.Nv;
    constexpr int SMEM = (NST * STG > 4 * 64 * BF3_EP) ? NST * STG : 4 * 64 * BF3_EP;
    __shared__ __attribute__((aligned(16))) __bf16 smem[SMEM];
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    const int wm = w / 2, wn = w % 2;
    const int r0 = I0 + 64 * wm, c0 = J0 + 64 * wn;
    bool active = (r0 < Mv) && (c0 < Nv);
    if (MODE != MODE_LEFT) active = active && (c0 >= r0);
    f4 acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};
    // Staging: thread t loads the 16-byte chunk t%4 of rows t/4 + RP*q, so
    // four consecutive lanes read one contiguous 64-byte row segment.  Rows
    // past the operand read row 0 instead: for the whitening products Np is
    // a multiple of 64, so such rows only feed inactive waves; for the Gram
    // they feed output rows >= m that are never read.  The loads are
    // therefore unconditional (no exec branches).
    static_assert(PA >= 1 && PA <= 2 && PB >= 1 && PB <= 2, "staging passes");
    const int lr = tid / 4, lc = tid % 4;
    const auto arow = [&](int q) { const int r = I0 + lr + RP * q; return r < arows ? r : 0; };
    const auto brow = [&](int q) { const int r = J0 + lr + RP * q; return r < brows ? r : 0; };
    // per-thread element offsets of the staged rows; the operand bases are
    // workgroup-uniform.  The whitening products' offsets stay below Np^2
    // (32 bits); the Gram's rows are whole members of 2 ntu TILE^2 elements,
    // so m ntu >= 2^18 (e.g. N ~ 2900 with 250 stored Q) needs 64 bits.
    // In the blocked layout chunk lc of row r at K step 0 is bf3_tiled(r,
    // 8 lc, Np), and K step k0 (a multiple of 32) adds 16 k0.
    using off_t = typename std::conditional<MODE == MODE_GRAM, long long, int>::type;
    const auto roff = [&](int r, long long ld) -> off_t {
        if constexpr (MODE == MODE_GRAM) return (off_t)r * (off_t)ld + 8 * lc;
        else return (off_t)bf3_tiled(r, 8 * lc, (int)ld);
    };
    const off_t oa0 = roff(arow(0), lda), oa1 = roff(arow(PA - 1), lda);
    const off_t ob0 = roff(brow(0), ldb), ob1 = roff(brow(PB - 1), ldb);
    constexpr int KSC = (MODE == MODE_GRAM) ? 1 : 16;   // element offset per unit of k
    const int wa0 = bf3_slot(lr, lc), wa1 = bf3_slot(lr + RP * (PA - 1), lc);
    const int wb0 = bf3_slot(lr, lc), wb1 = bf3_slot(lr + RP * (PB - 1), lc);
    // two staging register sets (P, Q): stage k+2 is in flight while k+1 waits
    // to be stored and k is multiplied
    uint4 Prah0, Pral0, Prah1, Pral1, Prbh0, Prbl0, Prbh1, Prbl1;
    [[maybe_unused]] uint4 Qrah0, Qral0, Qrah1, Qral1, Qrbh0, Qrbl0, Qrbh1, Qrbl1;
    // fragment reads: row 16x + lane%16 of the wave's 64-row slab, chunk lane/16
    const int rr = lane % 16, cq = lane / 16;
    const int aoff = bf3_slot(64 * wm + rr, cq), boff = bf3_slot(64 * wn + rr, cq);
#define QML_BF3_GLOADX(R, XA, XB, k0)                                            \
    R##rah0 = *(const uint4 *)((XA) + oa0 + KSC * (k0));                         \
    R##ral0 = *(const uint4 *)((XA) + alo + oa0 + KSC * (k0));                   \
    R##rbh0 = *(const uint4 *)((XB) + ob0 + KSC * (k0));                         \
    R##rbl0 = *(const uint4 *)((XB) + blo + ob0 + KSC * (k0));                   \
    if constexpr (PA > 1) {                                                      \
        R##rah1 = *(const uint4 *)((XA) + oa1 + KSC * (k0));                     \
        R##ral1 = *(const uint4 *)((XA) + alo + oa1 + KSC * (k0));               \
    }                                                                            \
    if constexpr (PB > 1) {                                                      \
        R##rbh1 = *(const uint4 *)((XB) + ob1 + KSC * (k0));                     \
        R##rbl1 = *(const uint4 *)((XB) + blo + ob1 + KSC * (k0));               \
    }
#define QML_BF3_GLOAD(R, k0) QML_BF3_GLOADX(R, Ahi, Bhi, k0)
#define QML_BF3_SSTORE(R, st)                                                    \
    {                                                                            \
        __bf16 *s0 = smem + (st) * STG;                                          \
        *(uint4 *)(s0 + wa0) = R##rah0;                                          \
        *(uint4 *)(s0 + PLA + wa0) = R##ral0;                                    \
        *(uint4 *)(s0 + 2 * PLA + wb0) = R##rbh0;                                \
        *(uint4 *)(s0 + 2 * PLA + PLB + wb0) = R##rbl0;                          \
        if constexpr (PA > 1) {                                                  \
            *(uint4 *)(s0 + wa1) = R##rah1;                                      \
            *(uint4 *)(s0 + PLA + wa1) = R##ral1;                                \
        }                                                                        \
        if constexpr (PB > 1) {                                                  \
            *(uint4 *)(s0 + 2 * PLA + wb1) = R##rbh1;                            \
            *(uint4 *)(s0 + 2 * PLA + PLB + wb1) = R##rbl1;                      \
        }                                                                        \
    }
#define QML_BF3_MFMA(st) bf3_mfma<NPROD>(smem + (st) * STG, aoff, boff, acc);
// One K sweep over n = (kend - kbeg)/BK stages, two steps per pass: the
// loads of stage s+2 are issued (into the register set just stored), stage
// s is multiplied out of LDS, and stage s+1 (loaded a whole step earlier)
// goes into the other LDS buffer.  Loads past the sweep repeat the last
// stage (unconditional, so the wait counts stay exact) and are not stored.
#define QML_BF3_SWEEP(DO_MFMA)                                                \
    {                                                                         \
        const int ns = (kend - kbeg) / BK;                                    \
        const int klast = kend - BK;                                          \
        QML_BF3_GLOAD(P, kbeg)                                                \
        QML_BF3_GLOAD(Q, min(kbeg + BK, klast))                               \
        QML_BF3_SSTORE(P, 0)                                                  \
        __syncthreads();                                                      \
        int s = 0;                                                            \
        for (; s + 1 < ns; s += 2) {                                          \
            const int k0 = kbeg + s * BK;                                     \
            QML_BF3_GLOAD(P, min(k0 + 2 * BK, klast))                         \
            if (DO_MFMA) QML_BF3_MFMA(0)                                      \
            QML_BF3_SSTORE(Q, 1)                                              \
            __syncthreads();                                                  \
            QML_BF3_GLOAD(Q, min(k0 + 3 * BK, klast))                         \
            if (DO_MFMA) QML_BF3_MFMA(1)                                      \
            if (s + 2 < ns) QML_BF3_SSTORE(P, 0)                              \
            __syncthreads();                                                  \
        }                                                                     \
        if (s < ns) {                                                         \
            if (DO_MFMA) QML_BF3_MFMA(0)                                      \
            __syncthreads();                                                  \
        }                                                                     \
    }
// Single LDS stage: stage s+1 is loaded into registers while stage s is
// multiplied, then stored once every wave has read stage s.  The last stage
// is multiplied outside the loop, so no load is left in flight into the
// epilogue.
#define QML_BF3_SWEEP1(DO_MFMA)                                               \
    {                                                                         \
        const int ns = (kend - kbeg) / BK;                                    \
        QML_BF3_GLOAD(P, kbeg)                                                \
        QML_BF3_SSTORE(P, 0)                                                  \
        __syncthreads();                                                      \
        for (int s = 0; s + 1 < ns; ++s) {                                    \
            QML_BF3_GLOAD(P, kbeg + (s + 1) * BK)                             \
            if (DO_MFMA) QML_BF3_MFMA(0)                                      \
            __syncthreads();                                                  \
            QML_BF3_SSTORE(P, 0)                                              \
            __syncthreads();                                                  \
        }                                                                     \
        if (DO_MFMA) QML_BF3_MFMA(0)                                          \
        __syncthreads();                                                      \
    }
#define QML_BF3_RUN                                                           \
    if constexpr (NST == 1) {                                                 \
        if (active) QML_BF3_SWEEP1(true)                                      \
        else QML_BF3_SWEEP1(false)                                            \
    } else {                                                                  \
        if (active) QML_BF3_SWEEP(true)                                       \
        else QML_BF3_SWEEP(false)                                             \
    }
// The right product's two terms as one sweep of 2n stages: P rows I
// against Li rows J, then Li rows I against P rows J (the same staged row
// offsets with the operand bases swapped), so the second term's first stage
// is in flight while the first term's last one is multiplied.
#define QML_BF3_SWEEP1_2T(DO_MFMA)                                            \
    {                                                                         \
        const int ns = (kend - kbeg) / BK;                                    \
        QML_BF3_GLOAD(P, kbeg)                                                \
        QML_BF3_SSTORE(P, 0)                                                  \
        __syncthreads();                                                      \
        for (int s = 0; s + 1 < 2 * ns; ++s) {                                \
            const int sn = s + 1;                                             \
            const bool second = sn >= ns;                                     \
            const __bf16 *xa = second ? Bhi : Ahi, *xb = second ? Ahi : Bhi;  \
            QML_BF3_GLOADX(P, xa, xb, kbeg + (second ? sn - ns : sn) * BK)    \
            if (DO_MFMA) QML_BF3_MFMA(0)                                      \
            __syncthreads();                                                  \
            QML_BF3_SSTORE(P, 0)                                              \
            __syncthreads();                                                  \
        }                                                                     \
        if (DO_MFMA) QML_BF3_MFMA(0)                                          \
        __syncthreads();                                                      \
    }
    if (kbeg < kend) {
        if constexpr (MODE == MODE_RIGHT && NST == 1) {
            if (active) QML_BF3_SWEEP1_2T(true)
            else QML_BF3_SWEEP1_2T(false)
        } else {
            QML_BF3_RUN
            static_assert(MODE != MODE_RIGHT || NST == 1, "the right product's two terms need one stage");
        }
    }
#undef QML_BF3_GLOAD
#undef QML_BF3_SSTORE
#undef QML_BF3_MFMA
#undef QML_BF3_SWEEP
#undef QML_BF3_SWEEP1
#undef QML_BF3_RUN
#undef QML_BF3_SWEEP1_2T
#undef QML_BF3_GLOADX
    if (!active) return;
    bf3_epilogue<MODE>(acc, c, wk, jlo, r0, c0, Cbase, smem + w * 64 * BF3_EP, lane);
}

// ------------------------------------------ split BF16, direct-to-LDS loads --
// The same products, tiles, LDS layout and MFMA order as k_gemm_bf3 (so the
// same results, bit for bit), with the operands staged by
// global_load_lds_dwordx4: each load moves 16 bytes per lane from global
// memory straight into LDS at (M0 base) + 16 x lane, with no register
// round trip, so NS stages can be in flight at once without registers held
// for them.  Wave w fills plane w of a stage (A hi, A lo, B hi, B lo), eight
// 16-row blocks of 1 KB.  For the whitening products a block is stored in
// slot order (bf3_tiled), so lane i loads the block's 16 bytes at 16 i; for
// the Gram lane i loads row i/4 of the block at the chunk that the swizzle
// places in slot i, (i%4) ^ bf3_swz(i/4).  Rows past the operand read row
// 0, as in k_gemm_bf3.
//
// Step s waits for its own stage (s_waitcnt vmcnt counting the stages issued
// after it), passes one barrier, issues stage s + NS - 1 into the buffer
// that step s - 1 read (every wave is past that step), then multiplies
// stage s.  The barrier is the bare s_barrier (no memory fence, which would
// wait for every load in flight), with compiler-level memory clobbers so
// that no LDS access moves across it.  NS = 2: two workgroups per CU (64 KB
// of LDS each); NS = 3: one (96 KB).
template <int NS>
__device__ __forceinline__ void bf3_wait_stage(bool last) {
    // s_waitcnt: vmcnt <= 8 (NS - 2) (the later stages' loads, eight per
    // wave each), or 0 on the last step; lgkmcnt 0; expcnt not waited on.
    // Encoding: vmcnt[3:0] and [15:14], expcnt[6:4], lgkmcnt[11:8].
    constexpr int V = 8 * (NS - 2);
    constexpr int ENC = (V & 15) | (7 << 4) | ((V >> 4) << 14);
    if (NS == 2 || last) __builtin_amdgcn_s_waitcnt(7 << 4);
    else __builtin_amdgcn_s_waitcnt(ENC);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <int MODE, int NS, int NPROD>
__global__ __launch_bounds__(256, NS == 2 ? 2 : 1)
void k_gemm_bf3d(const ChunkDev *cs, const Work *work, int nwork, const __bf16 *Abase, const __bf16 *Bbase,
                 void *Cbase, int jlo) {
    static_assert(NS == 2 || NS == 3, "two or three stages");
    const int li = xcd_logical(nwork);
    if (li >= nwork) return;
    const Work wk = work[li];
    const ChunkDev c = cs[wk.c];
    const int I0 = wk.I * BF3_BM, J0 = wk.J * BF3_BN;
    const Bf3Item it = bf3_item<MODE>(c, wk, jlo, Abase, Bbase);
    constexpr int SMEM = (NS * BF3_STG > 4 * 64 * BF3_EP) ? NS * BF3_STG : 4 * 64 * BF3_EP;
    __shared__ __attribute__((aligned(1024))) __bf16 smem[SMEM];
    const int tid = threadIdx.x, lane = tid % 64;
    const int w = __builtin_amdgcn_readfirstlane(tid / 64);
    const int wm = w / 2, wn = w % 2;
    const int r0 = I0 + 64 * wm, c0 = J0 + 64 * wn;
    bool active = (r0 < it.Mv) && (c0 < it.Nv);
    if (MODE != MODE_LEFT) active = active && (c0 >= r0);
    f4 acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};
    // this wave's plane: A (w < 2) or B rows, hi (w even) or lo
    const bool pa = w < 2;
    const int prow0 = pa ? I0 : J0, prows = pa ? it.arows : it.brows;
    const long long pld = pa ? it.lda : it.ldb, plo = (w & 1) ? (pa ? it.alo : it.blo) : 0;
    using off_t = typename std::conditional<MODE == MODE_GRAM, long long, int>::type;
    const int rl = lane / 4;
    const int ch = (lane % 4) ^ bf3_swz(rl);
    // blocked layout: the 16 rows of a block lie all inside the operand or
    // all past it (Np is a multiple of 64), and lane i reads slot i of the
    // block; K step k (a multiple of 32) adds 16 k
    off_t roff[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int r = prow0 + 16 * b + rl;
        if constexpr (MODE == MODE_GRAM) roff[b] = (off_t)(r < prows ? r : 0) * (off_t)pld + (off_t)plo + 8 * ch;
        else roff[b] = (off_t)bf3_tiled(r < prows ? prow0 + 16 * b : 0, 0, (int)pld) + (off_t)plo + 8 * lane;
    }
    constexpr int KSC = (MODE == MODE_GRAM) ? 1 : 16;
    const int kbeg = it.kbeg;
    const int ns = (it.kend - kbeg) / BK;
    const int total = (MODE == MODE_RIGHT) ? 2 * ns : ns;
    // stage s: operands (the right product's second term swaps them) and k
    const auto issue = [&](int s) {
        const bool second = MODE == MODE_RIGHT && s >= ns;
        const __bf16 *x = (pa != second) ? it.Ahi : it.Bhi;
        const int k = kbeg + (second ? s - ns : s) * BK;
        __bf16 *dst = smem + (s % NS) * BF3_STG + w * BF3_PL;
#pragma unroll
        for (int b = 0; b < 8; ++b)
            __builtin_amdgcn_global_load_lds((const void *)(x + roff[b] + KSC * k),
                                             (__attribute__((address_space(3))) void *)(dst + b * 16 * BK), 16, 0, 0);
    };
    const int rr = lane % 16, cq = lane / 16;
    const int aoff = bf3_slot(64 * wm + rr, cq), boff = bf3_slot(64 * wn + rr, cq);
    // K steps whose products are exact zeros for this wave are not
    // multiplied: Li[i][k] = 0 for k > i, the stored T[n][k] = 0 for k < n,
    // and P is lower triangular.  LEFT (rows i of Li, rows n of T) needs k in
    // [c0, r0 + 64); RIGHT (P or Li rows i in both terms) k < r0 + 64.  Past
    // the chunk's N pixels (D = 0) T, P and G vanish, so k < N and a wave
    // whose rows or columns all lie past N multiplies nothing.  The
    // accumulators start at +0 and no sum of exact zeros changes them, so the
    // results are those of the full sweep, bit for bit.
    const int kn32 = (c.N + BK - 1) / BK * BK;
    const bool pad = MODE != MODE_GRAM && (r0 >= c.N || c0 >= c.N);
    const int klo = (MODE == MODE_LEFT) ? c0 : INT_MIN;
    const int khi = (MODE == MODE_GRAM) ? INT_MAX : pad ? INT_MIN : min(r0 + 64, kn32);
    if (total > 0) {
#pragma unroll
        for (int s = 0; s < NS - 1; ++s)
            if (s < total) issue(s);
        for (int s = 0; s < total; ++s) {
            bf3_wait_stage<NS>(s + NS - 2 >= total);
            if (s + NS - 1 < total) issue(s + NS - 1);
            const int kk = kbeg + ((MODE == MODE_RIGHT && s >= ns) ? s - ns : s) * BK;
            if (active && kk + BK > klo && kk < khi) bf3_mfma<NPROD>(smem + (s % NS) * BF3_STG, aoff, boff, acc);
        }
        // every wave is done with the stages before the epilogue reuses LDS
        __builtin_amdgcn_s_waitcnt(7 << 4);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
    }
    if (!active) return;
    bf3_epilogue<MODE>(acc, c, wk, jlo, r0, c0, Cbase, smem + w * 64 * BF3_EP, lane);
}

template <int NPROD>
static void launch_bf3_np(hipStream_t st, int mode, int nst, const ChunkDev *chunks, const Work *work, int nwork,
                          const __bf16 *a, const __bf16 *b, void *C, int jlo) {
    const int grid = 8 * ((nwork + 7) / 8);
    // nst, the staging of the operands: 1 one LDS stage (three workgroups
    // per CU, the next stage in registers); 2 two LDS stages (two workgroups
    // per CU, a second register set two stages ahead; not for the right
    // product, whose two-term sweep uses 1 instead); 3 and 4 direct-to-LDS
    // loads with two (two workgroups per CU) and three (one) stages.  Each
    // output sums its K steps in the same order either way, so the results
    // are identical.
#define QML_BF3_LAUNCH(M, N) k_gemm_bf3<M, N, NPROD><<<grid, 256, 0, st>>>(chunks, work, nwork, a, b, C, jlo)
#define QML_BF3D_LAUNCH(M, N) k_gemm_bf3d<M, N, NPROD><<<grid, 256, 0, st>>>(chunks, work, nwork, a, b, C, jlo)
    if (mode == MODE_LEFT) {
        if (nst == 4) QML_BF3D_LAUNCH(MODE_LEFT, 3);
        else if (nst == 3) QML_BF3D_LAUNCH(MODE_LEFT, 2);
        else if (nst == 2) QML_BF3_LAUNCH(MODE_LEFT, 2);
        else QML_BF3_LAUNCH(MODE_LEFT, 1);
    } else if (mode == MODE_RIGHT) {
        if (nst == 4) QML_BF3D_LAUNCH(MODE_RIGHT, 3);
        else if (nst == 3) QML_BF3D_LAUNCH(MODE_RIGHT, 2);
        else QML_BF3_LAUNCH(MODE_RIGHT, 1);
    } else {
        if (nst == 4) QML_BF3D_LAUNCH(MODE_GRAM, 3);
        else if (nst == 3) QML_BF3D_LAUNCH(MODE_GRAM, 2);
        else if (nst == 1) QML_BF3_LAUNCH(MODE_GRAM, 1);
        else QML_BF3_LAUNCH(MODE_GRAM, 2);
    }
#undef QML_BF3_LAUNCH
#undef QML_BF3D_LAUNCH
}

void launch_bf3(hipStream_t st, int mode, int nprod, int nst, const ChunkDev *chunks, const Work *work, int nwork,
                const void *A, const void *B, void *C, int jlo) {
    if (nwork <= 0) return;
    const __bf16 *a = (const __bf16 *)A, *b = (const __bf16 *)B;
    if (nprod == 4)
        launch_bf3_np<4>(st, mode, nst, chunks, work, nwork, a, b, C, jlo);
    else
        launch_bf3_np<3>(st, mode, nst, chunks, work, nwork, a, b, C, jlo);
}

}  // namespace qml
