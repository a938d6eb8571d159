// Batched blocked Cholesky factorisation A = L L^T and triangular inverse
// Li = L^-1 of every chunk of a batch (gfx950, FP64).
//
// Right-looking with 64x64 blocks (Np is a multiple of 64).  The L slot of
// the matrix arena starts as a copy of A and is factorised in place.  Block
// column p takes three launches over the whole batch:
//   diag   : one workgroup per chunk factors the updated block (p,p) in LDS,
//            inverts it (Li[p,p] = L[p,p]^-1) and zeroes block row p right of
//            the diagonal in L and Li;
//   panel  : TRSM tiles L[i,p] = A[i,p] Li[p,p]^T for i > p, and the tiles of
//            block row p of the inverse, Li[p,j] = -Li[p,p] sum_{k=j}^{p-1}
//            L[p,k] Li[k,j] for j < p (rows k < p of Li are final by then);
//   update : trailing tiles A[i,j] -= L[i,p] L[j,p]^T for p < j <= i.
// The 64x64 tile products run on v_mfma_f64_16x16x4f64, four waves per
// workgroup, each owning a 32x32 quarter.  The panel and update kernels
// stage their operands in K slabs of 16 (17 KB of LDS, at most 128 VGPRs),
// so that they fit on a CU beside the whitening products of the other
// stream (two workgroups of 64 KB) instead of waiting for one to drain; the
// MFMAs still run over k in the same order, so the results are the same.  A non-positive pivot marks the
// chunk failed (status) and is replaced by 1, as in the reference's LAPACK
// error path the chunk's results are then discarded.
#include "qml_common.hpp"
#include "qml_kernels.hpp"

namespace qml {

typedef double fd4 __attribute__((ext_vector_type(4)));
constexpr int FB = 64, FLD = FB + 1;   // block size, LDS row pitch (doubles)

// Element (r, c) of the 64x64 tile held in acc by this lane: for u = 0..3,
// r = 32 wm + 16 x + 4 u + lane/16, c = 32 wn + 16 y + lane%16.
template <class F>
__device__ __forceinline__ void tile_each(const fd4 (&acc)[2][2], F f) {
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int wm = w / 2, wn = w % 2;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                f(wm * 32 + x * 16 + 4 * u + lane / 16, wn * 32 + y * 16 + lane % 16, acc[x][y][u]);
}

__device__ __forceinline__ void tile_zero(fd4 (&acc)[2][2]) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = fd4{0.0, 0.0, 0.0, 0.0};
}

// A 64x64 block in registers: thread t holds the column pairs 2 (e % 32)
// of rows e / 32, e = t + 256 q (16-byte global loads through a
// global-space pointer, so that they count in vmcnt only).
struct BlockRegs {
    dvec2 v[FB * FB / 512];
};
__device__ __forceinline__ void fetch_block(BlockRegs &R, const double *M, int Np, int r0, int c0) {
#pragma unroll
    for (int q = 0; q < FB * FB / 512; ++q) {
        const int e = threadIdx.x + 256 * q, r = e / 32, c2 = 2 * (e % 32);
        R.v[q] = *(GlobalD2 *)(M + (size_t)(r0 + r) * Np + c0 + c2);
    }
}
// s[r * FLD + c] = block(r, c), or s[c * FLD + r] transposed
__device__ __forceinline__ void put_block(double *s, const BlockRegs &R, bool transpose) {
#pragma unroll
    for (int q = 0; q < FB * FB / 512; ++q) {
        const int e = threadIdx.x + 256 * q, r = e / 32, c2 = 2 * (e % 32);
        if (transpose) {
            s[c2 * FLD + r] = R.v[q].x;
            s[(c2 + 1) * FLD + r] = R.v[q].y;
        } else {
            s[r * FLD + c2] = R.v[q].x;
            s[r * FLD + c2 + 1] = R.v[q].y;
        }
    }
}
// s[r * FLD + c] = M[(r0 + r) * Np + c0 + c] (or transposed) for a 64x64 block
__device__ __forceinline__ void load_block(double *s, const double *M, int Np, int r0, int c0, bool transpose) {
    BlockRegs R;
    fetch_block(R, M, Np, r0, c0);
    put_block(s, R, transpose);
}

// K slabs of 16: S[r * KLD + k] for r in [0, 64), k in [0, 16)
constexpr int KS = 16, KLD = KS + 1;

// acc += S_a S_b^T over one slab (k ascending, four MFMA steps)
__device__ __forceinline__ void tile_nt16(const double *sa, const double *sb, fd4 (&acc)[2][2]) {
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int wm = w / 2, wn = w % 2;
#pragma unroll
    for (int k = 0; k < KS; k += 4) {
        double a[2], b[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            a[q] = sa[(wm * 32 + q * 16 + lane % 16) * KLD + k + lane / 16];
            b[q] = sb[(wn * 32 + q * 16 + lane % 16) * KLD + k + lane / 16];
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
}

// Slab s of a 64x64 block in registers: 64 rows x 16 columns (row slab,
// M[r0 + r][c0 + 16 s + k]) or 16 rows x 64 columns (column slab for a
// transposed operand, M[r0 + 16 s + k][c0 + c]); two 16-byte loads a thread.
struct SlabRegs {
    dvec2 v[2];
};
__device__ __forceinline__ void fetch_slab(SlabRegs &R, const double *M, int Np, int r0, int c0, int s, bool transpose) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int e = threadIdx.x + 256 * q;
        if (transpose) {
            const int k = e / 32, c2 = 2 * (e % 32);
            R.v[q] = *(GlobalD2 *)(M + (size_t)(r0 + KS * s + k) * Np + c0 + c2);
        } else {
            const int r = e / 8, k2 = 2 * (e % 8);
            R.v[q] = *(GlobalD2 *)(M + (size_t)(r0 + r) * Np + c0 + KS * s + k2);
        }
    }
}
__device__ __forceinline__ void put_slab(double *S, const SlabRegs &R, bool transpose) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int e = threadIdx.x + 256 * q;
        if (transpose) {
            const int k = e / 32, c2 = 2 * (e % 32);
            S[c2 * KLD + k] = R.v[q].x;
            S[(c2 + 1) * KLD + k] = R.v[q].y;
        } else {
            const int r = e / 8, k2 = 2 * (e % 8);
            S[r * KLD + k2] = R.v[q].x;
            S[r * KLD + k2 + 1] = R.v[q].y;
        }
    }
}

// acc += A B^T over the four slabs of 64x64 operand blocks A (rows ra0,
// columns ca0 of Ma) and B (rows rb0, columns cb0 of Mb, or the transpose
// of the block at (rb0, cb0) when tb); the next slab is loaded into
// registers while one is multiplied.
__device__ __forceinline__ void block_nt(double *sa, double *sb, const double *Ma, int ra0, int ca0, const double *Mb,
                                         int rb0, int cb0, bool tb, int Np, fd4 (&acc)[2][2]) {
    SlabRegs ra, rb;
    fetch_slab(ra, Ma, Np, ra0, ca0, 0, false);
    fetch_slab(rb, Mb, Np, rb0, cb0, 0, tb);
#pragma unroll 1
    for (int s = 0; s < FB / KS; ++s) {
        put_slab(sa, ra, false);
        put_slab(sb, rb, tb);
        __syncthreads();
        if (s + 1 < FB / KS) {
            fetch_slab(ra, Ma, Np, ra0, ca0, s + 1, false);
            fetch_slab(rb, Mb, Np, rb0, cb0, s + 1, tb);
        }
        tile_nt16(sa, sb, acc);
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_chol_init(const ChunkDev *cs, double *mat) {
    const ChunkDev c = cs[blockIdx.y];
    const size_t n = (size_t)c.Np * c.Np;
    const double *A = mat + c.mat_off;
    double *L = mat + c.mat_off + n;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) L[e] = A[e];
}

__global__ __launch_bounds__(256) void k_chol_diag(const ChunkDev *cs, double *mat, int *status, int p) {
    const ChunkDev c = cs[blockIdx.x];
    const int Np = c.Np, tid = threadIdx.x, p0 = p * FB;
    if (p0 >= Np) return;
    double *L = mat + c.mat_off + (size_t)Np * Np;
    double *Li = mat + c.mat_off + 2 * (size_t)Np * Np;
    __shared__ double a[FB * FLD];
    __shared__ double v[FB * FLD];
    __shared__ int fail;
    if (tid == 0) fail = 0;
    load_block(a, L, Np, p0, p0, false);
    for (int e = tid; e < FB * FB; e += 256) v[(e / FB) * FLD + e % FB] = (e / FB == e % FB) ? 1.0 : 0.0;
    __syncthreads();
    // unblocked right-looking factorisation of the 64x64 block; alongside it,
    // V = L^-1 by forward elimination of the identity: at step k row k of V
    // (complete by then) is divided by L_kk, and every row i > k takes off
    // L_ik times it
    for (int k = 0; k < FB; ++k) {
        if (tid == 0) {
            const double d = a[k * FLD + k];
            if (!(d > 0.0)) fail = 1;
            a[k * FLD + k] = (d > 0.0) ? sqrt(d) : 1.0;
        }
        __syncthreads();
        const double dkk = a[k * FLD + k];
        if (tid > k && tid < FB) a[tid * FLD + k] /= dkk;
        if (tid >= FB && tid - FB <= k) v[k * FLD + tid - FB] /= dkk;
        __syncthreads();
        for (int e = tid; e < FB * FB; e += 256) {
            const int i = e / FB, j = e % FB;
            if (i > k) {
                if (j > k && j <= i) a[i * FLD + j] -= a[i * FLD + k] * a[j * FLD + k];
                else if (j <= k) v[i * FLD + j] -= a[i * FLD + k] * v[k * FLD + j];
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < FB * FB; e += 256) {
        const int i = e / FB, j = e % FB;
        const size_t o = (size_t)(p0 + i) * Np + p0 + j;
        L[o] = (j <= i) ? a[i * FLD + j] : 0.0;
        Li[o] = v[i * FLD + j];
    }
    // block row p right of the diagonal is zero in L and Li
    const int wr = Np - p0 - FB;
    for (size_t e = tid; e < (size_t)FB * wr; e += 256) {
        const size_t o = (size_t)(p0 + e / wr) * Np + p0 + FB + e % wr;
        L[o] = 0.0;
        Li[o] = 0.0;
    }
    if (tid == 0) {
        if (p == 0) status[c.gidx] = fail;
        else if (fail) status[c.gidx] = 1;
    }
}

__global__ __launch_bounds__(256, 4) void k_chol_panel(const ChunkDev *cs, double *mat, int p, int ntmax) {
    const ChunkDev c = cs[blockIdx.y];
    const int Np = c.Np, nt = Np / FB, p0 = p * FB, bx = blockIdx.x;
    if (p >= nt) return;
    double *L = mat + c.mat_off + (size_t)Np * Np;
    double *Li = mat + c.mat_off + 2 * (size_t)Np * Np;
    __shared__ double sa[FB * KLD];
    __shared__ double sb[FB * KLD];
    fd4 acc[2][2];
    tile_zero(acc);
    if (bx < ntmax) {
        // TRSM tile: L[i,p] = A[i,p] Li[p,p]^T
        const int i = p + 1 + bx;
        if (i >= nt) return;
        block_nt(sa, sb, L, i * FB, p0, Li, p0, p0, false, Np, acc);
        tile_each(acc, [&](int r, int q, double x) { L[(size_t)(i * FB + r) * Np + p0 + q] = x; });
        return;
    }
    // block row p of the inverse: T = sum_{k=j}^{p-1} L[p,k] Li[k,j]
    const int j = bx - ntmax;
    if (j >= p) return;
#pragma unroll 1
    for (int k = j; k < p; ++k) block_nt(sa, sb, L, p0, k * FB, Li, k * FB, j * FB, true, Np, acc);
    // Li[p,j] = -Li[p,p] T: T^T as the B operand, slab s holding rows
    // [16 s, 16 s + 16) of T, which the waves of row half s / 2 hold in
    // acc[s % 2] (rows 32 wm + 16 x + 4 u + lane / 16)
    fd4 t[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) t[x][y] = acc[x][y];
    tile_zero(acc);
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64, wm = w / 2, wn = w % 2;
    SlabRegs ra;
    fetch_slab(ra, Li, Np, p0, p0, 0, false);
#pragma unroll 1
    for (int s = 0; s < FB / KS; ++s) {
        put_slab(sa, ra, false);
        if (wm == s / 2) {
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    sb[(wn * 32 + y * 16 + lane % 16) * KLD + 4 * u + lane / 16] = t[s % 2][y][u];
        }
        __syncthreads();
        if (s + 1 < FB / KS) fetch_slab(ra, Li, Np, p0, p0, s + 1, false);
        tile_nt16(sa, sb, acc);
        __syncthreads();
    }
    tile_each(acc, [&](int r, int q, double x) { Li[(size_t)(p0 + r) * Np + j * FB + q] = -x; });
}

__global__ __launch_bounds__(256, 4) void k_chol_update(const ChunkDev *cs, double *mat, int p) {
    const ChunkDev c = cs[blockIdx.y];
    const int Np = c.Np, nt = Np / FB, p0 = p * FB;
    const int m = nt - p - 1;                 // trailing block rows of this chunk
    int t = blockIdx.x;
    if (m <= 0 || t >= m * (m + 1) / 2) return;
    // t -> (i, j), p < j <= i < nt, row-major over the lower triangle
    int i = 0;
    while (t > i) { t -= i + 1; ++i; }
    const int bi = p + 1 + i, bj = p + 1 + t;
    double *L = mat + c.mat_off + (size_t)Np * Np;
    __shared__ double sa[FB * KLD];
    __shared__ double sb[FB * KLD];
    fd4 acc[2][2];
    tile_zero(acc);
    block_nt(sa, sb, L, bi * FB, p0, L, bj * FB, p0, false, Np, acc);
    tile_each(acc, [&](int r, int q, double x) { L[(size_t)(bi * FB + r) * Np + bj * FB + q] -= x; });
}

void launch_cholesky(const LaunchCtx &L, int *status) {
    const int ntmax = L.np_max / FB;
    k_chol_init<<<dim3(64, L.nchunks), 256, 0, L.stream>>>(L.chunks, L.mat);
    for (int p = 0; p < ntmax; ++p) {
        k_chol_diag<<<L.nchunks, 256, 0, L.stream>>>(L.chunks, L.mat, status, p);
        const int gx = ntmax + p;   // TRSM tiles (< ntmax) and inverse tiles j < p
        k_chol_panel<<<dim3(gx, L.nchunks), 256, 0, L.stream>>>(L.chunks, L.mat, p, ntmax);
        const int m = ntmax - p - 1;
        if (m > 0) k_chol_update<<<dim3(m * (m + 1) / 2, L.nchunks), 256, 0, L.stream>>>(L.chunks, L.mat, p);
    }
}

// The inverse is formed inside the blocked factorisation.
void launch_trinv(const LaunchCtx &) {}

}  // namespace qml
