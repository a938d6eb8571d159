// Shared host/device definitions of the QML engine.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace qml {

constexpr int TILE = 64;          // GEMM tile and padding granularity
constexpr int MAX_NU = 4;         // continuum templates handled on the device
constexpr int MAX_ZB = 64;        // redshift bins
constexpr int MAX_NK = 64;        // k bands
constexpr int VEC_SLOTS = 8 + 4 * MAX_NU;  // FP64 vectors of length Np per chunk

constexpr double SPEED_OF_LIGHT = 299792.458;
constexpr double LYA_REST = 1215.67;

// Per-chunk input vectors, each Np long, at ChunkDev::in_off.
enum InSlot { IN_WAVE = 0, IN_DELTA = 1, IN_IVAR = 2, IN_D = 3, IN_NPRIME = 4, IN_U0 = 5,
              IN_SLOTS = 5 + MAX_NU };
// Per-chunk scratch vectors, each Np long, at ChunkDev::vec_off: x, y, the
// W_t, then for the solves the right-hand sides R (D delta, D U_t), T = Li R,
// Z = Li^T T, the residual E, F = Li E and G = Li^T F.
enum VecSlot { V_X = 0, V_Y = 1, V_W0 = 2, V_R0 = 2 + MAX_NU, V_T0 = 3 + 2 * MAX_NU, V_Z0 = 4 + 3 * MAX_NU,
               V_E = 5 + 4 * MAX_NU, V_F = 6 + 4 * MAX_NU, V_G = 7 + 4 * MAX_NU };
static_assert(V_G < VEC_SLOTS, "vector slots");

// Split-BF16 operand layout of the whitening products (Li, the stack, P):
// each Np x Np plane is cut into 1 KB blocks of 16 rows x 32 columns, block
// (r/16, k/32) at element ((r/16) (Np/32) + k/32) 512, and inside a block
// the 16-byte chunk c (columns 8c..8c+7) of row r sits at slot
// 4 (r%16) + (c ^ bf3_swz(r%16)).  That is the order in which the products
// stage a block in LDS, so one contiguous 1 KB load fills one LDS block.
__device__ __forceinline__ int bf3_swz(int r) { return ((r >> 1) ^ ((r >> 3) & 1)) & 3; }
__device__ __forceinline__ int bf3_tiled(int r, int k, int Np) {
    const int rr = r & 15;
    return (((r >> 4) * (Np >> 5) + (k >> 5)) << 9) + ((4 * rr + (((k >> 3) & 3) ^ bf3_swz(rr))) << 3) + (k & 7);
}

struct BinsDev {
    double zc[MAX_ZB];
    double ke[MAX_NK + 1];           // k-band edges (resolution-matrix mode)
    double dz;
    int nz, nk, method;
    int turn_off_sfid;
    int growth;        // redshift growth of the derivatives on
    double growth_b;   // its exponent B
};

struct TableDev {
    const double *qt;  // nk x nv, v contiguous (per-band slices, and rows read band by band)
    const double *q2;  // nv x nk pairs (q[v][k], q[v+1][k]): both nodes of a gather in one load
    const double *s;   // nv x nz, z contiguous
    int nv, nz;
    double v1, dv, z1, dz;
};

// Per chunk, per batch.  Offsets are in elements of the arena they index.
struct ChunkDev {
    int N, Np, nt;       // pixels, padded size, Np / TILE
    int n, m;            // stored Q count, stack members (n + 2: H and T)
    int kbu, nzq, zm0;   // KBIN_UPP, z bins of the stored Q, first z bin
    int table, nu;       // table index, templates in use
    int gidx;            // chunk id
    int ntu;             // upper TILE x TILE tiles of Np x Np
    int mt;              // member tiles, ceil(m / TILE)
    long long in_off;    // into inputs (doubles)
    long long mat_off;   // into the FP64 matrix arena: A, L, Li, S, each Np*Np
    long long vec_off;   // into the FP64 vector arena
    long long stk_off;   // into the stack arenas (compute-type elements), Np*Np per member of a slab
    long long g_off;     // into the G store (compute-type elements): member j's packed upper tiles
    long long pj_off;    // into the aux arena (doubles): P_j, m x Np x (1+MAX_NU)
    long long at_off;    // into the aux arena: projected template weights, m x MAX_NU x Np
    long long th_off;    // into the aux arena: theta-contracted Q tables, nzq x nvu
    int nvu;             // table rows reachable by this chunk's velocity separations
    long long sj_off;    // into the s arena (doubles): m x MAX_NU^2, then d (n)
    long long gram_off;  // into the Gram partials (doubles)
    long long lic_off;   // into the compute-type copy of Li
    // resolution-matrix mode (rm = 1): Q_k = R Qhat_k R^T and S = R Shat R^T
    // are materialised at qb_off (kbu + 1 matrices of Np x Np in the FP64
    // matrix arena, S last).  R is stored by rows at rin_off: N x rnb values,
    // row i covering fine columns rcol0 + i*ros .. + rnb - 1, then the rM
    // fine-grid wavelengths (units of lambda_Lya).
    int rm, rnb, ros, rcol0, rM;
    long long rin_off;   // into the resolution-matrix input arena (doubles)
    long long qb_off;    // into the FP64 matrix arena
};
// A double in the global address space: loads through it are global_load,
// counted by vmcnt only (a flat load would also hold up the LDS waits).
typedef const double __attribute__((address_space(1))) GlobalD;
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef const dvec2 __attribute__((address_space(1))) GlobalD2;

// Fine rows a 16-row tile of R touches: 15 ros + rnb <= RM_ROWS.
constexpr int RM_ROWS = 96;

struct ChunkOut {        // per chunk id, persistent
    long long fc_off;    // n_q x n_q
    long long dbt_off;   // 3 x n_q
    int nq, fis;
};

// GEMM work item: (local chunk, member or split, tile row, tile col)
struct Work {
    int c, j, I, J;
};

// ---------------------------------------------------------------- device ---

// DiscreteInterpolation1D/2D index rule (mathtools/discrete_interpolation.cpp
// :61-72,221-242): truncation toward zero, clamped to [0, n-2].
__device__ __forceinline__ int interp_index(double xx, int n) {
    int i = (int)xx;
    if (i < 0) i = 0;
    else if (i >= n - 1) i = n - 2;
    return i;
}

// Redshift bin weight (core/global_numbers.cpp:393-471), elementwise.
__device__ __forceinline__ double zbin_weight(const BinsDev &b, double z, int zm) {
    const double zc = b.zc[zm], dz = b.dz;
    if (b.method == 0 || b.nz == 1) {
        return (z >= zc - dz / 2 && z <= zc + dz / 2) ? 1.0 : 0.0;
    }
    if (zm == 0) {
        if (z < zc) return 1.0;
        if (z <= zc + dz) return 1. - (z - zc) / dz;
        return 0.0;
    }
    if (zm == b.nz - 1) {
        if (z < zc - dz) return 0.0;
        if (z <= zc) return 1. - (zc - z) / dz;
        return 1.0;
    }
    if (z >= zc - dz && z <= zc + dz) return 1. - fabs(z - zc) / dz;
    return 0.0;
}

// ((1 + z) / (1 + z_m))^B, out of line: inlined, pow's registers would count
// against every kernel that weights a stored Q, growth on or off.
static __device__ __attribute__((noinline)) double growth_factor(double z, double zc, double b) {
    return pow((z + 1.0) / (zc + 1.0), b);
}

// The factor every stored Q of z bin zm carries at pixel pair redshift z:
// the bin weight, times ((1 + z) / (1 + z_m))^B with redshift growth on
// (core/chunk_estimate.cpp:539-573, applied after the weights of :483-536).
__device__ __forceinline__ double zweight(const BinsDev &b, double z, int zm) {
    const double w = zbin_weight(b, z, zm);
    if (!b.growth || w == 0.0) return w;
    return w * growth_factor(z, b.zc[zm], b.growth_b);
}

// S_fid(z, v) by DiscreteInterpolation2D::evaluate with x = z (fast axis).
__device__ __forceinline__ double signal_interp(const TableDev &t, double z, double v) {
    const double xx = (z - t.z1) / t.dz, yy = (v - t.v1) / t.dv;
    const int nx = interp_index(xx, t.nz), ny = interp_index(yy, t.nv);
    const size_t ind = (size_t)nx + (size_t)t.nz * ny;
    const double dnx = xx - nx, dny = yy - ny;
    return t.s[ind] * (1 - dnx) * (1 - dny) + t.s[ind + 1] * (dnx) * (1 - dny)
         + t.s[ind + t.nz] * (1 - dnx) * (dny) + t.s[ind + t.nz + 1] * (dnx) * (dny);
}

// v_ij and z_ij of Chunk::_setVZMatrices (core/chunk_estimate.cpp:384-398),
// for any i, j (the reference fills j >= i and mirrors).
__device__ __forceinline__ void vz_pair(double li, double lj, int i, int j, double &v, double &z) {
    if (j >= i) {
        v = SPEED_OF_LIGHT * log(lj / li);
        z = sqrt(li * lj) - 1.;
    } else {
        v = SPEED_OF_LIGHT * log(li / lj);
        z = sqrt(lj * li) - 1.;
    }
}

}  // namespace qml
