// Resolution-matrix mode of the QML engine (gfx950, FP64).
//
// With a resolution matrix R (N x M, banded: row i has rnb values at
// matrix-grid columns rcol0 + i*ros ...) the reference builds, per chunk,
//   Q_k = R Qhat_k R^T,  Qhat_k(v) = (dk/pi) cos(k_c v) sin(dk v/2)/(dk v/2)
//                                    (dk/pi at v = 0),
//   S   = R Shat R^T,    Shat = the 2D signal table at (z, v),
// with v, z taken on the matrix's grid (core/chunk_estimate.cpp:384-480,
// mathtools/matrix_helper.cpp:877-902,1075-1084), then applies the z-bin
// weights on the pixel grid.  Here every Q_k (k < KBIN_UPP) and S is
// materialised once per iteration in the chunk's FP64 matrix arena (qb_off),
// in two launches per matrix over the whole batch:
//   k_rm_left : T = R Qhat for a 16-row x 64-column tile; the Qhat rows the
//               tile's R rows touch (15 ros + rnb of them) are evaluated into
//               LDS, then each thread forms four T entries;
//   k_rm_right: Q[i][j] = sum_t T[i][rcol0 + j ros + t] R[j][t] for j >= i,
//               mirrored, as Resolution::sandwich fills the upper triangle.
// The assembly, d-pass and stack kernels then read Q_k from these matrices
// instead of interpolating the derivative table.
#include "qml_common.hpp"
#include "qml_kernels.hpp"

namespace qml {

namespace {
constexpr double RM_PI = 3.14159265358979323846;

// Qhat_kn (kn >= 0) or Shat (kn < 0) between matrix-grid points a and b
__device__ __forceinline__ double rm_fine_value(const TableDev &t, const BinsDev &bins, int kn, double la,
                                                double lb, int a, int b) {
    double v, z;
    vz_pair(la, lb, a, b, v, z);
    if (kn < 0) return bins.turn_off_sfid ? 0.0 : signal_interp(t, z, v);
    const double kc = bins.kc[kn], dk = bins.dk[kn];
    if (v == 0) return dk / RM_PI;
    const double x = dk * v / 2;
    return (dk / RM_PI) * cos(kc * v) * (sin(x) / x);
}
}  // namespace

__global__ __launch_bounds__(256) void k_rm_left(const ChunkDev *cs, const double *rin, const TableDev *tabs,
                                                 BinsDev bins, int kn, double *rt) {
    const ChunkDev c = cs[blockIdx.z];
    if (!c.rm || kn >= c.kbu) return;
    const int N = c.N, M = c.rM, nb = c.rnb, os = c.ros, c0 = c.rcol0;
    const int i0 = blockIdx.y * 16, f0 = blockIdx.x * 64;
    if (i0 >= N || f0 >= M) return;
    const double *R = rin + c.rin_off;           // N x nb
    const double *grid = R + (size_t)N * nb;     // M wavelengths
    const TableDev t = tabs[c.table];
    const int iend = min(N, i0 + 16);
    const int amin = max(0, c0 + i0 * os), amax = min(M, c0 + (iend - 1) * os + nb);
    const int ra = amax - amin;                  // <= RM_ROWS (checked at registration)
    __shared__ double Qs[RM_ROWS][65];
    const int tid = threadIdx.x;
    for (int e = tid; e < ra * 64; e += 256) {
        const int r = e / 64, cc = e % 64, a = amin + r, b = f0 + cc;
        Qs[r][cc] = (b < M) ? rm_fine_value(t, bins, kn, grid[a], grid[b], a, b) : 0.0;
    }
    __syncthreads();
    const int r = tid / 16, cb = (tid % 16) * 4, i = i0 + r;
    if (i >= N) return;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int tt = 0; tt < nb; ++tt) {
        const int a = c0 + i * os + tt;
        if (a < amin || a >= amax) continue;     // outside the grid (diagonal storage edges)
        const double w = R[(size_t)i * nb + tt];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] += w * Qs[a - amin][cb + u];
    }
    double *T = rt + c.rt_off;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (f0 + cb + u < M) T[(size_t)i * M + f0 + cb + u] = acc[u];
}

__global__ __launch_bounds__(256) void k_rm_right(const ChunkDev *cs, const double *rin, const double *rt, int kn,
                                                  double *mat) {
    const ChunkDev c = cs[blockIdx.z];
    if (!c.rm || kn >= c.kbu) return;
    const int Np = c.Np, N = c.N;
    const int I0 = blockIdx.y * 16, J0 = blockIdx.x * 16;
    if (I0 >= Np || J0 >= Np || J0 + 15 < I0) return;
    const int i = I0 + threadIdx.x / 16, j = J0 + threadIdx.x % 16;
    if (j < i) return;
    double v = 0.0;
    if (i < N && j < N) {
        const double *R = rin + c.rin_off;
        const double *T = rt + c.rt_off;
        for (int tt = 0; tt < c.rnb; ++tt) {
            const int a = c.rcol0 + j * c.ros + tt;
            if (a < 0 || a >= c.rM) continue;
            v += T[(size_t)i * c.rM + a] * R[(size_t)j * c.rnb + tt];
        }
    }
    double *Q = mat + c.qb_off + (size_t)(kn < 0 ? c.kbu : kn) * Np * Np;
    Q[(size_t)i * Np + j] = v;
    Q[(size_t)j * Np + i] = v;
}

void launch_rmat(const LaunchCtx &L) {
    const dim3 gl((L.rm_M + 63) / 64, L.np_max / 16, L.nchunks), gr(L.np_max / 16, L.np_max / 16, L.nchunks);
    for (int kn = -1; kn < L.rm_kbu; ++kn) {
        k_rm_left<<<gl, 256, 0, L.stream>>>(L.chunks, L.rin, L.tables, L.bins, kn, L.rt);
        k_rm_right<<<gr, 256, 0, L.stream>>>(L.chunks, L.rin, L.rt, kn, L.mat);
    }
}

}  // namespace qml
