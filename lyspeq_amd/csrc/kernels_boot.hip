// Poisson bootstrap of the per-chunk results, FastBootstrap form
// (core/bootstrapper.hpp:183-221,282-375; core/one_qso_estimate.cpp:200-239,
// core/one_qso_estimate.hpp:55-60):
//
//   k_boot_collapse   theta_q = sum over the spectrum's chunks, in chunk order,
//                     of (d - b) - t at the chunk's fisher_index_start
//                     (OneQSOEstimate::collapseBootstrap)
//   k_boot_realize    R[b][i] = sum over spectra q, in queue order, of
//                     p[q][b] theta_q[i] (addBootPowerOnly: one dger per spectrum)
//   k_boot_mean       mean[i] = (sum over b of R[b][i]) / nboots, as sum * (1/n)
//   k_boot_cov        cov[i][j] = sum over b of (R[b][i] - mean[i])(R[b][j] -
//                     mean[j]), times 1/(nboots - 1) (dsyr per realisation)
//
// All FP64.  Every output is summed in the reference's order (chunks, spectra,
// realisations ascending), one thread per output element of a 64 x 64 tile,
// with the operands staged through LDS 16 terms at a time.  HBM traffic is
// small next to the chunk step (nboots x nspectra ints, nboots x T doubles).
#include "qml_common.hpp"
#include "qml_kernels.hpp"

namespace qml {

namespace {
constexpr int BT = 64;   // output tile edge
constexpr int KT = 16;   // terms staged per step
}

// One block per spectrum; its chunks are qchunk[qptr[q] .. qptr[q+1]).
__global__ __launch_bounds__(256) void k_boot_collapse(const ChunkOut *outs, const double *dbt, const int *qptr,
                                                       const int *qchunk, int T, double *theta) {
    const int q = blockIdx.x;
    const int c0 = qptr[q], c1 = qptr[q + 1];
    for (int i = threadIdx.x; i < T; i += blockDim.x) {
        double acc = 0.0;
        for (int k = c0; k < c1; ++k) {
            const ChunkOut o = outs[qchunk[k]];
            const int r = i - o.fis;
            if (r < 0 || r >= o.nq) continue;
            const double *v = dbt + o.dbt_off;
            acc += (v[r] - v[o.nq + r]) - v[2 * o.nq + r];
        }
        theta[(size_t)q * T + i] = acc;
    }
}

// R (nboots x T) = P^T Theta with P nqso x nboots (ints) and Theta nqso x T.
__global__ __launch_bounds__(256) void k_boot_realize(const int *P, const double *theta, int nqso, int nboots,
                                                      int T, double *R) {
    __shared__ double sp[KT][BT + 1], st[KT][BT + 1];
    const int b0 = blockIdx.y * BT, i0 = blockIdx.x * BT;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;   // 4 x 4 outputs per thread
    double acc[4][4] = {};
    for (int q0 = 0; q0 < nqso; q0 += KT) {
        for (int e = threadIdx.x; e < KT * BT; e += 256) {
            const int k = e / BT, c = e % BT;
            const int q = q0 + k;
            sp[k][c] = (q < nqso && b0 + c < nboots) ? (double)P[(size_t)q * nboots + b0 + c] : 0.0;
            st[k][c] = (q < nqso && i0 + c < T) ? theta[(size_t)q * T + i0 + c] : 0.0;
        }
        __syncthreads();
        const int kn = min(KT, nqso - q0);
        for (int k = 0; k < kn; ++k) {
            double p[4], t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                p[u] = sp[k][ty + 16 * u];
                t[u] = st[k][tx + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += p[u] * t[v];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int b = b0 + ty + 16 * u, i = i0 + tx + 16 * v;
            if (b < nboots && i < T) R[(size_t)b * T + i] = acc[u][v];
        }
}

__global__ __launch_bounds__(256) void k_boot_mean(const double *R, int nboots, int T, double *mean) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    double s = 0.0;
    for (int b = 0; b < nboots; ++b) s += R[(size_t)b * T + i];
    mean[i] = s * (1.0 / nboots);
}

// Upper 64 x 64 tiles of the covariance, mirrored into the lower triangle.
__global__ __launch_bounds__(256) void k_boot_cov(const double *R, const double *mean, int nboots, int T,
                                                  double *cov) {
    const int I = blockIdx.y, J = blockIdx.x;
    if (J < I) return;
    __shared__ double sa[KT][BT + 1], sb[KT][BT + 1];
    const int i0 = I * BT, j0 = J * BT;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    double acc[4][4] = {};
    for (int b0 = 0; b0 < nboots; b0 += KT) {
        for (int e = threadIdx.x; e < KT * BT; e += 256) {
            const int k = e / BT, c = e % BT;
            const int b = b0 + k;
            sa[k][c] = (b < nboots && i0 + c < T) ? R[(size_t)b * T + i0 + c] - mean[i0 + c] : 0.0;
            sb[k][c] = (b < nboots && j0 + c < T) ? R[(size_t)b * T + j0 + c] - mean[j0 + c] : 0.0;
        }
        __syncthreads();
        const int kn = min(KT, nboots - b0);
        for (int k = 0; k < kn; ++k) {
            double a[4], bb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = sa[k][ty + 16 * u];
                bb[u] = sb[k][tx + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * bb[v];
        }
        __syncthreads();
    }
    const double scale = 1.0 / (nboots - 1);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + ty + 16 * u, j = j0 + tx + 16 * v;
            if (i < T && j < T && j >= i) {
                const double c = acc[u][v] * scale;
                cov[(size_t)i * T + j] = c;
                cov[(size_t)j * T + i] = c;
            }
        }
}

void launch_bootstrap(hipStream_t st, const ChunkOut *outs, const double *dbt, const int *qptr, const int *qchunk,
                      int nqso, int T, const int *P, int nboots, double *theta, double *R, double *mean,
                      double *cov) {
    k_boot_collapse<<<nqso, 256, 0, st>>>(outs, dbt, qptr, qchunk, T, theta);
    const int tt = (T + BT - 1) / BT;
    k_boot_realize<<<dim3(tt, (nboots + BT - 1) / BT), 256, 0, st>>>(P, theta, nqso, nboots, T, R);
    k_boot_mean<<<(T + 255) / 256, 256, 0, st>>>(R, nboots, T, mean);
    k_boot_cov<<<dim3(tt, tt), 256, 0, st>>>(R, mean, nboots, T, cov);
}

}  // namespace qml
