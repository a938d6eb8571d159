// Poisson bootstrap of the per-chunk results.  FastBootstrap form
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
// The complete form (k_boot_collapse_fisher, k_boot_build, the chunk
// engine's batched Cholesky, k_boot_solve) follows below.
//
// All FP64.  Every output is summed in the reference's order (chunks, spectra,
// realisations ascending), one thread per output element of a 64 x 64 tile,
// with the operands staged through LDS 16 terms at a time.  HBM traffic is
// small next to the chunk step (nboots x nspectra ints, nboots x T doubles).
#include "qml_common.hpp"
#include "qml_kernels.hpp"

#include <mutex>

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

// ------------------------------------------------ complete bootstrap --
// The complete form (PoissonBootstrapper::_oneBoot, core/bootstrapper.hpp:
// 223-262; OneQSOEstimate::addBoot, core/one_qso_estimate.hpp:43-53): per
// realisation b, F_b = sum over spectra of p_qb F_q (upper) and x_b the
// solution of F_b x = sum p_qb theta_q (LAPACKE_safeSolveCho,
// mathtools/matrix_helper.cpp:385-396: empty diagonals set to 1, their x
// zeroed).

// Each spectrum's collapsed Fisher matrix (OneQSOEstimate::collapseBootstrap,
// core/one_qso_estimate.cpp:224-234): the upper triangle of ndim x ndim at
// fq_off[q], summed over its chunks in chunk order.  One block per spectrum.
__global__ __launch_bounds__(256) void k_boot_collapse_fisher(const ChunkOut *outs, const double *fc,
                                                              const int *qptr, const int *qchunk,
                                                              const int *qstart, const int *qdim,
                                                              const long long *fq_off, double *fq) {
    const int q = blockIdx.x;
    const int c0 = qptr[q], c1 = qptr[q + 1], is = qstart[q], nd = qdim[q];
    double *F = fq + fq_off[q];
    for (int e = threadIdx.x; e < nd * nd; e += blockDim.x) {
        const int i = e / nd, j = e - i * nd;
        double acc = 0.0;
        if (j >= i) {
            for (int k = c0; k < c1; ++k) {
                const ChunkOut o = outs[qchunk[k]];
                const int a = i + is - o.fis, b = j + is - o.fis;
                if (a < 0 || b >= o.nq) continue;   // a <= b
                acc += fc[o.fc_off + (size_t)a * o.nq + b];
            }
        }
        F[e] = acc;
    }
}

// F_b (full symmetric, Np x Np with identity padding, empty diagonals set to
// 1) at mat + b * 3 Np^2 and the right-hand side at rhs + b * Np, for the
// realisations b0 .. b0 + nb - 1 of the batch.  Grid (Np / 64, Np / 64, nb),
// upper tiles only.  Element (i, j), j >= i, sums over the spectra covering
// band i (band_q[band_ptr[i] ..], in queue order) that also cover j.
__global__ __launch_bounds__(256) void k_boot_build(const int *W, int nqso, int b0, int T, int Np,
                                                    const int *band_ptr, const int *band_q, const int *qstart,
                                                    const int *qdim, const long long *fq_off, const double *fq,
                                                    const double *theta, double *mat, double *rhs) {
    const int I = blockIdx.y, J = blockIdx.x, b = blockIdx.z;
    if (J < I) return;
    const int *w = W + (size_t)(b0 + b) * nqso;
    double *A = mat + (size_t)b * 3 * Np * Np;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int i = I * 64 + e / 64, j = J * 64 + e % 64;
        if (j < i) continue;
        double acc = 0.0;
        if (i < T && j < T) {
            for (int k = band_ptr[i]; k < band_ptr[i + 1]; ++k) {
                const int q = band_q[k];
                const int p = w[q];
                const int is = qstart[q];
                if (p == 0 || j >= is + qdim[q]) continue;
                acc += (double)p * fq[fq_off[q] + (size_t)(i - is) * qdim[q] + (j - is)];
            }
            if (i == j && acc == 0.0) acc = 1.0;
        } else if (i == j) {
            acc = 1.0;
        }
        A[(size_t)i * Np + j] = acc;
        A[(size_t)j * Np + i] = acc;
    }
    if (I == 0 && J == 0) {
        for (int i = threadIdx.x; i < Np; i += 256) {
            double acc = 0.0;
            if (i < T)
                for (int k = band_ptr[i]; k < band_ptr[i + 1]; ++k) {
                    const int q = band_q[k];
                    const int p = w[q];
                    if (p != 0) acc += (double)p * theta[(size_t)q * T + i];
                }
            rhs[(size_t)b * Np + i] = acc;
        }
    }
}

// x = Li^T (Li rhs) per realisation (one block each), written to R[b0 + b]
// (T values).  A band no weighted spectrum covers has a zero row and a zero
// right-hand side with its diagonal set to 1, so its x is zero, as
// safeSolveCho leaves it.
__global__ __launch_bounds__(256) void k_boot_solve(const double *mat, const double *rhs, int b0, int T, int Np,
                                                    double *R) {
    const int b = blockIdx.x;
    const double *Li = mat + (size_t)b * 3 * Np * Np + 2 * (size_t)Np * Np;
    extern __shared__ double sh[];
    double *y = sh;          // Np
    const double *r = rhs + (size_t)b * Np;
    for (int i = threadIdx.x; i < Np; i += 256) {
        double s = 0.0;
        for (int k = 0; k <= i; ++k) s += Li[(size_t)i * Np + k] * r[k];
        y[i] = s;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < T; i += 256) {
        double s = 0.0;
        for (int k = i; k < Np; ++k) s += Li[(size_t)k * Np + i] * y[k];
        R[(size_t)(b0 + b) * T + i] = s;
    }
}

void launch_boot_collapse_fisher(hipStream_t st, const ChunkOut *outs, const double *fc, const int *qptr,
                                 const int *qchunk, const int *qstart, const int *qdim, const long long *fq_off,
                                 int nqso, double *fq) {
    k_boot_collapse_fisher<<<nqso, 256, 0, st>>>(outs, fc, qptr, qchunk, qstart, qdim, fq_off, fq);
}

void launch_boot_build(hipStream_t st, const int *W, int nqso, int b0, int nb, int T, int Np, const int *band_ptr,
                       const int *band_q, const int *qstart, const int *qdim, const long long *fq_off,
                       const double *fq, const double *theta, double *mat, double *rhs) {
    const int t = Np / 64;
    k_boot_build<<<dim3(t, t, nb), 256, 0, st>>>(W, nqso, b0, T, Np, band_ptr, band_q, qstart, qdim, fq_off, fq,
                                                 theta, mat, rhs);
}

void launch_boot_solve(hipStream_t st, const double *mat, const double *rhs, int b0, int nb, int T, int Np,
                       double *R) {
    k_boot_solve<<<nb, 256, (size_t)Np * sizeof(double), st>>>(mat, rhs, b0, T, Np, R);
}

void launch_boot_collapse(hipStream_t st, const ChunkOut *outs, const double *dbt, const int *qptr,
                          const int *qchunk, int nqso, int T, double *theta) {
    k_boot_collapse<<<nqso, 256, 0, st>>>(outs, dbt, qptr, qchunk, T, theta);
}

void launch_boot_meancov(hipStream_t st, const double *R, int nboots, int T, double *mean, double *cov) {
    const int tt = (T + BT - 1) / BT;
    k_boot_mean<<<(T + 255) / 256, 256, 0, st>>>(R, nboots, T, mean);
    k_boot_cov<<<dim3(tt, tt), 256, 0, st>>>(R, mean, nboots, T, cov);
}

void launch_bootstrap(hipStream_t st, const ChunkOut *outs, const double *dbt, const int *qptr, const int *qchunk,
                      int nqso, int T, const int *P, int nboots, double *theta, double *R, double *mean,
                      double *cov) {
    launch_boot_collapse(st, outs, dbt, qptr, qchunk, nqso, T, theta);
    const int tt = (T + BT - 1) / BT;
    k_boot_realize<<<dim3(tt, (nboots + BT - 1) / BT), 256, 0, st>>>(P, theta, nqso, nboots, T, R);
    launch_boot_meancov(st, R, nboots, T, mean, cov);
}

// ---------------------------------------------------------------------------
// Median bootstrap (PoissonBootstrapper::_medianBootstrap,
// core/bootstrapper.hpp:301-346,376-390, with stats::medianOfUnsortedVector,
// mathtools/stats.cpp:8-28): per band the median over the realisations, the
// deviations from it, and per band pair (i <= j) the median of the products
// of the deviations, mirrored to the lower triangle.  A median of n values is
// the value of rank n/2 for odd n and the half-sum of ranks n/2 - 1 and n/2
// for even n, the ranks of the ascending order.
//
// Each median is a radix select over order-preserving 64-bit keys of the
// FP64 values: eight passes of 8 bits, most significant first, each a
// 256-bin LDS histogram of the values still matching the selected prefix and
// a block scan that picks the digit holding the rank; for even n one more
// pass counts the values at or below the selected one and finds the next
// larger value.  After the first two passes the keys of the selected bin
// (up to 4096) are copied to LDS and the rest of the select reads only them.  The selected values are the sorted array's, bit for bit;
// products are formed the same way in every pass (one FP64 multiply).
// Realisations are transposed first so that a band's values are contiguous.

namespace {

// Ascending keys are ascending values; -0 sorts below +0 (the two are equal
// values, so the median does not depend on it).
__device__ __forceinline__ unsigned long long order_key(double x) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double key_value(unsigned long long k) {
    const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}

template <bool PROD>
__device__ __forceinline__ unsigned long long median_term(const double *a, const double *b, int i) {
    return order_key(PROD ? a[i] * b[i] : a[i]);
}

constexpr int MED_CAP = 4096;   // keys of the median's 16-bit bin kept in LDS

struct SelectLds {
    unsigned int hist[256];
    unsigned int wave_tot[4];
    unsigned long long wave_min[4];
    unsigned int digit, rank, binsize, ncand;
    unsigned long long cand[MED_CAP];
};

// One radix pass: the 8-bit digit at `shift` of the keys key(i), i < count,
// that match prefix under mask; selects the digit whose bin holds rank r and
// leaves r as the rank within that bin and s.binsize as the bin's count.
template <class Key>
__device__ void radix_pass(Key key, int count, int shift, unsigned long long &prefix, unsigned long long &mask,
                           unsigned int &r, SelectLds &s) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    s.hist[tid] = 0;
    __syncthreads();
    for (int i = tid; i < count; i += 256) {
        const unsigned long long k = key(i);
        if ((k & mask) == prefix) atomicAdd(&s.hist[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    // inclusive scan of the 256 bins: within the wave, then across waves
    const unsigned int c = s.hist[tid];
    unsigned int incl = c;
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    if (lane == 63) s.wave_tot[wave] = incl;
    __syncthreads();
    for (int w = 0; w < wave; ++w) incl += s.wave_tot[w];
    const unsigned int excl = incl - c;
    if (excl <= r && r < incl) {   // exactly one bin holds the rank
        s.digit = (unsigned)tid;
        s.rank = r - excl;
        s.binsize = c;
    }
    __syncthreads();
    prefix |= (unsigned long long)s.digit << shift;
    mask |= 0xffull << shift;
    r = s.rank;
    __syncthreads();
}

// Block-wide sum of cnt and minimum of mn; every thread gets both.
__device__ void block_sum_min(unsigned int &cnt, unsigned long long &mn, SelectLds &s) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int d = 32; d >= 1; d >>= 1) {
        cnt += __shfl_xor(cnt, d, 64);
        const unsigned long long o = __shfl_xor(mn, d, 64);
        mn = o < mn ? o : mn;
    }
    if (lane == 0) {
        s.wave_tot[wave] = cnt;
        s.wave_min[wave] = mn;
    }
    __syncthreads();
    cnt = 0;
    mn = ~0ull;
    for (int w = 0; w < 4; ++w) {
        cnt += s.wave_tot[w];
        mn = s.wave_min[w] < mn ? s.wave_min[w] : mn;
    }
    __syncthreads();
}

// Among key(i), i < count: how many are <= sel, and the smallest above it.
template <class Key>
__device__ void count_le_min_gt(Key key, int count, unsigned long long sel, unsigned int &tot,
                                unsigned long long &gmin, SelectLds &s) {
    unsigned int cnt = 0;
    unsigned long long mn = ~0ull;
    for (int i = threadIdx.x; i < count; i += 256) {
        const unsigned long long k = key(i);
        if (k <= sel) ++cnt;
        else mn = k < mn ? k : mn;
    }
    block_sum_min(cnt, mn, s);
    tot = cnt;
    gmin = mn;
}

// The median of the n values a[i] (PROD false) or a[i] * b[i] (PROD true).
// Called by all 256 threads of the block; every thread returns the median.
// The first two passes read the values; when the selected 16-bit bin holds
// at most MED_CAP keys they are copied to LDS once and the remaining six
// passes (and the even-count step) read only those, else every pass reads
// the values again.
template <bool PROD>
__device__ double block_median(const double *a, const double *b, int n, SelectLds &s) {
    const unsigned int r_lo = (n % 2 == 0) ? (unsigned)(n / 2 - 1) : (unsigned)(n / 2);
    auto gkey = [a, b](int i) { return median_term<PROD>(a, b, i); };
    auto lkey = [&s](int i) { return s.cand[i]; };
    unsigned long long prefix = 0, mask = 0;
    unsigned int r = r_lo;
    radix_pass(gkey, n, 56, prefix, mask, r, s);
    radix_pass(gkey, n, 48, prefix, mask, r, s);
    const unsigned int nbin = s.binsize, r_bin = r;
    const bool in_lds = nbin <= (unsigned)MED_CAP;
    unsigned long long above = ~0ull;   // smallest key above the bin
    if (in_lds) {
        if (threadIdx.x == 0) s.ncand = 0;
        __syncthreads();
        unsigned int unused = 0;
        for (int i = threadIdx.x; i < n; i += 256) {
            const unsigned long long k = gkey(i);
            if ((k & mask) == prefix) {
                const unsigned int at = atomicAdd(&s.ncand, 1u);
                if (at < (unsigned)MED_CAP) s.cand[at] = k;   // at < nbin always
            } else if (k > prefix) {
                above = k < above ? k : above;
            }
        }
        block_sum_min(unused, above, s);
        for (int shift = 40; shift >= 0; shift -= 8) radix_pass(lkey, (int)nbin, shift, prefix, mask, r, s);
    } else {
        for (int shift = 40; shift >= 0; shift -= 8) radix_pass(gkey, n, shift, prefix, mask, r, s);
    }
    const double lo = key_value(prefix);
    if (n % 2 != 0) return lo;
    // even n: the value of rank n/2 is lo itself when more than n/2 values
    // are at or below it, else the smallest value above it
    unsigned int tot;
    unsigned long long gmin;
    double hi;
    if (in_lds) {
        // (r_lo - r_bin) values lie below the bin
        count_le_min_gt(lkey, (int)nbin, prefix, tot, gmin, s);
        hi = (tot >= r_bin + 2) ? lo : key_value(gmin < above ? gmin : above);
    } else {
        count_le_min_gt(gkey, n, prefix, tot, gmin, s);
        hi = (tot >= (unsigned)(n / 2 + 1)) ? lo : key_value(gmin);
    }
    return (lo + hi) / 2;
}
}  // namespace

// R (nboots x T, one realisation per row) to X (T x nboots), 64 x 64 tiles.
__global__ __launch_bounds__(256) void k_boot_transpose(const double *R, int nboots, int T, double *X) {
    __shared__ double tile[64][65];
    const int b0 = blockIdx.y * 64, i0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int rr = ty; rr < 64; rr += 4) {
        const int b = b0 + rr, i = i0 + tx;
        tile[rr][tx] = (b < nboots && i < T) ? R[(size_t)b * T + i] : 0.0;
    }
    __syncthreads();
    for (int rr = ty; rr < 64; rr += 4) {
        const int i = i0 + rr, b = b0 + tx;
        if (i < T && b < nboots) X[(size_t)i * nboots + b] = tile[tx][rr];
    }
}

// One block per band: median[i], then X[i][:] -= median[i] (_calcuateMedian,
// and the subtraction at the start of _calcuateMadCovariance).
__global__ __launch_bounds__(256) void k_boot_median(double *X, int nboots, double *median) {
    __shared__ SelectLds s;
    double *x = X + (size_t)blockIdx.x * nboots;
    const double m = block_median<false>(x, x, nboots, s);
    if (threadIdx.x == 0) median[blockIdx.x] = m;
    for (int i = threadIdx.x; i < nboots; i += 256) x[i] -= m;
}

// One block per band pair j >= i (blocks with j < i return at once):
// mad[i][j] = mad[j][i] = median over b of X[i][b] X[j][b].
__global__ __launch_bounds__(256) void k_boot_mad(const double *X, int nboots, int T, double *mad) {
    const int j = blockIdx.x, i = blockIdx.y;
    if (j < i) return;
    __shared__ SelectLds s;
    const double m = block_median<true>(X + (size_t)i * nboots, X + (size_t)j * nboots, nboots, s);
    if (threadIdx.x == 0) {
        mad[(size_t)i * T + j] = m;
        mad[(size_t)j * T + i] = m;
    }
}

void launch_boot_median(hipStream_t st, const double *R, int nboots, int T, double *X, double *median,
                        double *mad) {
    k_boot_transpose<<<dim3((T + 63) / 64, (nboots + 63) / 64), 256, 0, st>>>(R, nboots, T, X);
    k_boot_median<<<T, 256, 0, st>>>(X, nboots, median);
    k_boot_mad<<<dim3(T, T), 256, 0, st>>>(X, nboots, T, mad);
}

// Scratch of the allocation-free median bootstrap: the realisations
// transposed to band-major, BOOT_NOALLOC_MAX doubles (64 MiB) per device,
// held from the library's load on a device until the process ends.  Its
// address on each device comes from hipGetSymbolAddress at qml_create (with
// that device current); host code never passes the array's own name.
static __device__ __align__(256) double scratch[BOOT_NOALLOC_MAX];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[BOOT_NOALLOC_DEVICES] = {};
}

int boot_noalloc_setup(int device) {
    if (device < 0 || device >= BOOT_NOALLOC_DEVICES) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return (int)hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch_ptr[device] = (double *)ptr;
    return (int)hipSuccess;
}

double *boot_noalloc_scratch(int device) {
    if (device < 0 || device >= BOOT_NOALLOC_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace qml
