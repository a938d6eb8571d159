// The S/Q lookup tables on the device (SQLookupTable::computeTables,
// core/sq_table.cpp:201-323), for many (R, dv) pairs at once.
//
//   derivative: Q_kn(v) = (1/pi) int_{k1}^{k2} W(k)^2 cos(k v) dk, W(k) =
//               sinc(k dv/2) exp(-(k R)^2/2), R the 1-sigma resolution in
//               km/s (core/fiducial_cosmology.cpp:305-331).  The reference
//               integrates with QUADPACK's QAWO; here Filon's rule over 2M
//               panels, which integrates the cosine weight exactly and the
//               smooth W^2 to O(h^4).
//   signal:     the reference's real FFT of W^2 P_fid(k, z) on n = 2^(ceil(
//               log2(4 Nv)) + 1) points spaced dx = (Lv/(Nv-1))/2, divided by
//               the grid length, every second point kept (core/sq_table.cpp:
//               221-227,296-311; mathtools/real_field.cpp:40-92).  The kept
//               points are the inverse transform's cosine sums, evaluated here
//               directly: S(v_j) = (1/L) [F_0 + 2 sum_{0<m<n/2} F_m cos(k_m v_j)
//               + F_{n/2}], with the modes where W^2 < 1e-40 W^2(0) left out.
//
// One workgroup per (pair, band, 256 velocities) for the derivative and per
// (pair, z, 256 velocities) for the signal; the integrand at the Filon nodes,
// or a tile of 256 modes, sits in LDS and each thread sums one velocity with
// the cosine advanced by the three-term recurrence, re-seeded every 32 steps.
#include "qml_common.hpp"
#include "qml_kernels.hpp"

namespace qml {

namespace {

constexpr double TAB_PI = 3.14159265358979323846;

__device__ __forceinline__ double window_sq(double k, double dv, double R) {
    const double x = k * dv / 2.0;
    const double s = (x == 0.0) ? 1.0 : sin(x) / x;
    const double kr = k * R;
    const double w = s * exp(-kr * kr / 2.0);
    return w * w;
}

// Filon's alpha, beta, gamma at theta = t h (series below 0.02)
__device__ __forceinline__ void filon_coeffs(double th, double &a, double &b, double &g) {
    if (fabs(th) < 0.02) {
        const double t2 = th * th;
        a = 2.0 * th * t2 / 45.0 - 2.0 * th * t2 * t2 / 315.0 + 2.0 * th * t2 * t2 * t2 / 4725.0;
        b = 2.0 / 3.0 + 2.0 * t2 / 15.0 - 4.0 * t2 * t2 / 105.0 + 2.0 * t2 * t2 * t2 / 567.0;
        g = 4.0 / 3.0 - 2.0 * t2 / 15.0 + t2 * t2 / 210.0 - t2 * t2 * t2 / 11340.0;
        return;
    }
    double s, c;
    sincos(th, &s, &c);
    const double t3 = th * th * th;
    a = (th * th + th * s * c - 2.0 * s * s) / t3;
    b = 2.0 * (th * (1.0 + c * c) - 2.0 * s * c) / t3;
    g = 4.0 * (s - th * c) / t3;
}

constexpr int MAX_FILON_NODES = 2 * 128 + 1;

// pairs: (R_kms, dv) per table; out: npairs x nk x nv
__global__ __launch_bounds__(256) void k_deriv_table(const double *pairs, const double *kedges, int nk, int nv,
                                                     double dxv, int M, double *out) {
    const int kn = blockIdx.y, p = blockIdx.z;
    const double R = pairs[2 * p], dv = pairs[2 * p + 1];
    const double a = kedges[kn], b = kedges[kn + 1];
    const double h = (b - a) / (2 * M);
    __shared__ double f[MAX_FILON_NODES];
    for (int j = threadIdx.x; j <= 2 * M; j += 256) f[j] = window_sq(a + j * h, dv, R) / TAB_PI;
    __syncthreads();
    const int iv = blockIdx.x * 256 + threadIdx.x;
    if (iv >= nv) return;
    const double t = dxv * iv;
    double al, be, ga;
    filon_coeffs(t * h, al, be, ga);
    const double twoc = 2.0 * cos(t * h);
    double se = 0.0, so = 0.0, cprev = 0.0, c = 0.0;
    for (int j = 0; j <= 2 * M; ++j) {
        if (j % 32 == 0) {
            // re-seed: c = cos(t x_j), cprev = cos(t x_{j-1})
            c = cos(t * (a + j * h));
            cprev = cos(t * (a + (j - 1) * h));
        }
        const double term = f[j] * c;
        if (j & 1) so += term;
        else se += term;
        const double cn = twoc * c - cprev;
        cprev = c;
        c = cn;
    }
    // the end nodes carry half the even weight
    se -= 0.5 * (f[0] * cos(t * a) + f[2 * M] * cos(t * b));
    const double v = h * (al * (f[2 * M] * sin(t * b) - f[0] * sin(t * a)) + be * se + ga * so);
    out[((size_t)p * nk + kn) * nv + iv] = v;
}

// P_fid(k_m, z_i) (core/fiducial_cosmology.cpp:273-282) on the signal grid:
// pk[i * nmode + m], m = 0 .. nmode - 1
__global__ __launch_bounds__(256) void k_pfid_grid(double dk, int nmode, double z1, double dz, int nz, PD13 pd,
                                                   double *pk) {
    const int m = blockIdx.x * 256 + threadIdx.x, iz = blockIdx.y;
    if (m >= nmode || iz >= nz) return;
    const double k = m * dk, z = z1 + dz * iz;
    const double K0 = 0.009, Z0 = 3.0;
    const double q0 = k / K0 + 1e-15;
    const double x0 = (1.0 + z) / (1.0 + Z0);
    const double lq = log(q0);
    pk[(size_t)iz * nmode + m] = pd.A * TAB_PI / K0 * pow(q0, 2.0 + pd.n + pd.alpha * lq + pd.beta * log(x0)) *
                                 pow(x0, pd.B) / (1.0 + pd.lambda * k * k);
}

// out: npairs x nv x nz (z contiguous); mend[p]: last mode kept for pair p
__global__ __launch_bounds__(256) void k_signal_table(const double *pairs, const int *mend, const double *pk,
                                                      int nmode, int nhalf, double dk, double dxv, double inv_len,
                                                      int nv, int nz, double *out) {
    const int iz = blockIdx.y, p = blockIdx.z;
    const double R = pairs[2 * p], dv = pairs[2 * p + 1];
    const int me = mend[p];
    const double *pz = pk + (size_t)iz * nmode;
    __shared__ double F[256];
    const int iv = blockIdx.x * 256 + threadIdx.x;
    const double theta = dk * dxv * iv;   // k_m v_j = m theta
    const double twoc = 2.0 * cos(theta);
    double acc = 0.0;
    for (int m0 = 0; m0 <= me; m0 += 256) {
        const int m = m0 + threadIdx.x;
        double fm = 0.0;
        if (m <= me) fm = ((m == 0 || m == nhalf) ? 1.0 : 2.0) * window_sq(m * dk, dv, R) * pz[m];
        F[threadIdx.x] = fm;
        __syncthreads();
        const int cnt = min(256, me - m0 + 1);
        double c = 0.0, cprev = 0.0;
        for (int u = 0; u < cnt; ++u) {
            if (u % 32 == 0) {
                c = cos(theta * (m0 + u));
                cprev = cos(theta * (m0 + u - 1));
            }
            acc += F[u] * c;
            const double cn = twoc * c - cprev;
            cprev = c;
            c = cn;
        }
        __syncthreads();
    }
    if (iv < nv) out[((size_t)p * nv + iv) * nz + iz] = acc * inv_len;
}

}  // namespace

void launch_deriv_tables(hipStream_t st, const double *pairs, int npairs, const double *kedges, int nk, int nv,
                         double dxv, int M, double *out) {
    if (npairs <= 0 || nk <= 0 || nv <= 0) return;
    dim3 grid((nv + 255) / 256, nk, npairs);
    k_deriv_table<<<grid, 256, 0, st>>>(pairs, kedges, nk, nv, dxv, M, out);
}

void launch_pfid_grid(hipStream_t st, double dk, int nmode, double z1, double dz, int nz, PD13 pd, double *pk) {
    dim3 grid((nmode + 255) / 256, nz);
    k_pfid_grid<<<grid, 256, 0, st>>>(dk, nmode, z1, dz, nz, pd, pk);
}

void launch_signal_tables(hipStream_t st, const double *pairs, const int *mend, int npairs, const double *pk,
                          int nmode, int nhalf, double dk, double dxv, double inv_len, int nv, int nz, double *out) {
    if (npairs <= 0 || nv <= 0 || nz <= 0) return;
    dim3 grid((nv + 255) / 256, nz, npairs);
    k_signal_table<<<grid, 256, 0, st>>>(pairs, mend, pk, nmode, nhalf, dk, dxv, inv_len, nv, nz, out);
}

}  // namespace qml
