// Read-time preprocessing of a spectrum's resolution matrix (host code),
// done once per spectrum before it is cut into chunks, as the reference's
// reader does (core/one_qso_estimate.cpp:74-89):
//   qml_rmat_deconvolve  DiaMatrix::deconvolve   (mathtools/matrix_helper.cpp:653-682,
//                        RealField::deconvolveSinc, mathtools/real_field.cpp:80-111)
//   qml_rmat_oversample  Resolution::oversample  (mathtools/matrix_helper.cpp:1029-1073)
// Both work row by row on the diagonal storage (values ndiags x size,
// diagonal d holding R[c - (ndiags/2 - d)][c] at column c).  A row is read
// as ndiags values centred on the diagonal; where it runs off the matrix the
// missing values are the row's own values mirrored about its centre
// (DiaMatrix::_getRowIndices, mathtools/matrix_helper.cpp:608-620).
#include "lyspeq_qml.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace {

constexpr double PREP_PI = 3.14159265358979323846;
constexpr double PREP_EPS = 1e-15;   // MY_EPSILON_D

// storage index of row i, slot p (column i - h + p), or -1 off the matrix
inline long long dia_index(int size, int ndiags, int i, int p) {
    const int h = ndiags / 2, c = i - h + p;
    if (c < 0 || c >= size) return -1;
    return (long long)(ndiags - 1 - p) * size + c;
}

void dia_row(const double *values, int size, int ndiags, int i, double *row) {
    for (int p = 0; p < ndiags; ++p) {
        long long k = dia_index(size, ndiags, i, p);
        if (k < 0) k = dia_index(size, ndiags, i, ndiags - 1 - p);   // mirrored slot
        row[p] = (k < 0) ? 0.0 : values[k];
    }
}

// The smallest |x| above 1e-15 (1e-15 if none), nonzero_min_element
// (mathtools/matrix_helper.cpp:20-41).
double nonzero_min(const double *x, int n) {
    double s = std::fabs(x[0]);
    for (int i = 1; i < n; ++i) {
        const double t = std::fabs(x[i]);
        if (s < PREP_EPS) s = t;
        else if (t < s && t > PREP_EPS) s = t;
    }
    return s < PREP_EPS ? PREP_EPS : s;
}

// Natural cubic spline through (x0 + k h, y_k), k < n: second derivatives
// from the tridiagonal system with zero end values.
void natural_spline(const std::vector<double> &y, double h, std::vector<double> &m2) {
    const int n = (int)y.size();
    m2.assign(n, 0.0);
    if (n < 3) return;
    const int k = n - 2;                 // interior unknowns
    std::vector<double> cp(k), dp(k);
    for (int i = 0; i < k; ++i) {
        const double r = 6.0 * (y[i + 2] - 2.0 * y[i + 1] + y[i]) / (h * h);
        const double b = 4.0 - (i ? cp[i - 1] : 0.0);
        cp[i] = 1.0 / b;
        dp[i] = (r - (i ? dp[i - 1] : 0.0)) / b;
    }
    m2[k] = dp[k - 1];
    for (int i = k - 2; i >= 0; --i) m2[i + 1] = dp[i] - cp[i] * m2[i + 2];
}

double spline_eval(const std::vector<double> &y, const std::vector<double> &m2, double x0, double h, double x) {
    const int n = (int)y.size();
    int k = (int)std::floor((x - x0) / h);
    k = std::min(std::max(k, 0), n - 2);
    const double t = (x - (x0 + k * h)) / h, u = 1.0 - t;
    return u * y[k] + t * y[k + 1] + h * h / 6.0 * ((u * u * u - u) * m2[k] + (t * t * t - t) * m2[k + 1]);
}

}  // namespace

extern "C" int qml_rmat_deconvolve(int size, int ndiags, double m, double *values) {
    if (size < 1 || ndiags < 1 || ndiags % 2 == 0 || !values || !(m > 0)) return QML_ERR_ARG;
    const int pad = 5, n = ndiags + 2 * pad, nk = n / 2 + 1;
    std::vector<double> row(ndiags), x(n);
    std::vector<std::complex<double>> X(nk);
    for (int i = 0; i < size; ++i) {
        dia_row(values, size, ndiags, i, row.data());
        std::fill(x.begin(), x.end(), 0.0);
        std::copy(row.begin(), row.end(), x.begin() + pad);
        // real DFT, then divide mode j >= 1 by sinc(pi m j / n)
        for (int j = 0; j < nk; ++j) {
            std::complex<double> s = 0.0;
            for (int t = 0; t < n; ++t) s += x[t] * std::polar(1.0, -2.0 * PREP_PI * j * t / n);
            if (j > 0) {
                const double a = PREP_PI * m * j / n;
                s /= std::sin(a) / a;
            }
            X[j] = s;
        }
        // inverse of a real signal's DFT (the imaginary parts of modes 0 and n/2 are ignored)
        for (int t = 0; t < n; ++t) {
            double s = X[0].real();
            for (int j = 1; j < nk; ++j) {
                const double w = (2 * j == n) ? 1.0 : 2.0;
                s += w * (X[j] * std::polar(1.0, 2.0 * PREP_PI * j * t / n)).real();
            }
            x[t] = s / n;
        }
        for (int p = 0; p < ndiags; ++p) {
            const long long k = dia_index(size, ndiags, i, p);
            if (k >= 0) values[k] = x[p + pad];
        }
    }
    return QML_OK;
}

extern "C" int qml_rmat_oversample(int size, int ndiags, const double *values, int oversampling, double *rows) {
    if (size < 1 || ndiags < 3 || ndiags % 2 == 0 || oversampling < 1 || !values || !rows) return QML_ERR_ARG;
    const int noff = ndiags / 2, nel = 2 * noff * oversampling + 1;
    std::vector<double> row(ndiags), y(ndiags), m2;
    for (int i = 0; i < size; ++i) {
        dia_row(values, size, ndiags, i, row.data());
        // interpolate the log, after a shift that makes every value positive
        const double shift = *std::min_element(row.begin(), row.end()) - nonzero_min(row.data(), ndiags);
        for (int p = 0; p < ndiags; ++p) y[p] = std::log(row[p] - shift);
        natural_spline(y, 1.0, m2);
        double *out = rows + (size_t)i * nel;
        double sum = 0.0;
        for (int q = 0; q < nel; ++q) {
            out[q] = shift + std::exp(spline_eval(y, m2, -noff, 1.0, (double)q / oversampling - noff));
            sum += out[q];
        }
        const double isum = 1.0 / sum;
        for (int q = 0; q < nel; ++q) out[q] *= isum;
    }
    return QML_OK;
}
