// Newton-Raphson driver arithmetic on the host (see qml_driver.hpp).
#include "qml_driver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <stdexcept>
#include <string>

#include "qml_host.hpp"

namespace lyspeq_amd {

void normalizeFisher(double *F, int T) {
    for (long long i = 0; i < (long long)T * T; ++i) F[i] *= 0.5;
    for (int i = 0; i < T; ++i)
        for (int j = i + 1; j < T; ++j) F[(size_t)j * T + i] = F[(size_t)i * T + j];
}

std::vector<double> invertLU(const double *A, int n) {
    // Gauss-Jordan elimination with partial pivoting on [A | I]
    std::vector<double> a(A, A + (size_t)n * n), inv((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::fabs(a[(size_t)k * n + k]);
        for (int r = k + 1; r < n; ++r) {
            const double v = std::fabs(a[(size_t)r * n + k]);
            if (v > best) {
                best = v;
                p = r;
            }
        }
        if (best == 0.0) throw std::runtime_error("ERROR in LU decomposition.");
        if (p != k) {
            std::swap_ranges(a.begin() + (size_t)k * n, a.begin() + (size_t)(k + 1) * n, a.begin() + (size_t)p * n);
            std::swap_ranges(inv.begin() + (size_t)k * n, inv.begin() + (size_t)(k + 1) * n,
                             inv.begin() + (size_t)p * n);
        }
        const double piv = 1.0 / a[(size_t)k * n + k];
        double *ak = &a[(size_t)k * n], *ik = &inv[(size_t)k * n];
        for (int j = 0; j < n; ++j) {
            ak[j] *= piv;
            ik[j] *= piv;
        }
        for (int r = 0; r < n; ++r) {
            if (r == k) continue;
            double *ar = &a[(size_t)r * n], *ir = &inv[(size_t)r * n];
            const double f = ar[k];
            if (f == 0.0) continue;
            for (int j = k; j < n; ++j) ar[j] -= f * ak[j];
            for (int j = 0; j < n; ++j) ir[j] -= f * ik[j];
        }
    }
    return inv;
}

std::vector<double> symmetricSingularValues(const double *A, int n) {
    // Householder reduction to a symmetric tridiagonal (d, e): at step k the
    // reflector H = I - 2 v v^T (|v| = 1) zeroes column k below row k + 1,
    // and H A H = A - 2 v w^T - 2 w v^T with p = A v, w = p - (v^T p) v.
    std::vector<double> a(A, A + (size_t)n * n), d(n, 0.0), e(std::max(n - 1, 0), 0.0);
    std::vector<double> v(n), p(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;   // length of the column below the diagonal
        double norm = 0.0;
        for (int i = 0; i < m; ++i) norm += a[(size_t)(k + 1 + i) * n + k] * a[(size_t)(k + 1 + i) * n + k];
        norm = std::sqrt(norm);
        const double x0 = a[(size_t)(k + 1) * n + k];
        const double alpha = (x0 > 0) ? -norm : norm;
        double vn = 0.0;
        for (int i = 0; i < m; ++i) {
            v[i] = a[(size_t)(k + 1 + i) * n + k];
            if (i == 0) v[i] -= alpha;
            vn += v[i] * v[i];
        }
        d[k] = a[(size_t)k * n + k];
        if (vn == 0.0) {   // column already zero below the subdiagonal
            e[k] = x0;
            continue;
        }
        e[k] = alpha;
        vn = 1.0 / std::sqrt(vn);
        for (int i = 0; i < m; ++i) v[i] *= vn;
        double K = 0.0;
        for (int i = 0; i < m; ++i) {
            const double *row = &a[(size_t)(k + 1 + i) * n + k + 1];
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += row[j] * v[j];
            p[i] = s;
            K += v[i] * s;
        }
        for (int i = 0; i < m; ++i) p[i] -= K * v[i];
        for (int i = 0; i < m; ++i) {
            double *row = &a[(size_t)(k + 1 + i) * n + k + 1];
            for (int j = 0; j < m; ++j) row[j] -= 2.0 * (v[i] * p[j] + p[i] * v[j]);
        }
    }
    if (n >= 2) {
        d[n - 2] = a[(size_t)(n - 2) * n + n - 2];
        e[n - 2] = a[(size_t)(n - 1) * n + n - 2];
    }
    if (n >= 1) d[n - 1] = a[(size_t)(n - 1) * n + n - 1];

    // eigenvalues by bisection on the Sturm count: the number of negative
    // pivots of T - x I equals the number of eigenvalues below x
    double lo = 0.0, hi = 0.0;
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r);
        hi = std::max(hi, d[i] + r);
    }
    const double scale = std::max(std::fabs(lo), std::fabs(hi));
    const double tiny = std::numeric_limits<double>::min() / std::numeric_limits<double>::epsilon();
    auto count_below = [&](double x) {
        int c = 0;
        double q = 1.0;
        for (int i = 0; i < n; ++i) {
            const double e2 = (i > 0) ? e[i - 1] * e[i - 1] : 0.0;
            q = d[i] - x - ((i > 0) ? e2 / q : 0.0);
            if (q == 0.0) q = -tiny * std::max(scale, 1.0);
            if (q < 0.0) ++c;
        }
        return c;
    };
    std::vector<double> ev(n);
    const double eps = std::numeric_limits<double>::epsilon();
    for (int k = 0; k < n; ++k) {
        double l = lo - eps * scale - tiny, h = hi + eps * scale + tiny;
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (l + h);
            if (mid == l || mid == h) break;
            if (count_below(mid) > k)
                h = mid;
            else
                l = mid;
            if (h - l <= 2.0 * eps * std::max(std::fabs(l), std::fabs(h))) break;
        }
        ev[k] = 0.5 * (l + h);
    }
    for (double &x : ev) x = std::fabs(x);
    std::sort(ev.begin(), ev.end(), std::greater<double>());
    return ev;
}

FisherInverse stableInvertSym(const double *F, int T) {
    FisherInverse r;
    std::vector<double> s(F, F + (size_t)T * T);
    std::vector<int> empty;
    for (int i = 0; i < T; ++i)
        if (s[(size_t)i * T + i] == 0) {
            empty.push_back(i);
            s[(size_t)i * T + i] = 1.0;
        }
    r.dof = T - (int)empty.size();
    const std::vector<double> sv = symmetricSingularValues(s.data(), T);
    for (int i = T - 1; i > 0; --i) {
        if (sv[i - 1] / sv[i] > 8.0)
            r.damp = sv[i];
        else
            break;
    }
    r.damped = r.damp != 0.0;
    r.inverse = invertLU(s.data(), T);
    if (r.damped) {
        std::vector<double> ss((size_t)T * T, 0.0);
        for (int i = 0; i < T; ++i)
            for (int k = 0; k < T; ++k) {
                const double aik = s[(size_t)i * T + k];
                for (int j = 0; j < T; ++j) ss[(size_t)i * T + j] += aik * s[(size_t)k * T + j];
            }
        for (int i = 0; i < T; ++i) ss[(size_t)i * T + i] += r.damp;
        const std::vector<double> ai = invertLU(ss.data(), T);
        r.solver.assign((size_t)T * T, 0.0);
        for (int i = 0; i < T; ++i)
            for (int k = 0; k < T; ++k) {
                const double aik = ai[(size_t)i * T + k];
                for (int j = 0; j < T; ++j) r.solver[(size_t)i * T + j] += aik * s[(size_t)k * T + j];
            }
    } else {
        r.solver = r.inverse;
    }
    for (int i : empty) {
        r.solver[(size_t)i * T + i] = 0.0;
        r.inverse[(size_t)i * T + i] = 0.0;
    }
    return r;
}

PowerResult powerEstimates(const FisherInverse &fi, const double *dbt_sum, int T) {
    PowerResult p;
    p.weighted.assign((size_t)3 * T, 0.0);
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < T; ++i) {
            double s = 0.0;
            for (int j = 0; j < T; ++j) s += fi.solver[(size_t)i * T + j] * dbt_sum[(size_t)k * T + j];
            p.weighted[(size_t)k * T + i] = 0.5 * s;
        }
    p.theta.resize(T);
    p.sigma.resize(T);
    for (int i = 0; i < T; ++i) {
        p.theta[i] = p.weighted[i] - p.weighted[(size_t)T + i] - p.weighted[(size_t)2 * T + i];
        p.sigma[i] = std::sqrt(fi.inverse[(size_t)i * T + i]);
    }
    return p;
}

ConvergenceTest hasConverged(const double *current, const double *previous, const double *inverse,
                             const double *fisher, int T, int dof, double eps) {
    ConvergenceTest c;
    for (int i = 0; i < T; ++i) {
        const double p1 = std::fabs(current[i]), p2 = std::fabs(previous[i]);
        if (p1 == 0 && p2 == 0) continue;
        const double r = std::fabs(p1 - p2) / std::max(p1, p2);
        c.abs_mean += r / dof;
        c.abs_max = std::max(c.abs_max, r);
    }
    std::vector<double> t(T);
    for (int i = 0; i < T; ++i) t[i] = previous[i] - current[i];
    double r = 0.0;
    for (int i = 0; i < T; ++i) {
        const double e = inverse[(size_t)i * T + i];
        if (e <= 0) continue;
        r += t[i] * t[i] / e;
    }
    c.chi2_diag = std::sqrt(r / dof);
    // t^T F t from the upper triangle (my_cblas_dsymvdot with dsymv 'U')
    double full = 0.0;
    for (int i = 0; i < T; ++i) {
        double s = fisher[(size_t)i * T + i] * t[i];
        for (int j = i + 1; j < T; ++j) s += 2.0 * fisher[(size_t)i * T + j] * t[j];
        full += t[i] * s;
    }
    c.chi2_full = std::sqrt(std::fabs(full) / dof);
    c.converged = c.chi2_diag < eps || c.chi2_full < eps;
    return c;
}

void ScriptSmoother::operator()(const std::vector<double> &theta, const std::vector<double> &sigma,
                                std::vector<double> &fits) const {
    const int nk = (int)kcenters.size(), nz = (int)zcenters.size(), T = nk * nz;
    if ((int)theta.size() != T || (int)sigma.size() != T || (int)pfid.size() != T)
        throw std::invalid_argument("ScriptSmoother: theta, sigma and pfid need nz * nk values");
    const std::string fin = tmp_dir + "/tmp-power-" + tag + ".txt", fout = tmp_dir + "/tmp-smooth-" + tag + ".txt";
    FILE *f = std::fopen(fin.c_str(), "w");
    if (!f) throw std::runtime_error("ScriptSmoother: cannot write " + fin);
    std::fprintf(f, "%d %d\n", nz, nk);
    for (int z = 0; z < nz; ++z) std::fprintf(f, "%d ", z < (int)zcounts.size() ? zcounts[z] : 0);
    std::fprintf(f, "\n");
    for (int i = 0; i < T; ++i) {
        const int kn = i % nk, zm = i / nk;
        const double e = std::isnan(sigma[i]) ? 0.0 : sigma[i];
        std::fprintf(f, "%.3lf %e %e %e\n", zcenters[zm], kcenters[kn], theta[i] + pfid[i], e);
    }
    std::fclose(f);
    const std::string cmd = command + " " + fin + " " + fout + (interp_log ? " --interp_log" : "");
    const int rc = std::system(cmd.c_str());
    std::remove(fin.c_str());
    if (rc != 0) throw std::runtime_error("smoothing error");
    FILE *g = std::fopen(fout.c_str(), "r");
    if (!g) throw std::runtime_error("smoothing error");
    fits.assign(T, 0.0);
    for (int i = 0; i < T; ++i) {
        if (std::fscanf(g, "%le", &fits[i]) != 1) {
            std::fclose(g);
            throw std::runtime_error("Reading fit power values from tmp_fit_file!");
        }
        fits[i] -= pfid[i];
    }
    std::fclose(g);
    std::remove(fout.c_str());
}

PowerEstimate::PowerEstimate(Engine &eng, std::vector<OneQSOEstimate *> qsos, int iterations, double eps,
                             Smoother smooth)
    : eng_(&eng), qsos_(std::move(qsos)), iterations_(iterations), eps_(eps), smooth_(std::move(smooth)) {}

int PowerEstimate::iterate(std::vector<double> fits) {
    const int T = eng_->totalKzBins();
    if ((int)fits.size() != T) throw std::runtime_error("PowerEstimate::iterate: fits must hold TOTAL_KZ_BINS values");
    std::vector<double> previous(T, 0.0);
    history.clear();
    int it = 0;
    for (; it < iterations_; ++it) {
        fisher.assign((size_t)T * T, 0.0);
        std::vector<std::unique_ptr<double[]>> dbt;
        for (int k = 0; k < 3; ++k) {
            dbt.push_back(std::make_unique<double[]>(T));
            std::fill_n(dbt.back().get(), T, 0.0);
        }
        for (OneQSOEstimate *q : qsos_) q->oneQSOiteration(fits.data(), dbt, fisher.data());
        eng_->flush();
        normalizeFisher(fisher.data(), T);
        std::vector<double> flat((size_t)3 * T);
        for (int k = 0; k < 3; ++k) std::copy_n(dbt[k].get(), T, flat.begin() + (size_t)k * T);
        inverse = stableInvertSym(fisher.data(), T);
        if (it > 0) previous = power.theta;
        power = powerEstimates(inverse, flat.data(), T);
        history.push_back(power.theta);
        convergence = hasConverged(power.theta.data(), previous.data(), inverse.inverse.data(), fisher.data(), T,
                                   inverse.dof, eps_);
        if (convergence.converged || it == iterations_ - 1) {
            ++it;
            break;
        }
        if (smooth_)
            smooth_(power.theta, power.sigma, fits);
        else
            fits = power.theta;
    }
    return it;
}

}  // namespace lyspeq_amd
