// Runs the C++ Newton-Raphson driver arithmetic (lyspeq_amd/host/qml_driver.
// hpp) on matrices read from a file, for tests/test_host_driver_math.py.
//
//   driver_check invert   in out   in: n, A (n x n)           out: A^-1
//   driver_check svals    in out   in: n, A                   out: singular values
//   driver_check power    in out   in: T, fisher_sum, dbt     out: normalised F, solver,
//                                                                  inverse, theta, sigma,
//                                                                  weighted, damp, dof
//   driver_check converge in out   in: T, dof, eps, current, previous, inverse, fisher
//                                  out: abs_mean, abs_max, chi2_diag, chi2_full, converged
//   driver_check smooth in out command
//                                  in: nz, nk, interp_log, zcenters, kcenters, theta, sigma,
//                                      pfid (z-major)        out: the fits of ScriptSmoother
//                                  running `command` (temporary files in $TMPDIR or /tmp)
//
// Files: int32 counts and little-endian doubles.  Exit status 3 and the
// exception's message on stderr when the arithmetic throws.
#include "qml_driver.hpp"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

using namespace lyspeq_amd;

namespace {
struct In {
    FILE *f;
    int32_t i() {
        int32_t v;
        if (std::fread(&v, 4, 1, f) != 1) throw std::invalid_argument("short input");
        return v;
    }
    double d() {
        double v;
        if (std::fread(&v, 8, 1, f) != 1) throw std::invalid_argument("short input");
        return v;
    }
    std::vector<double> v(size_t n) {
        std::vector<double> x(n);
        if (n && std::fread(x.data(), 8, n, f) != n) throw std::invalid_argument("short input");
        return x;
    }
};
void put(FILE *o, const std::vector<double> &x) { std::fwrite(x.data(), 8, x.size(), o); }
void put(FILE *o, double x) { std::fwrite(&x, 8, 1, o); }
}  // namespace

int main(int argc, char **argv) {
    if (argc != 4 && !(argc == 5 && std::string(argv[1]) == "smooth")) {
        std::fprintf(stderr, "usage: %s invert|svals|power|converge in.bin out.bin | smooth in out command\n",
                     argv[0]);
        return 2;
    }
    const std::string cmd = argv[1];
    FILE *f = std::fopen(argv[2], "rb");
    if (!f) return 2;
    In in{f};
    std::vector<double> out;
    FILE *o = nullptr;
    try {
        if (cmd == "invert" || cmd == "svals") {
            const int n = in.i();
            const std::vector<double> a = in.v((size_t)n * n);
            out = (cmd == "invert") ? invertLU(a.data(), n) : symmetricSingularValues(a.data(), n);
            o = std::fopen(argv[3], "wb");
            put(o, out);
        } else if (cmd == "power") {
            const int T = in.i();
            std::vector<double> F = in.v((size_t)T * T);
            const std::vector<double> dbt = in.v((size_t)3 * T);
            normalizeFisher(F.data(), T);
            const FisherInverse fi = stableInvertSym(F.data(), T);
            const PowerResult p = powerEstimates(fi, dbt.data(), T);
            o = std::fopen(argv[3], "wb");
            put(o, F);
            put(o, fi.solver);
            put(o, fi.inverse);
            put(o, p.theta);
            put(o, p.sigma);
            put(o, p.weighted);
            put(o, fi.damp);
            put(o, (double)fi.dof);
        } else if (cmd == "converge") {
            const int T = in.i(), dof = in.i();
            const double eps = in.d();
            const std::vector<double> cur = in.v(T), prev = in.v(T), inv = in.v((size_t)T * T),
                                      fis = in.v((size_t)T * T);
            const ConvergenceTest c = hasConverged(cur.data(), prev.data(), inv.data(), fis.data(), T, dof, eps);
            o = std::fopen(argv[3], "wb");
            put(o, c.abs_mean);
            put(o, c.abs_max);
            put(o, c.chi2_diag);
            put(o, c.chi2_full);
            put(o, c.converged ? 1.0 : 0.0);
        } else if (cmd == "smooth") {
            const int nz = in.i(), nk = in.i(), lg = in.i();
            ScriptSmoother sm;
            sm.zcenters = in.v(nz);
            sm.kcenters = in.v(nk);
            const std::vector<double> theta = in.v((size_t)nz * nk), sigma = in.v((size_t)nz * nk);
            sm.pfid = in.v((size_t)nz * nk);
            sm.interp_log = lg != 0;
            sm.command = argv[4];
            const char *td = std::getenv("TMPDIR");
            if (td && *td) sm.tmp_dir = td;
            sm(theta, sigma, out);
            o = std::fopen(argv[3], "wb");
            put(o, out);
        } else {
            std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
            return 2;
        }
    } catch (const std::runtime_error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 3;
    }
    std::fclose(f);
    if (o) std::fclose(o);
    return 0;
}
