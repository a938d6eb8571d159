// Test driver of the C++ host layer (lyspeq_amd/host/qml_host.hpp): reads a
// workload written by tests/test_gpu_host_cpp.py, runs the reference's
// driver pattern (OneQSOEstimate per spectrum, then oneQSOiteration over the
// chunks and one Engine::flush per iteration) and writes the sums and the
// per-chunk results.
//
// Workload file (little-endian): int32 magic 0x574c4d51 ("QMLW"); binning:
// int32 nk, nz; f64 kband_edges[nk+1], zbin_centers[nz], z_bin_width;
// int32 method, logl_order, lam_order, turn_off_sfid; f64
// max_forest_length_v; int32 number_of_chunks; f64 max_pixel_length_v;
// int32 oversampling_factor; f64 resomat_deconvolution_m;
// int32 precision, iterations; int32 ntables, then per table int32 R, f64
// dv, int32 nv, f64 v_length, int32 nz, f64 deriv[nk*nv], f64 signal[nv*nz];
// int32 nspec, then per spectrum int32 R, int32 n, f64 wave[n], delta[n],
// ivar[n], int32 ndiags, f64 rmat[ndiags*n] (diagonal storage; none if
// ndiags is 0); f64 theta[nk*nz].
//
// Output file: int32 T, f64 fisher_sum[T*T], f64 dbt_sum[3*T] (the sums of
// the last iteration), int32 nchunks, then per chunk (in construction order)
// int32 spectrum, fisher_index_start, n_q, status, f64 fisher[n_q*n_q],
// f64 dbt[3*n_q].
//
// Options after the two file names:
//   graph     Engine::setGraph(true) before the first iteration
//   autotune  Engine::autotune at theta over caps {2, 5, 1000} first
//   nr        the Newton-Raphson loop (PowerEstimate, qml_driver.hpp) from
//             theta, each iteration's theta feeding the next, no smoother and
//             no early stop; the output is then int32 T, int32 iterations run,
//             f64 theta[iterations][T], f64 sigma[T] of the last iteration.
//   boot      after the iterations, Engine::bootstrap over the spectra (500
//             realisations, seed 7) and Engine::bootstrapMedian of its
//             realisations; out.bin.boot receives f64 mean[T], cov[T*T],
//             median[T], mad_cov[T*T].
#include "qml_driver.hpp"
#include "qml_host.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

using namespace lyspeq_amd;

namespace {
struct Reader {
    FILE *f;
    template <class T> T get() {
        T v;
        if (std::fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("short workload file");
        return v;
    }
    std::vector<double> vec(size_t n) {
        std::vector<double> v(n);
        if (n && std::fread(v.data(), sizeof(double), n, f) != n) throw std::runtime_error("short workload file");
        return v;
    }
};
}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s workload.bin out.bin [graph] [autotune] [nr]\n", argv[0]);
        return 2;
    }
    bool graph = false, tune = false, nr = false, boot = false;
    for (int a = 3; a < argc; ++a) {
        if (!std::strcmp(argv[a], "graph"))
            graph = true;
        else if (!std::strcmp(argv[a], "autotune"))
            tune = true;
        else if (!std::strcmp(argv[a], "nr"))
            nr = true;
        else if (!std::strcmp(argv[a], "boot"))
            boot = true;
        else {
            std::fprintf(stderr, "unknown option %s\n", argv[a]);
            return 2;
        }
    }
    try {
        FILE *f = std::fopen(argv[1], "rb");
        if (!f) throw std::runtime_error("cannot open workload");
        Reader r{f};
        if (r.get<int32_t>() != 0x574c4d51) throw std::runtime_error("bad magic");
        Binning b;
        const int nk = r.get<int32_t>(), nz = r.get<int32_t>();
        b.kband_edges = r.vec(nk + 1);
        b.zbin_centers = r.vec(nz);
        b.z_bin_width = r.get<double>();
        b.z_binning_method = r.get<int32_t>();
        b.cont_logl_order = r.get<int32_t>();
        b.cont_lam_order = r.get<int32_t>();
        b.turn_off_sfid = r.get<int32_t>() != 0;
        b.max_forest_length_v = r.get<double>();
        b.number_of_chunks = r.get<int32_t>();
        b.max_pixel_length_v = r.get<double>();
        b.oversampling_factor = r.get<int32_t>();
        b.resomat_deconvolution_m = r.get<double>();
        const int prec = r.get<int32_t>(), iters = r.get<int32_t>();
        Engine eng(b, prec);
        const int ntab = r.get<int32_t>();
        for (int t = 0; t < ntab; ++t) {
            const int R = r.get<int32_t>();
            const double dv = r.get<double>();
            const int nv = r.get<int32_t>();
            const double vl = r.get<double>();
            const int tz = r.get<int32_t>();
            const std::vector<double> dq = r.vec((size_t)nk * nv), sg = r.vec((size_t)nv * tz);
            eng.addTable(R, dv, nv, vl, tz, dq.data(), sg.data());
        }
        const int nspec = r.get<int32_t>();
        std::vector<std::unique_ptr<OneQSOEstimate>> qsos;
        for (int s = 0; s < nspec; ++s) {
            QSOSpectrum q;
            q.R_fwhm = r.get<int32_t>();
            const int n = r.get<int32_t>();
            q.wave = r.vec(n);
            q.delta = r.vec(n);
            q.ivar = r.vec(n);
            q.rmat_ndiags = r.get<int32_t>();
            q.rmat = r.vec((size_t)q.rmat_ndiags * n);
            q.fname = "spectrum-" + std::to_string(s);
            qsos.push_back(std::make_unique<OneQSOEstimate>(eng, q));
        }
        const int T = eng.totalKzBins();
        const std::vector<double> theta = r.vec(T);
        std::fclose(f);

        if (tune) {
            double ms = 0.0;
            const int cap = eng.autotune(theta.data(), {2, 5, 1000}, 1, &ms);
            std::printf("host driver: autotuned batch cap %d (%.2f ms)\n", cap, ms);
        }
        if (graph) eng.setGraph(true);
        if (nr) {
            std::vector<OneQSOEstimate *> ptrs;
            for (auto &q : qsos) ptrs.push_back(q.get());
            PowerEstimate pe(eng, ptrs, iters, 0.0);
            const int run = pe.iterate(theta);
            FILE *o = std::fopen(argv[2], "wb");
            if (!o) throw std::runtime_error("cannot open output");
            const int32_t hdr[2] = {T, run};
            std::fwrite(hdr, 4, 2, o);
            for (const auto &th : pe.history) std::fwrite(th.data(), 8, T, o);
            std::fwrite(pe.power.sigma.data(), 8, T, o);
            std::fclose(o);
            std::printf("host driver: Newton-Raphson, %d iteration(s), chi2 diag %.3e full %.3e\n", run,
                        pe.convergence.chi2_diag, pe.convergence.chi2_full);
            return 0;
        }

        std::vector<double> F;
        std::vector<std::unique_ptr<double[]>> dbt;
        for (int it = 0; it < iters; ++it) {
            F.assign((size_t)T * T, 0.0);
            dbt.clear();
            for (int k = 0; k < 3; ++k) {
                dbt.push_back(std::make_unique<double[]>(T));
                std::fill_n(dbt.back().get(), T, 0.0);
            }
            for (auto &q : qsos) q->oneQSOiteration(theta.data(), dbt, F.data());
            eng.flush();
        }

        if (boot) {
            std::vector<OneQSOEstimate *> queue;
            for (auto &q : qsos) queue.push_back(q.get());
            std::vector<double> mean(T), cov((size_t)T * T), reals((size_t)500 * T);
            eng.bootstrap(queue, 500, 7, mean.data(), cov.data(), reals.data());
            std::vector<double> median(T), mad((size_t)T * T);
            eng.bootstrapMedian(500, reals.data(), median.data(), mad.data());
            FILE *bo = std::fopen((std::string(argv[2]) + ".boot").c_str(), "wb");
            if (!bo) throw std::runtime_error("cannot open bootstrap output");
            std::fwrite(mean.data(), 8, T, bo);
            std::fwrite(cov.data(), 8, cov.size(), bo);
            std::fwrite(median.data(), 8, T, bo);
            std::fwrite(mad.data(), 8, mad.size(), bo);
            std::fclose(bo);
        }
        FILE *o = std::fopen(argv[2], "wb");
        if (!o) throw std::runtime_error("cannot open output");
        const int32_t T32 = T;
        std::fwrite(&T32, 4, 1, o);
        std::fwrite(F.data(), 8, F.size(), o);
        for (int k = 0; k < 3; ++k) std::fwrite(dbt[k].get(), 8, T, o);
        int32_t nch = 0;
        for (auto &q : qsos) nch += (int32_t)q->chunks.size();
        std::fwrite(&nch, 4, 1, o);
        for (int s = 0; s < (int)qsos.size(); ++s)
            for (auto &c : qsos[s]->chunks) {
                const int32_t hdr[4] = {s, c->fisher_index_start, c->N_Q_MATRICES, c->lastStatus()};
                std::fwrite(hdr, 4, 4, o);
                std::fwrite(c->fisher_matrix.get(), 8, (size_t)c->N_Q_MATRICES * c->N_Q_MATRICES, o);
                for (int k = 0; k < 3; ++k)
                    std::fwrite(c->dbt_estimate_before_fisher_vector[k].get(), 8, c->N_Q_MATRICES, o);
            }
        std::fclose(o);
        std::printf("host driver: %d spectra, %d chunks, T=%d, %d iteration(s)\n", nspec, (int)nch, T, iters);
        return 0;
    } catch (QMLError &e) {
        std::fprintf(stderr, "QMLError %d: %s\n", e.code, e.what());
        return 1;
    } catch (std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
