// Compiled against the drop-in headers exactly as the reference's driver
// includes them ("core/chunk_estimate.hpp", "core/one_qso_estimate.hpp").
// The static assertions pin every member the driver, the bootstrapper and
// the load balancer use to the reference's declarations
// (core/chunk_estimate.hpp:46-119, core/one_qso_estimate.hpp:16-61).
//
//   dropin_check cost <config>                 per-file and per-chunk cost estimates
//   dropin_check run <config> <tables_dir> <tables_base> <precision> <out>
//        every file through OneQSOEstimate, oneQSOiteration at theta = 0,
//        flush; then collapseBootstrap and addBootPowerOnly / addBoot with
//        the weights of the config; results as text in <out>
//   dropin_check bootio <base> <out>
//        the bootstrap output files written from fixed arrays (a chunks file
//        of three chunks, a realizations file); the arrays as text in <out>;
//        exit 1 if readBootstrapRealizations does not return them
//   dropin_check read binary|picca <file> <out>
//        one file through qio::QSOFile (readParameters, readData and, if the
//        file has one, readAllocResolutionMatrix); fields and arrays in <out>
//   dropin_check exposure <config> <tables_dir> <tables_base> <out_dir>
//        the first file's first chunk through a Chunk subclass built as the
//        reference's cross/exposure.hpp:19-70 builds itself (Chunk(),
//        _copyQSOFile, _setNQandFisherIndex, its own covariance buffer):
//        the signal matrix and the marginalisation of S + I, as raw doubles
//        in <out_dir>/ex_signal.bin, ex_cinv_in.bin, ex_cinv_marg.bin
//   dropin_check friend <config> <tables_dir> <tables_base> <out_dir>
//        the first file through TestOneQSOEstimate, the friend class of the
//        reference's tests/testSQCMatrices.cpp:24-95 (allocMemory,
//        _initMatrices and _setVZMatrices in its constructor; the signal
//        matrix into covariance_matrix; Q_0 weighted by its z bin), then the
//        whole stored set and the marginalisation of S + I; each matrix as
//        raw doubles in <out_dir>/<name>.bin
//
// <config>: whitespace-separated "key value..." lines: kedges <n> <values>,
// zcenters <n> <values>, zwidth, method, logl_order, lam_order, max_forest,
// nchunks, max_pixel, fast_bootstrap, nboots, weights <nqso*nboots ints>,
// format binary|picca, rmat 0|1, deconv <m>, osamp <factor>, growth <B>
// (RedshiftGrowthOn with exponent B),
// first_table <R> <dv> (the table every chunk uses with rmat 1),
// save_chunks <base> (_saveChunkResults' file after the flush), save_boot
// <base> (the realizations file of the bootstrap powers, with the Fisher sum
// standing in for the solver matrix), file <path> (one per spectrum).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "core/chunk_estimate.hpp"
#include "core/one_qso_estimate.hpp"
#include "io/bootstrap_file.hpp"
#include "lyspeq_dropin.hpp"

using dbt_t = std::vector<std::unique_ptr<double[]>>;

// --- Chunk (core/chunk_estimate.hpp:86-117)
static_assert(std::is_constructible<Chunk, const qio::QSOFile &, int, int>::value, "Chunk(qmaster, i1, i2)");
static_assert(!std::is_copy_constructible<Chunk>::value && !std::is_move_constructible<Chunk>::value,
              "Chunk is neither copied nor moved");
static_assert(std::is_same<decltype(&Chunk::getComputeTimeEst), double (*)(const qio::QSOFile &, int, int)>::value,
              "static double getComputeTimeEst(const qio::QSOFile&, int, int)");
static_assert(std::is_same<decltype(&Chunk::realSize), int (Chunk::*)() const>::value, "int realSize() const");
static_assert(std::is_same<decltype(&Chunk::size), int (Chunk::*)() const>::value, "int size() const");
static_assert(std::is_same<decltype(&Chunk::setInvertCovarianceMatrix), void (Chunk::*)(const double *)>::value,
              "void setInvertCovarianceMatrix(const double*)");
static_assert(std::is_same<decltype(&Chunk::computePSbeforeFvector), void (Chunk::*)()>::value,
              "void computePSbeforeFvector()");
static_assert(std::is_same<decltype(&Chunk::oneQSOiteration),
                           void (Chunk::*)(const double *, dbt_t &, double *)>::value,
              "void oneQSOiteration(const double*, vector<unique_ptr<double[]>>&, double*)");
static_assert(std::is_same<decltype(&Chunk::fprintfMatrices), void (Chunk::*)(const char *)>::value,
              "void fprintfMatrices(const char*)");
static_assert(std::is_same<decltype(&Chunk::getMinMemUsage), double (Chunk::*)()>::value, "double getMinMemUsage()");
static_assert(std::is_same<decltype(&Chunk::releaseFile), void (Chunk::*)()>::value, "void releaseFile()");
static_assert(std::is_same<decltype(Chunk::qFile), std::unique_ptr<qio::QSOFile>>::value, "qFile");
static_assert(std::is_same<decltype(Chunk::fisher_index_start), int>::value &&
                  std::is_same<decltype(Chunk::N_Q_MATRICES), int>::value &&
                  std::is_same<decltype(Chunk::ZBIN), int>::value && std::is_same<decltype(Chunk::ZBIN_LOW), int>::value &&
                  std::is_same<decltype(Chunk::ZBIN_UPP), int>::value && std::is_same<decltype(Chunk::KBIN_UPP), int>::value,
              "bin bookkeeping members");
static_assert(std::is_same<decltype(Chunk::dbt_estimate_before_fisher_vector), dbt_t>::value,
              "dbt_estimate_before_fisher_vector");
static_assert(std::is_same<decltype(Chunk::fisher_matrix), std::unique_ptr<double[]>>::value, "fisher_matrix");

// --- OneQSOEstimate (core/one_qso_estimate.hpp:16-61)
static_assert(std::is_constructible<OneQSOEstimate, const std::string &>::value, "OneQSOEstimate(const string&)");
static_assert(std::is_default_constructible<OneQSOEstimate>::value, "OneQSOEstimate()");
static_assert(std::is_move_constructible<OneQSOEstimate>::value && !std::is_copy_constructible<OneQSOEstimate>::value,
              "moved, not copied");
static_assert(std::is_same<decltype(&OneQSOEstimate::decideIndices), std::vector<int> (*)(int, double *)>::value,
              "static vector<int> decideIndices(int, double*)");
static_assert(std::is_same<decltype(&OneQSOEstimate::getComputeTimeEst), double (*)(std::string, int &, long &)>::value,
              "static double getComputeTimeEst(string, int&, long&)");
static_assert(std::is_same<decltype(&OneQSOEstimate::oneQSOiteration),
                           void (OneQSOEstimate::*)(const double *, dbt_t &, double *)>::value,
              "void oneQSOiteration(const double*, vector<unique_ptr<double[]>>&, double*)");
static_assert(std::is_same<decltype(&OneQSOEstimate::collapseBootstrap), void (OneQSOEstimate::*)()>::value,
              "void collapseBootstrap()");
static_assert(std::is_same<decltype(&OneQSOEstimate::addBoot), void (OneQSOEstimate::*)(int, double *, double *)>::value,
              "void addBoot(int, double*, double*)");
static_assert(std::is_same<decltype(&OneQSOEstimate::addBootPowerOnly),
                           void (OneQSOEstimate::*)(int, double *, double *)>::value,
              "void addBootPowerOnly(int, double*, double*)");
static_assert(std::is_same<decltype(OneQSOEstimate::istart), int>::value &&
                  std::is_same<decltype(OneQSOEstimate::ndim), int>::value,
              "istart, ndim");
static_assert(std::is_same<decltype(OneQSOEstimate::fisher_matrix), std::unique_ptr<double[]>>::value &&
                  std::is_same<decltype(OneQSOEstimate::theta_vector), std::unique_ptr<double[]>>::value,
              "fisher_matrix, theta_vector");
static_assert(std::is_same<decltype(OneQSOEstimate::chunks), std::vector<std::unique_ptr<Chunk>>>::value, "chunks");

namespace {
struct Config {
    lyspeq_amd::Binning b;
    bool fast_bootstrap = true, picca = false, rmat = false;
    std::pair<int, double> first_table{-1, 0.0};
    std::string save_chunks, save_boot;
    int nboots = 0;
    std::vector<int> weights;
    std::vector<std::string> files;
};

Config read_config(const char *path) {
    std::ifstream in(path);
    if (!in) {
        std::fprintf(stderr, "cannot open %s\n", path);
        std::exit(2);
    }
    Config c;
    c.b.kband_edges.clear();
    c.b.zbin_centers.clear();
    std::string key;
    while (in >> key) {
        if (key == "kedges" || key == "zcenters") {
            int n;
            in >> n;
            auto &v = key == "kedges" ? c.b.kband_edges : c.b.zbin_centers;
            v.resize(n);
            for (double &x : v) in >> x;
        } else if (key == "zwidth") {
            in >> c.b.z_bin_width;
        } else if (key == "method") {
            in >> c.b.z_binning_method;
        } else if (key == "logl_order") {
            in >> c.b.cont_logl_order;
        } else if (key == "lam_order") {
            in >> c.b.cont_lam_order;
        } else if (key == "max_forest") {
            in >> c.b.max_forest_length_v;
        } else if (key == "nchunks") {
            in >> c.b.number_of_chunks;
        } else if (key == "max_pixel") {
            in >> c.b.max_pixel_length_v;
        } else if (key == "fast_bootstrap") {
            int v;
            in >> v;
            c.fast_bootstrap = v > 0;
        } else if (key == "nboots") {
            in >> c.nboots;
        } else if (key == "weights") {
            int n;
            in >> n;
            c.weights.resize(n);
            for (int &w : c.weights) in >> w;
        } else if (key == "format") {
            std::string v;
            in >> v;
            c.picca = v == "picca";
        } else if (key == "rmat") {
            int v;
            in >> v;
            c.rmat = v > 0;
        } else if (key == "deconv") {
            in >> c.b.resomat_deconvolution_m;
        } else if (key == "osamp") {
            in >> c.b.oversampling_factor;
        } else if (key == "growth") {
            c.b.redshift_growth_on = true;
            in >> c.b.fiducial_redshift_power;
        } else if (key == "first_table") {
            in >> c.first_table.first >> c.first_table.second;
        } else if (key == "save_chunks") {
            in >> c.save_chunks;
        } else if (key == "save_boot") {
            in >> c.save_boot;
        } else if (key == "file") {
            std::string f;
            in >> f;
            c.files.push_back(f);
        } else {
            std::fprintf(stderr, "unknown key %s\n", key.c_str());
            std::exit(2);
        }
    }
    return c;
}

int cost(const Config &c) {
    for (const auto &f : c.files) {
        int zbin = -2;
        long tid = -2;
        const double t = OneQSOEstimate::getComputeTimeEst(f, zbin, tid);
        std::printf("file %s %.17g %d %ld\n", f.c_str(), t, zbin, tid);
        // the chunks of the same reading, one by one
        qio::QSOFile q(f, specifics::INPUT_QSO_FILE);
        q.readParameters();
        q.readData();
        q.cutBoundary(bins::Z_LOWER_EDGE, bins::Z_UPPER_EDGE);
        const std::vector<int> idx = OneQSOEstimate::decideIndices(q.size(), q.wave());
        for (size_t k = 0; k + 1 < idx.size(); ++k)
            std::printf("chunk %d %d %.17g\n", idx[k], idx[k + 1], Chunk::getComputeTimeEst(q, idx[k], idx[k + 1]));
    }
    return 0;
}

void write_vec(FILE *f, const char *tag, const double *v, size_t n) {
    std::fprintf(f, "%s %zu", tag, n);
    for (size_t i = 0; i < n; ++i) std::fprintf(f, " %.17g", v[i]);
    std::fprintf(f, "\n");
}

int run(const Config &c, const char *dir, const char *base, const char *prec, const char *out_path) {
    lyspeq_amd::dropin::Options o;
    const std::string p(prec);
    o.precision = p == "fp64" ? QML_PREC_FP64 : p == "fp32" ? QML_PREC_FP32 : p == "bf16x3" ? QML_PREC_BF16X3
                                                                                              : QML_PREC_BF16X4;
    lyspeq_amd::dropin::setup(o, lyspeq_amd::dropin::sqFileTables(dir, base, c.first_table));
    const int T = bins::TOTAL_KZ_BINS;
    std::vector<std::unique_ptr<OneQSOEstimate>> queue;
    for (const auto &f : c.files) queue.push_back(std::make_unique<OneQSOEstimate>(f));
    std::vector<double> theta(T, 0.0), fisher((size_t)T * T, 0.0);
    dbt_t dbt;
    for (int k = 0; k < 3; ++k) {
        dbt.push_back(std::make_unique<double[]>(T));
        std::fill_n(dbt.back().get(), T, 0.0);
    }
    for (auto &q : queue) q->oneQSOiteration(theta.data(), dbt, fisher.data());
    lyspeq_amd::dropin::flush();   // the driver's one new call
    FILE *out = std::fopen(out_path, "w");
    if (!out) return 2;
    write_vec(out, "fisher", fisher.data(), fisher.size());
    for (int k = 0; k < 3; ++k) write_vec(out, k == 0 ? "d" : k == 1 ? "b" : "t", dbt[k].get(), T);
    for (size_t s = 0; s < queue.size(); ++s)
        for (const auto &ch : queue[s]->chunks) {
            std::fprintf(out, "chunk %zu %d %d %d %d %d\n", s, ch->fisher_index_start, ch->N_Q_MATRICES, ch->size(),
                         ch->realSize(), ch->KBIN_UPP);
            write_vec(out, "cF", ch->fisher_matrix.get(), (size_t)ch->N_Q_MATRICES * ch->N_Q_MATRICES);
            for (int k = 0; k < 3; ++k) write_vec(out, "cdbt", ch->dbt_estimate_before_fisher_vector[k].get(), ch->N_Q_MATRICES);
        }
    if (!c.save_chunks.empty()) {   // _saveChunkResults (core/quadratic_estimate.cpp:33-55)
        ioh::BootstrapChunksFile bfile(c.save_chunks, 0);
        for (auto &q : queue)
            for (auto &ch : q->chunks)
                bfile.writeChunk(ch->dbt_estimate_before_fisher_vector[0].get(),
                                 ch->dbt_estimate_before_fisher_vector[1].get(),
                                 ch->dbt_estimate_before_fisher_vector[2].get(), ch->fisher_matrix.get(),
                                 ch->N_Q_MATRICES, ch->fisher_index_start, ch->qFile->id, ch->qFile->z_qso,
                                 ch->qFile->ra, ch->qFile->dec);
    }
    // the bootstrapper's use (core/bootstrapper.hpp): collapse every spectrum,
    // then the fast form's rank-one updates or the complete form's sums
    std::vector<OneQSOEstimate *> live;
    for (auto &q : queue)
        if (!q->chunks.empty()) live.push_back(q.get());
    for (auto *q : live) q->collapseBootstrap();
    if (c.nboots > 0 && (int)c.weights.size() == c.nboots * (int)live.size()) {
        std::vector<double> power((size_t)c.nboots * T, 0.0);
        std::vector<double> pcoeff(c.nboots);
        for (size_t s = 0; s < live.size(); ++s) {
            for (int b = 0; b < c.nboots; ++b) pcoeff[b] = c.weights[s * c.nboots + b];
            live[s]->addBootPowerOnly(c.nboots, pcoeff.data(), power.data());
        }
        write_vec(out, "bootpower", power.data(), power.size());
        if (!c.save_boot.empty())
            ioh::saveBootstrapRealizations(c.save_boot, power.data(), fisher.data(), c.nboots,
                                           bins::NUMBER_OF_K_BANDS, bins::NUMBER_OF_Z_BINS, specifics::FAST_BOOTSTRAP,
                                           "drop-in check");
        if (!specifics::FAST_BOOTSTRAP) {
            // first realisation's complete-form sums
            std::vector<double> tp(T, 0.0), tf((size_t)T * T, 0.0);
            for (size_t s = 0; s < live.size(); ++s) live[s]->addBoot(c.weights[s * c.nboots], tp.data(), tf.data());
            write_vec(out, "boot0power", tp.data(), tp.size());
            write_vec(out, "boot0fisher", tf.data(), tf.size());
        }
    }
    for (auto *q : live) std::fprintf(out, "collapsed %d %d %zu\n", q->istart, q->ndim, q->chunks.size());
    std::fclose(out);
    queue.clear();
    lyspeq_amd::dropin::reset();
    return 0;
}
}  // namespace

// the bootstrap files from fixed arrays: no bins, no engine
int bootio(const std::string &base, const char *out_path) {
    FILE *out = std::fopen(out_path, "w");
    if (!out) return 2;
    {
        ioh::BootstrapChunksFile bfile(base, 3);
        const int dims[3] = {4, 1, 6};
        for (int k = 0; k < 3; ++k) {
            const int n = dims[k];
            std::vector<double> p(n), nk(n), tk(n), F((size_t)n * n);
            for (int i = 0; i < n; ++i) {
                p[i] = std::sin(1.0 + i + 10 * k);
                nk[i] = 1.0 / (3.0 + i + k);
                tk[i] = -std::exp(-0.5 * i) * (k + 1);
            }
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) F[(size_t)i * n + j] = std::cos(0.3 * i - 0.7 * j + k) * 1e-3;
            const long id = 39627000000000000L + k;
            bfile.writeChunk(p.data(), nk.data(), tk.data(), F.data(), n, 7 * k, id, 2.1 + 0.35 * k, 0.1 * k,
                             -0.2 * k);
            std::fprintf(out, "chunk %d %d %ld\n", n, 7 * k, id);
            const double par[3] = {2.1 + 0.35 * k, 0.1 * k, -0.2 * k};
            write_vec(out, "par", par, 3);
            write_vec(out, "p", p.data(), n);
            write_vec(out, "n", nk.data(), n);
            write_vec(out, "t", tk.data(), n);
            write_vec(out, "F", F.data(), F.size());
        }
    }
    const unsigned nboots = 5;
    const int nk = 3, nz = 2, T = nk * nz;
    std::vector<double> pw((size_t)nboots * T), inv((size_t)T * T);
    for (size_t i = 0; i < pw.size(); ++i) pw[i] = std::sqrt(2.0 + i) * (i % 2 ? -1 : 1);
    for (size_t i = 0; i < inv.size(); ++i) inv[i] = 1.0 / (1.0 + i);
    const std::string name = ioh::saveBootstrapRealizations(base, pw.data(), inv.data(), nboots, nk, nz, true,
                                                            "fixed arrays");
    std::fprintf(out, "realizations %s\n", name.c_str() + 1);
    write_vec(out, "powers", pw.data(), pw.size());
    write_vec(out, "invf", inv.data(), inv.size());
    std::fclose(out);
    std::unique_ptr<double[]> rp, ri;
    unsigned rn = 0;
    int rk = 0, rz = 0;
    bool rf = false;
    ioh::readBootstrapRealizations(name.substr(1), rp, ri, rn, rk, rz, rf);
    if (rn != nboots || rk != nk || rz != nz || !rf) return 1;
    for (size_t i = 0; i < pw.size(); ++i)
        if (rp[i] != pw[i]) return 1;
    for (size_t i = 0; i < inv.size(); ++i)
        if (ri[i] != inv[i]) return 1;
    return 0;
}

// the reader alone: no bins, no engine
int read_one(const std::string &format, const char *fname, const char *out_path) {
    try {
        qio::QSOFile q(fname, format == "picca" ? qio::Picca : qio::Binary);
        q.readParameters();
        q.readData();
        FILE *out = std::fopen(out_path, "w");
        if (!out) return 2;
        std::fprintf(out, "id %ld\nR_fwhm %d\nexpid %d\nnight %d\nfiber %d\npetal %d\n", q.id, q.R_fwhm, q.expid,
                     q.night, q.fiber, q.petal);
        const double par[] = {q.z_qso, q.ra, q.dec, q.R_kms, q.snr, q.dv_kms, q.dlambda};
        write_vec(out, "params", par, 7);
        write_vec(out, "wave", q.wave(), q.size());
        write_vec(out, "delta", q.delta(), q.size());
        write_vec(out, "ivar", q.ivar(), q.size());
        if (q.pending_rmat) {
            q.readAllocResolutionMatrix();
            std::fprintf(out, "rmat_nelem %d\n", q.Rmat->getNElemPerRow());
            write_vec(out, "rmat", q.Rmat->matrix(), q.Rmat->getSize());
        }
        std::fclose(out);
    } catch (std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}

// --- glmemory (core/chunk_estimate.hpp:29-43)
static_assert(std::is_same<decltype(&glmemory::setMaxSizes), void (*)(int, int, int, bool)>::value,
              "void setMaxSizes(int, int, int, bool)");
static_assert(std::is_same<decltype(&glmemory::allocMemory), void (*)()>::value, "void allocMemory()");
static_assert(std::is_same<decltype(&glmemory::dealloc), void (*)()>::value, "void dealloc()");
static_assert(std::is_same<decltype(&glmemory::updateMemUsed), void (*)(double)>::value, "void updateMemUsed(double)");
static_assert(std::is_same<decltype(glmemory::max_size_2), int>::value, "int max_size_2");
static_assert(std::is_same<decltype(glmemory::covariance_matrix), std::unique_ptr<double[]>>::value &&
                  std::is_same<decltype(glmemory::temp_vector), std::unique_ptr<double[]>>::value &&
                  std::is_same<decltype(glmemory::stored_sfid), std::unique_ptr<double[]>>::value,
              "glmemory buffers");
static_assert(std::is_same<decltype(glmemory::stored_ikz_qi), std::vector<std::unique_ptr<double[]>>>::value &&
                  std::is_same<decltype(glmemory::temp_matrices), std::vector<std::unique_ptr<double[]>>>::value,
              "glmemory matrix lists");

// The friend class of tests/testSQCMatrices.cpp:24-36, with its constructor;
// the checks write the matrices out for comparison.
class TestOneQSOEstimate : public OneQSOEstimate {
public:
    TestOneQSOEstimate(const std::string &fname_qso) : OneQSOEstimate(fname_qso) {
        glmemory::allocMemory();
        chunks[0]->_initMatrices();
        chunks[0]->_setVZMatrices();
    }

    // the protected surface of core/chunk_estimate.hpp:48-82, as the friend test
    // and cross/exposure.hpp:19-109 use it
    static void surface() {
        static_assert(std::is_same<decltype(Chunk::DATA_SIZE_2), int>::value, "DATA_SIZE_2");
        static_assert(std::is_same<decltype(Chunk::covariance_matrix), double *>::value &&
                          std::is_same<decltype(Chunk::inverse_covariance_matrix), double *>::value &&
                          std::is_same<decltype(Chunk::_vmatrix), double *>::value &&
                          std::is_same<decltype(Chunk::_zmatrix), double *>::value &&
                          std::is_same<decltype(Chunk::_matrix_lambda), double *>::value,
                      "matrix pointers");
        static_assert(std::is_same<decltype(Chunk::temp_matrix), std::vector<double *>>::value, "temp_matrix");
        static_assert(std::is_same<decltype(Chunk::stored_ikz_qi), std::vector<std::pair<int, double *>>>::value,
                      "stored_ikz_qi");
        static_assert(std::is_same<decltype(&Chunk::_copyQSOFile), void (Chunk::*)(const qio::QSOFile &, int, int)>::value,
                      "_copyQSOFile");
        static_assert(std::is_same<decltype(&Chunk::_setNQandFisherIndex), void (Chunk::*)()>::value &&
                          std::is_same<decltype(&Chunk::_initMatrices), void (Chunk::*)()>::value &&
                          std::is_same<decltype(&Chunk::_setVZMatrices), void (Chunk::*)()>::value &&
                          std::is_same<decltype(&Chunk::_applyRedshiftInterp), void (Chunk::*)()>::value &&
                          std::is_same<decltype(&Chunk::_scaleDerivativesWithRedshiftGrowth), void (Chunk::*)()>::value &&
                          std::is_same<decltype(&Chunk::_addMarginalizations), void (Chunk::*)()>::value,
                      "builders without arguments");
        static_assert(std::is_same<decltype(&Chunk::_setFiducialSignalMatrix), void (Chunk::*)(double *)>::value &&
                          std::is_same<decltype(&Chunk::_setQiMatrix), void (Chunk::*)(double *, int)>::value,
                      "_setFiducialSignalMatrix(double*), _setQiMatrix(double*, int)");
    }

    static bool dump(const std::string &path, const double *a, size_t n) {
        FILE *f = std::fopen(path.c_str(), "wb");
        if (!f) return false;
        const bool ok = std::fwrite(a, sizeof(double), n, f) == n;
        std::fclose(f);
        return ok;
    }

    // tests/testSQCMatrices.cpp:38-65
    bool test_setFiducialSignalMatrix(const std::string &dir) {
        Chunk &c = *chunks[0];
        c._setFiducialSignalMatrix(c.covariance_matrix);
        return dump(dir + "/signal.bin", c.covariance_matrix, c.DATA_SIZE_2);
    }

    // tests/testSQCMatrices.cpp:67-95
    bool test_setQiMatrix(const std::string &dir) {
        Chunk &c = *chunks[0];
        c._setQiMatrix(c.stored_ikz_qi[0].second, 0);
        c._applyRedshiftInterp();
        return dump(dir + "/q0.bin", c.stored_ikz_qi[0].second, c.DATA_SIZE_2);
    }

    bool vz(const std::string &dir) {
        Chunk &c = *chunks[0];
        const size_t n2 = (size_t)c._matrix_n * c._matrix_n;
        return dump(dir + "/v.bin", c._vmatrix, n2) && dump(dir + "/z.bin", c._zmatrix, n2);
    }

    // every stored Q as oneQSOiteration builds them (core/chunk_estimate.cpp:860-882)
    bool stored(const std::string &dir) {
        Chunk &c = *chunks[0];
        for (auto &q : c.stored_ikz_qi) c._setQiMatrix(q.second, q.first);
        c._applyRedshiftInterp();
        if (specifics::REDSHIFT_GROWTH_ON) c._scaleDerivativesWithRedshiftGrowth();
        std::vector<double> all;
        for (auto &q : c.stored_ikz_qi) all.insert(all.end(), q.second, q.second + c.DATA_SIZE_2);
        return dump(dir + "/stored.bin", all.data(), all.size());
    }

    // _addMarginalizations on S + I (covariance_matrix holds S)
    bool marginalize(const std::string &dir) {
        Chunk &c = *chunks[0];
        c._setFiducialSignalMatrix(c.covariance_matrix);
        for (int i = 0; i < c.size(); ++i) c.covariance_matrix[(size_t)i * (c.size() + 1)] += 1.0;
        if (!dump(dir + "/cinv_in.bin", c.covariance_matrix, c.DATA_SIZE_2)) return false;
        c.inverse_covariance_matrix = c.covariance_matrix;
        c._addMarginalizations();
        return dump(dir + "/cinv_marg.bin", c.inverse_covariance_matrix, c.DATA_SIZE_2);
    }
};

// A Chunk subclass that builds itself as cross/exposure.hpp:23-31,38-45 does:
// the protected default constructor, _copyQSOFile and _setNQandFisherIndex,
// then its own covariance buffer after _initMatrices.
class TestExposure : public Chunk {
    std::unique_ptr<double[]> local_cov_mat;

public:
    TestExposure(const qio::QSOFile &qmaster, int i1, int i2) : Chunk() {
        _copyQSOFile(qmaster, i1, i2);
        _setNQandFisherIndex();
        glmemory::setMaxSizes(size(), size(), 2 * N_Q_MATRICES, false);
        N_Q_MATRICES = 0;
    }
    void initMatrices() {
        _initMatrices();
        local_cov_mat = std::make_unique<double[]>(DATA_SIZE_2);
        covariance_matrix = local_cov_mat.get();
    }
    bool check(const std::string &dir) {
        _setVZMatrices();
        _setFiducialSignalMatrix(covariance_matrix);
        if (!TestOneQSOEstimate::dump(dir + "/ex_signal.bin", covariance_matrix, DATA_SIZE_2)) return false;
        for (int i = 0; i < size(); ++i) covariance_matrix[(size_t)i * (size() + 1)] += 1.0;
        if (!TestOneQSOEstimate::dump(dir + "/ex_cinv_in.bin", covariance_matrix, DATA_SIZE_2)) return false;
        inverse_covariance_matrix = covariance_matrix;
        _addMarginalizations();
        return TestOneQSOEstimate::dump(dir + "/ex_cinv_marg.bin", inverse_covariance_matrix, DATA_SIZE_2);
    }
};

// OneQSOEstimate's reading of a file (its protected _readQsoFile)
struct FileReader : OneQSOEstimate {
    std::unique_ptr<qio::QSOFile> read(const std::string &f) { return _readQsoFile(f); }
};

int exposure_check(const Config &c, const char *dir, const char *base, const std::string &out_dir) {
    lyspeq_amd::dropin::Options o;
    lyspeq_amd::dropin::setup(o, lyspeq_amd::dropin::sqFileTables(dir, base, c.first_table));
    if (c.files.empty()) return 2;
    try {
        FileReader fr;
        auto q = fr.read(c.files[0]);
        if (!q) return 3;
        const std::vector<int> idx = OneQSOEstimate::decideIndices(q->size(), q->wave());
        TestExposure ex(*q, idx[0], idx[1]);
        glmemory::allocMemory();
        ex.initMatrices();
        const bool ok = ex.check(out_dir);
        glmemory::dealloc();
        return ok ? 0 : 3;
    } catch (std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 4;
    }
}

int friend_check(const Config &c, const char *dir, const char *base, const std::string &out_dir) {
    TestOneQSOEstimate::surface();
    lyspeq_amd::dropin::Options o;
    o.fiducial_redshift_power = c.b.fiducial_redshift_power;
    lyspeq_amd::dropin::setup(o, lyspeq_amd::dropin::sqFileTables(dir, base, c.first_table));
    if (c.files.empty()) return 2;
    try {
        TestOneQSOEstimate t(c.files[0]);
        const bool ok = t.vz(out_dir) && t.test_setFiducialSignalMatrix(out_dir) && t.test_setQiMatrix(out_dir) &&
                        t.stored(out_dir) && t.marginalize(out_dir);
        glmemory::dealloc();
        return ok ? 0 : 3;
    } catch (std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 4;
    }
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: dropin_check cost|run <config> ... | read binary|picca <file> <out>\n");
        return 2;
    }
    if (std::string(argv[1]) == "read" && argc == 5) return read_one(argv[2], argv[3], argv[4]);
    if (std::string(argv[1]) == "bootio" && argc == 4) return bootio(argv[2], argv[3]);
    const Config c = read_config(argv[2]);
    lyspeq_amd::dropin::setGlobals(c.b, c.rmat, c.fast_bootstrap);
    if (c.picca) specifics::INPUT_QSO_FILE = qio::Picca;
    const std::string mode(argv[1]);
    if (mode == "cost") return cost(c);
    if (mode == "run" && argc == 7) return run(c, argv[3], argv[4], argv[5], argv[6]);
    if (mode == "friend" && argc == 6) return friend_check(c, argv[3], argv[4], argv[5]);
    if (mode == "exposure" && argc == 6) return exposure_check(c, argv[3], argv[4], argv[5]);
    std::fprintf(stderr, "bad arguments\n");
    return 2;
}
