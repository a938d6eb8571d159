// C++ host layer over the QML engine's C ABI (include/lyspeq_qml.h).
//
// Mirrors the surface the reference's driver uses for the per-chunk QML
// step, so that a caller written against core/chunk_estimate.hpp and
// core/one_qso_estimate.hpp keeps its shape:
//
//   Chunk            core/chunk_estimate.hpp:46-119  (constructor bookkeeping,
//                    fisher_index_start, N_Q_MATRICES, ZBIN*, KBIN_UPP,
//                    dbt_estimate_before_fisher_vector, fisher_matrix,
//                    size(), realSize(), oneQSOiteration)
//   OneQSOEstimate   core/one_qso_estimate.hpp:16-61  (decideIndices, chunks,
//                    oneQSOiteration)
//
// The difference is when the work happens: Chunk::oneQSOiteration enqueues
// the chunk, and Engine::flush(), called once after the driver's chunk loop
// (before the 0.5 scaling of the Fisher sum, core/quadratic_estimate.cpp:456),
// runs the whole batch on the GPU and adds every chunk's F and d, b, t into
// the sums the chunks were enqueued with.  After flush() each Chunk's public
// result arrays hold its own F and d, b, t, as after the reference call.
//
// Errors: constructor failures throw std::runtime_error (as the reference
// constructor does, so OneQSOEstimate skips the chunk); engine failures throw
// QMLError, which carries the C-ABI status code.  A chunk whose covariance
// is not positive definite keeps zero results and is left out of the sums,
// as the reference's caught LAPACK failure (core/chunk_estimate.cpp:905-913).
#pragma once

#include "lyspeq_qml.h"

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace lyspeq_amd {

class QMLError : public std::runtime_error {
public:
    QMLError(int code, const std::string &what) : std::runtime_error(what), code(code) {}
    int code;
};

// The part of qio::QSOFile (io/qso_file.hpp) the chunk step reads, after
// the reader's preprocessing (flux to delta, boundary cut): wavelengths in
// Angstrom, delta, inverse variance, the resolution R_fwhm and the pixel
// width dv in km/s (table selection rounds it to 0.1 km/s).
struct QSOSpectrum {
    std::vector<double> wave, delta, ivar;
    int R_fwhm = 0;
    std::string fname;
    // resolution matrix as read from the file (qio::QSOFile::Rmat before
    // preprocessing), diagonal storage rmat_ndiags x size; empty without one
    std::vector<double> rmat;
    int rmat_ndiags = 0;
    int size() const { return (int)wave.size(); }
};

// A whole spectrum's resolution matrix after the reader's preprocessing
// (core/one_qso_estimate.cpp:74-89): QML_RMAT_DIA (nelem x size) or
// QML_RMAT_OVERSAMPLED (size x nelem rows).
struct ResolutionRows {
    int kind = 0;
    int nelem = 0;
    int oversampling = 1;
    std::vector<double> values;
};

// Binning and the specifics the chunk step depends on (bins:: and
// specifics:: globals of core/global_numbers.cpp).
struct Binning {
    std::vector<double> kband_edges;     // NUMBER_OF_K_BANDS + 1
    std::vector<double> zbin_centers;    // NUMBER_OF_Z_BINS
    double z_bin_width = 0.2;
    int z_binning_method = 1;            // 1 triangular, 0 top-hat
    int cont_logl_order = 1, cont_lam_order = -1;
    bool turn_off_sfid = false;
    double max_forest_length_v = 3.5e4;  // MaximumForestLength (before division)
    int number_of_chunks = 1;            // DynamicChunkNumber
    double max_pixel_length_v = 120.0;   // MaximumMeanPixelSize
    int oversampling_factor = -1;        // OversampleRmat: > 0 oversamples diagonal matrices
    double resomat_deconvolution_m = -1; // ResoMatrixDeconvolutionM: > 0 deconvolves them first
    bool redshift_growth_on = false;     // RedshiftGrowthOn > 0
    double fiducial_redshift_power = 3.591244;  // FiducialRedshiftPower (the growth exponent B)
    int nk() const { return (int)kband_edges.size() - 1; }
    int nz() const { return (int)zbin_centers.size(); }
    int totalKzBins() const { return nk() * nz(); }
    double maxForestLengthPerChunk() const { return max_forest_length_v / std::max(1, number_of_chunks); }
};

class Chunk;
class OneQSOEstimate;

// One device context: tables, chunks, and the batched iteration.
class Engine {
public:
    Engine(const Binning &bins, int precision = QML_PREC_FP64, int device = 0, double scratch_gb = 0.0,
           int max_batch_chunks = 0);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // readSQforR counterpart: derivative table nk x nv (v contiguous per
    // band), signal table nv x nz (z contiguous), both on the reference grid.
    // Returns RES_INDEX.
    int addTable(int R_fwhm, double dv_kms, int nv, double v_length, int nz, const double *deriv,
                 const double *signal);

    // Runs the batch for every chunk enqueued since the last flush and adds
    // its F and d, b, t into the sums they were enqueued with.  The first
    // call commits the engine (no chunk can be added after it).
    void flush();

    // Graph replay of the iterations (qml_set_graph): each flush after the
    // first replays one captured launch sequence.  Results are unchanged.
    void setGraph(bool on);
    // Commits the chunks constructed so far and picks the batch schedule by
    // timing iterations at ps_estimate under each cap on chunks per batch
    // (qml_autotune).  No chunk can be added afterwards.  Returns the cap.
    int autotune(const double *ps_estimate, const std::vector<int> &caps, int reps = 1,
                 double *best_ms = nullptr);

    // PoissonBootstrapper::run with FastBootstrap (qml_bootstrap) over the
    // driver's queue after the last flush: the spectra in queue order, each
    // with the chunks it still holds.  mean is TOTAL_KZ_BINS values, cov
    // TOTAL_KZ_BINS^2; realizations (nboots x TOTAL_KZ_BINS) may be null.
    void bootstrap(const std::vector<OneQSOEstimate *> &queue, int nboots, unsigned long long seed, double *mean,
                   double *cov, double *realizations = nullptr);
    // The same with FastBootstrap off (qml_bootstrap_complete): a Cholesky
    // solve per realisation; returns the number of valid realisations, whose
    // solutions fill the first rows of realizations (may be null).
    int bootstrapComplete(const std::vector<OneQSOEstimate *> &queue, int nboots, unsigned long long seed,
                          double *mean, double *cov, double *realizations = nullptr);
    // PoissonBootstrapper::run() on saved realisations (_medianBootstrap,
    // qml_bootstrap_median): realizations nboots x TOTAL_KZ_BINS, one per row;
    // median TOTAL_KZ_BINS values, mad_cov TOTAL_KZ_BINS^2.  Needs no chunks.
    void bootstrapMedian(int nboots, const double *realizations, double *median, double *mad_cov);

    // One matrix of the reference Chunk's protected builders for chunk c
    // (qml_chunk_matrix; which is a QML_CMAT_* value): writes n x n values to
    // out unless it is null and returns n.  Commits the engine first.
    int chunkMatrix(const Chunk &c, int which, int arg, double *out);
    // Chunk::_addMarginalizations on c's size^2 cinv, in place
    // (qml_chunk_marginalize).  Commits the engine first.
    void marginalize(const Chunk &c, double *cinv);

    int totalKzBins() const { return bins_.totalKzBins(); }
    const Binning &binning() const { return bins_; }
    qml_ctx *ctx() { return ctx_; }
    void timings(qml_timings *t);

private:
    friend class Chunk;
    int addChunk(const QSOSpectrum &q, int i1, int i2, qml_chunk_info *info, const ResolutionRows *rr);
    void enqueue(Chunk *c, const double *ps, std::vector<std::unique_ptr<double[]>> *dbt, double *fisher);
    void forget(Chunk *c);
    void check(int code) const;
    void commit();

    Binning bins_;
    qml_ctx *ctx_ = nullptr;
    bool committed_ = false;
    std::vector<Chunk *> chunks_;          // by engine chunk id; null once destroyed
    std::vector<int> active_;              // accumulation flag last sent to the device
    std::vector<Chunk *> queue_;
    std::vector<int> queue_sums_;          // 1 if the queued chunk enters the sums
    const double *ps_ = nullptr;
    std::vector<std::unique_ptr<double[]>> *dbt_ = nullptr;
    double *fisher_ = nullptr;
};

class Chunk {
public:
    // Chunk(const qio::QSOFile&, int i1, int i2): pixels [i1, i2) of q.  A
    // spectrum with a resolution matrix uses rr, its preprocessed matrix
    // (prepareResolution); without rr the chunk preprocesses q's itself.
    Chunk(Engine &eng, const QSOSpectrum &q, int i1, int i2, const ResolutionRows *rr = nullptr);
    ~Chunk();
    Chunk(const Chunk &) = delete;
    Chunk(Chunk &&) = delete;

    int fisher_index_start, N_Q_MATRICES;
    int ZBIN, ZBIN_LOW, ZBIN_UPP, KBIN_UPP;
    // 3 x N_Q_MATRICES: d, b, t; N_Q_MATRICES^2 Fisher (upper, unhalved)
    std::vector<std::unique_ptr<double[]>> dbt_estimate_before_fisher_vector;
    std::unique_ptr<double[]> fisher_matrix;

    int size() const { return info_.size; }
    int realSize() const { return info_.real_size; }
    const qml_chunk_info &info() const { return info_; }
    int engineId() const { return id_; }
    // QML_OK, or nonzero if the last flush found the covariance not positive definite
    int lastStatus() const { return status_; }

    // Enqueues this chunk for the next Engine::flush(): ps_estimate is the
    // full TOTAL_KZ_BINS vector, the sums are TOTAL_KZ_BINS^2 (upper
    // triangle) and 3 x TOTAL_KZ_BINS.
    void oneQSOiteration(const double *ps_estimate, std::vector<std::unique_ptr<double[]>> &dbt_sum_vector,
                         double *fisher_sum);
    // Enqueues this chunk for its own results only (it enters no sums).
    void enqueueResultsOnly(const double *ps_estimate);

    // Where flush() writes this chunk's results: fisher_matrix and
    // dbt_estimate_before_fisher_vector unless redirected here
    // (N_Q_MATRICES^2 and 3 arrays of N_Q_MATRICES).
    double *result_fisher = nullptr;
    double *result_dbt[3] = {nullptr, nullptr, nullptr};

private:
    friend class Engine;
    Engine *eng_;
    int id_ = -1;
    int status_ = 0;
    qml_chunk_info info_{};
};

class OneQSOEstimate {
public:
    // Splits q by decideIndices and constructs a Chunk for each piece; a
    // piece whose constructor throws is skipped (core/one_qso_estimate.cpp:97-126).
    OneQSOEstimate(Engine &eng, const QSOSpectrum &q);

    static std::vector<int> decideIndices(int size, const double *wave, const Binning &bins);

    // The reader's preprocessing of q's resolution matrix under the bins'
    // specifics: deconvolution if resomat_deconvolution_m > 0, then
    // oversampling if oversampling_factor > 0 (qml_rmat_deconvolve,
    // qml_rmat_oversample).  kind 0 if q has no matrix.
    static ResolutionRows prepareResolution(const QSOSpectrum &q, const Binning &bins);

    std::vector<std::unique_ptr<Chunk>> chunks;

    void oneQSOiteration(const double *ps_estimate, std::vector<std::unique_ptr<double[]>> &dbt_sum_vector,
                         double *fisher_sum);
};

}  // namespace lyspeq_amd
