// Drop-in replacement of the reference's core/chunk_estimate.hpp: class
// Chunk with the surface of core/chunk_estimate.hpp:46-119, backed by the
// MI355X QML engine (lyspeq_amd::Engine over include/lyspeq_qml.h).
//
// The difference is when the work happens.  oneQSOiteration enqueues the
// chunk (and its fit values and sums); lyspeq_amd::dropin::flush(), called
// once after the driver's chunk loop, runs every enqueued chunk on the GPU
// and adds their F and d, b, t into the sums.  After the flush each chunk's
// dbt_estimate_before_fisher_vector and fisher_matrix hold its own results,
// as after the reference call.  setInvertCovarianceMatrix +
// computePSbeforeFvector enqueue the chunk for its own results only.
//
// The protected matrix builders of the reference (_initMatrices,
// _setVZMatrices, _setFiducialSignalMatrix, _setQiMatrix,
// _applyRedshiftInterp, _addMarginalizations and the matrix pointers) have no
// counterpart: the matrices are built and kept on the device.  The
// bookkeeping the constructor does (_copyQSOFile, _findRedshiftBin,
// _setNQandFisherIndex, _setStoredMatrices, stored_ikz_qi with its indices)
// is here.
#pragma once

#include <memory>
#include <utility>
#include <vector>

#include "io/qso_file.hpp"

namespace lyspeq_amd {
class Chunk;
}

class Chunk {
protected:
    int DATA_SIZE_2;
    int _matrix_n, RES_INDEX;
    double LOWER_REDSHIFT, UPPER_REDSHIFT, MEDIAN_REDSHIFT, BIN_REDSHIFT;
    bool on_oversampling;
    // (i_kz, matrix): the stored Q indices; the matrices live on the device
    std::vector<std::pair<int, double *>> stored_ikz_qi;

    void _copyQSOFile(const qio::QSOFile &qmaster, int i1, int i2);
    void _findRedshiftBin();
    void _setNQandFisherIndex();
    void _setStoredMatrices();

    Chunk() : DATA_SIZE_2(0), _matrix_n(0), N_Q_MATRICES(0) {}

public:
    std::unique_ptr<qio::QSOFile> qFile;
    int fisher_index_start, N_Q_MATRICES;
    int ZBIN, ZBIN_LOW, ZBIN_UPP, KBIN_UPP;

    // 3 x N_Q_MATRICES: d, b, t; N_Q_MATRICES^2 Fisher (upper triangle, unhalved)
    std::vector<std::unique_ptr<double[]>> dbt_estimate_before_fisher_vector;
    std::unique_ptr<double[]> fisher_matrix;

    Chunk(const qio::QSOFile &qmaster, int i1, int i2);
    Chunk(Chunk &&rhs) = delete;
    Chunk(const Chunk &rhs) = delete;
    ~Chunk();

    // The reference's relative cost model (core/chunk_estimate.cpp:297-382),
    // used to balance chunks across ranks.
    static double getComputeTimeEst(const qio::QSOFile &qmaster, int i1, int i2);

    int realSize() const { return qFile ? qFile->realSize() : real_size_; }
    int size() const { return qFile ? qFile->size() : size_; }
    // Keeps the fit values for the next computePSbeforeFvector.
    void setInvertCovarianceMatrix(const double *ps_estimate);
    // Enqueues the chunk for its own d, b, t and Fisher matrix (no sums).
    void computePSbeforeFvector();

    // Enqueues the chunk: ps_estimate is the TOTAL_KZ_BINS fit, the sums are
    // 3 x TOTAL_KZ_BINS and TOTAL_KZ_BINS^2 (upper triangle).
    void oneQSOiteration(const double *ps_estimate, std::vector<std::unique_ptr<double[]>> &dbt_sum_vector,
                         double *fisher_sum);

    // Writes the chunk's d, b, t and Fisher matrix (the matrices themselves
    // stay on the device) and, with a resolution matrix, the matrix values.
    void fprintfMatrices(const char *fname_base);
    double getMinMemUsage();
    void releaseFile() { qFile.reset(); }

    // The engine-side chunk (for the bootstrap and the schedule).
    lyspeq_amd::Chunk *engineChunk() const { return impl_.get(); }

private:
    std::unique_ptr<lyspeq_amd::Chunk> impl_;
    const double *ps_ = nullptr;
    int size_ = 0, real_size_ = 0;
};
