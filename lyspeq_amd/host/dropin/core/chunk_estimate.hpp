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
// _applyRedshiftInterp, _scaleDerivativesWithRedshiftGrowth,
// _addMarginalizations) keep their names and the host buffers they fill
// (covariance_matrix, temp_matrix, stored_ikz_qi, _vmatrix, _zmatrix, from
// glmemory); the matrices are computed on the device from the chunk's
// registered pixels and tables (qml_chunk_matrix, qml_chunk_marginalize) and
// copied into those buffers.  The first such call commits the engine, after
// which no chunk can be constructed.  oneQSOiteration does not use them: the
// iteration keeps its matrices on the device.  The bookkeeping the
// constructor does (_copyQSOFile, _findRedshiftBin, _setNQandFisherIndex,
// _setStoredMatrices, stored_ikz_qi with its indices) is here too.
// an include guard, not #pragma once: the header may also be reached as a
// copy under another path (a build that shadows the reference's header)
#ifndef LYSPEQ_AMD_DROPIN_CHUNK_ESTIMATE_HPP
#define LYSPEQ_AMD_DROPIN_CHUNK_ESTIMATE_HPP

#include <memory>
#include <utility>
#include <vector>

#include "io/qso_file.hpp"

namespace lyspeq_amd {
class Chunk;
struct QSOSpectrum;
struct ResolutionRows;
}

// Host buffers of the protected builders (core/chunk_estimate.hpp:29-43):
// setMaxSizes records the largest chunk (every Chunk constructor calls it),
// allocMemory allocates covariance_matrix, two temp_matrices, stored_sfid
// (unless TURN_OFF_SFID) and max_nqdim stored_ikz_qi of max_size^2, the
// size vectors, and with an oversampled resolution matrix matrix_lambda,
// finer_matrix, v_matrix and z_matrix on its grid; dealloc frees them.  The
// interpolation tables of the reference live on the device.
namespace glmemory {
extern int max_size_2;
extern std::unique_ptr<double[]> covariance_matrix, temp_vector, weighted_data_vector, stored_sfid, matrix_lambda,
    finer_matrix, v_matrix, z_matrix;
extern std::vector<std::unique_ptr<double[]>> stored_ikz_qi, temp_matrices;

void setMaxSizes(int size, int matrix_n, int nqdim, bool onsamp);
void allocMemory();
void dealloc();
void updateMemUsed(double mem);
}  // namespace glmemory

class Chunk {
protected:
    int DATA_SIZE_2;
    int _matrix_n, RES_INDEX;
    double LOWER_REDSHIFT, UPPER_REDSHIFT, MEDIAN_REDSHIFT, BIN_REDSHIFT;
    // the matrix grid (finer with an oversampled resolution matrix) and the
    // caller's C^-1 for _addMarginalizations; neither is owned
    double *_matrix_lambda = nullptr, *inverse_covariance_matrix = nullptr;
    bool on_oversampling;
    double *_vmatrix = nullptr, *_zmatrix = nullptr;
    // DATA_SIZE x DATA_SIZE host buffers from glmemory (_initMatrices)
    double *covariance_matrix = nullptr;
    std::vector<double *> temp_matrix;
    // (i_kz, matrix): the stored Q indices and, after _initMatrices, their buffers
    std::vector<std::pair<int, double *>> stored_ikz_qi;

    void _copyQSOFile(const qio::QSOFile &qmaster, int i1, int i2);
    void _findRedshiftBin();
    void _setNQandFisherIndex();
    void _setStoredMatrices();

    // Points the matrices at glmemory's buffers (allocMemory first) and
    // registers the chunk with the engine if a subclass built it by hand.
    void _initMatrices();
    // sm = S_fid (R S R^T with a resolution matrix), size^2
    void _setFiducialSignalMatrix(double *sm);
    // _vmatrix, _zmatrix on the matrix grid (_matrix_n^2)
    void _setVZMatrices();
    // qi = Q of chunk index i_kz before the redshift weights (a copy of
    // stored_ikz_qi[kn] beyond the first z bin, as the reference)
    void _setQiMatrix(double *qi, int i_kz);
    // Multiplies each stored Q by its z bin's weights w_zm(z_ij) (uses temp_matrix[0])
    void _applyRedshiftInterp();
    // Multiplies each stored Q by ((1 + z_ij) / (1 + z_m))^B (uses temp_matrix[0])
    void _scaleDerivativesWithRedshiftGrowth();
    // inverse_covariance_matrix minus its projections on the continuum templates
    void _addMarginalizations();

    friend class TestOneQSOEstimate;

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
    // the master's pixels [i1, i2) (and resolution rows) from _copyQSOFile,
    // until the engine chunk is registered from them
    std::unique_ptr<lyspeq_amd::QSOSpectrum> src_;
    std::unique_ptr<lyspeq_amd::ResolutionRows> src_rmat_;
    const double *ps_ = nullptr;
    int size_ = 0, real_size_ = 0;
    lyspeq_amd::Chunk &engine_chunk();
};

#endif  // LYSPEQ_AMD_DROPIN_CHUNK_ESTIMATE_HPP
