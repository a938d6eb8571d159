// Drop-in replacement of the reference's core/one_qso_estimate.hpp: class
// OneQSOEstimate with the surface of core/one_qso_estimate.hpp:16-61 over the
// drop-in Chunk.  The file constructor reads the spectrum as the reference's
// _readQsoFile does (core/one_qso_estimate.cpp:52-95): parameters, the SNR
// cut, the data, flux to delta, outlier masking, the resolution matrix and
// its deconvolution and oversampling, the redshift boundary cut and the
// short-file test; then decideIndices and one Chunk per piece, skipping
// pieces whose constructor throws.
// an include guard, not #pragma once: the header may also be reached as a
// copy under another path (a build that shadows the reference's header)
#ifndef LYSPEQ_AMD_DROPIN_ONE_QSO_ESTIMATE_HPP
#define LYSPEQ_AMD_DROPIN_ONE_QSO_ESTIMATE_HPP

#include <memory>
#include <string>
#include <vector>

#include "core/chunk_estimate.hpp"

class OneQSOEstimate {
protected:
    std::string fname_qso;
    std::unique_ptr<qio::QSOFile> _readQsoFile(const std::string &f_qso);

public:
    int istart = 0, ndim = 0;
    std::unique_ptr<double[]> fisher_matrix, theta_vector;

    std::vector<std::unique_ptr<Chunk>> chunks;

    OneQSOEstimate() {}
    OneQSOEstimate(const std::string &f_qso);
    OneQSOEstimate(OneQSOEstimate &&rhs) = default;
    OneQSOEstimate(const OneQSOEstimate &rhs) = delete;

    static std::vector<int> decideIndices(int size, double *wave);
    static double getComputeTimeEst(std::string fname_qso, int &zbin, long &targetid);

    void oneQSOiteration(const double *ps_estimate, std::vector<std::unique_ptr<double[]>> &dbt_sum_vector,
                         double *fisher_sum);

    // After a flush: theta_vector = sum over the chunks of (d - b) - t at
    // their fisher_index_start, on [istart, istart + ndim); with
    // specifics::FAST_BOOTSTRAP off also fisher_matrix (ndim^2, upper).
    // The chunks are released, as in the reference.
    void collapseBootstrap();
    void addBoot(int p, double *temppower, double *tempfisher);
    void addBootPowerOnly(int nboots, double *pcoeff, double *temppower);
};

#endif  // LYSPEQ_AMD_DROPIN_ONE_QSO_ESTIMATE_HPP
