// Newton-Raphson driver arithmetic of OneDQuadraticPowerEstimate, in C++ on
// the host: what the reference does with the accumulated Fisher and d, b, t
// sums after its chunk loop, once per iteration (T x T FP64 work, T =
// TOTAL_KZ_BINS).
//
//   normalizeFisher   cblas_dscal(0.5) + copyUpperToLower
//                     (core/quadratic_estimate.cpp:456-457,
//                     mathtools/matrix_helper.cpp:81-94)
//   stableInvertSym   mxhelp::stableInvertSym (mathtools/matrix_helper.cpp:
//                     219-230,232-274,294-358): empty diagonals set to 1, the
//                     singular-value jump test (a value whose predecessor is
//                     more than 8 times larger sets the damping), LU inverse,
//                     damped solver (S S + damp I)^-1 S, empty diagonals
//                     zeroed again
//   powerEstimates    computePowerSpectrumEstimates (core/quadratic_estimate.
//                     cpp:290-318): weighted_i = 0.5 solver dbt_i, theta =
//                     weighted_d - weighted_b - weighted_t; sigma =
//                     sqrt(diag inverse) (:717-740)
//   hasConverged      hasConverged (core/quadratic_estimate.cpp:548-614)
//   PowerEstimate     the iteration loop of iterate() (:427-528) over an
//                     Engine and its OneQSOEstimate objects, one Engine::flush
//                     per iteration
//
// The matrix routines are this file's own (Gaussian elimination with partial
// pivoting; singular values of the symmetric matrix as the magnitudes of its
// eigenvalues, by Householder tridiagonalisation and Sturm-sequence
// bisection), not LAPACK, so results agree with the reference's to rounding.
// Failures throw std::runtime_error with the reference's messages.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace lyspeq_amd {

class Engine;
class OneQSOEstimate;

// F <- 0.5 F, then the upper triangle copied onto the lower (row-major T x T).
void normalizeFisher(double *F, int T);

// Inverse of a general n x n row-major matrix by Gaussian elimination with
// partial pivoting; throws "ERROR in LU decomposition." on an exactly zero
// pivot (LAPACK dgetrf's info > 0).
std::vector<double> invertLU(const double *A, int n);

// Singular values of a symmetric n x n matrix, descending.
std::vector<double> symmetricSingularValues(const double *A, int n);

struct FisherInverse {
    std::vector<double> solver;    // F^-1, or (F F + damp I)^-1 F when damped
    std::vector<double> inverse;   // F^-1
    int dof = 0;                   // NewDegreesOfFreedom: T minus empty diagonals
    double damp = 0.0;
    bool damped = false;
};
// mxhelp::stableInvertSym on a symmetric F (the normalised Fisher sum).
FisherInverse stableInvertSym(const double *F, int T);

struct PowerResult {
    std::vector<double> theta;      // T: current power estimate
    std::vector<double> sigma;      // T: sqrt(diag F^-1)
    std::vector<double> weighted;   // 3 x T: Fisher-weighted d, b, t
};
PowerResult powerEstimates(const FisherInverse &fi, const double *dbt_sum, int T);

struct ConvergenceTest {
    double abs_mean = 0, abs_max = 0;   // relative changes |p1 - p2| / max(p1, p2)
    double chi2_diag = 0, chi2_full = 0;
    bool converged = false;
};
// current and previous estimates, the inverse and the normalised Fisher sum,
// dof and CHISQ_CONVERGENCE_EPS.
ConvergenceTest hasConverged(const double *current, const double *previous, const double *inverse,
                             const double *fisher, int T, int dof, double eps);

// The reference's smoothing step (OneDQuadraticPowerEstimate::
// _smoothPowerSpectra, core/quadratic_estimate.cpp:353-396, with
// writeSpectrumEstimates :624-650 and _readScriptOutput :320-348): writes
// "nz nk", the z-bin counts, then z ("%.3lf"), k, P = theta + P_fid and
// e = sigma ("%e") per band, z-major, to <tmp_dir>/tmp-power-<tag>.txt; runs
// `<command> <in> <out>` (with " --interp_log" for the log mode) as a child
// process through the shell, as the reference runs smbivspline.py; reads one
// smoothed P per band from <tmp_dir>/tmp-smooth-<tag>.txt and subtracts P_fid.
// command defaults to this build's script (python3 -m lyspeq_amd.smooth,
// with the package on PYTHONPATH); the reference's "smbivspline.py" reads the
// same files.  Throws std::runtime_error("smoothing error") if the command
// fails, as the reference does.
struct ScriptSmoother {
    std::vector<double> kcenters;   // nk band centres
    std::vector<double> zcenters;   // nz bin centres
    std::vector<double> pfid;       // nz * nk fiducial power, z-major
    std::vector<int> zcounts;       // nz chunk counts per z bin (the file's second line)
    std::string command = "python3 -m lyspeq_amd.smooth";
    std::string tmp_dir = "/tmp";
    std::string tag = "0";          // the reference uses the MPI rank
    bool interp_log = false;        // SmoothLogkLogP
    void operator()(const std::vector<double> &theta, const std::vector<double> &sigma,
                    std::vector<double> &fits) const;
};

// The iterate() loop.  smooth maps (theta, sigma) to the power fits used by
// the next iteration (the reference runs py/smbivspline.py there: see
// ScriptSmoother); without one the next iteration uses theta itself.
class PowerEstimate {
public:
    using Smoother = std::function<void(const std::vector<double> &theta, const std::vector<double> &sigma,
                                        std::vector<double> &fits)>;
    PowerEstimate(Engine &eng, std::vector<OneQSOEstimate *> qsos, int iterations, double eps,
                  Smoother smooth = nullptr);
    // Runs up to `iterations` iterations from fits (TOTAL_KZ_BINS values,
    // zero for the first one); returns the number run.
    int iterate(std::vector<double> fits);

    std::vector<double> fisher;    // normalised Fisher sum of the last iteration
    FisherInverse inverse;
    PowerResult power;
    ConvergenceTest convergence;
    std::vector<std::vector<double>> history;   // theta of every iteration

private:
    Engine *eng_;
    std::vector<OneQSOEstimate *> qsos_;
    int iterations_;
    double eps_;
    Smoother smooth_;
};

}  // namespace lyspeq_amd
