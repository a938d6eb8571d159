// The part of the reference's core/global_numbers.hpp that the drop-in
// Chunk / OneQSOEstimate (chunk_estimate.hpp, one_qso_estimate.hpp) read:
// the same names, types and meanings (core/global_numbers.hpp:17-213).  In
// the reference tree the reference's own header is used instead and the
// values come from its config reader (readBins, readSpecifics,
// readProcess); outside it, lyspeq_amd::dropin::setGlobals fills them.
#pragma once

#include <string>
#include <vector>

constexpr double SPEED_OF_LIGHT = 299792.458, LYA_REST = 1215.67, MY_PI = 3.14159265358979323846,
                 ONE_SIGMA_2_FWHM = 2.35482004503, DOUBLE_EPSILON = 1e-15;

namespace qio {
enum ifileformat { Binary, Picca };
}

namespace process {
extern double MEMORY_ALLOC;
inline double getMemoryMB(int n) { return (double)sizeof(double) * n / 1048576.; }
void updateMemory(double deltamem);
}  // namespace process

namespace specifics {
extern bool TURN_OFF_SFID, USE_RESOLUTION_MATRIX, REDSHIFT_GROWTH_ON, USE_PRECOMPUTED_FISHER, FAST_BOOTSTRAP;
extern int CONT_LOGLAM_MARG_ORDER, CONT_LAM_MARG_ORDER, CONT_NVECS, NUMBER_OF_CHUNKS, OVERSAMPLING_FACTOR;
extern double RESOMAT_DECONVOLUTION_M, MIN_SNR_CUT;
extern qio::ifileformat INPUT_QSO_FILE;
extern double MAX_FOREST_LENGTH_V, MIN_FOREST_LENGTH_V, MAX_PIXEL_LENGTH_V;
}  // namespace specifics

namespace bins {
enum BinningMethod { TophatBinningMethod, TriangleBinningMethod };
extern int NUMBER_OF_K_BANDS, NUMBER_OF_Z_BINS, FISHER_SIZE, TOTAL_KZ_BINS;
extern std::vector<double> KBAND_EDGES, KBAND_CENTERS, ZBIN_CENTERS;
extern double Z_BIN_WIDTH, Z_LOWER_EDGE, Z_UPPER_EDGE;
extern BinningMethod Z_BINNING_METHOD;
// returns -1 if below, NUMBER_OF_Z_BINS if above
int findRedshiftBin(double z);
inline int getFisherMatrixIndex(int kn, int zm) { return kn + NUMBER_OF_K_BANDS * zm; }
inline void getFisherMatrixBinNoFromIndex(int ikz, int &kn, int &zm) {
    kn = ikz % NUMBER_OF_K_BANDS;
    zm = ikz / NUMBER_OF_K_BANDS;
}
}  // namespace bins

namespace conv {
// flux -> delta conversion of the reader (core/fiducial_cosmology.hpp); no
// conversion unless set
extern void (*convertFluxToDeltaF)(const double *, double *, double *, int);
}
