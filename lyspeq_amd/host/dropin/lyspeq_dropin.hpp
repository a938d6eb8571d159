// Set-up of the drop-in Chunk / OneQSOEstimate: one engine per process,
// created on first use from the bins:: and specifics:: globals (as the
// reference's chunks read them), with tables from a caller-given source.
#pragma once

#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "lyspeq_qml.h"
#include "qml_host.hpp"

namespace lyspeq_amd {
namespace dropin {

// One (R, dv) lookup table in the reference's layout: deriv Nk x nv (velocity
// contiguous per band), signal nv x nz (redshift contiguous).
struct TableData {
    int nv = 0, nz = 0;
    double v_length = 0;
    std::vector<double> deriv, signal;
};
// Returns false if there is no table for (R_fwhm, dv); dv is the chunk's
// median pixel width, rounded to 0.1 km/s.
using TableSource = std::function<bool(int R_fwhm, double dv, TableData &out)>;

struct Options {
    int precision = QML_PREC_FP64;
    int device = 0;
    double scratch_gb = 0.0;       // 0: the engine's default budget
    int max_batch_chunks = 0;      // 0: the engine's default cap
    double fiducial_redshift_power = 3.591244;   // growth exponent B with RedshiftGrowthOn
};

// Once, before the first chunk is constructed (after the config is read).
void setup(const Options &opts, TableSource tables);
// The process's engine (created on first use).
Engine &engine();
// Runs every chunk enqueued since the last flush and adds its results into
// the sums it was enqueued with: the driver's one new call, after its chunk
// loop and before the 0.5 scaling of the Fisher sum.
void flush();
// Drops the engine and the table source (chunks must be gone).
void reset();
// Uploads the table for (R, dv) once and returns its index, or -1 if the
// source has none (the Chunk constructor's RES_INDEX).
int ensureTable(int R, double dv);

// Reads an SQ lookup table file (the three-HDU FITS the reference's
// CreateSQLookUpTable writes; io/sq_lookup_table_file.cpp:36-171).
bool readSQTableFile(const std::string &path, TableData &out, std::string *err = nullptr);
// A source reading <dir>/<base>_R<R>_dv<dv, one decimal>.fits
// (sqhelper::SQTableFileNameConvention, io/sq_lookup_table_file.cpp:12-21).
// With resolution matrices every chunk uses the first (R, dv) pair of the
// reference's list (SQLookupTable's table 0, core/sq_table.cpp:165,189):
// pass it as first.
TableSource sqFileTables(const std::string &dir, const std::string &base,
                         std::pair<int, double> first = {-1, 0.0});

// Outside the reference tree: fills the bins:: and specifics:: globals of
// core/global_numbers.hpp from a Binning, as readBins / readSpecifics would
// from the config (KBAND_CENTERS as the edges' mid-points, Z_LOWER/UPPER_EDGE
// from the first and last bin centre, MAX_FOREST_LENGTH_V per chunk).
void setGlobals(const Binning &b, bool use_resolution_matrix = false, bool fast_bootstrap = true,
                double min_snr_cut = 0.0, double memory_alloc_mb = 1500.0);

}  // namespace dropin
}  // namespace lyspeq_amd
