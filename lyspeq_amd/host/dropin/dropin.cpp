// Drop-in Chunk and OneQSOEstimate (core/chunk_estimate.hpp,
// core/one_qso_estimate.hpp) over the MI355X QML engine, and the per-process
// engine they share.  Compiles against the reference's global_numbers.hpp
// and io/qso_file.hpp (or this build's mirrors of them).
#include "core/chunk_estimate.hpp"
#include "core/one_qso_estimate.hpp"
#include "lyspeq_dropin.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>

namespace lyspeq_amd {
namespace dropin {

namespace {
struct State {
    Options opts;
    TableSource tables;
    std::unique_ptr<Engine> engine;
    std::map<std::pair<int, long long>, int> loaded;   // (R, round(10 dv)) -> table index
    bool first_table = false;                          // resolution-matrix mode: table 0 loaded
};
State &state() {
    static State s;
    return s;
}

Binning binning_from_globals() {
    Binning b;
    b.kband_edges = bins::KBAND_EDGES;
    b.zbin_centers = bins::ZBIN_CENTERS;
    b.z_bin_width = bins::Z_BIN_WIDTH;
    b.z_binning_method = bins::Z_BINNING_METHOD == bins::TriangleBinningMethod ? 1 : 0;
    b.cont_logl_order = specifics::CONT_LOGLAM_MARG_ORDER;
    b.cont_lam_order = specifics::CONT_LAM_MARG_ORDER;
    b.turn_off_sfid = specifics::TURN_OFF_SFID;
    // specifics::MAX_FOREST_LENGTH_V is already per chunk
    b.number_of_chunks = std::max(1, specifics::NUMBER_OF_CHUNKS);
    b.max_forest_length_v = specifics::MAX_FOREST_LENGTH_V * b.number_of_chunks;
    b.max_pixel_length_v = specifics::MAX_PIXEL_LENGTH_V;
    b.oversampling_factor = specifics::OVERSAMPLING_FACTOR;
    b.resomat_deconvolution_m = specifics::RESOMAT_DECONVOLUTION_M;
    b.redshift_growth_on = specifics::REDSHIFT_GROWTH_ON;
    b.fiducial_redshift_power = state().opts.fiducial_redshift_power;
    return b;
}

void add_table(State &s, int R, double dv, const TableData &t) {
    if (t.nv < 2 || (int)t.deriv.size() != bins::NUMBER_OF_K_BANDS * t.nv ||
        (!specifics::TURN_OFF_SFID && (int)t.signal.size() != t.nv * t.nz))
        throw std::runtime_error("lookup table with inconsistent sizes");
    s.engine->addTable(R, dv, t.nv, t.v_length, t.nz, t.deriv.data(), t.signal.empty() ? nullptr : t.signal.data());
}

}  // namespace

// Uploads the table a chunk of (R, dv) uses, once, and returns its index
// (RES_INDEX), or -1 if the source has none.  With resolution matrices every
// chunk uses table 0, the first of the source's list (asked for as R = -1).
int ensureTable(int R, double dv) {
    State &s = state();
    if (!s.tables) throw std::runtime_error("lyspeq_amd::dropin::setup was not called (no table source)");
    engine();
    if (specifics::USE_RESOLUTION_MATRIX) {
        if (!s.first_table) {
            TableData t;
            if (!s.tables(-1, 0.0, t)) return -1;
            add_table(s, 0, 0.0, t);
            s.first_table = true;
        }
        return 0;
    }
    const double rounded = std::round(dv * 10) / 10;   // SQLookupTable::findSpecResIndex
    const auto key = std::make_pair(R, (long long)std::llround(rounded * 10));
    auto it = s.loaded.find(key);
    if (it != s.loaded.end()) return it->second;
    TableData t;
    if (!s.tables(R, rounded, t)) return -1;
    add_table(s, R, rounded, t);
    const int idx = (int)s.loaded.size();
    s.loaded[key] = idx;
    return idx;
}

void setup(const Options &opts, TableSource tables) {
    State &s = state();
    if (s.engine) throw std::logic_error("lyspeq_amd::dropin::setup: the engine already exists");
    s.opts = opts;
    s.tables = std::move(tables);
}

Engine &engine() {
    State &s = state();
    if (!s.engine) {
        s.engine = std::make_unique<Engine>(binning_from_globals(), s.opts.precision, s.opts.device,
                                            s.opts.scratch_gb, s.opts.max_batch_chunks);
    }
    return *s.engine;
}

void flush() {
    if (state().engine) state().engine->flush();
}

void reset() {
    State &s = state();
    s.engine.reset();
    s.tables = nullptr;
    s.loaded.clear();
    s.first_table = false;
}

namespace {
// FITS primitives for the SQ table layout: 2880-byte blocks of 80-byte cards,
// BITPIX = -64 images stored big-endian
struct Hdu {
    std::map<std::string, std::string> cards;
    std::vector<double> data;
};

bool read_hdus(const std::string &path, std::vector<Hdu> &out, std::string *err) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    std::vector<char> buf;
    char tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    size_t off = 0;
    const size_t B = 2880;
    while (off + B <= buf.size()) {
        Hdu h;
        bool end = false;
        while (!end) {
            if (off + B > buf.size()) {
                if (err) *err = "truncated FITS header";
                return false;
            }
            for (size_t c = 0; c < B; c += 80) {
                const std::string card(buf.data() + off + c, 80);
                std::string key = card.substr(0, 8);
                key.erase(key.find_last_not_of(' ') + 1);
                if (key == "END") {
                    end = true;
                    break;
                }
                if (card.compare(8, 2, "= ") != 0) continue;
                std::string val = card.substr(10);
                const size_t slash = val.find('/');
                if (val.find('\'') == std::string::npos && slash != std::string::npos) val = val.substr(0, slash);
                val.erase(0, val.find_first_not_of(' '));
                val.erase(val.find_last_not_of(' ') + 1);
                if (!val.empty() && val[0] == '\'') {
                    const size_t q = val.find('\'', 1);
                    val = val.substr(1, q == std::string::npos ? std::string::npos : q - 1);
                    val.erase(val.find_last_not_of(' ') + 1);
                }
                h.cards[key] = val;
            }
            off += B;
        }
        const int naxis = h.cards.count("NAXIS") ? std::stoi(h.cards["NAXIS"]) : 0;
        size_t nel = naxis > 0 ? 1 : 0;
        const size_t left = (buf.size() - off) / 8;   // doubles left in the file
        for (int a = 1; a <= naxis; ++a) {
            const std::string k = "NAXIS" + std::to_string(a);
            const long long d = h.cards.count(k) ? std::stoll(h.cards[k]) : 0;
            if (d < 0 || (d > 0 && nel > left / (size_t)d)) {
                if (err) *err = "malformed or truncated FITS data";
                return false;
            }
            nel *= (size_t)d;
        }
        const int bitpix = h.cards.count("BITPIX") ? std::stoi(h.cards["BITPIX"]) : -64;
        if (nel > 0) {
            if (bitpix != -64) {
                if (err) *err = "only BITPIX = -64 images are read";
                return false;
            }
            if (nel > left) {
                if (err) *err = "truncated FITS data";
                return false;
            }
            h.data.resize(nel);
            for (size_t i = 0; i < nel; ++i) {
                unsigned char b[8];
                for (int k = 0; k < 8; ++k) b[k] = (unsigned char)buf[off + i * 8 + 7 - k];
                std::memcpy(&h.data[i], b, 8);
            }
        }
        off += (nel * 8 + B - 1) / B * B;
        out.push_back(std::move(h));
    }
    return true;
}
}  // namespace

// The SQ table file (io/sq_lookup_table_file.cpp:36-171): METADATA keys
// NVPTS, NZPTS, LENV; DERIVATIVE (k-band major, velocity contiguous);
// SIGNAL (redshift contiguous).
bool readSQTableFile(const std::string &path, TableData &out, std::string *err) {
    std::vector<Hdu> hdus;
    if (!read_hdus(path, hdus, err)) return false;
    const Hdu *meta = nullptr, *der = nullptr, *sig = nullptr;
    for (const Hdu &h : hdus) {
        auto it = h.cards.find("EXTNAME");
        if (it == h.cards.end()) continue;
        if (it->second == "METADATA") meta = &h;
        if (it->second == "DERIVATIVE") der = &h;
        if (it->second == "SIGNAL") sig = &h;
    }
    if (!meta || !der || !sig) {
        if (err) *err = "SQ table without METADATA, DERIVATIVE or SIGNAL";
        return false;
    }
    try {
        out.nv = std::stoi(meta->cards.at("NVPTS"));
        out.nz = std::stoi(meta->cards.at("NZPTS"));
        out.v_length = std::stod(meta->cards.at("LENV"));
    } catch (std::exception &) {
        if (err) *err = "SQ table METADATA without NVPTS, NZPTS or LENV";
        return false;
    }
    out.deriv = der->data;
    out.signal = sig->data;
    if ((int)out.signal.size() != out.nv * out.nz || out.deriv.size() % out.nv != 0) {
        if (err) *err = "SQ table with inconsistent sizes";
        return false;
    }
    return true;
}

TableSource sqFileTables(const std::string &dir, const std::string &base, std::pair<int, double> first) {
    return [dir, base, first](int R, double dv, TableData &out) {
        if (R < 0) {   // resolution matrices: the first (R, dv) of the list
            if (first.first < 0) return false;
            R = first.first;
            dv = first.second;
        }
        char name[64];
        std::snprintf(name, sizeof name, "_R%d_dv%.1f.fits", R, dv);
        std::string err;
        return readSQTableFile(dir + "/" + base + name, out, &err);
    };
}

}  // namespace dropin
}  // namespace lyspeq_amd

using lyspeq_amd::dropin::engine;

namespace {
void log_err(const std::string &msg) { std::fprintf(stderr, "%s", msg.c_str()); }

// _getMaxKindex (core/chunk_estimate.cpp:33-39): first k-band centre at or
// above 0.85 k_Nyquist
int max_k_index(double knyq) {
    auto it = std::lower_bound(bins::KBAND_CENTERS.begin(), bins::KBAND_CENTERS.end(), 0.85 * knyq);
    return (int)std::distance(bins::KBAND_CENTERS.begin(), it);
}
}  // namespace

// ---------------------------------------------------------------- glmemory --
// The reference's global buffers (core/chunk_estimate.cpp:45-139), for the
// protected builders; two temp matrices, the reference's minimum.
namespace glmemory {
int max_size = 0, max_size_2 = 0, max_matrix_n = 0, max_nqdim = 0;
bool on_oversampling = false;
double memUsed = 0;
std::unique_ptr<double[]> covariance_matrix, temp_vector, weighted_data_vector, stored_sfid, matrix_lambda,
    finer_matrix, v_matrix, z_matrix;
std::vector<std::unique_ptr<double[]>> stored_ikz_qi, temp_matrices;

void setMaxSizes(int size, int matrix_n, int nqdim, bool onsamp) {
    max_size = std::max(size, max_size);
    max_size_2 = max_size * max_size;
    max_matrix_n = std::max(matrix_n, max_matrix_n);
    max_nqdim = std::max(nqdim, max_nqdim);
    on_oversampling |= onsamp;
}

void allocMemory() {
    const size_t s2 = (size_t)max_size * max_size, hs = (size_t)max_matrix_n * max_matrix_n;
    constexpr int ntemp = 2;
    const double mb = 8.0 / 1048576.;
    double mem = mb * (2.0 * max_size + (1 + ntemp + !specifics::TURN_OFF_SFID + max_nqdim) * (double)s2);
    if (on_oversampling) mem += mb * (max_matrix_n + 3.0 * hs);
    updateMemUsed(mem);
    temp_vector = std::make_unique<double[]>(max_size);
    weighted_data_vector = std::make_unique<double[]>(max_size);
    covariance_matrix = std::make_unique<double[]>(s2);
    temp_matrices.clear();
    for (int i = 0; i < ntemp; ++i) temp_matrices.push_back(std::make_unique<double[]>(s2));
    if (!specifics::TURN_OFF_SFID) stored_sfid = std::make_unique<double[]>(s2);
    stored_ikz_qi.clear();
    for (int i = 0; i < max_nqdim; ++i) stored_ikz_qi.push_back(std::make_unique<double[]>(s2));
    if (on_oversampling) {
        matrix_lambda = std::make_unique<double[]>(max_matrix_n);
        finer_matrix = std::make_unique<double[]>(hs);
        v_matrix = std::make_unique<double[]>(hs);
        z_matrix = std::make_unique<double[]>(hs);
    }
}

void dealloc() {
    temp_vector.reset();
    weighted_data_vector.reset();
    covariance_matrix.reset();
    temp_matrices.clear();
    stored_sfid.reset();
    stored_ikz_qi.clear();
    matrix_lambda.reset();
    finer_matrix.reset();
    v_matrix.reset();
    z_matrix.reset();
    process::updateMemory(memUsed);
    memUsed = 0;
}

void updateMemUsed(double mem) {
    memUsed += mem;
    process::updateMemory(-mem);
}
}  // namespace glmemory

// ------------------------------------------------------------------- Chunk --
Chunk::Chunk(const qio::QSOFile &qmaster, int i1, int i2) : DATA_SIZE_2(0), _matrix_n(0), N_Q_MATRICES(0) {
    _copyQSOFile(qmaster, i1, i2);
    // the reference reads ZBIN_CENTERS[ZBIN_LOW / ZBIN_UPP] next; its reader
    // has cut every spectrum to the bins, which a chunk outside them breaks
    if (ZBIN_LOW < 0 || ZBIN_UPP >= bins::NUMBER_OF_Z_BINS)
        throw std::out_of_range("Chunk: the chunk lies outside the redshift bins");
    _setNQandFisherIndex();
    process::updateMemory(-getMinMemUsage());
    stored_ikz_qi.reserve(N_Q_MATRICES);
    for (int i_kz = 0; i_kz < N_Q_MATRICES; ++i_kz) {
        int kn, zm;
        bins::getFisherMatrixBinNoFromIndex(i_kz + fisher_index_start, kn, zm);
        if (kn < KBIN_UPP) stored_ikz_qi.push_back(std::make_pair(i_kz, nullptr));
    }
    _setStoredMatrices();
    for (int k = 0; k < 3; ++k) {
        dbt_estimate_before_fisher_vector.push_back(std::make_unique<double[]>(N_Q_MATRICES));
        std::fill_n(dbt_estimate_before_fisher_vector.back().get(), N_Q_MATRICES, 0.0);
    }
    fisher_matrix = std::make_unique<double[]>((size_t)N_Q_MATRICES * N_Q_MATRICES);
    std::fill_n(fisher_matrix.get(), (size_t)N_Q_MATRICES * N_Q_MATRICES, 0.0);
    glmemory::setMaxSizes(size(), _matrix_n, (int)stored_ikz_qi.size(), on_oversampling);

    engine_chunk();
    if (impl_->fisher_index_start != fisher_index_start || impl_->N_Q_MATRICES != N_Q_MATRICES ||
        impl_->KBIN_UPP != KBIN_UPP)
        throw std::logic_error("Chunk: the engine's bins differ from the constructor's");
    impl_->result_fisher = fisher_matrix.get();
    for (int k = 0; k < 3; ++k) impl_->result_dbt[k] = dbt_estimate_before_fisher_vector[k].get();
    size_ = qFile->size();
    real_size_ = qFile->realSize();
}

// The engine's chunk, registered on first use from the pixels _copyQSOFile
// kept: the master's [i1, i2) in Angstrom (the engine cuts the masked
// boundary and decides the same bins) with the master's preprocessed
// resolution matrix.
lyspeq_amd::Chunk &Chunk::engine_chunk() {
    if (!impl_) {
        if (!src_) throw std::logic_error("Chunk: no spectrum to register (_copyQSOFile was not called)");
        impl_ = std::make_unique<lyspeq_amd::Chunk>(engine(), *src_, 0, src_->size(), src_rmat_.get());
        src_.reset();
        src_rmat_.reset();
    }
    return *impl_;
}

Chunk::~Chunk() { process::updateMemory(getMinMemUsage()); }

// Chunk::_copyQSOFile (core/chunk_estimate.cpp:168-214)
void Chunk::_copyQSOFile(const qio::QSOFile &qmaster, int i1, int i2) {
    qFile = std::make_unique<qio::QSOFile>(qmaster, i1, i2);
    // the engine's input: the same pixels of the master (wave in Angstrom)
    // and its preprocessed resolution matrix's rows for them
    src_ = std::make_unique<lyspeq_amd::QSOSpectrum>();
    src_->wave.assign(qmaster.wave() + i1, qmaster.wave() + i2);
    src_->delta.assign(qmaster.delta() + i1, qmaster.delta() + i2);
    src_->ivar.assign(qmaster.ivar() + i1, qmaster.ivar() + i2);
    src_->R_fwhm = qmaster.R_fwhm;
    src_->fname = qmaster.fname;
    src_rmat_.reset();
    if (specifics::USE_RESOLUTION_MATRIX && qmaster.Rmat) {
        const mxhelp::Resolution &R = *qmaster.Rmat;
        const int n = i2 - i1, nel = R.getNElemPerRow(), size = qmaster.size();
        src_rmat_ = std::make_unique<lyspeq_amd::ResolutionRows>();
        lyspeq_amd::ResolutionRows &rr = *src_rmat_;
        rr.nelem = nel;
        if (R.isDiaMatrix()) {
            rr.kind = QML_RMAT_DIA;
            rr.values.resize((size_t)nel * n);
            for (int d = 0; d < nel; ++d)
                std::copy_n(R.matrix() + (size_t)d * size + i1, n, rr.values.data() + (size_t)d * n);
        } else {
            rr.kind = QML_RMAT_OVERSAMPLED;
            rr.oversampling = (R.getNCols() - nel + 1) / size;
            rr.values.assign(R.matrix() + (size_t)i1 * nel, R.matrix() + (size_t)i2 * nel);
        }
    }
    const double lenv = qFile->getLengthV();
    const int minnopix = (int)(lenv / specifics::MAX_PIXEL_LENGTH_V);
    if (qFile->realSize() < minnopix || lenv < specifics::MIN_FOREST_LENGTH_V) {
        const double meandv = lenv / (qFile->realSize() + 1e-12);
        std::ostringstream err;
        err.precision(1);
        err << "Short chunk with realsize " << qFile->realSize() << '/' << qFile->size() << " and velocity length "
            << std::scientific << lenv << '/' << specifics::MIN_FOREST_LENGTH_V << " || mean dv: " << meandv << '/'
            << specifics::MAX_PIXEL_LENGTH_V;
        throw std::runtime_error(err.str());
    }
    if (specifics::USE_RESOLUTION_MATRIX) {
        RES_INDEX = 0;
        if (!qFile->Rmat) throw std::runtime_error("Chunk: UseResoMatrix is on but the spectrum has no matrix");
        _matrix_n = qFile->Rmat->getNCols();
        if (lyspeq_amd::dropin::ensureTable(qFile->R_fwhm, qFile->dv_kms) < 0)
            throw std::out_of_range("Chunk: no lookup table for the fiducial signal");
    } else {
        RES_INDEX = lyspeq_amd::dropin::ensureTable(qFile->R_fwhm, qFile->dv_kms);
        if (RES_INDEX == -1) {
            std::ostringstream err;
            err << "SpecRes R: " << qFile->R_fwhm << ", dv: " << qFile->dv_kms << " not found in tables!";
            throw std::out_of_range(err.str());
        }
        _matrix_n = size();
    }
    DATA_SIZE_2 = size() * size();
    on_oversampling = specifics::USE_RESOLUTION_MATRIX && !qFile->Rmat->isDiaMatrix();
    qFile->readMinMaxMedRedshift(LOWER_REDSHIFT, UPPER_REDSHIFT, MEDIAN_REDSHIFT);
    KBIN_UPP = max_k_index(MY_PI / qFile->dv_kms);
    _findRedshiftBin();
    std::for_each(qFile->wave(), qFile->wave() + size(), [](double &w) { w /= LYA_REST; });
}

// Chunk::_findRedshiftBin (core/chunk_estimate.cpp:218-229)
void Chunk::_findRedshiftBin() {
    ZBIN = bins::findRedshiftBin(MEDIAN_REDSHIFT);
    ZBIN_LOW = bins::findRedshiftBin(LOWER_REDSHIFT);
    ZBIN_UPP = bins::findRedshiftBin(UPPER_REDSHIFT);
    BIN_REDSHIFT = (ZBIN >= 0 && ZBIN < bins::NUMBER_OF_Z_BINS) ? bins::ZBIN_CENTERS[ZBIN] : 0.0;
}

// Chunk::_setNQandFisherIndex (core/chunk_estimate.cpp:233-250)
void Chunk::_setNQandFisherIndex() {
    N_Q_MATRICES = ZBIN_UPP - ZBIN_LOW + 1;
    fisher_index_start = bins::getFisherMatrixIndex(0, ZBIN_LOW);
    if (bins::Z_BINNING_METHOD == bins::TriangleBinningMethod) {
        if (LOWER_REDSHIFT < bins::ZBIN_CENTERS[ZBIN_LOW] && ZBIN_LOW != 0) {
            ++N_Q_MATRICES;
            fisher_index_start -= bins::NUMBER_OF_K_BANDS;
        }
        if (bins::ZBIN_CENTERS[ZBIN_UPP] < UPPER_REDSHIFT && ZBIN_UPP < bins::NUMBER_OF_Z_BINS - 1) ++N_Q_MATRICES;
    }
    N_Q_MATRICES *= bins::NUMBER_OF_K_BANDS;
}

// Chunk::_setStoredMatrices (core/chunk_estimate.cpp:258-282): the matrices
// live on the device, whose scratch budget the engine checks when the chunk
// is registered ("Not all matrices are stored." comes from there).
void Chunk::_setStoredMatrices() {}

// ------------------------------------------------- the protected builders --
// Chunk::_initMatrices (core/chunk_estimate.cpp:917-957)
void Chunk::_initMatrices() {
    if (!glmemory::covariance_matrix || glmemory::max_size_2 < DATA_SIZE_2 || glmemory::temp_matrices.size() < 2 ||
        glmemory::stored_ikz_qi.size() < stored_ikz_qi.size() || (on_oversampling && !glmemory::v_matrix))
        throw std::logic_error("Chunk::_initMatrices: glmemory::allocMemory has not been called for this chunk");
    covariance_matrix = glmemory::covariance_matrix.get();
    temp_matrix.resize(glmemory::temp_matrices.size());
    for (size_t i = 0; i < temp_matrix.size(); ++i) temp_matrix[i] = glmemory::temp_matrices[i].get();
    for (size_t i = 0; i < stored_ikz_qi.size(); ++i) stored_ikz_qi[i].second = glmemory::stored_ikz_qi[i].get();
    if (on_oversampling) {
        // the oversampled grid, units of lambda_Lya (as the engine's chunk has it)
        _matrix_lambda = glmemory::matrix_lambda.get();
        const double fine_dlambda = qFile->dlambda / LYA_REST / specifics::OVERSAMPLING_FACTOR;
        const int disp = qFile->Rmat->getNElemPerRow() / 2;
        for (int i = 0; i < _matrix_n; ++i) _matrix_lambda[i] = qFile->wave()[0] + (i - disp) * fine_dlambda;
        _vmatrix = glmemory::v_matrix.get();
        _zmatrix = glmemory::z_matrix.get();
    } else {
        _matrix_lambda = qFile->wave();
        _vmatrix = temp_matrix[0];
        _zmatrix = temp_matrix[1];
    }
    engine_chunk();
}

namespace {
// one matrix of the engine chunk into out, which holds n^2 values
void fetch_matrix(lyspeq_amd::Chunk &c, int which, int arg, double *out, int n, const char *who) {
    if (!out) throw std::logic_error(std::string(who) + ": null matrix (call _initMatrices first)");
    const int m = engine().chunkMatrix(c, which, arg, nullptr);
    if (m != n) throw std::logic_error(std::string(who) + ": the engine's matrix size differs from the chunk's");
    engine().chunkMatrix(c, which, arg, out);
}
}  // namespace

// Chunk::_setVZMatrices (core/chunk_estimate.cpp:384-398)
void Chunk::_setVZMatrices() {
    fetch_matrix(engine_chunk(), QML_CMAT_V, 0, _vmatrix, _matrix_n, "Chunk::_setVZMatrices");
    fetch_matrix(engine_chunk(), QML_CMAT_Z, 0, _zmatrix, _matrix_n, "Chunk::_setVZMatrices");
}

// Chunk::_setFiducialSignalMatrix (core/chunk_estimate.cpp:400-426)
void Chunk::_setFiducialSignalMatrix(double *sm) {
    fetch_matrix(engine_chunk(), QML_CMAT_SFID, 0, sm, size(), "Chunk::_setFiducialSignalMatrix");
}

// Chunk::_setQiMatrix (core/chunk_estimate.cpp:429-480): beyond the first z
// bin the band's matrix is the stored one of the first
void Chunk::_setQiMatrix(double *qi, int i_kz) {
    int kn, zm;
    bins::getFisherMatrixBinNoFromIndex(i_kz + fisher_index_start, kn, zm);
    if (i_kz > kn) {
        if (kn >= (int)stored_ikz_qi.size() || !stored_ikz_qi[kn].second)
            throw std::logic_error("Chunk::_setQiMatrix: the first z bin's matrix is not set");
        std::copy_n(stored_ikz_qi[kn].second, DATA_SIZE_2, qi);
    } else {
        fetch_matrix(engine_chunk(), QML_CMAT_Q, i_kz, qi, size(), "Chunk::_setQiMatrix");
    }
}

namespace {
// stored_ikz_qi[KBIN_UPP iz ..] *= the z bin's factor matrix, for each z bin
// of the chunk (core/chunk_estimate.cpp:483-573)
void scale_stored(lyspeq_amd::Chunk &c, int which, std::vector<std::pair<int, double *>> &stored, int kbin_upp,
                  int fisher_index_start, double *tmp, int n, const char *who) {
    if (kbin_upp <= 0) return;
    const int Nz = (int)stored.size() / kbin_upp;
    const size_t n2 = (size_t)n * n;
    for (int iz = 0; iz < Nz; ++iz) {
        const int zm = fisher_index_start / bins::NUMBER_OF_K_BANDS + iz;
        fetch_matrix(c, which, zm, tmp, n, who);
        for (int k = kbin_upp * iz; k < kbin_upp * (iz + 1); ++k) {
            double *q = stored[k].second;
            if (!q) throw std::logic_error(std::string(who) + ": stored matrices not set (call _initMatrices first)");
            for (size_t e = 0; e < n2; ++e) q[e] *= tmp[e];
        }
    }
}
}  // namespace

void Chunk::_applyRedshiftInterp() {
    scale_stored(engine_chunk(), QML_CMAT_ZWEIGHT, stored_ikz_qi, KBIN_UPP, fisher_index_start,
                 temp_matrix.empty() ? nullptr : temp_matrix[0], size(), "Chunk::_applyRedshiftInterp");
}

void Chunk::_scaleDerivativesWithRedshiftGrowth() {
    scale_stored(engine_chunk(), QML_CMAT_GROWTH, stored_ikz_qi, KBIN_UPP, fisher_index_start,
                 temp_matrix.empty() ? nullptr : temp_matrix[0], size(), "Chunk::_scaleDerivativesWithRedshiftGrowth");
}

// Chunk::_addMarginalizations (core/chunk_estimate.cpp:596-665)
void Chunk::_addMarginalizations() {
    if (!inverse_covariance_matrix)
        throw std::logic_error("Chunk::_addMarginalizations: inverse_covariance_matrix is not set");
    engine().marginalize(engine_chunk(), inverse_covariance_matrix);
}

double Chunk::getMinMemUsage() { return process::getMemoryMB((N_Q_MATRICES + 1) * N_Q_MATRICES); }

// Chunk::getComputeTimeEst (core/chunk_estimate.cpp:297-382): the reference's
// relative cost model (NERSC Perlmutter, -c 2)
double Chunk::getComputeTimeEst(const qio::QSOFile &qmaster, int i1, int i2) {
    try {
        qio::QSOFile qtemp(qmaster, i1, i2);
        const double lenv = qtemp.getLengthV();
        const int minnopix = (int)(lenv / specifics::MAX_PIXEL_LENGTH_V);
        if (qtemp.realSize() < minnopix || lenv < specifics::MIN_FOREST_LENGTH_V) return 0;
        double z1, z2, zm;
        qtemp.readMinMaxMedRedshift(z1, z2, zm);
        const int zlo = bins::findRedshiftBin(z1), zup = bins::findRedshiftBin(z2);
        if (zlo > bins::NUMBER_OF_Z_BINS - 1 || zup < 0) return 0;
        int nq = zup - zlo + 1;
        const double agemm = 19., asymv = 0.45, adot = 1.0;
        const double x = qtemp.size() / 100.;
        const double one_dgemm = agemm * std::pow(x, 3), one_dsymv = asymv * std::pow(x, 2),
                     one_ddot = adot * std::pow(x, 2);
        int fidx = bins::getFisherMatrixIndex(0, zlo);
        // (the bin-centre reads are guarded: the reference's reader has cut
        // every spectrum to the bins, so they are never out of range there)
        if (bins::Z_BINNING_METHOD == bins::TriangleBinningMethod) {
            if (zlo >= 0 && z1 < bins::ZBIN_CENTERS[zlo] && zlo != 0) {
                ++nq;
                fidx -= bins::NUMBER_OF_K_BANDS;
            }
            if (zup < bins::NUMBER_OF_Z_BINS && bins::ZBIN_CENTERS[zup] < z2 && zup < bins::NUMBER_OF_Z_BINS - 1)
                ++nq;
        }
        nq *= bins::NUMBER_OF_K_BANDS;
        const int kncut = max_k_index(MY_PI / qtemp.dv_kms);
        int real_nq = 0;
        for (int i_kz = 0; i_kz < nq; ++i_kz) {
            int kn, zb;
            bins::getFisherMatrixBinNoFromIndex(i_kz + fidx, kn, zb);
            if (kn < kncut) ++real_nq;
        }
        const int n_m_combo = real_nq + 1 + 1;   // without FISHER_OPTIMIZATION
        double res = real_nq * (one_dgemm + one_ddot * n_m_combo) + (real_nq + 1) * one_dsymv;
        if (!specifics::TURN_OFF_SFID) res += one_dgemm + real_nq * one_ddot;
        if (specifics::USE_RESOLUTION_MATRIX) {
            const int ndiags = 11;
            double extra = ndiags * one_ddot * real_nq;
            const int osamp = specifics::OVERSAMPLING_FACTOR;
            extra *= (osamp > 0) ? osamp * (osamp + 1) : 2;
            res += extra;
        }
        return res;
    } catch (std::exception &e) {
        log_err("Chunk::getComputeTimeEst::" + std::string(e.what()) + ". Skipping " + qmaster.fname + ".\n");
        return 0;
    }
}

void Chunk::setInvertCovarianceMatrix(const double *ps_estimate) { ps_ = ps_estimate; }

void Chunk::computePSbeforeFvector() {
    if (!ps_) throw std::logic_error("Chunk::computePSbeforeFvector: call setInvertCovarianceMatrix first");
    impl_->enqueueResultsOnly(ps_);
}

void Chunk::oneQSOiteration(const double *ps_estimate, std::vector<std::unique_ptr<double[]>> &dbt_sum_vector,
                            double *fisher_sum) {
    impl_->oneQSOiteration(ps_estimate, dbt_sum_vector, fisher_sum);
}

namespace {
void write_text_matrix(const std::string &path, const double *a, int nr, int nc) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("Cannot open file: " + path);
    std::fprintf(f, "%d %d\n", nr, nc);
    for (int i = 0; i < nr; ++i) {
        for (int j = 0; j < nc; ++j) std::fprintf(f, "%14le ", a[(size_t)i * nc + j]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}
}  // namespace

void Chunk::fprintfMatrices(const char *fname_base) {
    const std::string base(fname_base);
    write_text_matrix(base + "-fisher.txt", fisher_matrix.get(), N_Q_MATRICES, N_Q_MATRICES);
    std::vector<double> dbt((size_t)3 * N_Q_MATRICES);
    for (int k = 0; k < 3; ++k)
        std::copy_n(dbt_estimate_before_fisher_vector[k].get(), N_Q_MATRICES, dbt.data() + (size_t)k * N_Q_MATRICES);
    write_text_matrix(base + "-dbt.txt", dbt.data(), 3, N_Q_MATRICES);
    if (qFile && qFile->Rmat) {
        const mxhelp::Resolution &R = *qFile->Rmat;
        const int nel = R.getNElemPerRow();
        write_text_matrix(base + "-resolution.txt", R.matrix(), R.getSize() / nel, nel);
    }
}

// ---------------------------------------------------------- OneQSOEstimate --
namespace {
constexpr int MIN_PIXELS_IN_FILE = 100;
}

// OneQSOEstimate::decideIndices (core/one_qso_estimate.cpp:17-50)
std::vector<int> OneQSOEstimate::decideIndices(int size, double *wave) {
    int nchunks = 1;
    std::vector<double> vel(size);
    const double w0 = wave[0];
    for (int i = 0; i < size; ++i) vel[i] = SPEED_OF_LIGHT * std::log(wave[i] / w0);
    double vmax = vel[size - 1];
    if (specifics::NUMBER_OF_CHUNKS > 1) {
        nchunks = (int)(nchunks + vmax / specifics::MAX_FOREST_LENGTH_V);
        vmax = std::min(specifics::MAX_FOREST_LENGTH_V, vmax / nchunks);
    }
    std::vector<int> indices;
    indices.reserve(nchunks + 1);
    indices.push_back(0);
    int lo = 0;
    for (int i = 0; i < nchunks; ++i) {
        const int jj = (int)(std::upper_bound(vel.begin() + lo, vel.end(), vmax) - vel.begin());
        indices.push_back(jj);
        if (jj == size) break;
        lo = jj;
        const double v0 = vel[lo];
        for (int k = lo; k < size; ++k) vel[k] -= v0;
    }
    return indices;
}

// OneQSOEstimate::_readQsoFile (core/one_qso_estimate.cpp:52-95)
std::unique_ptr<qio::QSOFile> OneQSOEstimate::_readQsoFile(const std::string &f_qso) {
    auto q = std::make_unique<qio::QSOFile>(f_qso, specifics::INPUT_QSO_FILE);
    q->readParameters();
    if (q->snr < specifics::MIN_SNR_CUT) throw std::runtime_error("OneQSOEstimate::_readQsoFile::Low SNR");
    q->readData();
    conv::convertFluxToDeltaF(q->wave(), q->delta(), q->ivar(), q->size());
    const int nout = q->maskOutliers();
    if (nout > 0)
        log_err("WARNING::OneQSOEstimate::_readQsoFile::Found " + std::to_string(nout) + " outlier pixels in " +
                f_qso + ".\n");
    if (specifics::USE_RESOLUTION_MATRIX) {
        q->readAllocResolutionMatrix();
        if (specifics::RESOMAT_DECONVOLUTION_M > 0) q->Rmat->deconvolve(specifics::RESOMAT_DECONVOLUTION_M);
        if (specifics::OVERSAMPLING_FACTOR > 0) q->Rmat->oversample(specifics::OVERSAMPLING_FACTOR, q->dlambda);
    }
    q->closeFile();
    q->cutBoundary(bins::Z_LOWER_EDGE, bins::Z_UPPER_EDGE);
    if (q->realSize() < MIN_PIXELS_IN_FILE) throw std::runtime_error("OneQSOEstimate::_readQsoFile::Short file");
    return q;
}

// OneQSOEstimate::OneQSOEstimate (core/one_qso_estimate.cpp:97-126)
OneQSOEstimate::OneQSOEstimate(const std::string &f_qso) : fname_qso(f_qso) {
    try {
        auto q = _readQsoFile(f_qso);
        const std::vector<int> idx = decideIndices(q->size(), q->wave());
        const int n = (int)idx.size() - 1;
        chunks.reserve(n);
        for (int nc = 0; nc < n; ++nc) {
            try {
                chunks.push_back(std::make_unique<Chunk>(*q, idx[nc], idx[nc + 1]));
            } catch (lyspeq_amd::QMLError &) {
                throw;   // engine failures are not chunk rejections
            } catch (std::exception &e) {
                log_err("OneQSOEstimate::OneQSOEstimate::" + std::string(e.what()) + " Skipping chunk " +
                        std::to_string(nc + 1) + "/" + std::to_string(n) + " of " + f_qso + ".\n");
            }
        }
    } catch (lyspeq_amd::QMLError &) {
        throw;
    } catch (std::exception &e) {
        log_err(std::string(e.what()) + " in " + f_qso + ".\n");
    }
}

// OneQSOEstimate::getComputeTimeEst (core/one_qso_estimate.cpp:128-166)
double OneQSOEstimate::getComputeTimeEst(std::string fname, int &zbin, long &targetid) {
    zbin = -1;
    targetid = -1;
    try {
        qio::QSOFile q(fname, specifics::INPUT_QSO_FILE);
        q.readParameters();
        targetid = q.id;
        if (q.snr < specifics::MIN_SNR_CUT) return 0;
        q.readData();
        q.cutBoundary(bins::Z_LOWER_EDGE, bins::Z_UPPER_EDGE);
        if (q.realSize() < MIN_PIXELS_IN_FILE) return 0;
        double z1, z2, zm;
        q.readMinMaxMedRedshift(z1, z2, zm);
        zbin = bins::findRedshiftBin(zm);
        const std::vector<int> idx = decideIndices(q.size(), q.wave());
        double res = 0;
        for (size_t nc = 0; nc + 1 < idx.size(); ++nc) res += Chunk::getComputeTimeEst(q, idx[nc], idx[nc + 1]);
        return res;
    } catch (std::exception &e) {
        log_err("OneQSOEstimate::getComputeTimeEst::" + std::string(e.what()) + ". Skipping " + fname + ".\n");
        return 0;
    }
}

// OneQSOEstimate::oneQSOiteration (core/one_qso_estimate.cpp:168-188)
void OneQSOEstimate::oneQSOiteration(const double *ps_estimate,
                                     std::vector<std::unique_ptr<double[]>> &dbt_sum_vector, double *fisher_sum) {
    for (auto &chunk : chunks) {
        try {
            chunk->oneQSOiteration(ps_estimate, dbt_sum_vector, fisher_sum);
        } catch (lyspeq_amd::QMLError &) {
            throw;
        } catch (std::exception &e) {
            log_err("OneQSOEstimate::oneQSOiteration::" + std::string(e.what()) + " Skipping " +
                    chunk->qFile->fname + ".\n");
            chunk.reset();
        }
    }
    chunks.erase(std::remove_if(chunks.begin(), chunks.end(), [](const auto &x) { return !x; }), chunks.end());
}

// OneQSOEstimate::collapseBootstrap (core/one_qso_estimate.cpp:191-231), after
// the flush that filled the chunks' results
void OneQSOEstimate::collapseBootstrap() {
    int fend = 0;
    istart = bins::TOTAL_KZ_BINS;
    for (const auto &c : chunks) {
        istart = std::min(istart, c->fisher_index_start);
        fend = std::max(fend, c->fisher_index_start + c->N_Q_MATRICES);
    }
    ndim = fend - istart;
    theta_vector = std::make_unique<double[]>(std::max(ndim, 0));
    std::fill_n(theta_vector.get(), std::max(ndim, 0), 0.0);
    for (const auto &c : chunks) {
        const int off = c->fisher_index_start - istart;
        double *pk = c->dbt_estimate_before_fisher_vector[0].get();
        const double *nk = c->dbt_estimate_before_fisher_vector[1].get();
        const double *tk = c->dbt_estimate_before_fisher_vector[2].get();
        for (int i = 0; i < c->N_Q_MATRICES; ++i) pk[i] += -1 * nk[i];
        for (int i = 0; i < c->N_Q_MATRICES; ++i) pk[i] += -1 * tk[i];
        for (int i = 0; i < c->N_Q_MATRICES; ++i) theta_vector[off + i] += pk[i];
    }
    if (!specifics::FAST_BOOTSTRAP) {
        fisher_matrix = std::make_unique<double[]>((size_t)ndim * ndim);
        std::fill_n(fisher_matrix.get(), (size_t)ndim * ndim, 0.0);
        for (const auto &c : chunks) {
            const int off = c->fisher_index_start - istart, nq = c->N_Q_MATRICES;
            for (int i = 0; i < nq; ++i)
                for (int j = i; j < nq; ++j)
                    fisher_matrix[j + (size_t)i * ndim + (size_t)(ndim + 1) * off] += c->fisher_matrix[j + (size_t)i * nq];
        }
    }
    chunks.clear();
}

// OneQSOEstimate::addBoot (core/one_qso_estimate.hpp:43-53)
void OneQSOEstimate::addBoot(int p, double *temppower, double *tempfisher) {
    const int T = bins::TOTAL_KZ_BINS;
    double *outfisher = tempfisher + (size_t)(T + 1) * istart;
    for (int i = 0; i < ndim; ++i)
        for (int j = i; j < ndim; ++j) outfisher[j + (size_t)i * T] += p * fisher_matrix[j + (size_t)i * ndim];
    for (int i = 0; i < ndim; ++i) temppower[istart + i] += p * theta_vector[i];
}

// OneQSOEstimate::addBootPowerOnly (core/one_qso_estimate.hpp:55-60): the
// rank-one update temppower[b][istart + i] += pcoeff[b] theta[i]
void OneQSOEstimate::addBootPowerOnly(int nboots, double *pcoeff, double *temppower) {
    const int T = bins::TOTAL_KZ_BINS;
    for (int b = 0; b < nboots; ++b)
        for (int i = 0; i < ndim; ++i) temppower[(size_t)b * T + istart + i] += pcoeff[b] * theta_vector[i];
}
