/*
 * lyspeq_qml.h -- C ABI of the MI355X per-chunk quadratic maximum-likelihood
 * (QML) engine.
 *
 * The engine replaces the body of lyspeq's Chunk / OneQSOEstimate
 * (core/chunk_estimate.hpp:46-119, core/one_qso_estimate.hpp:16-61): per
 * chunk it assembles the covariance from the S/Q lookup tables, factorises
 * it, applies continuum marginalisation and computes d_k, b_k, t_k and the
 * unhalved per-chunk Fisher matrix F_kk' = Tr(C^-1 Q_k C^-1 Q_k'), then adds
 * them into the driver's global arrays exactly as Chunk::oneQSOiteration
 * does (core/chunk_estimate.cpp:893-903).  Chunks are deferred: a driver
 * registers all of them once, then calls qml_iterate() once per
 * Newton-Raphson iteration where OneDQuadraticPowerEstimate::iterate runs
 * its chunk loop (core/quadratic_estimate.cpp:437-447), i.e. before the
 * 0.5 scaling at core/quadratic_estimate.cpp:456.
 *
 * Conventions: plain pointers and sizes, row-major doubles, no exceptions
 * across the boundary.  Every call returns QML_OK (0) or a negative status;
 * qml_last_error() returns the message of the calling thread's last failure.
 * One context is driven by one host thread.
 */
#ifndef LYSPEQ_QML_H
#define LYSPEQ_QML_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define QML_OK 0
#define QML_ERR_ARG -1          /* invalid argument or call order */
#define QML_ERR_HIP -2          /* HIP runtime failure or no device */
#define QML_ERR_SHORT_CHUNK -3  /* Chunk ctor's short-chunk test failed
                                   (core/chunk_estimate.cpp:172-187) */
#define QML_ERR_NO_TABLE -4     /* (R, dv) not among the tables
                                   (core/chunk_estimate.cpp:198-206) */
#define QML_ERR_EMPTY -5        /* all pixels masked (io/qso_file.cpp:150-153) */
#define QML_ERR_MEMORY -6       /* the chunk's matrices exceed the device scratch budget:
                                   Chunk::_setStoredMatrices' "Not all matrices are
                                   stored." (core/chunk_estimate.cpp:258-282); the
                                   chunk is not registered */
#define QML_ERR_LIMIT -7        /* a size beyond this build's limits */

/* Per-chunk status after an iteration (qml_get_chunk_result). */
#define QML_CHUNK_OK 0
#define QML_CHUNK_NOT_POSITIVE 1  /* factorisation failed: accumulation
                                     skipped, chunk kept, as for a LAPACK
                                     failure at core/chunk_estimate.cpp:905-913 */

/* Arithmetic of the dense products (W_k, Fisher, bias contractions). The
 * factorisation, the data-vector solve, d_k and all accumulation are FP64
 * in every mode. */
#define QML_PREC_FP64 0
#define QML_PREC_FP32 1    /* FP32 MFMA, FP32 accumulation */
#define QML_PREC_BF16X3 2  /* split-BF16 (hi*hi + hi*lo + lo*hi) MFMA, FP32 accumulation */
#define QML_PREC_BF16X4 3  /* split-BF16 with lo*lo as well: about half the error
                              of BF16X3 for a few percent more time */

typedef struct qml_ctx qml_ctx;

typedef struct {
    int device;            /* HIP device ordinal */
    int precision;         /* QML_PREC_* */
    double scratch_gb;     /* device scratch budget for one batch of chunks; 0 = default */
    int max_batch_chunks;  /* cap on chunks per batch; 0 = default */
} qml_opts;

/* bins:: and the specifics:: keys the chunk path reads
 * (core/global_numbers.hpp:17-22,55-213, core/global_numbers.cpp:78-168). */
typedef struct {
    int nk;                    /* NUMBER_OF_K_BANDS */
    int nz;                    /* NUMBER_OF_Z_BINS */
    const double *kband_edges; /* nk + 1 values (KBAND_EDGES) */
    const double *zbin_centers;/* nz values (ZBIN_CENTERS) */
    double z_bin_width;        /* Z_BIN_WIDTH */
    int z_binning_method;      /* 0 top hat, 1 triangular (RedshiftBinningMethod) */
    int cont_logl_order;       /* ContinuumLogLambdaMargOrder */
    int cont_lam_order;        /* ContinuumLambdaMargOrder */
    int turn_off_sfid;         /* TurnOffBaseline > 0 */
    double max_forest_length_v;/* MaximumForestLength / DynamicChunkNumber */
    double max_pixel_length_v; /* MaximumMeanPixelSize */
    int redshift_growth_on;    /* RedshiftGrowthOn > 0 (core/global_numbers.cpp:127): every stored
                                  Q of z bin m is scaled by ((1+z_ij)/(1+z_m))^B
                                  (Chunk::_scaleDerivativesWithRedshiftGrowth,
                                  core/chunk_estimate.cpp:539-573) */
    double fiducial_redshift_power; /* B, FiducialRedshiftPower (core/fiducial_cosmology.cpp:204);
                                  read only when redshift_growth_on */
} qml_bins;

/* One chunk: the pixel subset [i1, i2) of a spectrum after the file-level
 * preprocessing of OneQSOEstimate::_readQsoFile.  Pointers are read during
 * the call only. */
typedef struct {
    const double *wave;   /* observed wavelength, Angstrom */
    const double *delta;  /* flux fluctuations */
    const double *ivar;   /* inverse variance, 0 = masked */
    int size;             /* number of pixels */
    int R_fwhm;           /* spectrograph resolution (integer R) */
} qml_chunk_in;

/* A chunk's resolution matrix, in the reference's storage
 * (mathtools/matrix_helper.hpp:122-216), for the pixel range of the
 * qml_chunk_in it comes with (the engine cuts it with the masked boundary,
 * io/qso_file.cpp:82-85).
 *   QML_RMAT_DIA: mxhelp::DiaMatrix, values nelem x size; diagonal d holds
 *     R[c - (nelem/2 - d)][c] at column c.  Q and S are built on the pixel
 *     grid and sandwiched (Resolution::sandwich, DiaMatrix case).
 *   QML_RMAT_OVERSAMPLED: mxhelp::OversampledMatrix, values size x nelem;
 *     row i holds R[i][i*oversampling + a], on a fine grid of
 *     size*oversampling + nelem - 1 wavelengths spaced by the chunk's median
 *     dlambda / oversampling and starting nelem/2 fine steps before its first
 *     pixel (core/chunk_estimate.cpp:936-944). */
#define QML_RMAT_DIA 1
#define QML_RMAT_OVERSAMPLED 2
typedef struct {
    int kind;             /* QML_RMAT_* */
    int nelem;            /* diagonals or elements per row; odd */
    int oversampling;     /* QML_RMAT_OVERSAMPLED: factor >= 1 */
    const double *values; /* read during the call only */
} qml_rmat;

/* What the Chunk constructor derived (core/chunk_estimate.cpp:141-283). */
typedef struct {
    int size;               /* pixels after the masked-boundary cut */
    int real_size;          /* unmasked pixels */
    int res_index;          /* table index (RES_INDEX) */
    int kbin_upp;           /* KBIN_UPP */
    int zbin, zbin_low, zbin_upp;
    int n_q_matrices;       /* N_Q_MATRICES */
    int fisher_index_start; /* fisher_index_start */
    int n_stored;           /* stored Q_k count (kn < KBIN_UPP) */
    int n_templates;        /* continuum templates kept after the SVD cut */
    double dv_kms;          /* median pixel width */
    double lower_z, upper_z, median_z;
} qml_chunk_info;

/* Per-phase device time of the last qml_iterate, milliseconds, in the
 * categories of mytime:: (core/global_numbers.cpp:484-502). */
typedef struct {
    double assemble_ms;   /* v/z, S_fid, Q_k interpolation, C assembly  (T_Sfid + T_Q) */
    double invert_ms;     /* factorisation, inverse, solve, marginalisation (T_Cinv) */
    double weight_ms;     /* d_k pass, projected Q_k generation, W products (T_Qmod) */
    double fisher_ms;     /* Gram contraction and reductions (T_F) */
    double accumulate_ms; /* cross-chunk accumulation and copy-out */
    double total_ms;      /* first to last device event of the iteration */
    long long n_batches;
    /* device time per kernel, summed over batches, indexed by QML_K_* */
    double kernel_ms[13];
} qml_timings;

/* Indices into qml_timings.kernel_ms. */
#define QML_K_ASSEMBLE 0
#define QML_K_CHOLESKY 1
#define QML_K_TRINV 2
#define QML_K_SOLVE 3
#define QML_K_DPASS 4
#define QML_K_SMALL 5
#define QML_K_XGEN 6
#define QML_K_WHITEN_LEFT 7
#define QML_K_WHITEN_RIGHT 8
#define QML_K_GRAM 9
#define QML_K_FINALIZE 10
#define QML_K_ACCUMULATE 11
#define QML_K_RMAT 12       /* resolution-matrix sandwiches R Qhat R^T (part of assemble_ms) */

/* Context lifetime.  qml_create selects the device and creates a stream. */
int qml_create(qml_ctx **ctx, const qml_opts *opts);
int qml_destroy(qml_ctx *ctx);
const char *qml_last_error(void);

/* Binning; must precede tables and chunks.  Replaces the bins::/specifics::
 * globals read by Chunk. */
int qml_set_bins(qml_ctx *ctx, const qml_bins *bins);

/* One (R, dv) S/Q table, as SQLookupTable holds it after readTables()
 * (core/sq_table.cpp:325-368): deriv is nk x nv with v contiguous
 * (the DERIVATIVE HDU's byte order), signal is nv x nz with z contiguous
 * (the SIGNAL HDU's).  The v grid is v_j = j * v_length/(nv-1); the z grid is
 * ZBIN_CENTERS[0] - Z_BIN_WIDTH + j * Z_BIN_WIDTH*(NZBINS+1)/(nz-1).
 * Returns the table index in *r_index (the order of calls, like the rows of
 * the R-dv list read by ioh::readListRdv). */
int qml_add_table(qml_ctx *ctx, int R_fwhm, double dv_kms, int nv, double v_length,
                  int nz, const double *deriv, const double *signal, int *r_index);

/* Registers a chunk (the work of Chunk::Chunk).  On QML_ERR_SHORT_CHUNK,
 * QML_ERR_NO_TABLE, QML_ERR_EMPTY or QML_ERR_MEMORY the chunk is not
 * registered: the caller skips it, as OneQSOEstimate does
 * (core/one_qso_estimate.cpp:115-125).
 * *chunk_id receives the chunk's index. */
int qml_add_chunk(qml_ctx *ctx, const qml_chunk_in *in, qml_chunk_info *info, int *chunk_id);

/* Uploads all registered chunks, plans the size-bucketed batches and
 * allocates device scratch.  Inputs stay resident in device memory. */
int qml_commit(qml_ctx *ctx);

/* qml_add_chunk with a resolution matrix (specifics::USE_RESOLUTION_MATRIX):
 * the chunk uses table 0 for S_fid (RES_INDEX = 0), Q_k is the analytic band
 * window (dk/pi) cos(k_c v) sinc(dk v/2) on the matrix's grid, and both are
 * sandwiched by R (core/chunk_estimate.cpp:189-193,400-480).  Limit:
 * 15*oversampling + nelem <= 96 (diagonal storage: oversampling 1), e.g. 11
 * diagonals oversampled x3 (31 elements per row) give 76.  Beyond it the
 * call returns QML_ERR_LIMIT. */
int qml_add_chunk_rmat(qml_ctx *ctx, const qml_chunk_in *chunk, const qml_rmat *rmat, qml_chunk_info *info,
                       int *chunk_id);

/* Read-time preprocessing of a whole spectrum's resolution matrix, before it
 * is cut into chunks (core/one_qso_estimate.cpp:74-89).  Host code; no
 * context or device is needed.  values is the diagonal storage (ndiags x
 * size, as QML_RMAT_DIA); a row running off the matrix is completed with its
 * own values mirrored about the diagonal (DiaMatrix::getRow,
 * mathtools/matrix_helper.cpp:608-638).
 *
 * qml_rmat_deconvolve: DiaMatrix::deconvolve(m) (mathtools/matrix_helper.cpp:
 * 653-682): each row, zero-padded by 5 on both sides, has its Fourier modes
 * j >= 1 divided by sinc(pi m j / (ndiags + 10)) (RealField::deconvolveSinc,
 * mathtools/real_field.cpp:95-111); values is updated in place.  m > 0.
 *
 * qml_rmat_oversample: Resolution::oversample(oversampling) (mathtools/
 * matrix_helper.cpp:1029-1073): each row, shifted by (its minimum - its
 * smallest nonzero magnitude), is interpolated in log by a natural cubic
 * spline at oversampling points per pixel, shifted back and normalised to
 * unit sum.  rows receives size x (2 (ndiags/2) oversampling + 1) values, the
 * QML_RMAT_OVERSAMPLED storage of the whole spectrum; a chunk's rows are the
 * slice for its pixels. */
int qml_rmat_deconvolve(int size, int ndiags, double m, double *values);
int qml_rmat_oversample(int size, int ndiags, const double *values, int oversampling, double *rows);

/* The S/Q lookup tables of (R, dv) pairs, computed on the device: the work of
 * SQLookupTable::computeTables (core/sq_table.cpp:201-323), i.e. of the
 * reference's table generator, for many pairs at once.  Grids and layouts as
 * there and as qml_add_table takes them: v_n = n L_V / (N_V - 1); z_m = z_c0 -
 * dz + m dz (nz + 1) / (N_Z - 1) from the bins of qml_set_bins (call it first).
 *   deriv : npairs x nk x N_V, Q_kn(v) = (1/pi) int_band W(k)^2 cos(k v) dk with
 *           W = sinc(k dv/2) exp(-(k R)^2/2), R = c / R_fwhm / 2.35482004503 km/s
 *           (core/fiducial_cosmology.cpp:305-331); Filon's rule over
 *           2 filon_panels panels per band (the reference: QAWO);
 *   signal: npairs x N_V x N_Z, the reference's real-FFT evaluation of
 *           W^2 P_fid(k, z) (core/sq_table.cpp:221-227,296-311), P_fid the
 *           PD13 fit with Lorentzian (core/fiducial_cosmology.cpp:273-282).
 * Either output may be NULL. */
typedef struct {
    int nv;             /* NUMBER_OF_VELOCITY_POINTS (N_V >= 2) */
    double v_length;    /* VELOCITY_LENGTH, km/s */
    int nz;             /* NUMBER_OF_Z_POINTS (N_Z >= 2) */
    double pd13[6];     /* FIDUCIAL_PD13_PARAMS: A, n, alpha, B, beta, lambda */
    int filon_panels;   /* M: 2M Filon panels per k band, 1..128; 0 = 64 */
} qml_table_spec;
int qml_make_tables(qml_ctx *ctx, const qml_table_spec *spec, int npairs, const int *R_fwhm, const double *dv_kms,
                    double *deriv, double *signal);

/* One Newton-Raphson iteration's chunk loop.  theta is TOTAL_KZ_BINS values
 * (the ps_estimate argument of Chunk::oneQSOiteration).  Adds every chunk's
 * unhalved Fisher matrix into the upper triangle of fisher_sum (T x T,
 * row-major) and d, b, t into dbt_sum (3 x T), exactly like
 * core/chunk_estimate.cpp:893-903. */
int qml_iterate(qml_ctx *ctx, const double *theta, double *fisher_sum, double *dbt_sum);

/* The last iteration's result for one chunk: fisher is n_q x n_q (upper
 * triangle filled, like Chunk::fisher_matrix), dbt is 3 x n_q (like
 * Chunk::dbt_estimate_before_fisher_vector).  Either pointer may be NULL. */
int qml_get_chunk_result(qml_ctx *ctx, int chunk_id, double *fisher, double *dbt, int *status);

int qml_get_timings(qml_ctx *ctx, qml_timings *t);

/* The matrices of Chunk's protected builders for one committed chunk,
 * computed on the device from the chunk's registered pixels and tables and
 * written to out (row-major n x n, both triangles; the reference fills the
 * upper one and mirrors the matrices it reads in full).  which:
 *   QML_CMAT_V        v_ij = c ln(l_j / l_i) for j >= i   Chunk::_setVZMatrices
 *   QML_CMAT_Z        z_ij = sqrt(l_i l_j) - 1              (core/chunk_estimate.cpp:384-398)
 *                     on the matrix grid: the oversampled grid of an
 *                     oversampled resolution matrix (n = its column count,
 *                     core/chunk_estimate.cpp:929-944), else the pixels
 *   QML_CMAT_SFID     the fiducial signal matrix, R S R^T with a resolution
 *                     matrix          Chunk::_setFiducialSignalMatrix (:400-426)
 *   QML_CMAT_Q        Q of chunk-local index arg = i_kz before the redshift
 *                     weights, R Q R^T with a resolution matrix
 *                                     Chunk::_setQiMatrix (:429-480)
 *   QML_CMAT_ZWEIGHT  the weights w_zm(z_ij) of absolute z bin arg = zm that
 *                     Chunk::_applyRedshiftInterp multiplies in (:483-536)
 *   QML_CMAT_GROWTH   ((1 + z_ij) / (1 + z_zm))^B of z bin arg = zm, B the
 *                     bins' growth exponent
 *                     Chunk::_scaleDerivativesWithRedshiftGrowth (:539-573)
 * The pixel-grid matrices have n = the chunk's size.  With out NULL only *n
 * is set.  QML_ERR_ARG for a chunk without signal table (SFID), an i_kz
 * outside the chunk or of a band at or above KBIN_UPP (Q), or a z bin
 * outside the bins.  Not for use while an iteration is running on ctx. */
#define QML_CMAT_V 0
#define QML_CMAT_Z 1
#define QML_CMAT_SFID 2
#define QML_CMAT_Q 3
#define QML_CMAT_ZWEIGHT 4
#define QML_CMAT_GROWTH 5
int qml_chunk_matrix(qml_ctx *ctx, int chunk_id, int which, int arg, double *out, int *n);

/* Chunk::_addMarginalizations (core/chunk_estimate.cpp:596-665) on the
 * device: cinv (the chunk's size squared, row-major; the upper triangle is
 * read) loses each of the chunk's orthonormal continuum templates by a
 * Sherman-Morrison update, cinv <- cinv - y y^T / (u^T y) with y = cinv u,
 * in place.  The templates are those the chunk was registered with (the
 * SVD-orthonormalised set, small singular values dropped). */
int qml_chunk_marginalize(qml_ctx *ctx, int chunk_id, double *cinv);

/* Leaves chunk chunk_id out of (active = 0) or back in (1) the sums that
 * qml_iterate accumulates; its own results are still computed.  The
 * counterpart of the reference dropping a Chunk object from
 * OneQSOEstimate::chunks (core/one_qso_estimate.cpp:176-197). */
int qml_set_chunk_active(qml_ctx *ctx, int chunk_id, int active);

/* Stage overlap of qml_iterate (default on): the factorisation and solves of
 * batch b+1 run on one HIP stream while the dense products of batch b run on
 * a second.  Off serialises both stages on one stream, so that the per-kernel
 * durations in qml_timings are those of each kernel alone.  Results are
 * identical either way.  No counterpart in the reference (its chunk loop is
 * serial per MPI rank). */
int qml_set_pipeline(qml_ctx *ctx, int on);

/* Graph replay of qml_iterate (default off).  On, the next qml_iterate
 * captures its whole launch sequence (the theta upload, both stages of every
 * batch on their two streams, the accumulation and the copy-out of the sums)
 * into a HIP graph, and every later iteration replays that graph with the new
 * theta, so the host issues one launch per iteration.  The graph is captured
 * again when qml_set_pipeline changes the stage overlap, and released by
 * qml_set_graph(ctx, 0) or qml_destroy.  Results are identical to those of
 * eager launches.  A replayed iteration reports total_ms only in its
 * qml_timings (the per-stage events are not recorded).  No counterpart in the
 * reference. */
int qml_set_graph(qml_ctx *ctx, int on);

/* Poisson bootstrap of the last iteration's per-chunk results, in the
 * reference's FastBootstrap form (specifics::FAST_BOOTSTRAP, the default;
 * PoissonBootstrapper::_fastBootstrap and _meanBootstrap, core/bootstrapper.hpp:
 * 183-221,282-375,392-399; OneQSOEstimate::collapseBootstrap and
 * addBootPowerOnly, core/one_qso_estimate.cpp:200-239,
 * core/one_qso_estimate.hpp:55-60).
 *   qso_of_chunk: for every registered chunk, its spectrum's position in the
 *     driver's queue (0 .. nqso-1), or -1 to leave the chunk out.  Every
 *     spectrum 0 .. nqso-1 is one the reference's queue still holds.
 *   Spectrum q's collapsed vector theta_q (ndim values at istart) is the sum,
 *   in chunk order, of its chunks' (d - b) - t at their fisher_index_start.
 *   Its weights are draws q nboots .. (q+1) nboots - 1 of qml_poisson_draws(seed).
 *   realizations (nboots x TOTAL_KZ_BINS, may be NULL): R[b] = sum over q of
 *     p_qb theta_q (not normalised: 0.5 solver R[b] is a power realisation);
 *   mean (T) and cov (T x T): the mean of the realisations and their sample
 *     covariance with 1/(nboots - 1), as _calcuateMean / _calcuateCovariance.
 *     The reference then writes 0.25 cov as the bootstrap Fisher matrix and
 *     sandwiches it with the solver (_saveData, _sandwichInvFisher).
 * Weights, sums and their order follow the reference; the seed it uses is the
 * MPI rank (0 on one process). */
int qml_bootstrap(qml_ctx *ctx, int nboots, unsigned long long seed, int nqso, const int *qso_of_chunk,
                  double *realizations, double *mean, double *cov);

/* The bootstrap's complete form (FastBootstrap off; PoissonBootstrapper::
 * _oneBoot, core/bootstrapper.hpp:223-262, OneQSOEstimate::addBoot,
 * core/one_qso_estimate.hpp:43-53): for every realisation b, one weight per
 * spectrum in queue order (draws b nqso .. (b+1) nqso - 1 of
 * qml_poisson_draws(seed), or weights[b nqso + q] if weights is not NULL),
 * F_b = sum of p F_q over the spectra with p != 0 (F_q the spectrum's chunk
 * Fisher matrices summed at their fisher_index_start) and the power
 * realisation x_b solving F_b x = sum of p theta_q by Cholesky, bands without
 * data left at zero (LAPACKE_safeSolveCho, mathtools/matrix_helper.cpp:
 * 385-396).  A realisation whose F_b is not positive definite is invalid.
 *   realizations (nboots x T, may be NULL): the valid x_b, in order, first
 *     *nvalid rows; valid (nboots, may be NULL): 1 for a valid realisation;
 *   mean, cov: over the valid realisations, as for qml_bootstrap.
 * With unit weights every x_b is the driver's power estimate
 * F_sum^-1 (d - b - t) of the included chunks. */
int qml_bootstrap_complete(qml_ctx *ctx, int nboots, unsigned long long seed, int nqso, const int *qso_of_chunk,
                           const int *weights, double *realizations, int *valid, int *nvalid, double *mean,
                           double *cov);

/* The median bootstrap of saved realisations (PoissonBootstrapper::
 * _medianBootstrap, core/bootstrapper.hpp:301-346,376-390, reached through
 * PoissonBootstrapper::run() as lyspeqBootStats.cpp:35-36 calls it; medians
 * by stats::medianOfUnsortedVector, mathtools/stats.cpp:8-28).
 *   realizations: nboots x T, one realisation per row (host);
 *   median (T): per band the median over the realisations: the middle value,
 *     or the half-sum of the two middle values for even nboots;
 *   mad_cov (T x T): for j >= i the median over the realisations of
 *     (R[b][i] - median[i]) (R[b][j] - median[j]), mirrored to j < i.
 * The selected values are the sorted arrays' bit for bit.  Needs a context
 * only for its device and stream (no chunks or commit); T <= 65535. */
int qml_bootstrap_median(qml_ctx *ctx, int nboots, int T, const double *realizations, double *median,
                         double *mad_cov);

/* qml_bootstrap_median without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 *   d_realizations (nboots x T), d_median (T), d_mad_cov (T x T): device
 *     memory of the context's device, owned by the caller; results as
 *     qml_bootstrap_median's, bit for bit.
 * The transposed realisations live in a static device array of 2^23 doubles
 * (64 MiB) per device, set up at qml_create: nboots x T above 8388608 (e.g.
 * 20000 x 455) is rejected with QML_ERR_LIMIT.  Every call on one device
 * shares that array, so calls that overlap on one device (two threads, or two
 * contexts) race.  Runs on the context's stream and returns when the results
 * are in place. */
int qml_bootstrap_median_noalloc(qml_ctx *ctx, int nboots, int T, const double *d_realizations,
                                 double *d_median, double *d_mad_cov);

/* n unit-mean Poisson draws of PoissonRNG(seed) (core/bootstrapper.hpp:19-37):
 * std::poisson_distribution<int> over std::mt19937_64 seeded once.  With the
 * same C++ standard library as the reference's build the draws are the
 * reference's.  Host code; no context or device is needed. */
int qml_poisson_draws(unsigned long long seed, long long n, int *out);

/* Autotuned batch schedule.  For each candidate cap on chunks per batch
 * (caps[0..ncaps-1]), the committed chunks are planned again into
 * size-bucketed batches (sorted by padded size, cut at the cap or at the
 * scratch budget of qml_opts, and with or without a break at every change of
 * padded size), and reps + 1 iterations are run at theta, the first as
 * warm-up; the plan with the shortest iteration is kept.  In the
 * split-BF16 precisions the staging of the three products' operands is tuned
 * next, on that plan, one product at a time (left, right, Gram;
 * qml_set_kernel_stages), each choice timed the same way.  The chosen cap and
 * the best time (ms) are returned; qml_get_schedule reports the whole choice.
 * The per-chunk results do not depend on any of these choices.  The
 * accumulated sums of these iterations are discarded.  No counterpart in the
 * reference (its chunk loop is serial per MPI rank). */
int qml_autotune(qml_ctx *ctx, const double *theta, const int *caps, int ncaps, int reps, int *best_cap,
                 double *best_ms);

/* Staging of the split-BF16 products' operands into LDS, for the left and
 * right whitening products and the Gram:
 *   1  through registers, one LDS stage (the next stage waits in registers;
 *      three workgroups per CU);
 *   2  through registers, two LDS stages (two workgroups per CU); not for
 *      the right product;
 *   3  direct-to-LDS loads (global_load_lds), two stages (two per CU);
 *   4  direct-to-LDS loads, three stages (one per CU).
 * Defaults 3, 3 and 3.  Results are identical for every choice.  No
 * counterpart in the reference. */
int qml_set_kernel_stages(qml_ctx *ctx, int left, int right, int gram);

/* Streams each chunk's stack of weighted matrices in slabs of `members`
 * (0, the default: the whole stack at once).  The stack and its left
 * products then take members x Np^2 elements each per chunk instead of
 * (n + 2) x Np^2; each slab's whitened matrices go to a store of packed upper
 * tiles, and one Gram runs over the store.  A chunk's device scratch is
 * bounded by the slab instead of growing with the stored Q count (N ~ 2000
 * with ~150 stored Q: about 4 GB in FP64 instead of 10.6).  The results are
 * identical.  Set before registering chunks (their memory check uses it) or
 * any time after qml_commit (the batches are planned again).  The reference's
 * counterpart is its bounded number of stored Q matrices
 * (Chunk::_setStoredMatrices, core/chunk_estimate.cpp:258-282). */
int qml_set_stack_slab(qml_ctx *ctx, int members);

/* The current launch schedule. */
typedef struct {
    int max_batch_chunks;  /* cap on chunks per batch of the current plan */
    int n_batches;         /* batches of the current plan */
    int stages_left;       /* operand staging of the split-BF16 left product (qml_set_kernel_stages) */
    int stages_gram;       /* ... of the split-BF16 Gram */
    int pipeline;          /* two-stream stage overlap (qml_set_pipeline) */
    int graph;             /* graph replay (qml_set_graph) */
    int stack_slab;        /* members per stack slab, 0 for the whole stack (qml_set_stack_slab) */
    double max_chunk_gb;   /* the largest chunk's device scratch in the current plan, GiB */
    int stages_right;      /* operand staging of the split-BF16 right product */
    int np_buckets;        /* 1 if a batch also ends where the padded size changes (qml_autotune) */
} qml_schedule;
int qml_get_schedule(qml_ctx *ctx, qml_schedule *schedule);

/* Batch planning: 1 also ends a batch where the chunks' padded size changes
 * (every batch then holds one size), 0 (the default) only at the cap on
 * chunks per batch and at the scratch budget.  Set before qml_commit or any
 * time after (the batches are planned again).  qml_autotune tries both.
 * Results are identical.  No counterpart in the reference. */
int qml_set_np_buckets(qml_ctx *ctx, int on);

/* Batch `batch` (0 .. n_batches - 1, in launch order) of the current plan:
 * its chunk count, the range of its chunks' padded sizes (the plan's
 * Np bucket edges: chunks are planned in decreasing Np) and its largest
 * stack member count. */
typedef struct {
    int n_chunks;
    int np_min, np_max;
    int members_max;
} qml_batch_info;
int qml_get_batch(qml_ctx *ctx, int batch, qml_batch_info *info);

/* Number of chunks registered, and TOTAL_KZ_BINS. */
int qml_num_chunks(qml_ctx *ctx);
int qml_total_kz_bins(qml_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* LYSPEQ_QML_H */
