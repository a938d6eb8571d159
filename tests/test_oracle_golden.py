"""Pins the CPU oracle against the reference's own golden vectors and against
closed forms the mathematics fixes."""
import numpy as np
import pytest

import goldencase as gc
from oracle.bins import Bins, redshift_weight
from oracle.chunk import Chunk, TableSet, LapackError
from oracle.driver import power_estimates
from oracle import interp
from oracle.tables import derivative_table_qawo, signal_table_fft


def _golden_chunk(**over):
    kw = dict(gc.BINS_KW)
    kw.update(over)
    b = Bins(**kw)
    t = gc.table()
    ts = [TableSet(70000, 20.0, t.nvpoints, t.v_length, t.nzpoints, t.deriv, t.signal)]
    sp = gc.spectrum()
    return b, Chunk(b, ts, sp.wave, sp.flux, sp.ivar, sp.R_fwhm)


def test_chunk_metadata():
    b, ch = _golden_chunk()
    assert ch.N == 488 and ch.n_q == 10 and ch.fisher_index_start == 0
    assert ch.kbin_upp == 5 and ch.stored == list(range(10))


def test_signal_and_q0_matrices_all_entries():
    # tests/testSQCMatrices.cpp:38-95 compares row 0 only; here all 488^2
    b, ch = _golden_chunk()
    v, z = ch.vz_matrices()
    s_gold, q0_gold = gc.chunk_matrices()
    assert gc.all_close(ch.signal_matrix(v, z), s_gold)
    assert gc.all_close(ch.q_matrices(v, z)[0], q0_gold)


def test_end_to_end_fisher_theta_sigma_dbt():
    b, ch = _golden_chunk()
    T = b.total_kz_bins
    F, dbt = np.zeros((T, T)), np.zeros((3, T))
    assert ch.one_iteration(np.zeros(T), F, dbt)
    res = power_estimates(F, dbt)
    rows = gc.detailed_rows()
    assert gc.all_close(res["fisher"], gc.golden_fisher())
    assert gc.all_close(res["inverse"], gc.golden_inverse_fisher())
    assert gc.all_close(rows[:, 0], b.zbin_centers[np.arange(T) // b.nk])
    assert gc.all_close(rows[:, 1], b.kband_edges[np.arange(T) % b.nk])
    assert gc.all_close(rows[:, 3], b.kband_centers[np.arange(T) % b.nk])
    assert gc.all_close(rows[:, 5], res["theta"])
    assert gc.all_close(rows[:, 7], res["sigma"])
    for i in range(3):
        assert gc.all_close(rows[:, 8 + i], res["weighted"][i])
        assert gc.all_close(rows[:, 11 + i], dbt[i])


def test_tables_against_golden_fits():
    b = Bins(**gc.BINS_KW)
    t = gc.table()
    s = signal_table_fft(b, 70000, 20.0, t.nvpoints, t.v_length, t.nzpoints, gc.PD13)
    assert gc.all_close(s, t.signal)
    q = derivative_table_qawo(b, 70000, 20.0, t.nvpoints, t.v_length)
    assert gc.all_close(q, t.deriv)


def test_lu_inverse_against_truth(golden_dir):
    from scipy.linalg import lapack
    d = np.load(f"{golden_dir}/lu_inverse.npz")
    lu, piv, info = lapack.dgetrf(d["matrix"])
    inv, info2 = lapack.dgetri(lu, piv)
    assert info == 0 and info2 == 0
    assert gc.all_close(inv, d["inverse"])


def test_svd_jobu_o_against_truth(golden_dir):
    # tests/cblas_tests.cpp:285-302: col-major 6x5, jobu='O'
    from scipy.linalg import lapack
    d = np.load(f"{golden_dir}/cblas_inline.npz")
    a = d["matrix_for_SVD_A"].reshape(5, 6).T          # col-major m=6, n=5
    u, s, vt, info = lapack.dgesvd(a.copy(order="F"), compute_uv=1, full_matrices=0)
    assert info == 0
    got = u.T.reshape(-1)                              # back to the col-major buffer
    assert gc.all_close(got, d["truth_SVD_A"])


def test_redshift_weights_partition_of_unity():
    b = Bins(**dict(gc.BINS_KW, nz=5))
    z = np.linspace(b.zbin_centers[0] - 0.05, b.zbin_centers[-1] + 0.05, 2001)
    w = sum(redshift_weight(b, z, m) for m in range(b.nz))
    assert np.allclose(w, 1.0, rtol=0, atol=1e-12)
    zc, dz = b.zbin_centers[2], b.z_bin_width
    wi = redshift_weight(b, np.array([zc - dz, zc - dz / 2, zc, zc + dz / 4, zc + 1.01 * dz]), 2)
    assert np.allclose(wi, [0.0, 0.5, 1.0, 0.75, 0.0], atol=1e-12)
    bt = Bins(**dict(gc.BINS_KW, nz=3, method=0))
    wt = redshift_weight(bt, np.array([1.999, 2.099, 2.101]), 1)
    assert np.array_equal(wt, [1.0, 1.0, 0.0])


def test_interp_closed_forms():
    y = np.array([1.0, 3.0, 2.0, 5.0])
    assert np.allclose(interp.evaluate_1d([0, 1, 2, 3], 0.0, 1.0, y), y)
    assert np.isclose(interp.evaluate_1d(0.25, 0.0, 1.0, y), 1.5)
    assert np.isclose(interp.evaluate_1d(4.0, 0.0, 1.0, y), 8.0)     # extrapolates
    assert np.isclose(interp.evaluate_1d(-1.0, 0.0, 1.0, y), -1.0)
    zz = np.arange(12.0) ** 2                                          # nx=4, ny=3
    f = interp.evaluate_2d(1.5, 0.5, 0.0, 1.0, 0.0, 1.0, zz, 4, 3)
    g = 0.25 * (zz[1] + zz[2] + zz[5] + zz[6])
    assert np.isclose(f, g)


def test_special_case_noise_only_closed_form():
    """S off, theta = 0, no marginalisation: C = N, so
    F_kk' = Tr(N^-1 Q_k N^-1 Q_k'), b_k = sum_i ivar_i Q_k,ii, t_k = 0."""
    b, ch = _golden_chunk(turn_off_sfid=True)
    T = b.total_kz_bins
    F, dbt = np.zeros((T, T)), np.zeros((3, T))
    assert ch.one_iteration(np.zeros(T), F, dbt)
    v, z = ch.vz_matrices()
    qs = ch.q_matrices(v, z)
    iv = ch.ivar
    for a in range(len(qs)):
        wa = iv[:, None] * qs[a]
        assert np.isclose(dbt[1, a], np.sum(iv * np.diag(qs[a])), rtol=1e-12)
        for c in range(a, len(qs)):
            wc = iv[:, None] * qs[c]
            assert np.isclose(F[a, c], np.sum(wa * wc.T), rtol=1e-11)
    assert np.all(dbt[2] == 0)


def test_marginalisation_invariants():
    """C_marg^-1 V = 0 and equality with C^-1 - C^-1 V (V^T C^-1 V)^-1 V^T C^-1."""
    b, ch = _golden_chunk(cont_logl_order=2, cont_lam_order=1)
    v, z = ch.vz_matrices()
    s = ch.signal_matrix(v, z)
    qs = ch.q_matrices(v, z)
    theta = np.zeros(b.total_kz_bins)
    b0 = Bins(**dict(gc.BINS_KW))
    ch.bins = b0
    cinv = ch.inverse_covariance(s, qs, theta)
    ch.bins = b
    cm = ch.add_marginalizations(cinv)
    V = ch.templates().T
    assert np.max(np.abs(cm @ V)) < 1e-10 * np.max(np.abs(cinv @ V))
    # the projector onto the templates' span, written with an orthonormal
    # basis of that span (the 4 templates here are nearly degenerate, which
    # makes V^T C^-1 V itself too ill-conditioned to solve with directly)
    U, _ = np.linalg.qr(V)
    x = cinv @ U
    wood = cinv - x @ np.linalg.solve(U.T @ x, x.T)
    assert np.max(np.abs(cm - wood)) < 1e-9 * np.max(np.abs(cinv))


def test_lu_vs_cholesky_floor():
    """The SPD (Cholesky) form of the same inverse agrees with the LU form to
    the cond*eps floor; this is the floor the GPU FP64 path is held to."""
    b, ch = _golden_chunk(cont_logl_order=1, cont_lam_order=-1)
    T = b.total_kz_bins
    F1, d1 = np.zeros((T, T)), np.zeros((3, T))
    F2, d2 = np.zeros((T, T)), np.zeros((3, T))
    ch.one_iteration(np.zeros(T), F1, d1, mode="lu")
    ch.one_iteration(np.zeros(T), F2, d2, mode="cholesky")
    assert np.max(np.abs(F1 - F2) / np.abs(F1).max()) < 1e-12
    assert np.max(np.abs(d1 - d2) / np.abs(d1).max(axis=1, keepdims=True)) < 1e-11


def _gauss_jordan_longdouble(a):
    n = a.shape[0]
    m = np.concatenate([a.astype(np.longdouble), np.eye(n, dtype=np.longdouble)], axis=1)
    for c in range(n):
        p = c + int(np.argmax(np.abs(m[c:, c])))
        m[[c, p]] = m[[p, c]]
        m[c] /= m[c, c]
        for r in range(n):
            if r != c:
                m[r] -= m[r, c] * m[c]
    return m[:, n:]


def test_brute_force_tiny_chunk_long_double():
    """N=12 pixels: the oracle's LU-based C^-1 and F against a long-double
    Gauss-Jordan inverse and the trace definition."""
    b = Bins(**dict(gc.BINS_KW, max_forest_length_v=400.0, max_pixel_length_v=1e9))
    t = gc.table()
    ts = [TableSet(70000, 20.0, t.nvpoints, t.v_length, t.nzpoints, t.deriv, t.signal)]
    sp = gc.spectrum()
    ch = Chunk(b, ts, sp.wave, sp.flux, sp.ivar, sp.R_fwhm, 100, 112)
    assert ch.N == 12
    v, z = ch.vz_matrices()
    s = ch.signal_matrix(v, z)
    qs = ch.q_matrices(v, z)
    theta = np.zeros(b.total_kz_bins)
    cinv = ch.inverse_covariance(s, qs, theta)
    c = s.astype(np.longdouble) + np.diag(1.0 / ch.ivar.astype(np.longdouble))
    exact = _gauss_jordan_longdouble(c)
    assert np.max(np.abs(cinv - exact.astype(np.float64))) < 1e-12 * np.max(np.abs(cinv))
    T = b.total_kz_bins
    F, dbt = np.zeros((T, T)), np.zeros((3, T))
    ch.one_iteration(theta, F, dbt)
    for a in range(len(qs)):
        for cc in range(a, len(qs)):
            wa = exact @ qs[a].astype(np.longdouble)
            wc = exact @ qs[cc].astype(np.longdouble)
            ref = float(np.trace(wa @ wc))
            assert abs(F[ch.stored[a], ch.stored[cc]] - ref) <= 1e-11 * abs(F).max()


# ---------------------------------------------------------------- resolution matrix
def _rmat_chunk(kind="dia", nelem=1, values=None, scale=1.0):
    from lyspeq_amd import synth
    import parity as P
    from oracle.chunk import Chunk as OChunk, ResoMatrix
    w = synth.rmat_config(nchunks=1, n_range=(110, 110), kind="dia")
    c = w.chunks[0]
    n = len(c["wave"])
    if values is None:
        values = np.ones((nelem, n)) if kind == "dia" else None
    rm = ResoMatrix(kind, scale * values, nelem, 1 if kind == "dia" else 3)
    b = P.oracle_bins(w.binning)
    ivar = np.ones(n)   # no masked boundary: the matrix is used as given
    return OChunk(b, P.oracle_tables(w), c["wave"], c["delta"], ivar, c["R"], rmat=rm), b


def test_rmat_identity_reproduces_the_analytic_band_window():
    """R = I: Q_kn is the analytic window (dk/pi) cos(kc v) sinc(dk v / 2) on the
    pixel grid (core/chunk_estimate.cpp:444-454)."""
    ch, b = _rmat_chunk("dia", 1)
    v, z = ch.vz_matrices()
    vv = v + v.T - np.diag(np.diag(v))
    for kn in (0, 7, 20, 34):
        q = ch.q_base(v, kn)
        kc, dk = b.kband_centers[kn], b.kband_edges[kn + 1] - b.kband_edges[kn]
        with np.errstate(invalid="ignore", divide="ignore"):
            x = dk * vv / 2
            want = np.where(vv == 0, dk / np.pi, dk / np.pi * np.cos(kc * vv) * np.sin(x) / x)
        np.testing.assert_allclose(q, want, rtol=1e-13, atol=1e-16)


def test_rmat_band_windows_telescope():
    """sum_kn Qhat_kn(v) = (sin(k_end v) - sin(k_0 v)) / (pi v): pins k_c and dk
    (centres at the arithmetic mid-points, core/global_numbers.cpp:240-242)."""
    ch, b = _rmat_chunk("dia", 1)
    v, _ = ch.vz_matrices()
    vv = v + v.T - np.diag(np.diag(v))
    tot = sum(ch.q_base(v, kn) for kn in range(b.nk))
    k0, k1 = b.kband_edges[0], b.kband_edges[-1]
    off = vv != 0
    np.testing.assert_allclose(tot[off], (np.sin(k1 * vv[off]) - np.sin(k0 * vv[off])) / (np.pi * vv[off]),
                               rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(np.diag(tot), (k1 - k0) / np.pi, rtol=1e-13)


def test_rmat_sandwich_is_quadratic_in_r_and_samples_with_a_delta_row():
    """Q(2R) = 4 Q(R) for a Gaussian R; an oversampled R whose rows are a unit
    spike at the row's centre samples Qhat at the pixels themselves."""
    from lyspeq_amd import synth
    w = synth.rmat_config(nchunks=1, n_range=(110, 110), kind="dia")
    vals = w.chunks[0]["rmat"]["values"]
    ch1, _ = _rmat_chunk("dia", 11, vals)
    ch2, _ = _rmat_chunk("dia", 11, vals, scale=2.0)
    v, _ = ch1.vz_matrices()
    np.testing.assert_allclose(ch2.q_base(v, 5), 4.0 * ch1.q_base(v, 5), rtol=1e-12, atol=1e-18)
    n = len(w.chunks[0]["wave"])
    spike = np.zeros((n, 31))
    spike[:, 15] = 1.0
    chs, b = _rmat_chunk("oversampled", 31, spike)
    chi, _ = _rmat_chunk("dia", 1)
    for kn in (3, 30):
        np.testing.assert_allclose(chs.q_base(v, kn), chi.q_base(v, kn), rtol=1e-7, atol=1e-12)


def test_rmat_storage_against_reference_matrix_products(golden_dir):
    """The oracle's reading of both resolution-matrix storages against the
    products the reference's own matrix tests hold (tests/cblas_tests.cpp:
    303-427 OversampledMatrix multiplyLeft / sandwich, 429-530 DiaMatrix
    multiply in the four side/transpose combinations)."""
    from oracle.chunk import ResoMatrix
    d = np.load(f"{golden_dir}/resolution_inline.npz")
    R = ResoMatrix("dia", d["diamatrix_diagonals"].reshape(5, 7), 5).dense(7)
    B = d["diamatrix_multiplier_B"].reshape(7, 7)
    for name, got in (("LN", R @ B), ("LT", R.T @ B), ("RN", B @ R), ("RT", B @ R.T)):
        np.testing.assert_allclose(got, d[f"truth_diamatrix_{name}_multiplication"].reshape(7, 7),
                                   rtol=0, atol=1e-12)
    # the row view: row i holds R[i][i - 2 .. i + 2]; entries outside the matrix are not used
    rows = d["diamatrix_rows"].reshape(7, 5)
    for i in range(7):
        for t in range(5):
            if 0 <= i - 2 + t < 7:
                assert R[i, i - 2 + t] == rows[i, t]
    Ro = ResoMatrix("oversampled", d["oversamp_values"].reshape(7, 5), 5, 2).dense(7)
    A = d["oversample_multiplier_A"].reshape(18, 18)
    assert Ro.shape == (7, 18)
    # the truths are printed to 6 significant digits
    np.testing.assert_allclose(Ro @ A, d["truth_oversample_left_multiplication"].reshape(7, 18), rtol=1e-5)
    np.testing.assert_allclose(Ro @ A @ Ro.T, d["truth_oversample_right_multiplication"].reshape(7, 7),
                               rtol=1e-5)
    # the banded row-by-row sandwich the oracle's chunks use (tests/cblas_tests.cpp:417-418
    # checks OversampledMatrix::sandwich against the same truth)
    Rs = ResoMatrix("oversampled", d["oversamp_values"].reshape(7, 5), 5, 2)
    np.testing.assert_allclose(Rs.sandwich(A, 7), d["truth_oversample_right_multiplication"].reshape(7, 7),
                               rtol=1e-5)
    # and the dense product on random symmetric fine matrices, oversampling 1-3
    rng = np.random.default_rng(17)
    for os_, nel, n in ((1, 5, 9), (2, 5, 7), (3, 31, 23)):
        ncols = n * os_ + nel - 1
        a = rng.normal(size=(ncols, ncols))
        a = a + a.T
        Rr = ResoMatrix("oversampled", rng.uniform(0, 1, (n, nel)), nel, os_)
        want = Rr.dense(n) @ a @ Rr.dense(n).T
        np.testing.assert_allclose(Rr.sandwich(a, n), want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_rmat_rows_with_mirrored_edges_against_reference_rows(golden_dir):
    """DiaMatrix::getRow, edge slots included: rows 0-1 and 5-6 of the
    reference's 7x7, 5-diagonal matrix run off it and are completed by
    mirroring (tests/cblas_tests.cpp:431-447)."""
    from oracle.chunk import ResoMatrix
    d = np.load(f"{golden_dir}/resolution_inline.npz")
    R = ResoMatrix("dia", d["diamatrix_diagonals"].reshape(5, 7), 5)
    rows = np.array([R.dia_row(i) for i in range(7)])
    assert np.array_equal(rows, d["diamatrix_rows"].reshape(7, 5))


def test_rmat_oversampling_spline_and_invariants():
    """Resolution::oversample: the spline is scipy's natural cubic spline;
    at the original points (every osamp-th output) the shifted log
    interpolation returns the row itself, so ratios there are the row's; rows
    sum to one; osamp = 1 only normalises."""
    from scipy.interpolate import CubicSpline
    from oracle.chunk import ResoMatrix, natural_cubic_spline
    rng = np.random.default_rng(5)
    x = np.arange(11) - 5.0
    for _ in range(3):
        y = rng.normal(size=11)
        xe = np.linspace(-5, 5, 61)
        np.testing.assert_allclose(natural_cubic_spline(x, y, xe), CubicSpline(x, y, bc_type="natural")(xe),
                                   rtol=0, atol=1e-14)
    n = 40
    sig = rng.uniform(0.8, 1.6, n)
    offs = np.arange(11) - 5
    dense_rows = np.exp(-0.5 * (offs[None, :] / sig[:, None]) ** 2)
    dense_rows[:, 0] -= 2e-3                     # a negative tail value exercises the shift
    vals = np.zeros((11, n))
    for i in range(n):
        for p in range(11):
            c = i - 5 + p
            if 0 <= c < n:
                vals[10 - p, c] = dense_rows[i, p]
    R = ResoMatrix("dia", vals, 11)
    O = R.oversample(3, dlambda=0.8)
    assert O.values.shape == (n, 31) and O.nelem == 31 and O.oversampling == 3
    np.testing.assert_allclose(O.values.sum(axis=1), 1.0, rtol=1e-14)
    for i in (0, 7, n - 1):
        row = R.dia_row(i)
        knots = O.values[i, ::3]
        np.testing.assert_allclose(knots / knots[5], row / row[5], rtol=1e-12)
    O1 = R.oversample(1)
    for i in (3, 20):
        np.testing.assert_allclose(O1.values[i], R.dia_row(i) / R.dia_row(i).sum(), rtol=1e-13)
    # the grid spacing does not change the interpolant
    np.testing.assert_allclose(R.oversample(3, dlambda=0.3).values, O.values, rtol=1e-12)


def test_rmat_deconvolution_against_a_direct_dft():
    """DiaMatrix::deconvolve on a delta row: the interior row becomes the
    inverse DFT of 1 / sinc(pi m j / n) over the 21-point padded row, summed
    here mode by mode over j = -10..10 with complex exponentials; m -> 0
    leaves the matrix unchanged."""
    from oracle.chunk import ResoMatrix
    n, nel, pad = 30, 11, 5
    vals = np.zeros((nel, n))
    vals[5, :] = 1.0                            # the identity: delta rows
    for m in (1.0, 1.7):
        D = ResoMatrix("dia", vals, nel).deconvolve(m)
        nf = nel + 2 * pad
        t = np.arange(nf)
        x = np.zeros(nf)
        x[pad + 5] = 1.0
        want = np.zeros(nf, complex)
        for j in range(-(nf // 2), nf // 2 + 1):
            Xj = np.sum(x * np.exp(-2j * np.pi * j * t / nf))
            s = 1.0 if j == 0 else np.sin(np.pi * m * abs(j) / nf) / (np.pi * m * abs(j) / nf)
            want += Xj / s * np.exp(2j * np.pi * j * t / nf)
        want = (want / nf).real
        got = D.dia_row(15)
        np.testing.assert_allclose(got, want[pad:pad + nel], rtol=0, atol=1e-13)
    np.testing.assert_allclose(ResoMatrix("dia", vals, nel).deconvolve(1e-6).values, vals, atol=1e-11)


def test_rmat_preprocessing_through_the_c_abi_matches_the_oracle():
    """qml_rmat_deconvolve and qml_rmat_oversample (host code behind the C
    ABI; no device is used) against the oracle on an 11-diagonal Gaussian
    matrix with jittered widths, edges included."""
    from lyspeq_amd.qml import rmat_deconvolve, rmat_oversample
    from oracle.chunk import ResoMatrix
    from lyspeq_amd import synth
    w = synth.rmat_config(nchunks=1, n_range=(180, 180), kind="dia", with_tables=False)
    vals = np.asarray(w.chunks[0]["rmat"]["values"])
    ro = ResoMatrix("dia", vals, 11)
    dec = rmat_deconvolve(vals, 1.0)
    np.testing.assert_allclose(dec, ro.deconvolve(1.0).values, rtol=0, atol=1e-14)
    for osamp in (1, 3, 4):
        got = rmat_oversample(dec, osamp)
        want = ResoMatrix("dia", dec, 11).oversample(osamp).values
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)


@pytest.mark.parametrize("power", [3.591244, 1.0, -2.5])
def test_redshift_growth_is_a_per_pixel_congruence(power):
    """With 1 + z_ij = sqrt(l_i l_j) (l in units of the Lya wavelength), the
    growth factor ((1 + z_ij)/(1 + z_m))^B separates into g_i g_j with
    g_i = (l_i / (1 + z_m))^(B/2): every growth-scaled Q of z bin m is
    G_m Q G_m.  A wrong pair redshift (arithmetic mean), a bin centre without
    the 1 +, or a wrong exponent breaks the identity; B = 0 must reproduce
    the unscaled matrices exactly."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    import parity as P
    from lyspeq_amd import synth
    w = synth.desi_config(nchunks=6)
    tabs = P.oracle_tables(w)
    c = w.chunks[0]
    b_off = P.oracle_bins(w.binning)
    w.binning.redshift_growth_on = True
    w.binning.fiducial_redshift_power = power
    b_on = P.oracle_bins(w.binning)
    w.binning.fiducial_redshift_power = 0.0
    b_zero = P.oracle_bins(w.binning)
    ch_off = Chunk(b_off, tabs, c["wave"], c["delta"], c["ivar"], c["R"])
    ch_on = Chunk(b_on, tabs, c["wave"], c["delta"], c["ivar"], c["R"])
    ch_zero = Chunk(b_zero, tabs, c["wave"], c["delta"], c["ivar"], c["R"])
    v, z = ch_off.vz_matrices()
    q_off, q_on, q_zero = ch_off.q_matrices(v, z), ch_on.q_matrices(v, z), ch_zero.q_matrices(v, z)
    nk = b_off.nk
    assert len(q_on) == len(q_off) > 0
    nzq = len(q_off) // ch_off.kbin_upp
    assert nzq >= 1
    for iz in range(nzq):
        zm = ch_off.fisher_index_start // nk + iz
        g = np.power(ch_off.wave / (1.0 + b_off.zbin_centers[zm]), power / 2.0)
        for s in range(iz * ch_off.kbin_upp, (iz + 1) * ch_off.kbin_upp):
            ref = g[:, None] * q_off[s] * g[None, :]
            scale = np.max(np.abs(ref))
            assert scale > 0
            assert np.max(np.abs(q_on[s] - ref)) <= 1e-13 * scale
            assert np.array_equal(q_zero[s], q_off[s])


def test_reader_preprocessing_on_the_golden_spectrum():
    """The reader's masking and cuts leave the reference's test spectrum
    whole (its golden run holds all 488 pixels), and the chunk built from
    read_spectrum's output reproduces the golden Fisher matrix."""
    from oracle import reader
    b = Bins(**gc.BINS_KW)
    sp = gc.spectrum()
    out = reader.read_spectrum(b, sp.wave, sp.flux, sp.ivar, sp.snr)
    assert out is not None
    w, d, iv = out
    assert w.size == 488 and np.array_equal(w, sp.wave) and np.array_equal(d, sp.flux)
    t = gc.table()
    ts = [TableSet(70000, 20.0, t.nvpoints, t.v_length, t.nzpoints, t.deriv, t.signal)]
    ch = Chunk(b, ts, w, d, iv, sp.R_fwhm)
    T = b.total_kz_bins
    F, dbt = np.zeros((T, T)), np.zeros((3, T))
    assert ch.one_iteration(np.zeros(T), F, dbt)
    assert gc.all_close(power_estimates(F, dbt)["fisher"], gc.golden_fisher())


def test_reader_outlier_mask_and_cuts_closed_forms():
    """QSOFile::maskOutliers (io/qso_file.cpp:189-230) on constructed data:
    with sigma = 0.1 and a MAD below 1 the threshold is 5 max(sigma, 1) = 5
    about [-1, 1/F(lambda)], so -6.5 and F^-1 + 5.5 are masked, -5.5 and
    F^-1 + 4.5 kept; with sigma = 2 (ivar 1/4) the threshold doubles.  The
    median of an even count is the half-sum of the middle two.
    cutBoundary (io/qso_file.cpp:159-187) keeps [lower_bound(l1),
    upper_bound(l2)) and then trims masked pixels at both ends."""
    from oracle import reader
    from oracle.bins import LYA_REST
    n = 40
    wave = np.linspace(3500.0, 3600.0, n)
    finv = np.exp(1.663e-3 * (wave / LYA_REST) ** 4.107)
    delta = np.zeros(n)
    ivar = np.full(n, 100.0)
    delta[3], delta[5] = -6.5, -5.5
    delta[7], delta[9] = finv[7] + 5.5, finv[9] + 4.5
    ivar[11], delta[11] = 0.25, -1.0 - 9.5   # sigma 2: threshold 10, kept
    ivar[13], delta[13] = 0.25, -1.0 - 10.5  # masked
    d2, iv2, nm = reader.mask_outliers(wave, delta, ivar)
    assert nm == 3
    assert set(np.flatnonzero(iv2 == 0)) == {3, 7, 13}
    assert np.all(d2[[3, 7, 13]] == 0) and np.array_equal(d2[[5, 9, 11]], delta[[5, 9, 11]])
    assert reader._median(np.array([4.0, 1.0, 3.0, 2.0])) == 2.5 and reader._median(np.array([3.0, 1.0, 2.0])) == 2.0
    b = Bins(**gc.BINS_KW)
    l1, l2 = LYA_REST * (1 + b.z_lower_edge), LYA_REST * (1 + b.z_upper_edge)
    wave = np.linspace(l1 - 20.0, l2 + 20.0, 200)
    iv = np.ones(200)
    a0, e0 = int(np.searchsorted(wave, l1, "left")), int(np.searchsorted(wave, l2, "right"))
    iv[a0:a0 + 3] = 0.0
    iv[e0 - 2:e0] = 0.0
    assert reader.cut_boundary(b, wave, np.zeros(200), iv) == (a0 + 3, e0 - 2)
    assert reader.cut_boundary(b, wave, np.zeros(200), np.zeros(200)) is None
    assert reader.cut_boundary(b, wave + 5000.0, np.zeros(200), iv) is None
