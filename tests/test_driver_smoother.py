"""The power smoothing between Newton-Raphson iterations
(lyspeq_amd/driver.py smooth_power: py/smbivspline.py's weighted smoothing
spline, restated around the same scipy routine), on the CPU.  The
reference's script is not run; it is pinned by what a smoothing spline with
s = len(w) must do: reproduce a bicubic surface, and pull noisy estimates
towards the smooth truth."""
import numpy as np

from lyspeq_amd.driver import smooth_power


def _grid():
    kc = np.concatenate([np.linspace(0.0005, 0.0195, 20), np.geomspace(0.021, 0.05, 15)])
    zc = 2.0 + 0.2 * np.arange(13)
    return kc, zc


def _surface(kc, zc):
    """A bicubic polynomial in (z, k), positive over the grid."""
    z = np.repeat(zc, kc.size)
    k = np.tile(kc, zc.size)
    x = (z - 3.2) / 1.2
    y = (k - 0.025) / 0.025
    return 10.0 + 2.0 * x + 1.5 * y - 0.8 * x * y + 0.6 * y ** 2 + 0.2 * x ** 3 - 0.3 * y ** 3 + 0.1 * x * y ** 2


def test_a_bicubic_surface_is_reproduced():
    """P bicubic in (z, k), and ln P bicubic in (ln(1+z), ln k) for the log
    mode: the spline space holds the surface, so the fit reproduces it to the
    6 significant digits the script reads."""
    kc, zc = _grid()
    truth = _surface(kc, zc)
    pfid = 0.5 * truth
    sigma = 0.01 * truth
    fits = smooth_power(truth - pfid, sigma, kc, zc, pfid)
    assert np.max(np.abs(fits - (truth - pfid)) / truth) < 2e-5
    z = np.repeat(zc, kc.size)
    k = np.tile(kc, zc.size)
    u, v = np.log(1 + z), np.log(k)
    ltruth = np.exp(1.0 + 0.5 * u - 0.2 * v + 0.05 * u * v - 0.03 * v ** 2 + 0.002 * v ** 3)
    fits = smooth_power(0.5 * ltruth, 0.01 * ltruth, kc, zc, 0.5 * ltruth, interp_log=True)
    assert np.max(np.abs(fits - 0.5 * ltruth) / ltruth) < 2e-5


def test_noisy_estimates_move_towards_the_smooth_truth():
    kc, zc = _grid()
    truth = _surface(kc, zc)
    pfid = 0.9 * truth
    sigma = 0.05 * truth
    rng = np.random.default_rng(11)
    noisy = truth + sigma * rng.standard_normal(truth.size)
    fits = smooth_power(noisy - pfid, sigma, kc, zc, pfid)
    raw = np.sqrt(np.mean(((noisy - truth) / sigma) ** 2))
    smooth = np.sqrt(np.mean(((fits + pfid - truth) / sigma) ** 2))
    assert raw > 0.8 and smooth < 0.5 * raw


def test_log_mode_drops_non_positive_bands_and_empty_bands_are_skipped():
    kc, zc = _grid()
    truth = _surface(kc, zc)
    pfid = 0.9 * truth
    sigma = 0.05 * truth
    theta = truth - pfid
    theta[[3, 40, 100]] = -2.0 * pfid[[3, 40, 100]]      # P < 0 at three bands
    sigma[7] = 0.0                                        # a band without data
    fits = smooth_power(theta, sigma, kc, zc, pfid, interp_log=True)
    assert np.all(np.isfinite(fits)) and np.all(fits + pfid > 0)
    assert np.max(np.abs(fits - (truth - pfid))[[3, 40, 100]] / truth[[3, 40, 100]]) < 0.05
    fits = smooth_power(theta, sigma, kc, zc, pfid)
    assert np.all(np.isfinite(fits))


def test_power_estimates_reproduce_the_golden_run():
    """The product driver's power_estimates (normalise F, invert it, theta =
    F^-1 (d - b - t)/2, sigma, the weighted d/b/t) on the sums of the
    reference's test chunk: the golden Fisher, inverse Fisher, theta, sigma
    and weighted columns (tests/truth/test_it1_*.dat), at the reference's
    own tolerance (tests/compareTestResults.py:8-9)."""
    import goldencase as gc
    from lyspeq_amd.driver import power_estimates
    from oracle.bins import Bins
    from oracle.chunk import Chunk, TableSet
    b = Bins(**gc.BINS_KW)
    t = gc.table()
    ts = [TableSet(70000, 20.0, t.nvpoints, t.v_length, t.nzpoints, t.deriv, t.signal)]
    sp = gc.spectrum()
    ch = Chunk(b, ts, sp.wave, sp.flux, sp.ivar, sp.R_fwhm)
    T = b.total_kz_bins
    F, dbt = np.zeros((T, T)), np.zeros((3, T))
    assert ch.one_iteration(np.zeros(T), F, dbt)
    res = power_estimates(F, dbt)
    rows = gc.detailed_rows()
    assert gc.all_close(res["fisher"], gc.golden_fisher())
    assert gc.all_close(res["inverse"], gc.golden_inverse_fisher())
    assert gc.all_close(rows[:, 5], res["theta"]) and gc.all_close(rows[:, 7], res["sigma"])
    for i in range(3):
        assert gc.all_close(rows[:, 8 + i], res["weighted"][i])


def test_the_spline_is_scipys_where_scipy_completes_and_log_mode_completes():
    """_smoothing_spline equals SmoothBivariateSpline value for value where
    scipy's one workspace retry suffices (the noisy surface).  On the bands
    of the log-mode test in the linear mode (three negative P, one band
    without data), scipy's retry falls short (FITPACK ier = 1299, a warning
    and an unfinished fit); the grown workspace completes the fit: surfit
    returns normally, with the residual sum fp at s."""
    import warnings
    from scipy.interpolate import SmoothBivariateSpline
    from lyspeq_amd.driver import _smoothing_spline
    kc, zc = _grid()
    truth = _surface(kc, zc)
    sigma = 0.05 * truth
    rng = np.random.default_rng(11)
    noisy = truth + sigma * rng.standard_normal(truth.size)
    z, k = np.repeat(zc, kc.size), np.tile(kc, zc.size)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        ref = SmoothBivariateSpline(z, k, noisy, w=1.0 / sigma, s=len(z))
        mine = _smoothing_spline(z, k, noisy, 1.0 / sigma, len(z))
    assert np.array_equal(ref(z, k, grid=False), mine(z, k, grid=False))
    pfid = 0.9 * truth
    p = truth.copy()
    p[[3, 40, 100]] = -pfid[[3, 40, 100]]
    w = 1.0 / sigma
    keep = np.ones(z.size, bool)
    keep[7] = False
    with warnings.catch_warnings(record=True) as wl:
        warnings.simplefilter("always")
        SmoothBivariateSpline(z[keep], k[keep], p[keep], w=w[keep], s=int(keep.sum()))
    assert any("ier=" in str(x.message) for x in wl)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        spl = _smoothing_spline(z[keep], k[keep], p[keep], w[keep], int(keep.sum()))
        fits = smooth_power(p - pfid, np.where(keep, sigma, 0.0), kc, zc, pfid)
    assert spl.ier == 0 and abs(spl.fp - keep.sum()) < 1e-3 * keep.sum()
    assert np.all(np.isfinite(fits))
