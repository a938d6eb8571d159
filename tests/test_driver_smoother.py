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
