"""HIP engine vs the CPU oracle on the same chunks (needs an MI355X)."""
import os
import sys

import numpy as np
import pytest

import goldencase as gc
import parity as P
from lyspeq_amd import synth
from lyspeq_amd.qml import Binning, QMLEngine

pytestmark = pytest.mark.gpu

# FP64 path: the Cholesky/whitened form vs the oracle's LU form, to the bar
# of parity.assert_fp64 (normalised errors below 1e-10, elementwise errors of
# the entries above 1e-6 of the largest below 1e-9).


def _golden_workload(**over):
    kw = dict(gc.BINS_KW)
    kw.update(over)
    b = Binning(**kw)
    t = gc.table()
    sp = gc.spectrum()
    w = synth.Workload("golden", b, nv=t.nvpoints, v_length=t.v_length, nz_table=t.nzpoints,
                       pd13=gc.PD13)
    w.pairs = [(70000, 20.0)]
    w.deriv = t.deriv[None]
    w.signal = t.signal[None]
    w.chunks.append(dict(wave=sp.wave, delta=sp.flux, ivar=sp.ivar, R=sp.R_fwhm, dv_table=20.0))
    return w


def _print_case(name):
    print(f"[case] {name}", flush=True)


def test_golden_chunk_fp64_matches_reference_goldens():
    _print_case("golden chunk fp64")
    w = _golden_workload()
    T = w.binning.total_kz_bins
    eng, out, F, dbt = P.run_engine(w, np.zeros(T), [0])
    from oracle.driver import power_estimates
    res = power_estimates(F, dbt)
    assert gc.all_close(res["fisher"], gc.golden_fisher())
    rows = gc.detailed_rows()
    assert gc.all_close(rows[:, 5], res["theta"])
    assert gc.all_close(rows[:, 7], res["sigma"])
    for i in range(3):
        assert gc.all_close(rows[:, 11 + i], dbt[i])
    oout, oF, od = P.run_oracle(w, np.zeros(T), [0])
    P.assert_fp64(F, dbt, oF, od, "golden chunk")


@pytest.mark.parametrize("orders", [(1, -1), (2, 1)])
def test_golden_chunk_marginalised_masked_theta_fp64(orders):
    _print_case(f"golden chunk marginalised {orders}")
    w = _golden_workload(cont_logl_order=orders[0], cont_lam_order=orders[1])
    rng = np.random.default_rng(5)
    c = w.chunks[0]
    c["ivar"] = c["ivar"].copy()
    c["ivar"][1 + rng.choice(c["ivar"].size - 2, 30, replace=False)] = 0.0
    th = P.theta_perturbed(w.binning, gc.PD13, 0.1, 3)
    eng, out, F, dbt = P.run_engine(w, th, [0])
    oout, oF, od = P.run_oracle(w, th, [0])
    # four templates (orders (2, 1)) span s_min / s_0 ~ 4e-8: the reference's
    # dgesvd fixes their span only to ~1e-9 of F (the floor)
    floor, _, sx = P.oracle_floor(w, th, [0], oout)
    P.assert_fp64(F, dbt, oF, od, f"golden chunk marginalised {orders}", floor=floor, exact=sx)


def test_cpu_config_64_chunks_fp64():
    _print_case("cpu-64x256 fp64")
    w = synth.cpu_reference_config(nchunks=64)
    T = w.binning.total_kz_bins
    idx = list(range(len(w.chunks)))
    eng, out, F, dbt = P.run_engine(w, np.zeros(T), idx)
    oout, oF, od = P.run_oracle(w, np.zeros(T), idx)
    assert all(out[i][2] and oout[i][2] for i in out)
    P.assert_chunks_fp64(out, oout, "cpu-64x256")
    P.assert_fp64(F, dbt, oF, od, "cpu-64x256 sums")


def test_desi_sample_fp64_multiple_tables_and_zbins():
    _print_case("desi sample fp64")
    w = synth.desi_config(nchunks=24)
    T = w.binning.total_kz_bins
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 11)
    idx = list(range(len(w.chunks)))
    eng, out, F, dbt = P.run_engine(w, th, idx, max_batch_chunks=7)
    oout, oF, od = P.run_oracle(w, th, idx)
    assert len(out) >= 20
    P.assert_chunks_fp64(out, oout, "desi sample")
    P.assert_fp64(F, dbt, oF, od, "desi sample sums")


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16x4"])
def test_mixed_precision_against_fp64(precision):
    _print_case(f"desi sample {precision}")
    w = synth.desi_config(nchunks=24)
    T = w.binning.total_kz_bins
    idx = list(range(len(w.chunks)))
    _, _, F64, d64 = P.run_engine(w, np.zeros(T), idx, precision="fp64")
    _, _, Fm, dm = P.run_engine(w, np.zeros(T), idx, precision=precision)
    P.assert_mixed(Fm, dm, F64, d64, precision, f"desi sample {precision}")


@pytest.mark.parametrize("config,precision,nchunks", [("desi", "bf16x3", 64), ("hires", "bf16x4", 24)])
def test_p1d_within_a_thousandth_of_sigma(config, precision, nchunks):
    """The mixed path's P1D against the FP64 path's, through the driver's
    arithmetic (0.5 scaling, stable inverse, theta = 0.5 F^-1 (d - b - t)):
    max |dP| / sigma must stay below 1e-3 (the high-resolution batch runs in
    split-BF16x4, whose lo*lo product keeps it there)."""
    _print_case(f"P1D {config}-{nchunks} {precision} vs fp64")
    w = synth.desi_config(nchunks=nchunks) if config == "desi" else synth.hires_config(nchunks=nchunks)
    T = w.binning.total_kz_bins
    idx = list(range(len(w.chunks)))
    _, _, F64, d64 = P.run_engine(w, np.zeros(T), idx, precision="fp64", scratch_gb=64)
    _, _, Fm, dm = P.run_engine(w, np.zeros(T), idx, precision=precision, scratch_gb=64)
    dev = P.p1d_deviation(Fm, dm, F64, d64)
    print(f"[mixed] P1D {config}-{nchunks} {precision}: max |dP|/sigma {dev:.2e}", flush=True)
    assert dev < P.P1D_TOL[(config, precision)]


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3", "bf16x4"])
def test_poisoned_scratch_gives_identical_results(precision, monkeypatch):
    """Scratch arenas and LDS filled with NaN bytes (LYSPEQ_QML_POISON=1)
    must not change a single bit of the results: every value the iteration
    reads is one it wrote (chunks of the sample have all-padding tiles, whose
    LDS slices are never staged)."""
    _print_case(f"poisoned scratch {precision}")
    w = synth.desi_config(nchunks=24)
    T = w.binning.total_kz_bins
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 5)
    idx = list(range(len(w.chunks)))
    eng0, out0, F0, d0 = P.run_engine(w, th, idx, precision=precision)
    eng0.close()
    monkeypatch.setenv("LYSPEQ_QML_POISON", "1")
    eng1, out1, F1, d1 = P.run_engine(w, th, idx, precision=precision)
    F2, d2 = np.zeros((T, T)), np.zeros((3, T))
    eng1.set_pipeline(False)
    eng1.iterate(th, F2, d2)
    eng1.close()
    assert np.isfinite(F1).all() and np.isfinite(d1).all()
    assert np.array_equal(F0, F1) and np.array_equal(d0, d1)
    assert np.array_equal(F0, F2) and np.array_equal(d0, d2)
    for i in out0:
        assert np.array_equal(out0[i][0], out1[i][0]) and np.array_equal(out0[i][1], out1[i][1])


def test_hires_chunks_fp64_match_oracle():
    """High-resolution workload (Nk = 50, 0.5 km/s tables, masked runs) at
    N ~ 850, where the oracle still finishes in seconds."""
    _print_case("hires N~850 fp64 vs oracle")
    w = synth.hires_config(nchunks=2, n_range=(820, 880))
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 21)
    idx = list(range(len(w.chunks)))
    eng, out, F, dbt = P.run_engine(w, th, idx)
    eng.close()
    oout, oF, od = P.run_oracle(w, th, idx)
    assert len(out) == 2
    P.assert_chunks_fp64(out, oout, "hires N~850")
    P.assert_fp64(F, dbt, oF, od, "hires N~850 sums")


def test_hires_full_size_mixed_precision_against_fp64():
    """N ~ 2000 (Np = 2048, n up to ~250 stored Q): the mixed paths against
    the FP64 path at the largest chunk size of the high-resolution workload."""
    _print_case("hires N~2000 mixed vs fp64")
    w = synth.hires_config(nchunks=2, n_range=(1950, 2000))
    T = w.binning.total_kz_bins
    idx = list(range(len(w.chunks)))
    _, _, F64, d64 = P.run_engine(w, np.zeros(T), idx, precision="fp64", scratch_gb=64)
    errs = {}
    for precision in ("fp32", "bf16x3", "bf16x4"):
        _, _, Fm, dm = P.run_engine(w, np.zeros(T), idx, precision=precision, scratch_gb=64)
        errs[precision] = P.assert_mixed(Fm, dm, F64, d64, precision, f"hires N~2000 {precision}", "hires")
    # the lo*lo product removes about half of the split-BF16 error
    assert errs["bf16x4"]["F_norm"] < 0.75 * errs["bf16x3"]["F_norm"]
    assert errs["bf16x4"]["b_norm"] < 0.9 * errs["bf16x3"]["b_norm"]


@pytest.mark.parametrize("kind", ["dia", "oversampled"])
def test_resolution_matrix_chunks_fp64_match_oracle(kind):
    """Chunks with their own resolution matrix (11 diagonals, Gaussian rows
    with jittered widths; kept diagonal or oversampled x3): Q_k = R Qhat_k R^T
    and S = R Shat R^T on the matrix's grid, 3 continuum templates."""
    _print_case(f"resolution matrix {kind} fp64 vs oracle")
    w = synth.rmat_config(nchunks=6, n_range=(100, 170), kind=kind)
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 31)
    idx = list(range(len(w.chunks)))
    eng, out, F, dbt = P.run_engine(w, th, idx)
    eng.close()
    oout, oF, od = P.run_oracle(w, th, idx)
    assert len(out) >= 5
    floor, xout, sx = P.oracle_floor(w, th, idx, oout)
    P.assert_chunks_fp64(out, oout, f"resolution matrix {kind}", floor=floor, exact=xout)
    P.assert_fp64(F, dbt, oF, od, f"resolution matrix {kind} sums", floor=floor, exact=sx)


def test_resolution_matrix_read_time_preprocessing_fp64_matches_oracle():
    """Diagonal-storage matrices deconvolved (m = 1) and oversampled x3 at
    read time, then registered as oversampled rows (core/one_qso_estimate.cpp:
    74-89): the engine takes rows from qml_rmat_deconvolve/qml_rmat_oversample,
    the oracle preprocesses its own copy."""
    import copy
    from lyspeq_amd.qml import rmat_deconvolve, rmat_oversample
    _print_case("resolution matrix deconvolved + oversampled at read time, fp64 vs oracle")
    w = synth.rmat_config(nchunks=6, n_range=(100, 170), kind="dia")
    wo = copy.copy(w)
    wo.chunks = [dict(c) for c in w.chunks]
    for c, co in zip(w.chunks, wo.chunks):
        dia = np.asarray(c["rmat"]["values"])
        c["rmat"] = dict(kind="oversampled", nelem=31, oversampling=3,
                         values=rmat_oversample(rmat_deconvolve(dia, 1.0), 3))
        ro = P.oracle_rmat(co).deconvolve(1.0).oversample(3)
        co["rmat"] = dict(kind="oversampled", nelem=31, oversampling=3, values=ro.values)
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 37)
    idx = list(range(len(w.chunks)))
    eng, out, F, dbt = P.run_engine(w, th, idx)
    eng.close()
    oout, oF, od = P.run_oracle(wo, th, idx)
    assert len(out) >= 5
    floor, xout, sx = P.oracle_floor(wo, th, idx, oout)
    P.assert_chunks_fp64(out, oout, "preprocessed resolution matrix", floor=floor, exact=xout)
    P.assert_fp64(F, dbt, oF, od, "preprocessed resolution matrix sums", floor=floor, exact=sx)


def test_resolution_matrix_full_range_mixed_precision_against_fp64():
    """Mixed precisions against FP64 across the whole log-U[100, 1200] size
    range of the resolution-matrix workload (oversampled x3)."""
    _print_case("resolution matrix N in [100,1200] mixed vs fp64")
    w = synth.rmat_config(nchunks=16, n_range=(100, 1200))
    T = w.binning.total_kz_bins
    idx = list(range(len(w.chunks)))
    assert max(len(c["wave"]) for c in w.chunks) > 1000
    _, _, F64, d64 = P.run_engine(w, np.zeros(T), idx, precision="fp64")
    for precision in ("fp32", "bf16x3", "bf16x4"):
        _, _, Fm, dm = P.run_engine(w, np.zeros(T), idx, precision=precision)
        P.assert_mixed(Fm, dm, F64, d64, precision, f"resolution matrix full range {precision}", "rmat")


@pytest.mark.parametrize("form", ["dia", "os3"])
def test_resolution_matrix_large_chunks_fp64_match_oracle(form, golden_dir):
    """N = 828, 981, 1190 (fine grids up to M = 3600): the oracle's results
    are kept in tests/golden/rmat_large.npz (make_rmat_large.py).  "dia": the
    11-diagonal matrices as read; "os3": deconvolved (m = 1) and oversampled
    x3 through the C ABI's read-time preprocessing, against the oracle's own
    preprocessing."""
    from lyspeq_amd.qml import rmat_deconvolve, rmat_oversample
    sys.path.insert(0, golden_dir)
    import make_rmat_large as M
    _print_case(f"resolution matrix N in [800,1200] {form} fp64 vs oracle fixture")
    d = np.load(os.path.join(golden_dir, "rmat_large.npz"))
    w = M.workload()
    assert len(w.chunks) == int(d["nchunks"])
    for i, c in enumerate(w.chunks):
        c.update(wave=d[f"wave{i}"], delta=d[f"delta{i}"], ivar=d[f"ivar{i}"], R=int(d[f"R{i}"]))
        dia = d[f"dia{i}"]
        if form == "dia":
            c["rmat"] = dict(kind="dia", nelem=11, oversampling=1, values=dia)
        else:
            c["rmat"] = dict(kind="oversampled", nelem=31, oversampling=3,
                             values=rmat_oversample(rmat_deconvolve(dia, 1.0), 3))
    th = d["theta"]
    idx = list(range(len(w.chunks)))
    eng, out, F, dbt = P.run_engine(w, th, idx)
    eng.close()
    assert sorted(out) == idx and all(out[i][2] for i in idx)
    oout = {i: (d[f"{form}_F{i}"], d[f"{form}_dbt{i}"]) for i in idx}
    xout = {i: (d[f"{form}_Fx{i}"], d[f"{form}_dbtx{i}"]) for i in idx}
    floor = P.worst_errors(oout, xout)
    P.assert_chunks_fp64(out, oout, f"resolution matrix large {form}", floor=floor, exact=xout)
    assert max(len(c["wave"]) for c in w.chunks) >= 1150


def _with_growth(w, power=3.591244):
    w.binning.redshift_growth_on = True
    w.binning.fiducial_redshift_power = power
    return w


@pytest.mark.parametrize("config", ["desi", "rmat"])
def test_redshift_growth_fp64_matches_oracle(config):
    """RedshiftGrowthOn: every stored Q of z bin m carries ((1+z_ij)/(1+z_m))^B
    on top of the bin weight, in the table and resolution-matrix modes."""
    _print_case(f"redshift growth {config} fp64 vs oracle")
    if config == "desi":
        w = _with_growth(synth.desi_config(nchunks=12))
    else:
        w = _with_growth(synth.rmat_config(nchunks=6, n_range=(100, 170)))
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 41)
    idx = list(range(len(w.chunks)))
    eng, out, F, dbt = P.run_engine(w, th, idx, max_batch_chunks=5)
    eng.close()
    oout, oF, od = P.run_oracle(w, th, idx)
    assert len(out) >= 5
    floor, xout, sx = P.oracle_floor(w, th, idx, oout) if config == "rmat" else (None, None, None)
    P.assert_chunks_fp64(out, oout, f"redshift growth {config}", floor=floor, exact=xout)
    P.assert_fp64(F, dbt, oF, od, f"redshift growth {config} sums", floor=floor, exact=sx)
    # growth changes the results (the test is not vacuous)
    w.binning.redshift_growth_on = False
    _, _, F0, _ = P.run_engine(w, th, idx)
    assert P.rel_err(F0, F) > 1e-3


def test_redshift_growth_mixed_precision_against_fp64():
    _print_case("redshift growth mixed vs fp64")
    w = _with_growth(synth.desi_config(nchunks=24))
    T = w.binning.total_kz_bins
    idx = list(range(len(w.chunks)))
    _, _, F64, d64 = P.run_engine(w, np.zeros(T), idx, precision="fp64")
    for precision in ("fp32", "bf16x3"):
        _, _, Fm, dm = P.run_engine(w, np.zeros(T), idx, precision=precision)
        P.assert_mixed(Fm, dm, F64, d64, precision, f"redshift growth {precision}", "growth")


def test_gram_offsets_beyond_two_to_the_31_elements():
    """A high-resolution chunk of N = 3783 pixels (Np = 3840, 60 tiles of 64,
    1830 packed upper tiles) with 200 stored Q (m = 202 members): the
    split-BF16 Gram reads member rows of 2 x 1830 x 4096 bf16, so its row
    offsets reach (m - 1) x 1830 x 8192 ~ 3.0e9 elements, beyond 32 bits.
    BF16x3 against the FP64 path (64-bit offsets throughout) on the same
    chunk, at the high-resolution workload's mixed-precision bar; wrapped
    offsets would read other members' tiles."""
    print("[case] Gram offsets beyond 2^31 elements (N = 3783, m = 202)", flush=True)
    w = synth.hires_config(nchunks=6, n_range=(3760, 3800))
    assert len(w.chunks[2]["wave"]) == 3783
    T = w.binning.total_kz_bins
    th = np.zeros(T)
    e64, o64, F64, d64 = P.run_engine(w, th, [2], precision="fp64", scratch_gb=96)
    e64.close()
    em, om, Fm, dm = P.run_engine(w, th, [2], precision="bf16x3", scratch_gb=96)
    info = em.schedule()
    em.close()
    assert list(o64) == [2] and list(om) == [2] and o64[2][2] and om[2][2]
    print(f"largest chunk {info['max_chunk_gb']:.1f} GB", flush=True)
    P.assert_mixed(Fm, dm, F64, d64, "bf16x3", "N = 3783, m = 202", workload="hires")
