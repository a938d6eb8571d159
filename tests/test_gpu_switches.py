"""The binning and specifics switches in combination (needs an MI355X).

Triangular or top-hat z bins (RedshiftBinningMethod), the fiducial signal on
or off (TurnOffBaseline), continuum templates in ln lambda and lambda
(ContinuumLogLambdaMargOrder, ContinuumLambdaMargOrder), in the table mode
and both resolution-matrix storages: per-chunk results and sums in FP64
against the oracle, to the bar of parity.assert_fp64 (widened to the
oracle's own rounding floor where that exceeds it).
"""
import numpy as np
import pytest

import parity as P
from lyspeq_amd import synth

pytestmark = pytest.mark.gpu


def _theta(w, sfid_off, seed):
    """Without the fiducial signal the covariance is sum theta_k Q_k + N, so
    theta stays positive (P_fid times 1 +- 10 %); with it, the usual
    perturbation around zero."""
    if not sfid_off:
        return P.theta_perturbed(w.binning, w.pd13, 0.05, seed)
    b = w.binning
    pf = np.array([[synth.pd13_power(k, z, *w.pd13) for k in b.kband_centers()] for z in b.zbin_centers()])
    rng = np.random.default_rng(seed)
    return (pf * (1.0 + 0.1 * rng.uniform(-1.0, 1.0, pf.shape))).reshape(-1)


@pytest.mark.parametrize("mode", ["table", "rmat-dia", "rmat-oversampled"])
@pytest.mark.parametrize("method,sfid_off,orders", [(0, False, (1, -1)), (0, True, (0, 1)),
                                                     (1, True, (-1, -1)), (1, False, (0, 2))])
def test_switch_combinations_fp64_match_oracle(mode, method, sfid_off, orders):
    label = f"{mode}, {'top hat' if method == 0 else 'triangular'}, sfid {'off' if sfid_off else 'on'}, " \
            f"orders {orders}"
    print(f"[case] switches: {label}", flush=True)
    if mode == "table":
        w = synth.desi_config(nchunks=6)
    else:
        w = synth.rmat_config(nchunks=5, n_range=(100, 200), kind=mode.split("-")[1])
    w.binning.method = method
    w.binning.turn_off_sfid = sfid_off
    w.binning.cont_logl_order, w.binning.cont_lam_order = orders
    th = _theta(w, sfid_off, 17)
    idx = list(range(len(w.chunks)))
    eng, out, F, dbt = P.run_engine(w, th, idx)
    eng.close()
    oout, oF, od = P.run_oracle(w, th, idx)
    assert len(out) >= 3
    assert all(o[2] for o in out.values()), "every chunk factorises"
    # resolution-matrix windows and three nearly collinear templates carry
    # the reference's own rounding above the flat bar: measure it
    if mode != "table" or orders == (0, 2):
        floor, xout, sx = P.oracle_floor(w, th, idx, oout)
    else:
        floor, xout, sx = None, None, None
    P.assert_chunks_fp64(out, oout, label, floor=floor, exact=xout)
    P.assert_fp64(F, dbt, oF, od, f"{label} sums", floor=floor, exact=sx)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("mode", ["table", "rmat-oversampled"])
@pytest.mark.parametrize("method,sfid_off,orders", [(0, True, (0, 1)), (1, False, (0, 2))])
def test_switch_combinations_mixed_precision_against_fp64(precision, mode, method, sfid_off, orders):
    """The mixed paths on the same switch combinations, against FP64, to the
    bars of the DESI and resolution-matrix workloads (parity.MIXED_TOL)."""
    label = f"{mode}, {'top hat' if method == 0 else 'triangular'}, sfid {'off' if sfid_off else 'on'}, " \
            f"orders {orders}, {precision}"
    print(f"[case] switches mixed: {label}", flush=True)
    if mode == "table":
        w = synth.desi_config(nchunks=24)
    else:
        w = synth.rmat_config(nchunks=12, n_range=(100, 400), kind="oversampled")
    w.binning.method = method
    w.binning.turn_off_sfid = sfid_off
    w.binning.cont_logl_order, w.binning.cont_lam_order = orders
    th = _theta(w, sfid_off, 23)
    idx = list(range(len(w.chunks)))
    eng, out64, F64, d64 = P.run_engine(w, th, idx)
    eng.close()
    eng, outm, Fm, dm = P.run_engine(w, th, idx, precision=precision)
    eng.close()
    assert set(out64) == set(outm) and all(o[2] for o in outm.values())
    P.assert_mixed(Fm, dm, F64, d64, precision, label, workload="desi" if mode == "table" else "rmat")
