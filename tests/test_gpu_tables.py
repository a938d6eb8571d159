"""The device S/Q table generator (qml_make_tables) against the reference's
golden tables, the oracle's restatement (QAWO derivative, real-FFT signal)
and the workload synthesiser's numpy Filon/FFT tables (needs an MI355X)."""
import numpy as np
import pytest

import goldencase as gc
from oracle.bins import Bins
from oracle.tables import derivative_table_qawo, signal_table_fft
from lyspeq_amd import synth
from lyspeq_amd.qml import Binning, QMLEngine, QMLError

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_golden_tables_match_the_reference_fits_and_the_oracle():
    print("[case] golden tables R=70000 dv=20", flush=True)
    t = gc.table()
    eng = QMLEngine(Binning(**gc.BINS_KW), precision="fp64")
    q, s = eng.make_tables([(70000, 20.0)], t.nvpoints, t.v_length, t.nzpoints, gc.PD13)
    eng.close()
    # the reference test's own tolerance (rel 1e-5 / abs 1e-8)
    assert gc.all_close(q[0], t.deriv)
    assert gc.all_close(s[0], t.signal)
    ob = Bins(**gc.BINS_KW)
    so = signal_table_fft(ob, 70000, 20.0, t.nvpoints, t.v_length, t.nzpoints, gc.PD13)
    qo = derivative_table_qawo(ob, 70000, 20.0, t.nvpoints, t.v_length)
    es, eq = _rel(s[0], so), _rel(q[0], qo)
    print(f"golden: signal vs FFT oracle {es:.2e}, derivative vs QAWO oracle {eq:.2e}", flush=True)
    assert es < 1e-11          # the same sums, summed in another order
    assert eq < 1e-6           # Filon against QAWO at epsrel 1e-7


def test_desi_pairs_match_the_numpy_tables_and_qawo():
    print("[case] desi pairs tables", flush=True)
    w = synth.desi_config(nchunks=16, with_tables=False)
    pairs = w.pairs[:6]
    eng = QMLEngine(w.binning, precision="fp64")
    q, s = eng.make_tables(pairs, w.nv, w.v_length, w.nz_table, w.pd13, filon_panels=32)
    qn = synth.derivative_tables(w.binning.kband_edges(), pairs, w.nv, w.v_length, M=32)
    sn = synth.signal_tables(w.binning, pairs, w.nv, w.v_length, w.nz_table, w.pd13)
    eq, es = _rel(q, qn), _rel(s, sn)
    print(f"desi pairs: derivative vs numpy Filon {eq:.2e}, signal vs numpy FFT {es:.2e}", flush=True)
    assert eq < 1e-10 and es < 1e-10
    # Filon at the default 64 panels against QAWO on a coarse velocity grid
    ob = Bins(k0=w.binning.k0, nklin=w.binning.nklin, dklin=w.binning.dklin, nklog=w.binning.nklog,
              dklog=w.binning.dklog, z0=w.binning.z0, z_bin_width=w.binning.z_bin_width, nz=w.binning.nz)
    nvc = 93
    qc, _ = eng.make_tables(pairs[:2], nvc, w.v_length, w.nz_table, w.pd13, signal=False)
    for i, (R, dv) in enumerate(pairs[:2]):
        qo = derivative_table_qawo(ob, R, dv, nvc, w.v_length)
        e = _rel(qc[i], qo)
        print(f"pair {R}, {dv}: derivative vs QAWO {e:.2e}", flush=True)
        assert e < 1e-6
    eng.close()


def test_filon_converges_and_arguments_are_checked():
    print("[case] tables convergence and errors", flush=True)
    t = gc.table()
    eng = QMLEngine(Binning(**gc.BINS_KW), precision="fp64")
    q64, _ = eng.make_tables([(3000, 40.0)], t.nvpoints, t.v_length, t.nzpoints, gc.PD13, signal=False)
    q128, _ = eng.make_tables([(3000, 40.0)], t.nvpoints, t.v_length, t.nzpoints, gc.PD13,
                              filon_panels=128, signal=False)
    assert _rel(q64, q128) < 1e-9
    with pytest.raises(QMLError):
        eng.make_tables([(0, 20.0)], t.nvpoints, t.v_length, t.nzpoints, gc.PD13)
    with pytest.raises(QMLError):
        eng.make_tables([(70000, 20.0)], 1, t.v_length, t.nzpoints, gc.PD13)
    eng.close()
