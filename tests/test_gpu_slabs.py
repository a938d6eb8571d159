"""Stack slabs (qml_set_stack_slab): each chunk's weighted matrices streamed
a slab of members at a time must give the unslabbed results bit for bit
(every output sums in the same order), and bound a chunk's scratch so that
the high-resolution workload's largest chunks fit budgets the whole stack
does not (needs an MI355X)."""
import numpy as np
import pytest

import parity as P
from lyspeq_amd import synth
from lyspeq_amd.qml import QMLEngine, QMLError

pytestmark = pytest.mark.gpu


def _same(a, b):
    _, oa, Fa, da = a
    _, ob, Fb, db = b
    assert np.array_equal(Fa, Fb) and np.array_equal(da, db)
    assert set(oa) == set(ob)
    for i in oa:
        assert np.array_equal(oa[i][0], ob[i][0]) and np.array_equal(oa[i][1], ob[i][1])


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3", "bf16x4"])
def test_slabbed_stack_is_bit_identical(precision):
    """Slabs of 16 (within a z bin's run of members), 7 (slabs that straddle
    z bins), and 1; DESI chunks of several sizes and z-bin counts, batches
    of mixed member counts."""
    print(f"[case] stack slabs {precision}", flush=True)
    w = synth.desi_config(nchunks=24)
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 13)
    idx = list(range(len(w.chunks)))
    ref = P.run_engine(w, th, idx, precision=precision, max_batch_chunks=9)
    whole = ref[0].schedule()["max_chunk_gb"]
    ref[0].close()
    for slab in (16, 7, 1):
        got = P.run_engine(w, th, idx, precision=precision, max_batch_chunks=9, stack_slab=slab)
        sch = got[0].schedule()
        got[0].close()
        assert sch["stack_slab"] == slab and sch["max_chunk_gb"] < whole
        _same(ref, got)


def test_slabbed_resolution_matrix_chunks_and_graph_replay():
    """Resolution-matrix chunks (Q_k from the sandwiches) in slabs, eager and
    replayed from a captured graph, against the unslabbed eager run."""
    print("[case] stack slabs, resolution matrices, graph replay", flush=True)
    w = synth.rmat_config(nchunks=8, n_range=(100, 400))
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 17)
    idx = list(range(len(w.chunks)))
    for precision in ("fp64", "bf16x3"):
        ref = P.run_engine(w, th, idx, precision=precision)
        ref[0].close()
        eng, out, F, dbt = P.run_engine(w, th, idx, precision=precision, stack_slab=16)
        _same(ref, (eng, out, F, dbt))
        eng.set_graph(True)
        T = eng.T
        for _ in range(3):
            F2, d2 = np.zeros((T, T)), np.zeros((3, T))
            eng.iterate(th, F2, d2)
            assert np.array_equal(F2, ref[2]) and np.array_equal(d2, ref[3])
        eng.close()


def test_high_resolution_chunks_fit_a_small_budget_only_in_slabs():
    """N ~ 2000 with 13 z bins of Nk = 50: in FP64 the whole stack needs
    ~10 GB per chunk and is refused under a 7 GB budget; in slabs of 16 the
    chunk needs ~4 GB (the slab's stack and left products, the store of
    packed whitened matrices) and its results equal the unslabbed ones."""
    print("[case] high-resolution N~2000 fp64 in slabs under 7 GB", flush=True)
    w = synth.hires_config(nchunks=2, n_range=(1950, 2000))
    T = w.binning.total_kz_bins
    idx = list(range(len(w.chunks)))
    th = np.zeros(T)
    ref = P.run_engine(w, th, idx, precision="fp64", scratch_gb=64)
    whole = ref[0].schedule()["max_chunk_gb"]
    ref[0].close()
    assert whole > 7.0 and len(ref[1]) == len(idx)
    # unslabbed, the largest chunk is refused when it is added, as the
    # reference's constructor refuses it ("Not all matrices are stored"), and
    # the spectrum's other chunks go on without it
    eng = QMLEngine(w.binning, precision="fp64", scratch_gb=7)
    for i, (R, dv) in enumerate(w.pairs):
        eng.add_table(R, dv, w.deriv[i], w.signal[i], w.v_length)
    codes = []
    for c in w.chunks:
        try:
            eng.add_chunk(c["wave"], c["delta"], c["ivar"], c["R"])
            codes.append(0)
        except QMLError as e:
            assert "Not all matrices are stored" in str(e)
            codes.append(e.code)
    eng.close()
    assert -6 in codes   # QML_ERR_MEMORY
    got = P.run_engine(w, th, idx, precision="fp64", scratch_gb=7, stack_slab=16)
    assert len(got[1]) == len(idx)
    per_chunk = got[0].schedule()["max_chunk_gb"]
    got[0].close()
    print(f"largest chunk: {whole:.2f} GB whole stack, {per_chunk:.2f} GB in slabs of 16", flush=True)
    assert per_chunk < 7.0
    _same(ref, got)
