"""Graph replay and the autotuned batch schedule against eager launches of the
default schedule (needs an MI355X).  Neither changes what is computed for a
chunk: each kernel sums every output in an order fixed by the chunk alone, so
the results must be bit-identical."""
import numpy as np
import pytest

import parity as P
from lyspeq_amd import synth
from lyspeq_amd.qml import QMLEngine

pytestmark = pytest.mark.gpu


def _print_case(name):
    print(f"[case] {name}", flush=True)


def _engine(w, precision, **kw):
    eng = QMLEngine(w.binning, precision=precision, **kw)
    accepted = synth.load_engine(eng, w, list(range(len(w.chunks))))
    eng.commit()
    return eng, accepted


def _run(eng, accepted, theta):
    T = eng.T
    F, d = np.zeros((T, T)), np.zeros((3, T))
    eng.iterate(theta, F, d)
    per = {cid: eng.chunk_result(cid) for _, cid in accepted}
    return F, d, per


def _same(a, b):
    Fa, da, pa = a
    Fb, db, pb = b
    assert np.array_equal(Fa, Fb) and np.array_equal(da, db)
    assert pa.keys() == pb.keys()
    for k in pa:
        assert np.array_equal(pa[k][0], pb[k][0]) and np.array_equal(pa[k][1], pb[k][1])
        assert pa[k][2] == pb[k][2]


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_graph_replay_matches_eager(precision):
    """Replays with a new theta each time, eager iterations and host
    synchronisations between them, a new capture after the stage overlap is
    switched off, and a chunk left out of the sums after capture."""
    _print_case(f"graph replay {precision}")
    w = synth.desi_config(nchunks=24)
    tha = P.theta_perturbed(w.binning, w.pd13, 0.05, 3)
    thb = P.theta_perturbed(w.binning, w.pd13, 0.08, 4)
    # seven chunks per batch: several batches, both arena sets, both streams
    eng, acc = _engine(w, precision, max_batch_chunks=7)
    ea = _run(eng, acc, tha)
    eb = _run(eng, acc, thb)
    assert not np.array_equal(ea[0], eb[0])
    eng.set_graph(True)
    _same(_run(eng, acc, tha), ea)     # capture and first replay
    _same(_run(eng, acc, thb), eb)     # replay with another theta
    _same(_run(eng, acc, thb), eb)
    assert eng.timings()["total_ms"] > 0
    eng.set_graph(False)
    _same(_run(eng, acc, tha), ea)     # eager between replays
    eng.set_graph(True)
    _same(_run(eng, acc, thb), eb)     # a new capture
    _same(_run(eng, acc, tha), ea)
    eng.set_pipeline(False)            # captured again on one stream
    _same(_run(eng, acc, thb), eb)
    _same(_run(eng, acc, tha), ea)
    eng.set_pipeline(True)
    cid0 = acc[0][1]
    eng.set_chunk_active(cid0, False)  # read by the replay from device memory
    g = _run(eng, acc, tha)
    eng.set_graph(False)
    e = _run(eng, acc, tha)
    _same(g, e)
    assert not np.array_equal(g[0], ea[0])
    eng.close()
    print(f"graph replay {precision}: bit-identical to eager", flush=True)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_autotuned_schedule_gives_identical_results(precision):
    _print_case(f"autotune {precision}")
    w = synth.desi_config(nchunks=24)
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 8)
    eng, acc = _engine(w, precision)   # default cap: the 24 chunks in one batch
    e0 = _run(eng, acc, th)
    best, ms = eng.autotune(th, caps=(5, 24, 9), reps=1)
    print(f"autotune {precision}: best cap {best}, {ms:.2f} ms", flush=True)
    assert best in (5, 24, 9) and ms > 0
    _same(_run(eng, acc, th), e0)
    for cap in (3, 11):                # each schedule forced by a one-candidate tune
        b, _ = eng.autotune(th, caps=(cap,), reps=1)
        assert b == cap
        _same(_run(eng, acc, th), e0)
    eng.set_graph(True)                # a graph of the tuned plan
    _same(_run(eng, acc, th), e0)
    _same(_run(eng, acc, th), e0)
    eng.close()


@pytest.mark.parametrize("precision", ["bf16x3", "bf16x4"])
def test_kernel_stage_variants_give_identical_results(precision):
    """The split-BF16 products' operand staging: through registers with one
    or two LDS stages, and direct-to-LDS loads with two or three, for the
    left product, the right product and the Gram.  Every choice, eager, in
    stack slabs of 7 and replayed from a graph captured again after each
    change, must give the default's results bit for bit."""
    _print_case(f"kernel stage variants {precision}")
    w = synth.desi_config(nchunks=24)
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 9)
    eng, acc = _engine(w, precision, max_batch_chunks=10)
    s = eng.schedule()
    assert (s["stages_left"], s["stages_right"], s["stages_gram"]) == (3, 3, 3) and s["n_batches"] >= 2
    e0 = _run(eng, acc, th)
    for st in ((1, 1, 2), (2, 1, 1), (1, 3, 3), (4, 4, 4), (3, 4, 1), (4, 3, 2), (1, 3, 4)):
        print(f"staging {st}", flush=True)
        eng.set_kernel_stages(*st)
        _same(_run(eng, acc, th), e0)
    eng.set_stack_slab(7)
    for st in ((3, 3, 3), (4, 4, 4)):
        eng.set_kernel_stages(*st)
        _same(_run(eng, acc, th), e0)
    eng.set_stack_slab(0)
    eng.set_graph(True)
    for st in ((1, 1, 2), (3, 3, 3), (4, 4, 4)):
        eng.set_kernel_stages(*st)
        _same(_run(eng, acc, th), e0)
        _same(_run(eng, acc, th), e0)
    for bad in ((5, 1, 1), (1, 2, 1), (1, 1, 0)):
        with pytest.raises(Exception):
            eng.set_kernel_stages(*bad)
    best, ms = eng.autotune(th, caps=(10,), reps=1)
    s = eng.schedule()
    print(f"autotuned staging {precision}: left {s['stages_left']} right {s['stages_right']} "
          f"gram {s['stages_gram']}, {ms:.2f} ms", flush=True)
    assert s["stages_left"] in (1, 2, 3, 4) and s["stages_right"] in (1, 3, 4)
    assert s["stages_gram"] in (1, 2, 3, 4) and s["max_batch_chunks"] == 10
    _same(_run(eng, acc, th), e0)
    eng.close()


def test_autotune_rejects_bad_candidates():
    _print_case("autotune arguments")
    w = synth.desi_config(nchunks=4)
    eng, acc = _engine(w, "fp64")
    T = eng.T
    with pytest.raises(Exception):
        eng.autotune(np.zeros(T), caps=(0,))
    with pytest.raises(Exception):
        eng.autotune(np.zeros(T), caps=())
    eng.close()


@pytest.mark.parametrize("precision", ["bf16x3", "bf16x4"])
def test_stack_slabs_set_after_commit_with_graph_replay(precision):
    """Slabs chosen on a committed engine (the batches planned again), with
    both kernel-stage settings, eager and replayed, then back to the whole
    stack: bit-identical to the first plan's results every time."""
    _print_case(f"stack slabs after commit {precision}")
    w = synth.desi_config(nchunks=24)
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 8)
    eng, acc = _engine(w, precision, max_batch_chunks=9)
    ref = _run(eng, acc, th)
    for stages in ((1, 1, 2), (2, 3, 1)):
        eng.set_kernel_stages(*stages)
        eng.set_stack_slab(7)
        assert eng.schedule()["stack_slab"] == 7
        _same(_run(eng, acc, th), ref)
        eng.set_graph(True)
        _same(_run(eng, acc, th), ref)
        _same(_run(eng, acc, th), ref)
        eng.set_graph(False)
        eng.set_stack_slab(0)
        _same(_run(eng, acc, th), ref)
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_batches_report_np_edges_and_bucketed_plans_give_identical_results(precision):
    """qml_get_batch: the plan's batches cover every chunk once, in
    decreasing padded size, at most the cap each.  autotune tries each cap
    with and without a batch break at every change of padded size; with it,
    every batch holds one padded size.  Results are bit-identical under both
    plans."""
    _print_case(f"batch edges and Np buckets {precision}")
    w = synth.desi_config(nchunks=40)
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 12)
    eng, acc = _engine(w, precision, max_batch_chunks=16)
    e0 = _run(eng, acc, th)
    bs = eng.batches()
    print("batches:", bs, flush=True)
    assert sum(b["n_chunks"] for b in bs) == len(acc) and all(b["n_chunks"] <= 16 for b in bs)
    assert all(b["np_min"] <= b["np_max"] for b in bs)
    assert all(bs[i]["np_min"] >= bs[i + 1]["np_max"] for i in range(len(bs) - 1))
    assert eng.schedule()["np_buckets"] == 0
    best, ms = eng.autotune(th, caps=(16, 64), reps=1)
    s = eng.schedule()
    print(f"autotuned cap {best}, np_buckets {s['np_buckets']}, {ms:.2f} ms", flush=True)
    _same(_run(eng, acc, th), e0)
    for on in (1, 0, 1):
        eng.set_np_buckets(on)
        bb = eng.batches()
        assert eng.schedule()["np_buckets"] == on and sum(b["n_chunks"] for b in bb) == len(acc)
        if on:
            assert all(b["np_min"] == b["np_max"] for b in bb)
        _same(_run(eng, acc, th), e0)
    eng.close()
