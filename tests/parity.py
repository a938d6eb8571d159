"""Helpers that run the CPU oracle and the HIP engine on the same chunks."""
import numpy as np

from oracle.bins import Bins
from oracle.chunk import Chunk as OracleChunk, ChunkError, ResoMatrix, TableSet
from lyspeq_amd.qml import Binning, QMLEngine
from lyspeq_amd import synth


def oracle_bins(b: Binning) -> Bins:
    return Bins(k0=b.k0, nklin=b.nklin, dklin=b.dklin, nklog=b.nklog, dklog=b.dklog, z0=b.z0,
                z_bin_width=b.z_bin_width, nz=b.nz, klast=b.klast, method=b.method,
                cont_logl_order=b.cont_logl_order, cont_lam_order=b.cont_lam_order,
                turn_off_sfid=b.turn_off_sfid, max_forest_length_v=b.max_forest_length_v,
                number_of_chunks=b.number_of_chunks, max_pixel_length_v=b.max_pixel_length_v,
                redshift_growth_on=b.redshift_growth_on, fiducial_redshift_power=b.fiducial_redshift_power)


def oracle_tables(w: synth.Workload):
    return [TableSet(R, dv, w.nv, w.v_length, w.nz_table, w.deriv[i], w.signal[i].reshape(-1))
            for i, (R, dv) in enumerate(w.pairs)]


def oracle_rmat(c):
    """The workload chunk's resolution matrix for the oracle, or None."""
    if "rmat" not in c:
        return None
    rm = c["rmat"]
    return ResoMatrix(rm["kind"], np.asarray(rm["values"]), rm["nelem"], rm["oversampling"])


def theta_perturbed(b: Binning, pd13, scale, seed):
    """theta_k = scale * P_fid(k_c, z_m) * N(0,1)."""
    rng = np.random.default_rng(seed)
    kc = b.kband_centers()
    zc = b.zbin_centers()
    pf = np.array([[synth.pd13_power(k, z, *pd13) for k in kc] for z in zc]).reshape(-1)
    return scale * pf * rng.standard_normal(pf.size)


def run_oracle(w: synth.Workload, theta, idx, exact=False):
    """Per-chunk (F, dbt) of the oracle for workload chunks idx, plus the sums
    (exact=True: the oracle's extended-precision v/z and template basis, for
    measuring a case's rounding floor)."""
    b = oracle_bins(w.binning)
    tabs = oracle_tables(w)
    T = b.total_kz_bins
    Fs, ds = np.zeros((T, T)), np.zeros((3, T))
    out = {}
    for i in idx:
        c = w.chunks[i]
        try:
            ch = OracleChunk(b, tabs, c["wave"], c["delta"], c["ivar"], c["R"], rmat=oracle_rmat(c), exact=exact)
        except ChunkError:
            continue
        ok = ch.one_iteration(theta, Fs, ds)
        out[i] = (ch.fisher.copy(), ch.dbt.copy(), ok, ch)
    return out, Fs, ds


def run_engine(w: synth.Workload, theta, idx, precision="fp64", **kw):
    eng = QMLEngine(w.binning, precision=precision, **kw)
    accepted = synth.load_engine(eng, w, idx)
    eng.commit()
    T = eng.T
    Fs, ds = np.zeros((T, T)), np.zeros((3, T))
    eng.iterate(theta, Fs, ds)
    out = {}
    for i, cid in accepted:
        F, dbt, st = eng.chunk_result(cid)
        out[i] = (F, dbt, st == 0)
    return eng, out, Fs, ds


def rel_err(a, b):
    """max |a - b| / max |b| (zero when both vanish)."""
    a, b = np.asarray(a), np.asarray(b)
    den = np.max(np.abs(b))
    return 0.0 if den == 0 else float(np.max(np.abs(a - b)) / den)


def fisher_norm_err(Fa, Fb):
    """max |dF_kk'| / sqrt(F_kk F_k'k') over the upper triangle."""
    d = np.sqrt(np.abs(np.diag(Fb)))
    den = np.outer(d, d)
    iu = np.triu_indices(Fb.shape[0])
    m = den[iu] > 0
    return float(np.max(np.abs(Fa - Fb)[iu][m] / den[iu][m])) if m.any() else 0.0


def vec_norm_err(va, vb, Fb):
    """max |dv_k| / sqrt(F_kk) over the bands with F_kk > 0: a d/b/t error in
    units of the band's own 1-sigma (theta = F^-1 (d - b - t))."""
    d = np.sqrt(np.abs(np.diag(Fb)))
    m = d > 0
    return float(np.max(np.abs(np.asarray(va) - np.asarray(vb))[m] / d[m])) if m.any() else 0.0


def elem_err(a, b, floor):
    """max |a - b| / |b| over the entries with |b| >= floor * max |b| (upper
    triangle for a square matrix)."""
    a, b = np.asarray(a), np.asarray(b)
    if b.ndim == 2:
        iu = np.triu_indices(b.shape[0])
        a, b = a[iu], b[iu]
    big = np.abs(b) >= floor * np.max(np.abs(b)) if b.size else np.zeros(0, bool)
    big &= b != 0
    return float(np.max(np.abs(a[big] - b[big]) / np.abs(b[big]))) if big.any() else 0.0


# FP64 bars.  Normalised: F by sqrt(F_kk F_k'k'), d/b/t by sqrt(F_kk).
# Elementwise: relative error of every entry at least ELEM_FLOOR times the
# largest of its quantity.  The rounding floor of the two inversions (the
# oracle's LU against its own Cholesky form, DESI chunks) is about 1e-14
# normalised and 1e-12 elementwise at this floor.
FP64_TOL = 1e-10
ELEM_FLOOR = 1e-6
ELEM_TOL = 1e-9


def errors(Fa, dbta, Fb, dbtb, floor=ELEM_FLOOR):
    """Normalised and elementwise errors of (Fa, dbta) against (Fb, dbtb)."""
    e = dict(F_norm=fisher_norm_err(Fa, Fb), F_elem=elem_err(Fa, Fb, floor), F_max=rel_err(Fa, Fb))
    for k, nm in enumerate("dbt"):
        e[f"{nm}_norm"] = vec_norm_err(dbta[k], dbtb[k], Fb)
        e[f"{nm}_elem"] = elem_err(dbta[k], dbtb[k], floor)
    return e


def fmt(e):
    return " ".join(f"{k} {v:.1e}" for k, v in e.items())


# Where the reference's own rounding moves a case's results by more than the
# flat bar (resolution-matrix windows, whose phase k v carries the rounding of
# v = c ln(l_j / l_i); template sets with s_min / s_0 far below 1, whose span
# dgesvd fixes only to about eps s_0 / s_min), the case's floor is the
# oracle's deviation from its own extended-precision run.  The engine must
# then lie no farther from that run than FLOOR_FACTOR floors, and hence no
# farther from the oracle than FLOOR_FACTOR + 1.
FLOOR_FACTOR = 3.0


def _check(e, what, tol, elem_tol, floor, factor, label):
    def bar(k):
        base = tol if k.endswith("_norm") else elem_tol
        return max(base, factor * floor[k]) if floor and k in floor else base
    bad = {k: (v, bar(k)) for k, v in e.items()
           if (k.endswith("_norm") or k.endswith("_elem")) and not v < bar(k)}
    bad.update({k: (v, None) for k, v in e.items() if not np.isfinite(v)})
    assert not bad, f"{what}: FP64 parity bar missed{label} (value, bar): {bad}"


def assert_fp64(Fa, dbta, Fb, dbtb, what="", tol=FP64_TOL, elem_tol=ELEM_TOL, floor=None, exact=None):
    """The FP64 parity bar: every normalised error below tol and every
    elementwise error (entries above the floor) below elem_tol.  With the
    case's rounding floor and the oracle's extended-precision run (Fx, dbtx)
    as exact, the bars widen to FLOOR_FACTOR + 1 floors against the oracle
    and FLOOR_FACTOR floors against that run."""
    e = errors(Fa, dbta, Fb, dbtb)
    print(f"[fp64 parity] {what}: {fmt(e)}", flush=True)
    if floor:
        print(f"[fp64 parity] {what} floor (oracle vs its extended-precision run): {fmt(floor)}", flush=True)
    _check(e, what, tol, elem_tol, floor, FLOOR_FACTOR + 1, " against the oracle")
    if exact is not None:
        ex = errors(Fa, dbta, exact[0], exact[1])
        print(f"[fp64 parity] {what} engine vs the oracle's extended-precision run: {fmt(ex)}", flush=True)
        _check(ex, what, tol, elem_tol, floor, FLOOR_FACTOR, " against the extended-precision run")
    return e


def worst_errors(out, oout):
    """The worst of each error over the chunks both sides accepted."""
    assert set(out) == set(oout), "accepted chunks differ"
    worst = {}
    for i in sorted(out):
        for k, v in errors(out[i][0], out[i][1], oout[i][0], oout[i][1]).items():
            worst[k] = max(worst.get(k, 0.0), v)
    return worst


def oracle_floor(w, theta, idx, oout=None):
    """(floor, extended-precision per-chunk results, their sums): the case's
    rounding floor is the worst per-chunk deviation of the oracle from its
    extended-precision run (oracle.chunk.Chunk exact=True)."""
    if oout is None:
        oout = run_oracle(w, theta, idx)[0]
    xout, Fx, dx = run_oracle(w, theta, idx, exact=True)
    return worst_errors(oout, xout), xout, (Fx, dx)


def assert_chunks_fp64(out, oout, what="", tol=FP64_TOL, elem_tol=ELEM_TOL, floor=None, exact=None):
    """assert_fp64 on every chunk both sides accepted (the worst of each
    error over the chunks); exact: the oracle's extended-precision per-chunk
    results.  Returns the worst errors against the oracle."""
    assert set(out) == set(oout), f"{what}: accepted chunks differ"
    worst = worst_errors(out, oout)
    print(f"[fp64 parity] {what} worst per chunk over {len(out)}: {fmt(worst)}", flush=True)
    if floor:
        print(f"[fp64 parity] {what} floor (oracle vs its extended-precision run): {fmt(floor)}", flush=True)
    _check(worst, what, tol, elem_tol, floor, FLOOR_FACTOR + 1, " per chunk against the oracle")
    if exact is not None:
        wx = worst_errors(out, exact)
        print(f"[fp64 parity] {what} engine vs the oracle's extended-precision run: {fmt(wx)}", flush=True)
        _check(wx, what, tol, elem_tol, floor, FLOOR_FACTOR, " per chunk against the extended-precision run")
    return worst


def p1d_deviation(Fa, dbta, Fb, dbtb):
    """max |theta_a - theta_b| / sigma_b over the bands with sigma_b > 0, both
    power estimates from the driver's arithmetic (0.5 scaling, the stable
    inverse, theta = 0.5 F^-1 (d - b - t))."""
    from oracle.driver import power_estimates
    ra, rb = power_estimates(Fa.copy(), dbta), power_estimates(Fb.copy(), dbtb)
    s = rb["sigma"]
    m = np.isfinite(s) & (s > 0)
    return float(np.max(np.abs(ra["theta"] - rb["theta"])[m] / s[m]))


# Mixed-precision bars against the FP64 path, per workload and precision:
# about ten times the errors measured on MI355X for the same chunks (the
# kernels are deterministic).  Normalised F by sqrt(F_kk F_k'k'), b and t by
# sqrt(F_kk); d is FP64 in every mode and must agree to rounding.  Single
# high-resolution chunks carry a noise bias b far above sqrt(F_kk), so their
# b bar is the loosest.
MIXED_TOL = {
    ("desi", "fp32"): dict(F_norm=4e-5, b_norm=2e-4, t_norm=3e-5),
    ("desi", "bf16x3"): dict(F_norm=3e-4, b_norm=5e-4, t_norm=4e-4),
    ("desi", "bf16x4"): dict(F_norm=2e-4, b_norm=3e-4, t_norm=2e-4),
    ("hires", "fp32"): dict(F_norm=6e-5, b_norm=5e-3, t_norm=4e-4),
    ("hires", "bf16x3"): dict(F_norm=1.2e-3, b_norm=1.2e-2, t_norm=4e-3),
    ("hires", "bf16x4"): dict(F_norm=6e-4, b_norm=8e-3, t_norm=2e-3),
    ("rmat", "fp32"): dict(F_norm=3e-5, b_norm=7e-5, t_norm=3e-5),
    ("rmat", "bf16x3"): dict(F_norm=1.2e-4, b_norm=3e-4, t_norm=4e-4),
    ("rmat", "bf16x4"): dict(F_norm=1e-4, b_norm=2e-4, t_norm=3e-4),
    ("growth", "fp32"): dict(F_norm=7e-5, b_norm=1.5e-4, t_norm=3e-5),
    ("growth", "bf16x3"): dict(F_norm=1.3e-4, b_norm=3e-4, t_norm=4e-4),
}
# P1D bar, max |dP| / sigma: 1e-3, or ten times the measured value where
# that is lower (DESI sample: 4.9e-5; high-resolution sample in BF16x4:
# 4.6e-4, so the 1e-3 bar itself)
P1D_TOL = {("desi", "bf16x3"): 5e-4, ("hires", "bf16x4"): 1e-3}


def assert_mixed(Fm, dm, F64, d64, precision, what="", workload="desi"):
    """Errors of a mixed-precision run against the FP64 run on the same
    chunks: normalised F, d, b, t, and b/t relative to their largest entry."""
    e = dict(F_norm=fisher_norm_err(Fm, F64),
             d_norm=vec_norm_err(dm[0], d64[0], F64), d_max=rel_err(dm[0], d64[0]),
             b_norm=vec_norm_err(dm[1], d64[1], F64), b_max=rel_err(dm[1], d64[1]),
             t_norm=vec_norm_err(dm[2], d64[2], F64), t_max=rel_err(dm[2], d64[2]))
    print(f"[mixed] {what}: {fmt(e)}", flush=True)
    assert np.isfinite(Fm).all() and np.isfinite(dm).all(), f"{what}: non-finite results"
    assert e["d_max"] < 1e-12, f"{what}: d_k is FP64 in every precision"
    bad = {k: (e[k], tol) for k, tol in MIXED_TOL[(workload, precision)].items() if not e[k] < tol}
    assert not bad, f"{what}: mixed-precision bar missed (value, bar): {bad}"
    return e
