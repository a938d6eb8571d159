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


def run_oracle(w: synth.Workload, theta, idx):
    """Per-chunk (F, dbt) of the oracle for workload chunks idx, plus the sums."""
    b = oracle_bins(w.binning)
    tabs = oracle_tables(w)
    T = b.total_kz_bins
    Fs, ds = np.zeros((T, T)), np.zeros((3, T))
    out = {}
    for i in idx:
        c = w.chunks[i]
        try:
            ch = OracleChunk(b, tabs, c["wave"], c["delta"], c["ivar"], c["R"], rmat=oracle_rmat(c))
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
