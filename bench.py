"""Benchmark of the per-chunk QML step on one MI355X.

One step = one Newton-Raphson iteration's chunk loop over the whole batch
(every chunk's covariance, factorisation, marginalisation, d/b/t and Fisher,
accumulated into the global F and d/b/t on the host), with all inputs and
tables already resident in device memory.  The headline value is the
mixed-precision (split-BF16x3 MFMA) throughput on the DESI-like batch; the
FP64 path runs on the same batch for the speed-up and the error figures.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision bf16x3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from lyspeq_amd import synth                                   # noqa: E402
from lyspeq_amd.driver import power_estimates, smooth_power    # noqa: E402
from lyspeq_amd.qml import QMLEngine, KERNEL_NAMES             # noqa: E402

METRIC = "chunks/s (F,d,b) mixed-prec vs FP64 on 1 MI355X; max rel err of F and P1D"
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15, "bf16x4": 2.5e15}
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def chunk_shapes(eng):
    out = []
    for info in eng.chunk_infos:
        out.append((info.size, info.n_stored, info.n_templates))
    return out


def algorithmic_flops(shapes):
    """FLOP per chunk of the reference algorithm:
    2nN^3 (W_k) + 2N^3 (S_fid C^-1) + N^3 (inverse) + n(n+1)N^2 (Fisher)
    + 8nN^2 (assembly, d, b, t) + 4 n_t N^2 (marginalisation)."""
    tot = 0.0
    for N, n, nt in shapes:
        N = float(N)
        tot += 2 * n * N ** 3 + 2 * N ** 3 + N ** 3 + n * (n + 1) * N ** 2 + 8 * n * N ** 2 + 4 * nt * N ** 2
    return tot


def kernel_work(shapes, prec):
    """Algorithmic work of each dense kernel at the real chunk sizes
    (m = n + 2 stack members: the stored Q_k, the noise and the signal)."""
    es = 8 if prec == "fp64" else 4
    w = {"whiten_left": 0.0, "whiten_right": 0.0, "gram": 0.0, "xgen_bytes": 0.0}
    for N, n, nt in shapes:
        m = n + 2
        N = float(N)
        if prec in ("bf16x3", "bf16x4"):
            # X_j = M + M^T: P = Li M (lower x lower), G = P Li^T + Li P^T (upper half)
            w["whiten_left"] += m * N * (N + 1) * (N + 2) / 3.0
            w["whiten_right"] += 2.0 * m * N * (N + 1) * (N + 2) / 3.0
        else:
            w["whiten_left"] += m * N * N * (N + 1)               # lower-triangular x dense
            w["whiten_right"] += m * N * (N + 1) * (N + 2) / 3.0  # upper half, triangular operand
        w["gram"] += m * (m + 1) / 2.0 * N * (N + 1)          # symmetric pairs over packed halves
        w["xgen_bytes"] += m * N * N * es                      # projected Q stack written once
    return w


def fiducial_power(w):
    """P_fid(k_c, z_m) of the workload's bands, in the driver's band order."""
    kc, zc = w.binning.kband_centers(), w.binning.zbin_centers()
    return np.array([[synth.pd13_power(k, z, *w.pd13) for k in kc] for z in zc]).reshape(-1)


def run_engine(w, precision, steps, warmup, scratch_gb, graph=False, tune_caps=(), nr=0, nr_floor=0.0,
               nr_smooth=False, boots=0, stack_slab=0, stages=None, max_batch=0):
    eng = QMLEngine(w.binning, precision=precision, scratch_gb=scratch_gb, stack_slab=stack_slab,
                    max_batch_chunks=max_batch)
    t0 = time.perf_counter()
    accepted = synth.load_engine(eng, w)
    eng.commit()
    log(f"[{precision}] {len(accepted)} chunks registered and committed in {time.perf_counter() - t0:.1f} s")
    T = eng.T
    theta = np.zeros(T)
    if stages and precision not in ("fp64", "fp32"):
        eng.set_kernel_stages(*stages)
    schedule = dict(eng.schedule(), autotuned=False)
    if tune_caps:
        t0 = time.perf_counter()
        cap, tms = eng.autotune(theta, caps=tune_caps, reps=1)
        schedule = dict(eng.schedule(), autotuned=True, candidates=list(tune_caps), tuned_ms=tms,
                        tuning_s=time.perf_counter() - t0)
        log(f"[{precision}] autotuned batch cap {cap}, Np buckets {schedule['np_buckets']}, operand staging "
            f"left {schedule['stages_left']} right {schedule['stages_right']} gram {schedule['stages_gram']} "
            f"({tms:.1f} ms) in {schedule['tuning_s']:.1f} s")
    # the plan's batches: chunk count and padded-size range (Np bucket edges)
    schedule["batches"] = [[b["n_chunks"], b["np_min"], b["np_max"]] for b in eng.batches()]
    if graph:
        eng.set_graph(True)
    for i in range(warmup):
        F, d = np.zeros((T, T)), np.zeros((3, T))
        eng.iterate(theta, F, d)
        log(f"[{precision}] warmup {i + 1}/{warmup} done")
    times = []
    for i in range(steps):
        F, d = np.zeros((T, T)), np.zeros((3, T))
        t0 = time.perf_counter()
        eng.iterate(theta, F, d)
        times.append(time.perf_counter() - t0)
        log(f"[{precision}] step {i + 1}/{steps}: {times[-1] * 1e3:.1f} ms")
    # one more (untimed) iteration with the two stages serialised on one
    # stream: its per-kernel event durations are those of each kernel alone
    eng.set_graph(False)
    eng.set_pipeline(False)
    F2, d2 = np.zeros((T, T)), np.zeros((3, T))
    t0 = time.perf_counter()
    eng.iterate(theta, F2, d2)
    serial_ms = (time.perf_counter() - t0) * 1e3
    eng.set_pipeline(True)
    log(f"[{precision}] serialised iteration for kernel timings: {serial_ms:.1f} ms")
    if not (np.array_equal(F2, F) and np.array_equal(d2, d)):
        raise RuntimeError("serialised and overlapped iterations differ")
    tm = eng.timings()
    shapes = chunk_shapes(eng)
    # Newton-Raphson chain (untimed): each iteration's theta from the
    # previous one's sums by the driver arithmetic, no smoother
    chain = None
    if nr > 0:
        th = np.zeros(T)
        chain = dict(ms=[], theta=[], sigma=[], rejected=[], damp=[])
        for i in range(nr):
            Fi, di = np.zeros((T, T)), np.zeros((3, T))
            t0 = time.perf_counter()
            eng.iterate(th, Fi, di)
            chain["ms"].append((time.perf_counter() - t0) * 1e3)
            res = power_estimates(Fi, di)
            th = res["theta"]
            chain["theta"].append(th)
            if nr_smooth:      # the reference's smoothing of the fits (py/smbivspline.py)
                th = smooth_power(th, res["sigma"], w.binning.kband_centers(), w.binning.zbin_centers(),
                                  fiducial_power(w))
            elif nr_floor > 0:   # keep S + theta Q a covariance without smoothing
                th = np.maximum(th, -nr_floor * fiducial_power(w))
            chain["sigma"].append(res["sigma"])
            chain["rejected"].append(sum(eng.chunk_result(cid)[2] != 0 for _, cid in accepted))
            chain["damp"].append(float(res["damp"]))
            log(f"[{precision}] Newton-Raphson iteration {i + 1}/{nr}: {chain['ms'][-1]:.1f} ms, "
                f"{chain['rejected'][-1]} chunks rejected, damping {chain['damp'][-1]:.3g}")
    boot = None
    if boots > 0:   # untimed in the metric: the FastBootstrap and the complete form on the last sums
        eng.iterate(np.zeros(T), np.zeros((T, T)), np.zeros((3, T)))
        qso = np.arange(len(accepted), dtype=np.int32)
        t0 = time.perf_counter()
        eng.bootstrap(boots, qso, len(accepted), seed=0, realizations=False)
        t_fast = time.perf_counter() - t0
        t0 = time.perf_counter()
        X, valid, _, _ = eng.bootstrap_complete(boots, qso, len(accepted), seed=0)
        t_full = time.perf_counter() - t0
        R, _, _ = eng.bootstrap(boots, qso, len(accepted), seed=0, realizations=True)
        t0 = time.perf_counter()
        eng.bootstrap_median(R)
        t_median = time.perf_counter() - t0
        boot = {"realisations": boots, "spectra": len(accepted), "fast_s": t_fast, "complete_s": t_full,
                "complete_valid": int(valid.sum()), "median_s": t_median}
        log(f"[{precision}] bootstrap of {boots}: fast {t_fast:.2f} s, complete {t_full:.2f} s "
            f"({boot['complete_valid']} valid), median covariance of the fast realisations {t_median:.3f} s")
    eng.close()
    return dict(F=F, dbt=d, times=times, timings=tm, n=len(accepted), shapes=shapes, serial_ms=serial_ms,
                schedule=schedule, chain=chain, boot=boot)


def copy_calibration(mb: int):
    """A plain device-to-device copy of mb MiB (the blit kernel of the HIP
    runtime), run twice: the reference for scaling the FETCH_SIZE and
    WRITE_SIZE counters of the engine's kernels to bytes."""
    import torch
    n = mb * 1024 * 1024
    x = torch.ones(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n, "copies": 2}), flush=True)


TRAFFIC_DIR = os.path.join(ROOT, "profiles")


def measured_traffic(kernel: str, precision: str, workload: str):
    """HBM bytes per launch of the dominant kernel from the committed PMC
    summary (profiles/traffic_<kernel>_<precision>_<workload>.json), if one
    was recorded for this configuration."""
    path = os.path.join(TRAFFIC_DIR, f"traffic_{kernel}_{precision}_{workload}.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    return float(t["hbm_bytes_per_launch"]), os.path.relpath(path, ROOT)


def measured_mfma_busy(kernel: str, precision: str):
    """MFMA-busy fraction of a dense product from the committed counter
    summary (profiles/mfma_util_products_desi-4096.json), if recorded."""
    path = os.path.join(TRAFFIC_DIR, "mfma_util_products_desi-4096.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    v = t.get("mfma_busy", {}).get(precision, {}).get(kernel)
    return v, (os.path.relpath(path, ROOT) if v is not None else None)


def _noop(_):
    return 0


def _cpu_worker_init():
    """A worker of the rank-emulation CPU baseline: two BLAS threads, as the
    reference runs each MPI rank ("-c 2")."""
    global _WORKER_LIMITS
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from threadpoolctl import threadpool_limits
    _WORKER_LIMITS = threadpool_limits(2)
    import oracle.chunk  # noqa: F401  (imported before any timed job)


def _cpu_worker_job(args):
    b, tabs, chunks = args
    from oracle.chunk import Chunk as OChunk, ChunkError
    T = b.total_kz_bins
    F, d, th = np.zeros((T, T)), np.zeros((3, T)), np.zeros(T)
    done = 0
    for c in chunks:
        try:
            ch = OChunk(b, tabs, c["wave"], c["delta"], c["ivar"], c["R"])
        except ChunkError:
            continue
        ch.one_iteration(th, F, d)
        done += 1
    return done


# host cores a one-GPU run may use: the box allots 16 per GPU, whatever
# nproc shows (its other cores belong to the other GPUs' jobs)
CPU_SHARE = 16


def _cores():
    """(nproc, cores in this process's affinity mask, cores used)."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    return nproc, aff, max(1, min(CPU_SHARE, aff))


def start_cpu_pool():
    """Workers of the rank-emulation baseline (cores used / 2 processes),
    forked by a fork server that is started before the GPU is touched, so no
    worker comes from a process that has initialised the device."""
    import multiprocessing as mp
    n = max(1, _cores()[2] // 2)
    pool = mp.get_context("forkserver").Pool(n, initializer=_cpu_worker_init)
    pool.map(_noop, range(n))
    return pool, n


def _cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), model)
    except OSError:
        pass
    try:
        from threadpoolctl import threadpool_info
        blas = sorted({f"{p.get('internal_api')}" for p in threadpool_info()})
    except Exception:
        blas = []
    return model, blas


def cpu_baseline(w, budget_s=20.0, max_chunks=48, pool=None):
    """The CPU oracle (numpy + LAPACK/BLAS) on the first chunks of the same
    batch, timed in two modes: (A) chunks one after another with threaded
    BLAS on all host cores, until budget_s of work is spent; (B) the same
    chunks spread over worker processes of two BLAS threads each, as the
    reference runs its MPI ranks.  The better mode is the baseline."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import parity as P
    from oracle.chunk import Chunk as OChunk, ChunkError
    try:
        from threadpoolctl import threadpool_info
        threads = max([p.get("num_threads", 1) for p in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    b = P.oracle_bins(w.binning)
    tabs = P.oracle_tables(w)
    T = b.total_kz_bins
    F, d = np.zeros((T, T)), np.zeros((3, T))
    th = np.zeros(T)
    done = 0
    used = []
    phases = {}
    t0 = time.perf_counter()
    for c in w.chunks[:max_chunks]:
        used.append(c)
        try:
            ch = OChunk(b, tabs, c["wave"], c["delta"], c["ivar"], c["R"])
        except ChunkError:
            continue
        ch.one_iteration(th, F, d, times=phases)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    # the driver's F inversion (T_Finv) once on the sample's sum
    from oracle.driver import power_estimates
    tf = time.perf_counter()
    power_estimates(F.copy(), d)
    phases["f_inv"] = time.perf_counter() - tf
    model, blas = _cpu_info()
    nproc, aff, _ = _cores()
    modes = {"A_serial_threaded_blas": {"chunks_per_s": done / el, "threads": int(threads), "seconds": el,
                                        "phase_seconds": {k: round(v, 3) for k, v in phases.items()}}}
    best, best_mode, cores = done / el, "A", int(threads)
    if pool is not None:
        p, n = pool
        need = {(c["R"], c.get("dv_table")) for c in used}
        sub = [t for t in tabs if (t.R, t.dv) in need] or tabs
        jobs = [(b, sub, used[k::n]) for k in range(n)]
        t1 = time.perf_counter()
        doneb = sum(p.map(_cpu_worker_job, jobs))
        elb = time.perf_counter() - t1
        modes["B_rank_emulation"] = {"chunks_per_s": doneb / elb, "workers": n, "threads_per_worker": 2,
                                     "seconds": elb}
        if doneb / elb > best:
            best, best_mode, cores = doneb / elb, "B", 2 * n
    return dict(value=best, unit="chunks/s", cores=cores, kind="port", mode=best_mode, modes=modes,
                cpu_model=model, blas=blas, nproc=nproc, affinity_cores=aff, cpu_share=CPU_SHARE,
                phases="mode A seconds per mytime category (core/global_numbers.cpp:476-502): set_qs = T_Q, "
                       "set_sfid = T_Sfid, c_inv = T_Cinv, set_modqs = T_Qmod, set_fisher = T_F, "
                       "f_inv = T_Finv; dbt_rest the b, t contractions",
                sample=f"first {done} chunks of the same batch, FP64 LU oracle; mode A {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16x3", choices=["fp32", "bf16x3", "bf16x4"])
    ap.add_argument("--config", default="desi", choices=list(synth.CONFIGS))
    ap.add_argument("--nchunks", type=int, default=0)
    ap.add_argument("--scratch-gb", type=float, default=96.0)
    ap.add_argument("--launch", default="eager", choices=["eager", "graph"],
                    help="eager launches, or the iteration captured once into a HIP graph and replayed")
    ap.add_argument("--autotune", default="", metavar="CAPS",
                    help="comma-separated caps on chunks per batch for the autotuned schedule "
                         "(tuned before the warm-up, outside the timed steps)")
    ap.add_argument("--nr", type=int, default=0, metavar="N",
                    help="after the timed steps, run N Newton-Raphson iterations (theta fed forward, no "
                         "smoother) per precision and report the final P1D deviation of the mixed chain")
    ap.add_argument("--nr-floor", type=float, default=0.0, metavar="F",
                    help="feed theta forward as max(theta, -F P_fid) (a stand-in for the reference's smoother)")
    ap.add_argument("--nr-smooth", action="store_true",
                    help="feed the smoothed fits forward, as the reference's driver does")
    ap.add_argument("--bootstrap", type=int, default=0, metavar="N",
                    help="after the timed steps, time both bootstrap forms with N realisations "
                         "(every chunk its own spectrum)")
    ap.add_argument("--save-sums", default="", metavar="NPZ",
                    help="write the FP64 and mixed Fisher and d, b, t sums to this .npz")
    ap.add_argument("--stack-slab", type=int, default=-1, metavar="M",
                    help="stream each chunk's stack in slabs of M members (0: whole stack; default 16 for "
                         "the high-resolution workload, else 0)")
    ap.add_argument("--stages", default="", metavar="L,R,G",
                    help="operand staging of the split-BF16 left, right and Gram products (1, 2: through "
                         "registers with one or two LDS stages; 3, 4: direct-to-LDS loads with two or three "
                         "stages); default: the engine's")
    ap.add_argument("--max-batch", type=int, default=0, metavar="N",
                    help="cap on chunks per batch (0: the engine's default, 256)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain MIB-MiB device copy (PMC counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; sharding chunks over several GPUs (one engine per "
              "rank, an all-reduce of the sums) is not built, so --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    # the CPU baseline's worker processes start before anything touches the GPU
    pool = start_cpu_pool() if not args.no_cpu_baseline else None

    t0 = time.perf_counter()
    kw = {"nchunks": args.nchunks} if args.nchunks else {}
    w = synth.CONFIGS[args.config](**kw)
    log(f"workload {w.name}: {len(w.chunks)} chunks, {len(w.pairs)} (R,dv) tables, "
        f"generated in {time.perf_counter() - t0:.1f} s")

    caps = tuple(int(x) for x in args.autotune.split(",") if x.strip())
    slab = args.stack_slab if args.stack_slab >= 0 else (16 if args.config == "hires" else 0)
    opt = dict(graph=args.launch == "graph", tune_caps=caps, nr=args.nr, nr_floor=args.nr_floor,
               nr_smooth=args.nr_smooth, boots=args.bootstrap, stack_slab=slab,
               stages=tuple(int(x) for x in args.stages.split(",")) if args.stages else None,
               max_batch=args.max_batch)
    res64 = None
    if not args.no_fp64:
        res64 = run_engine(w, "fp64", args.steps, args.warmup, args.scratch_gb, **opt)
    resm = run_engine(w, args.precision, args.steps, args.warmup, args.scratch_gb, **opt)

    nch = resm["n"]
    ms = 1e3 * float(np.mean(resm["times"]))
    value = nch / (ms / 1e3)
    cfg = {"workload": f"{w.name}: {w.description}", "chunks": nch, "tables": len(w.pairs),
           "precision": args.precision, "theta": "0 (first iteration)", "launch": args.launch,
           "schedule": resm["schedule"]}
    if res64:
        ms64 = 1e3 * float(np.mean(res64["times"]))
        cfg["fp64_chunks_per_s"] = nch / (ms64 / 1e3)
        cfg["fp64_ms_per_step"] = ms64
        cfg["speedup_vs_fp64"] = ms64 / ms
        cfg["fp64_schedule"] = res64["schedule"]
        F64, d64 = res64["F"], res64["dbt"]
        Fm, dm = resm["F"], resm["dbt"]
        dF = np.sqrt(np.abs(np.diag(F64)))
        den = np.outer(dF, dF)
        iu = np.triu_indices(F64.shape[0])
        msk = den[iu] > 0
        errs = {"F_normalized": float(np.max(np.abs(Fm - F64)[iu][msk] / den[iu][msk]))}
        big = np.abs(F64[iu]) > 1e-6 * np.abs(F64).max()
        errs["F_elementwise"] = float(np.max(np.abs(Fm - F64)[iu][big] / np.abs(F64[iu][big])))
        for k, name in enumerate("dbt"):
            errs[name] = float(np.max(np.abs(dm[k] - d64[k])) / max(np.abs(d64[k]).max(), 1e-300))
        p64 = power_estimates(F64, d64)
        pm = power_estimates(Fm, dm)
        ok = np.isfinite(p64["sigma"]) & (p64["sigma"] > 0)
        dev = np.abs(pm["theta"] - p64["theta"])[ok] / p64["sigma"][ok]
        errs["P1D_max_dev_over_sigma"] = float(dev.max())
        errs["P1D_rel"] = float(np.max(np.abs(pm["theta"] - p64["theta"])[ok]) /
                                np.max(np.abs(p64["theta"][ok])))
        cfg["max_rel_err_vs_fp64"] = errs
        if args.nr and res64["chain"] and resm["chain"]:
            c64, cm = res64["chain"], resm["chain"]
            devs = []
            for a, b, sg in zip(cm["theta"], c64["theta"], c64["sigma"]):   # each iteration's own sigma
                okc = np.isfinite(sg) & (sg > 0)
                devs.append(float(np.max(np.abs(a - b)[okc] / sg[okc])))
            sg = c64["sigma"][-1]
            okc = np.isfinite(sg) & (sg > 0)
            cfg["newton_raphson"] = {"iterations": args.nr, "smoothed_fits": args.nr_smooth,
                                     "theta_floor_times_pfid": args.nr_floor,
                                     "fp64_ms": c64["ms"], "mixed_ms": cm["ms"],
                                     "P1D_max_dev_over_sigma_per_iteration": devs,
                                     "rejected_chunks_fp64": c64["rejected"], "rejected_chunks_mixed": cm["rejected"],
                                     "damping_fp64": c64["damp"], "damping_mixed": cm["damp"],
                                     "theta_rms_over_sigma_final": float(np.sqrt(np.mean(
                                         (c64["theta"][-1][okc] / sg[okc]) ** 2)))}
            log("Newton-Raphson chain:", json.dumps(cfg["newton_raphson"]))
        if args.save_sums:
            np.savez(args.save_sums, F64=F64, dbt64=d64, Fm=Fm, dbtm=dm)
        log("errors vs fp64:", json.dumps(errs))
        log("fp64 kernel ms:", json.dumps({k: round(v, 2) for k, v in res64["timings"]["kernel_ms"].items()}))
    log(f"{args.precision} kernel ms:", json.dumps({k: round(v, 2) for k, v in resm["timings"]["kernel_ms"].items()}))

    # roofline of the dominant kernel of the headline (mixed) run
    km = resm["timings"]["kernel_ms"]
    nb = max(1, resm["timings"]["n_batches"])
    dom = max(km, key=lambda k: km[k])
    work = kernel_work(resm["shapes"], args.precision)
    if dom in ("whiten_left", "whiten_right", "gram"):
        ach = work[dom] / (km[dom] / 1e3)
        traffic, tsrc = measured_traffic(dom, args.precision, w.name) if not args.nchunks else (None, None)
        roof = {"kernel": dom, "bound": "mfma", "achieved": ach / 1e12, "peak": PEAK[args.precision] / 1e12,
                "unit": "TFLOP/s", "frac": ach / PEAK[args.precision], "traffic": traffic,
                "launches": nb, "avg_launch_ms": km[dom] / nb}
        busy, bsrc = measured_mfma_busy(dom, args.precision)
        if bsrc:
            # issued MFMA-pipe occupancy from counters (three MFMAs per split product)
            roof["mfma_busy_issued"] = busy
            roof["mfma_busy_source"] = bsrc
        if tsrc:
            roof["traffic_source"] = tsrc
            # the same launches against the HBM roof: at split-BF16 operand
            # planes (4 bytes per element) the whitening products are close to it
            roof["hbm_achieved_GBps"] = traffic / (roof["avg_launch_ms"] / 1e3) / 1e9
            roof["hbm_frac"] = roof["hbm_achieved_GBps"] * 1e9 / HBM_PEAK
    else:
        byt = work["xgen_bytes"] if dom == "xgen" else None
        ach = (byt / (km[dom] / 1e3)) if byt else None
        roof = {"kernel": dom, "bound": "hbm", "achieved": (ach / 1e9) if ach else None, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": (ach / HBM_PEAK) if ach else None, "traffic": None,
                "launches": nb, "avg_launch_ms": km[dom] / nb}
    if resm.get("boot"):
        cfg["bootstrap"] = resm["boot"]
    step_flops = algorithmic_flops(resm["shapes"])
    cfg["step_algorithmic_tflop"] = step_flops / 1e12
    cfg["step_tflops_achieved"] = step_flops / (ms / 1e3) / 1e12

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(w, pool=pool)
        log("cpu baseline:", json.dumps(cpu))
    if pool is not None:
        pool[0].close()
        pool[0].join()

    out = {"metric": METRIC, "value": value, "unit": "chunks/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": args.precision, "data": "synthetic", "config": cfg,
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
