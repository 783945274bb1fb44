#!/usr/bin/env python3
"""Benchmark: time to an FP64 true relative residual of 1e-8 with ILU-FGMRES
on the 128^3 convection-diffusion problem, mixed precision vs FP64, plus the
ParILU sweep's achieved HBM bandwidth.

Workload (one step = one complete solve with A and b resident in HBM):
  A: 3-D 7-point first-order upwind convection-diffusion, 128^3 grid
     (n = 2,097,152, nnz = 14,581,760), cell Peclet (2, 1, 0.5), Dirichlet.
  b: N(0,1), Philox4x32-10 seed 0;  x0 = 0.
  step = ParILU(0) setup from A (3 synchronous sweeps) + FGMRES to
         ||b - A x||_2 / ||b||_2 <= 1e-8 (verified in FP64 on the device).
  mixed path: outer Householder FGMRES(10) in FP64 (reflectors, Givens,
     least squares, SpMV with A in FP64) preconditioned by 10 steps of FP32
     Householder GMRES with FP32 ParILU factors and Jacobi-5 triangular
     solves (FP64 flexible iterative refinement of an FP32 solve).
  FP64 path: Householder FGMRES(50), FP64 factors, Jacobi-5 triangular
     solves (the faster of the level-scheduled and Jacobi-5 FP64 variants).

Prints one JSON line.  --gpus N with N > 1 is refused: this build is a
single-GPU implementation.
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

METRIC = "time-to-1e-8 residual: mixed vs FP64 ILU-FGMRES; ParILU sweep HBM GB/s"
HBM_PEAK_GBS = 8000.0


def make_opts(L, **kw):
    o = L.default_opts()
    o.rtol = 1e-8
    o.parilu_sweeps = 3
    o.restart_max = 100
    for k, v in kw.items():
        setattr(o, k, v)
    return o


MIXED = dict(search_max=10, inner_iters=10, trsv=1, jacobi_sweeps=5, prec_precond=1)
FP64 = dict(search_max=50, trsv=1, jacobi_sweeps=5, prec_precond=0)
# the mixed path's nested structure entirely in FP64: outer FGMRES(10) x
# inner FP64 GMRES(10) with FP64 factors
FP64_NESTED = dict(search_max=10, inner_iters=10, inner_fp64=1, trsv=1, jacobi_sweeps=5, prec_precond=0)
# the mixed path with BF16 factor storage (per-row power-of-two scales)
MIXED_BF16 = dict(MIXED, prec_precond=2)
# the reference's own method on the GPU: FP64 FGMRES(50), FP64 factors,
# exact triangular solves
FP64_EXACT = dict(FP64, trsv=0)
EXACT_STEPS = 3


def timed_solves(L, S, db, dx, zeros, k, flush=None):
    lib = L.lib()
    out = []
    for _ in range(k):
        dx.upload(zeros)
        if flush:
            flush()
        L.check(lib.lace_device_synchronize())
        t0 = time.perf_counter()
        S.setup()
        st = S.solve_dev(db, dx)
        L.check(lib.lace_device_synchronize())
        out.append((time.perf_counter() - t0, st))
    return out


def host_cpus():
    """Logical CPUs this process may run on, and the host's topology from
    lscpu (model, sockets, physical cores per socket, threads per core)."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    topo = {"usable_logical_cpus": usable, "host_logical_cpus": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "cpu_model", "Socket(s)": "sockets",
                "Core(s) per socket": "cores_per_socket", "Thread(s) per core": "threads_per_core"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                topo[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, ValueError):
        pass
    if "cpu_model" not in topo:
        topo["cpu_model"] = _cpu_model()
    # physical cores among the usable CPUs: distinct (package, core) pairs
    try:
        cores = set()
        for c in os.sched_getaffinity(0):
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(base + "physical_package_id") as f1, open(base + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        topo["usable_physical_cores"] = len(cores)
    except (AttributeError, OSError):
        pass
    return topo


def cpu_baseline(row, col, val, b, threads):
    """The reference algorithm on the host cores: in-place OpenMP ParILU v0.3
    (3 sweeps) then Householder FGMRES(50) with exact triangular solves, to
    1e-8 relative (the oracle's restatement; the reference's MKL build
    cannot run here).  OpenMP threads are bound one per physical core where
    the usable CPUs allow it (OMP_PLACES=cores, OMP_PROC_BIND=close, set
    before the oracle's OpenMP runtime starts)."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PLACES"] = "cores"
    os.environ["OMP_PROC_BIND"] = "close"
    import oracle as O
    topo = host_cpus()
    A = O.CSR(len(row) - 1, len(row) - 1, row, col, val)
    t0 = time.perf_counter()
    st, sweeps, _, _ = O.parilu_v03_omp(A, 3)
    Lf, Uf = st.factors()
    t1 = time.perf_counter()
    r = O.fgmres_householder(A, b, m=50, restart_max=100, rtol=1e-8, tol_type=1, choice=0,
                             Lf=Lf, Uf=Uf, nthreads=threads)
    t2 = time.perf_counter()
    rel = float(np.linalg.norm(b - A.scipy() @ r.x) / np.linalg.norm(b))
    out = {"value": round(t2 - t0, 3), "unit": "s", "cores": threads, "kind": "port",
           "sample": (f"full solve on the same 128^3 problem: ParILU v0.3 in-place OpenMP "
                      f"({sweeps} sweeps, {t1 - t0:.2f} s) + Householder FGMRES(50) with exact "
                      f"substitution, {r.searches} iterations to rel. residual {rel:.2e}"),
           "omp": {"OMP_NUM_THREADS": threads, "OMP_PLACES": "cores", "OMP_PROC_BIND": "close"},
           **topo}
    # cores: the physical cores the threads run on (one thread per core
    # where the usable CPUs hold that many cores, else SMT siblings share)
    phys = topo.get("usable_physical_cores") or topo["usable_logical_cpus"]
    out["cores"] = min(threads, phys)
    out["threads"] = threads
    # the deterministic one-thread run (in-place sweeps in row order) on a
    # bounded sample: the three sweeps and one FGMRES(50) cycle, scaled to
    # the iteration count of the full solve above
    O.set_threads(1)
    try:
        t0 = time.perf_counter()
        st1, _, _, _ = O.parilu_v03_omp(A, 3)
        L1, U1 = st1.factors()
        t1 = time.perf_counter()
        r1 = O.fgmres_householder(A, b, m=50, restart_max=1, rtol=1e-8, tol_type=1, choice=0,
                                  Lf=L1, Uf=U1, nthreads=1)
        t2 = time.perf_counter()
    finally:
        O.set_threads(0)
    per_it = (t2 - t1) / max(1, r1.searches)
    out["one_thread"] = {
        "value": round((t1 - t0) + per_it * r.searches, 3), "unit": "s", "cores": 1,
        "sample": (f"3 in-place sweeps ({t1 - t0:.2f} s) + one FGMRES(50) cycle ({r1.searches} "
                   f"iterations, {t2 - t1:.2f} s) on one thread, the FGMRES time scaled to the "
                   f"{r.searches} iterations of the full solve")}
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


# kernel classes by the kernel each launch runs on the 128^3 operator (the
# FP32 streaming sweep; Jacobi sweeps and products on the diagonal form)
SWEEP_KERNEL = "parilu_stream4_sweep_kernel<float>"
JACOBI_KERNEL = "tri_jacobi_dia_vec_kernel<float>"
SPMV64_KERNEL = "spmv_dia_kernel<double, float, double>"
SPMV32_KERNEL = "spmv_dia_kernel<float, float, float>"
# kernel class -> per-launch timing regions of the library; each launch
# records its algorithmic bytes (every element of the storage form it reads,
# and its vectors, once)
KERNEL_CLASSES = {
    SWEEP_KERNEL: ["parilu_sweep"],
    JACOBI_KERNEL: ["trsv_jacobi_sweep_lower", "trsv_jacobi_sweep_upper"],
    # several sweeps per launch (option jacobi_pipe=1)
    "tri_jacobi_dia_pipe_kernel<float>": ["trsv_jacobi_pipe_lower", "trsv_jacobi_pipe_upper"],
    SPMV64_KERNEL: ["spmv"],
    SPMV32_KERNEL: ["spmv_f32a"],
}
HOT_REGIONS = [r for rs in KERNEL_CLASSES.values() for r in rs]


def pmc_traffic():
    """HBM bytes per launch of each kernel class from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy
    kernel of known size), or None."""
    p = os.path.join(ROOT, "profiles", "hbm_traffic_convdiff128.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        data = json.load(f)
    return {k: v.get("bytes_per_launch") for k, v in data.get("kernels", {}).items()}


FLUSH_BYTES = 1 << 30
# events around the first four and then every k-th launch of each hot
# region in the timed steps
TIMING_SAMPLE = 16
# bytes read (and written) by the reference copy: 2 x 120 MiB ~ one FP32 sweep
COPY_BYTES = 120 << 20


def make_flush(L):
    """A copy of 512 MiB into another 512 MiB buffer: it touches 1 GiB of
    HBM, four times the 256 MB Infinity Cache, so the kernel launched next
    finds none of its data in the Infinity Cache or L2."""
    lib = L.lib()
    half = FLUSH_BYTES // 2
    src = L.DevVec(half // 8, np.float64)
    dst = L.DevVec(half // 8, np.float64)

    def flush():
        L.check(lib.lace_stream_copy(dst.ptr, src.ptr, half))
        L.check(lib.lace_device_synchronize())
    flush.buffers = (src, dst)
    return flush


def cache_rates(L, H, n, flush, reps=10):
    """Per-launch duration and rate of each hot kernel class run on its own:
    warm (reps launches back to back, data left in the caches by the
    previous launch) and cold (each launch after a flush), the median over
    the reps calls (a call that meets a stall of the device's queue, as
    after the release of large allocations, does not move it).  Bytes per
    call are those the launches recorded for the storage form they read."""
    lib = L.lib()
    rng = np.random.default_rng(0)
    Ps = L.ParILU(H, L.LACE_F32)
    Pj = L.ParILU(H, L.LACE_F32)
    Pj.sweeps(3)
    r = L.DevVec(n, np.float32, rng.standard_normal(n).astype(np.float32))
    z = L.DevVec(n, np.float32)
    A64 = L.DevCSRMatrix(H, L.LACE_F64)
    A32 = L.DevCSRMatrix(H, L.LACE_F32)
    x64 = L.DevVec(n, np.float64, rng.standard_normal(n))
    y64 = L.DevVec(n, np.float64)
    x32 = L.DevVec(n, np.float32, rng.standard_normal(n).astype(np.float32))
    y32 = L.DevVec(n, np.float32)
    csrc = L.DevVec(COPY_BYTES // 8, np.float64)
    cdst = L.DevVec(COPY_BYTES // 8, np.float64)
    cases = [
        (SWEEP_KERNEL, ["parilu_sweep"], lambda: Ps.sweeps(1)),
        # Jacobi-2: one timed off-diagonal sweep each for L and U
        (JACOBI_KERNEL, ["trsv_jacobi_sweep_lower", "trsv_jacobi_sweep_upper"],
         lambda: L.check(lib.lace_parilu_apply(Pj.h, L.TRSV_JACOBI, 2, 0, r.ptr, z.ptr))),
        # FP64 A, x and y (the solve's outer SpMV reads an FP32 x)
        (SPMV64_KERNEL, ["spmv"], lambda: L.check(lib.lace_csr_spmv(A64.h, 1.0, x64.ptr, 0.0, y64.ptr))),
        (SPMV32_KERNEL, ["spmv_f32a"], lambda: L.check(lib.lace_csr_spmv(A32.h, 1.0, x32.ptr, 0.0, y32.ptr))),
        # a plain copy moving as many bytes as the FP32 sweep: what a
        # streaming kernel of this length reaches
        ("stream_copy_kernel", ["stream_copy"],
         lambda: L.check(lib.lace_stream_copy(cdst.ptr, csrc.ptr, COPY_BYTES))),
    ]
    out = {}
    for name, regions, fn in cases:
        fn()  # first call: one-time conversions outside the timings
        L.check(lib.lace_device_synchronize())
        per = {}
        for label in ("warm", "cold"):
            L.check(lib.lace_timing_only(",".join(regions).encode()))
            calls_us, nbytes = [], 0.0
            for _ in range(reps):
                if label == "cold":
                    flush()
                L.check(lib.lace_timing_reset())
                L.check(lib.lace_timing_enable(1))
                fn()
                L.check(lib.lace_timing_enable(0))
                # each call starts on an idle queue, so a launch's start
                # event never waits behind the previous call's kernels
                L.check(lib.lace_device_synchronize())
                # one launch per region per call: the call's kernel time
                # and bytes are the sums over its regions (an option can
                # route a call to other regions: those record none)
                timed = [rg for rg in regions if L.timing(rg)[1] > 0]
                if timed:
                    calls_us.append(sum(L.timing(rg)[0] for rg in timed) * 1e3)
                    nbytes = sum(L.timing_work(rg)[1] for rg in timed)
            L.check(lib.lace_timing_only(None))
            if not calls_us:
                per[f"{label}_us"] = None
                continue
            us = float(np.median(calls_us))
            per[f"{label}_us"] = round(us, 2)
            per[f"{label}_GBps"] = round(nbytes / (us * 1e-6) / 1e9, 1)
        per["bytes_per_call"] = round(nbytes)
        out[name] = per
    return out


# time-to-solution broken down by what the GPU does: the ParILU setup, the
# preconditioner applies (ILU triangular solves), the sparse products and the
# orthogonalisation (outer Householder, inner classical Gram-Schmidt)
BREAKDOWN = {
    "setup": ["setup"],
    "preconditioner": ["precond"],
    "spmv": ["spmv", "spmv_f32a", "spmv_residual"],
    "orthogonalization": ["hh_start", "householder", "hh_finish", "cgs_start", "cgs_step"],
}


def breakdown(L, S, db, dx, zeros):
    """One more solve with every launch of the category regions timed
    (events around each region on the library stream): milliseconds per
    category, the step's time with those events, and the rest (casts, the
    host's convergence checks, event overhead)."""
    lib = L.lib()
    regions = [r for rs in BREAKDOWN.values() for r in rs]
    L.check(lib.lace_timing_reset())
    L.check(lib.lace_timing_only(",".join(regions).encode()))
    L.check(lib.lace_timing_enable(1))
    res = timed_solves(L, S, db, dx, zeros, 1)
    L.check(lib.lace_timing_enable(0))
    L.check(lib.lace_timing_only(None))
    out = {f"{cat}_ms": round(sum(L.timing(r)[0] for r in rs), 3) for cat, rs in BREAKDOWN.items()}
    total = res[0][0] * 1e3
    out["total_ms"] = round(total, 3)
    out["other_ms"] = round(total - sum(v for k, v in out.items() if k != "total_ms"), 3)
    return out


def copy_rate(L, nbytes, reps=10):
    """GB/s of a plain streaming copy moving nbytes (reading and writing
    nbytes / 2 each), warm, the median of reps event-timed launches: the
    rate a streaming kernel of that length reaches on this device (short
    launches do not reach the long-copy rate, and working sets beyond the
    256 MB Infinity Cache run at the HBM rate)."""
    lib = L.lib()
    half = int(nbytes) // 2 // 16 * 16
    src = L.DevVec(half // 8 + 2, np.float64)
    dst = L.DevVec(half // 8 + 2, np.float64)
    L.check(lib.lace_stream_copy(dst.ptr, src.ptr, half))
    L.check(lib.lace_timing_only(b"stream_copy"))
    ts = []
    for _ in range(reps):
        L.check(lib.lace_timing_reset())
        L.check(lib.lace_timing_enable(1))
        L.check(lib.lace_stream_copy(dst.ptr, src.ptr, half))
        L.check(lib.lace_timing_enable(0))
        L.check(lib.lace_device_synchronize())
        ms, cnt = L.timing("stream_copy")
        ts.append(ms / cnt)
    L.check(lib.lace_timing_only(None))
    return round(2.0 * half / (float(np.median(ts)) * 1e-3) / 1e9, 1)


def _region_rate(L, name, nbytes, fn, reps):
    """Duration of the named region per call of fn (HIP events on the
    library stream), the median over reps calls (a call that meets a stall
    of the device queue does not move it), and the rate nbytes / duration."""
    lib = L.lib()
    L.check(lib.lace_timing_only(name.encode()))
    ts = []
    for _ in range(reps):
        L.check(lib.lace_timing_reset())
        L.check(lib.lace_timing_enable(1))
        fn()
        L.check(lib.lace_timing_enable(0))
        L.check(lib.lace_device_synchronize())
        ms, cnt = L.timing(name)
        if cnt:
            ts.append(ms / cnt * 1e-3)
    L.check(lib.lace_timing_only(None))
    avg = float(np.median(ts))
    return avg, nbytes / avg / 1e9


def bench_reorder(args) -> int:
    """Reverse Cuthill-McKee as a pre-step on the 3-D convection-diffusion
    problem, in the reference's ordering (data_sparse_reorder 2, every quirk
    of src/cuthill_mckee.cpp kept) and the textbook one (lace_reorder_rcm,
    data_sparse_reorder 4), against natural order: the host reordering time,
    bandwidth, level counts of the factors, the nonlinear ILU residual after
    each of 5 synchronous FP32 sweeps, the storage forms the products and
    sweeps get, and the mixed solve (the headline configuration)."""
    import ctypes as C
    import lace_amd as L
    lib = L.lib()
    L.check(lib.lace_init(0), "lace_init")
    N = args.size
    row, col, val = L.gen_convdiff3d(N, N, N)
    n = len(row) - 1
    b = L.gen_normal(0, n)
    Hn = L.HostCSR(row, col, val)
    out = {"metric": "RCM pre-step: bandwidth, level counts, ParILU convergence and mixed solve "
                     "vs natural order",
           "config": {"workload": f"3-D 7-point upwind convection-diffusion {N}^3", "n": n,
                      "nnz": int(row[-1])}, "orders": {}}
    ri = np.repeat(np.arange(n), np.diff(row))
    orders = [("natural", None), ("reference_rcm", 2), ("textbook_rcm", 4)]
    for name, reorder in orders:
        if reorder is None:
            H, rhs, P, t_ord = Hn, b, np.arange(n, dtype=np.int32), 0.0
        else:
            M = L.DMatrix()
            L.check(lib.data_zmconvert(Hn.m, C.byref(M), L.Magma_CSR, L.Magma_CSR))
            P = np.zeros(n, np.int32)
            Pinv = np.zeros(n, np.int32)
            ip = C.POINTER(C.c_int)
            t0 = time.perf_counter()
            L.check(lib.data_sparse_reorder(C.byref(M), P.ctypes.data_as(ip), Pinv.ctypes.data_as(ip), reorder),
                    "data_sparse_reorder")
            t_ord = time.perf_counter() - t0
            H = L.HostCSR(*L.owned_to_numpy(M))
            rhs = b[P]
        pos = np.empty_like(P)
        pos[P] = np.arange(n)
        F = L.ParILU(H, L.LACE_F32)
        lev = F.levels()
        nl = []
        for _ in range(5):
            F.sweeps(1)
            nl.append(F.nonlinres())
        S = L.Solver(H, make_opts(L, **MIXED))
        db = L.DevVec(n, np.float64, rhs)
        dx = L.DevVec(n, np.float64)
        ok = all(np.isfinite(nl))
        res = timed_solves(L, S, db, dx, np.zeros(n), args.warmup + args.steps) if ok else []
        res = res[args.warmup:]
        st = res[-1][1] if res else None
        A = L.DevCSRMatrix(H, L.LACE_F64)
        out["orders"][name] = {
            "reorder_seconds_host": round(t_ord, 3),
            "bandwidth": int(np.abs(pos[ri] - pos[col]).max()),
            "levels_L_U": list(lev),
            "forms_A_L_U": [list(A.forms()), list(F.forms(False)), list(F.forms(True))],
            "parilu_f32_nonlinres_sweeps_1_to_5": [float(f"{v:.4e}") for v in nl],
            "mixed_solve_ms": round(float(np.mean([t for t, _ in res])) * 1e3, 3) if res else None,
            "outer_iterations": st.iterations if st else None,
            "inner_iterations": st.inner_iterations if st else None,
            "rel_true_residual": st.rel_true_residual if st else None,
            "converged": st.converged if st else 0}
    print(json.dumps(out))
    good = out["orders"]["natural"]["converged"] == 1 and out["orders"]["textbook_rcm"]["converged"] == 1
    return 0 if good else 1


def bench_bcsr(args, compact=False):
    """Config 3: BCSR 5x5 blocks on the 64^3 7-point block grid
    (n = 1,310,720; 1,810,432 blocks = 45,260,800 entries); entries
    0.1 N(0,1) (seed 1), diagonal blocks + 5 I, b ~ N(0,1) (seed 2).
    FP64 FGMRES(40) to 1e-8 with block ParILU(0) factors (3 sweeps) stored
    in FP32 or BF16 (FP32 iteration, Jacobi-5 block solves); the error of
    each solution against an FP64 solve to 1e-13 is reported."""
    import lace_amd as L
    lib = L.lib()
    L.check(lib.lace_init(0), "lace_init")
    N = args.size if args.size != 128 else 64
    row, col, val = L.gen_bcsr7(N, N, N, 5, seed=1, sigma=0.1, shift=5.0)
    H = L.HostBCSR(row, col, val, 5)
    n = H.n
    b = L.gen_normal(2, n)
    db = L.DevVec(n, np.float64, b)
    dx = L.DevVec(n, np.float64)
    zeros = np.zeros(n)
    out = {"mode": "bcsr", "workload": f"config 3: BCSR 5x5, 7-point block grid {N}^3",
           "n": n, "blocks": int(row[-1])}
    # sweeps and SpMV bandwidth
    variants = ((0, ""), (2, "_merge"), (3, "_tables"), (4, "_tables_lds"), (5, "_tables_lds_lu"),
                (6, "_wave"))
    for name, prec in (("f32", L.LACE_F32), ("bf16", L.LACE_BF16), ("f64", L.LACE_F64)):
        # forms of the paired sweep: 0 the default, 2 the device-side merge of
        # the two block rows, 3 set-up tables, 4 tables + LDS staging, 5
        # tables + LDS staging of L, U and U_JJ only, 6 whole-word wave runs
        for variant, tag in (variants[:1] if compact else variants):
            P = L.BILU(H, prec)
            if variant:
                P.set_variant(variant)
            P.sweeps(2)
            avg, gbs = _region_rate(L, "bilu_sweep", P.sweep_bytes(), lambda: P.sweeps(1), 10)
            out[f"bilu_sweep_{name}{tag}"] = {"us": round(avg * 1e6, 2), "GBps": round(gbs, 1),
                                              "bytes": P.sweep_bytes()}
            del P
    A = L.DevBCSR(H, L.LACE_F64)
    x = L.DevVec(n, np.float64, b)
    y = L.DevVec(n, np.float64)
    avg, gbs = _region_rate(L, "bcsr_spmv", A.bytes(),
                            lambda: L.check(lib.lace_bcsr_spmv(A.h, x.ptr, y.ptr)), 20)
    out["bcsr_spmv_f64"] = {"us": round(avg * 1e6, 2), "GBps": round(gbs, 1), "bytes": A.bytes()}

    def solver(prec, rtol, m=40):
        o = make_opts(L, search_max=m, trsv=L.TRSV_JACOBI, jacobi_sweeps=5, prec_precond=prec)
        o.rtol = rtol
        return L.Solver(H, o)

    ref = solver(L.LACE_F64, 1e-13)
    timed_solves(L, ref, db, dx, zeros, 1)
    xref = dx.download()
    for name, prec in (("f32", L.LACE_F32), ("bf16", L.LACE_BF16), ("f64", L.LACE_F64)):
        S = solver(prec, 1e-8)
        timed_solves(L, S, db, dx, zeros, args.warmup)
        res = timed_solves(L, S, db, dx, zeros, args.steps)
        t = float(np.mean([r[0] for r in res]))
        st = res[-1][1]
        xs = dx.download()
        out[f"solve_{name}"] = {
            "s": round(t, 5), "iterations": st.iterations, "converged": st.converged,
            "rel_true_residual": st.rel_true_residual,
            "err_vs_fp64_solution": float(np.linalg.norm(xs - xref) / np.linalg.norm(xref))}
        del S
    sw = out["bilu_sweep_f32"]
    out["roofline"] = {"kernel": "block ParILU sweep (FP32)", "bound": "hbm", "achieved": sw["GBps"],
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(sw["GBps"] / HBM_PEAK_GBS, 4),
                       "bytes_per_launch": sw["bytes"], "avg_launch_us": sw["us"],
                       "copy_same_bytes_GBps": copy_rate(L, sw["bytes"])}
    out["all_converged"] = all(out[f"solve_{k}"]["converged"] == 1 and
                               out[f"solve_{k}"]["rel_true_residual"] <= 1e-8 for k in ("f32", "bf16", "f64"))
    return out


MFMA_F64_PEAK_TFS = 78.6
BLOCK_REGIONS = ["tsgemm_tn", "tsgemm_nn", "spmm32_f32x", "spmm32_f32a", "spmm32_residual",
                 "block_jacobi_sweep_lower", "block_jacobi_sweep_upper", "block_jacobi_pair_lower",
                 "block_jacobi_pair_upper", "inner_dots",
                 "inner_update", "block_inner"]


def bench_trsv(args):
    """Exact triangular solves (the level-scheduled ParILU apply, L then U)
    on the 128^3 convection-diffusion factors after 3 sweeps, FP32 and FP64:
    the persistent kernel with blocks taken box by box (default) and as runs
    of consecutive rows, and the level-ordered sync-free kernel, each timed
    by HIP events on the library stream; the level counts of L and U, the
    time per level, and whether all kernels give the same bits."""
    import ctypes as C
    import lace_amd as L
    lib = L.lib()
    L.check(lib.lace_init(0), "lace_init")
    N = args.size
    row, col, val = L.gen_convdiff3d(N, N, N)
    n = len(row) - 1
    H = L.HostCSR(row, col, val)
    rng = np.random.default_rng(0)
    out = {"mode": "trsv", "workload": f"exact L and U solves of the 3-sweep ParILU factors, "
           f"7-point convection-diffusion {N}^3", "n": n}
    reps = max(args.steps, 5)
    ok = True
    for name, prec, dt in (("f32", L.LACE_F32, np.float32), ("f64", L.LACE_F64, np.float64)):
        P = L.ParILU(H, prec)
        P.sweeps(3)
        nl, nu = P.levels()
        r = L.DevVec(n, dt, rng.standard_normal(n).astype(dt))
        z = L.DevVec(n, dt)
        res = {"levels_L": nl, "levels_U": nu}
        zs = {}
        res["tiles_L"], res["tiles_U"] = P.tiles(0), P.tiles(1)
        for mode, tiles, tag, regions in ((1, 1, "persistent", ["trsv_exact_lower", "trsv_exact_upper"]),
                                          (1, 0, "persistent_rows", ["trsv_exact_lower", "trsv_exact_upper"]),
                                          (0, 1, "syncfree", ["trsv_syncfree"])):
            L.check(lib.lace_set_option(b"trsv_exact", mode))
            L.check(lib.lace_set_option(b"trsv_exact_tiles", tiles))
            apply = lambda: L.check(lib.lace_parilu_apply(P.h, L.TRSV_LEVEL, 0, 0, r.ptr, z.ptr))
            apply()
            L.check(lib.lace_device_synchronize())
            L.check(lib.lace_timing_reset())
            L.check(lib.lace_timing_only(",".join(regions).encode()))
            L.check(lib.lace_timing_enable(1))
            t0 = time.perf_counter()
            for _ in range(reps):
                apply()
            L.check(lib.lace_device_synchronize())
            wall = (time.perf_counter() - t0) / reps
            L.check(lib.lace_timing_enable(0))
            L.check(lib.lace_timing_only(None))
            ms = sum(L.timing(rg)[0] for rg in regions if L.timing(rg)[1] > 0) / reps
            ev = C.c_int()
            L.check(lib.lace_syncfree_error(P.h, C.byref(ev)))
            zs[tag] = z.download()
            res[tag] = {"ms_kernels": round(ms, 3), "ms_call": round(wall * 1e3, 3),
                        "us_per_level": round(ms * 1e3 / (nl + nu), 3), "wait_error": ev.value}
            ok = ok and ev.value == 0 and bool(np.isfinite(zs[tag]).all())
        res["bit_identical"] = bool(np.array_equal(zs["persistent"], zs["syncfree"]) and
                                    np.array_equal(zs["persistent"], zs["persistent_rows"]))
        ok = ok and res["bit_identical"]
        out[name] = res
        del P
    L.check(lib.lace_set_option(b"trsv_exact", 1))
    L.check(lib.lace_set_option(b"trsv_exact_tiles", 1))
    out["all_ok"] = ok
    return out


def bench_block(args):
    """Config 5: 32 N(0,1) right-hand sides (Philox seeds 0-31) on the 128^3
    convection-diffusion matrix.  One step = FP32 ParILU setup (3 sweeps) +
    block Householder FGMRES(30) for the 32 right-hand sides to 1e-8 per
    column (FP32 factors, Jacobi-5, FP64 reflectors with the MFMA Gram and
    panel products) + one adjoint block solve A^T Lambda = X - X_t for the
    32 quadratic losses 0.5 ||x_c - x_t,c||^2 (x_t,c ~ N(0,1), seeds
    1000-1031) + the gradient gather of their sum."""
    import lace_amd as L
    lib = L.lib()
    L.check(lib.lace_init(0), "lace_init")
    N = args.size
    row, col, val = L.gen_convdiff3d(N, N, N)
    n, nnz = len(row) - 1, int(row[-1])
    H = L.HostCSR(row, col, val)
    B = np.stack([L.gen_normal(c, n) for c in range(32)], axis=1)
    Xt = np.stack([L.gen_normal(1000 + c, n) for c in range(32)], axis=1)
    o = make_opts(L, search_max=30, trsv=L.TRSV_JACOBI, jacobi_sweeps=5, prec_precond=L.LACE_F32,
                  inner_iters=args.block_inner)
    o.restart_max = 20
    S = L.Solver(H, o)
    m = n * 32
    dB = L.DevVec(m, np.float64, B.reshape(-1))
    dXt = L.DevVec(m, np.float64, Xt.reshape(-1))
    dX = L.DevVec(m, np.float64)
    dG = L.DevVec(m, np.float64)
    dLam = L.DevVec(m, np.float64)
    dth = L.DevVec(nnz, np.float64)
    sync = lambda: L.check(lib.lace_device_synchronize())

    def step():
        L.check(lib.lace_daxpby(m, 0.0, None, 0.0, None, dX.ptr))
        L.check(lib.lace_daxpby(m, 0.0, None, 0.0, None, dLam.ptr))
        sync()
        t0 = time.perf_counter()
        S.setup()
        sync()
        t1 = time.perf_counter()
        stf = S.solve_block_dev(dB, dX, 0)
        t2 = time.perf_counter()
        L.check(lib.lace_daxpby(m, 1.0, dX.ptr, -1.0, dXt.ptr, dG.ptr))
        sta = S.solve_block_dev(dG, dLam, 1)
        L.check(lib.lace_solver_grad_values_block(S.h, dLam.ptr, dX.ptr, dth.ptr))
        sync()
        t3 = time.perf_counter()
        return t3 - t0, t1 - t0, t2 - t1, t3 - t2, stf, sta

    for _ in range(args.warmup):
        step()
    L.check(lib.lace_timing_reset())
    L.check(lib.lace_timing_only(",".join(BLOCK_REGIONS).encode()))
    L.check(lib.lace_timing_enable(1))
    res = [step() for _ in range(args.steps)]
    L.check(lib.lace_timing_enable(0))
    L.check(lib.lace_timing_only(None))
    kernels = {}
    for r in BLOCK_REGIONS:
        ms, cnt = L.timing(r)
        if cnt == 0:
            continue
        fl, by = L.timing_work(r)
        sec = ms * 1e-3
        kernels[r] = {"launches": cnt, "total_ms": round(ms, 3), "avg_us": round(ms / cnt * 1e3, 2)}
        if by > 0:
            kernels[r].update({"GBps": round(by / sec / 1e9, 1),
                               "hbm_frac": round(by / sec / 1e9 / HBM_PEAK_GBS, 4),
                               "TFLOPs": round(fl / sec / 1e12, 2)})
        if r.startswith("tsgemm"):
            kernels[r]["mfma_frac"] = round(fl / sec / 1e12 / MFMA_F64_PEAK_TFS, 4)
    t = float(np.mean([r[0] for r in res]))
    stf, sta = res[-1][4], res[-1][5]
    ok = all(r[4].converged == 1 and r[5].converged == 1 and r[4].max_rel_true_residual <= 1e-8
             and r[5].max_rel_true_residual <= 1e-8 for r in res)
    # FP64 oracle error of two columns: FP64 single-vector solves to 1e-12
    X = dX.download().reshape(n, 32)
    errs = {}
    o64 = make_opts(L, search_max=50, trsv=L.TRSV_JACOBI, jacobi_sweeps=5, prec_precond=L.LACE_F64)
    o64.rtol = 1e-12
    ref = L.Solver(H, o64)
    ref.setup()
    for c in (0, 31):
        xr, st = ref.solve(B[:, c])
        errs[str(c)] = float(np.linalg.norm(X[:, c] - xr) / np.linalg.norm(xr))
    tn = kernels.get("tsgemm_tn", {})
    out = {
        "mode": "block",
        "metric": "32-RHS forward + adjoint block ILU-FGMRES(30) to 1e-8, s per step",
        "value": round(t, 4), "unit": "s", "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": False, "dtype": "f32 preconditioner / f64 block FGMRES",
        "config": {"workload": f"config 5: 32 right-hand sides, convection-diffusion {N}^3",
                   "n": n, "nnz": nnz, "nrhs": 32, "restart": 30, "parilu_sweeps": 3,
                   "trsv": "Jacobi-5 (FP32 factors)", "rhs_seeds": "0-31", "x_t_seeds": "1000-1031",
                   "preconditioner": (f"{args.block_inner} batched FP32 GMRES steps per column, "
                                      "FP32 ILU / Jacobi-5" if args.block_inner
                                      else "FP32 ILU / Jacobi-5")},
        "solves_per_s": round(64 / t, 2),
        "setup_s": round(float(np.mean([r[1] for r in res])), 4),
        "forward_s": round(float(np.mean([r[2] for r in res])), 4),
        "adjoint_and_gradient_s": round(float(np.mean([r[3] for r in res])), 4),
        "forward": {"block_steps": stf.iterations, "restarts": stf.restarts,
                    "panel_passes": stf.panel_passes, "shifted_panels": stf.shifted_panels,
                    "max_rel_true_residual": stf.max_rel_true_residual},
        "adjoint": {"block_steps": sta.iterations, "restarts": sta.restarts,
                    "panel_passes": sta.panel_passes, "shifted_panels": sta.shifted_panels,
                    "max_rel_true_residual": sta.max_rel_true_residual},
        "err_vs_fp64_solution": errs,
        "all_converged": ok,
        "roofline": {"kernel": "tsgemm_tn (MFMA FP64 Gram)", "bound": "hbm",
                     "achieved": tn.get("GBps"), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": tn.get("hbm_frac"), "traffic": None,
                     "mfma_TFLOPs": tn.get("TFLOPs"), "mfma_peak_TFLOPs": MFMA_F64_PEAK_TFS,
                     "mfma_frac": tn.get("mfma_frac")},
        "kernels": kernels,
    }
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="solve mode: skip the compact config-3 (BCSR) and config-5 (32 RHS) summaries")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--block-inner", type=int, default=10,
                    help="block mode: inner FP32 GMRES steps per outer block step (0: ILU only)")
    ap.add_argument("--mode", choices=["solve", "kernels", "bcsr", "block", "reorder", "trsv"], default="solve",
                    help="kernels: one mixed-path step plus one 256 MiB copy kernel "
                         "(the workload profiled for HBM traffic counters); bcsr: the "
                         "block-sparse 5x5 workload on the 64^3 block grid; block: 32 "
                         "right-hand sides, forward and adjoint block FGMRES; reorder: the "
                         "RCM pre-step against natural order; trsv: the exact L and U "
                         "solves, persistent and level-ordered kernels")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="library kernel-shape option (lace_set_option) applied before the run; "
                         "repeatable; every choice but inner_ortho gives bit-identical results")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this implementation runs on one "
              "GPU (the path would shard by z-slabs of the grid with one plane halo per sweep or "
              "product; that is not built)",
              file=sys.stderr)
        return 2
    if args.set:
        import lace_amd as L
        for kv in args.set:
            name, _, value = kv.partition("=")
            L.check(L.lib().lace_set_option(name.encode(), int(value)), f"option {name}")
    if args.mode == "bcsr":
        out = bench_bcsr(args)
        print(json.dumps(out))
        return 0 if out["all_converged"] else 1
    if args.mode == "block":
        out = bench_block(args)
        print(json.dumps(out))
        return 0 if out["all_converged"] else 1
    if args.mode == "reorder":
        return bench_reorder(args)
    if args.mode == "trsv":
        out = bench_trsv(args)
        print(json.dumps(out))
        return 0 if out["all_ok"] else 1

    import lace_amd as L
    lib = L.lib()
    L.check(lib.lace_init(0), "lace_init")
    N = args.size
    row, col, val = L.gen_convdiff3d(N, N, N)
    n, nnz = len(row) - 1, int(row[-1])
    b = L.gen_normal(0, n)
    H = L.HostCSR(row, col, val)
    if args.mode == "kernels":
        # one mixed-path step (every hot kernel as in the timed step) plus
        # one 256 MiB copy, the traffic-counter calibration
        S = L.Solver(H, make_opts(L, **MIXED))
        dbk = L.DevVec(n, np.float64, b)
        dxk = L.DevVec(n, np.float64)
        res = timed_solves(L, S, dbk, dxk, np.zeros(n), 1)
        src = L.DevVec(1 << 25, np.float64)
        dst = L.DevVec(1 << 25, np.float64)
        L.check(lib.lace_stream_copy(dst.ptr, src.ptr, 1 << 28))
        L.check(lib.lace_device_synchronize())
        print(json.dumps({"mode": "kernels", "step_s": res[0][0],
                          "iterations": res[0][1].iterations, "copy_bytes": 1 << 28}))
        return 0
    db = L.DevVec(n, np.float64, b)
    dx = L.DevVec(n, np.float64)
    zeros = np.zeros(n)

    mixed = L.Solver(H, make_opts(L, **MIXED))
    fp64 = L.Solver(H, make_opts(L, **FP64))
    fp64n = L.Solver(H, make_opts(L, **FP64_NESTED))
    mixed16 = L.Solver(H, make_opts(L, **MIXED_BF16))
    for S in (mixed, fp64, fp64n, mixed16):
        timed_solves(L, S, db, dx, zeros, args.warmup)

    # timed region: mixed steps with the hot kernels timed by HIP events on
    # the library stream (only these regions carry events), every
    # TIMING_SAMPLE-th launch of each region
    L.check(lib.lace_timing_reset())
    L.check(lib.lace_set_option(b"timing_sample", TIMING_SAMPLE))
    L.check(lib.lace_timing_only(",".join(HOT_REGIONS).encode()))
    L.check(lib.lace_timing_enable(1))
    res_mixed = timed_solves(L, mixed, db, dx, zeros, args.steps)
    L.check(lib.lace_timing_enable(0))
    L.check(lib.lace_set_option(b"timing_sample", 1))
    live = {r: L.timing(r) for r in HOT_REGIONS}
    calls = {r: L.timing_calls(r) for r in HOT_REGIONS}
    work = {r: L.timing_work(r) for r in HOT_REGIONS}
    res_fp64 = timed_solves(L, fp64, db, dx, zeros, args.steps)
    x_fp64 = dx.download()
    # one more FP64 FGMRES(50) solve with the Householder multi-dot U_j^T w
    # (j + 1 columns of 16.8 MB and w, every Arnoldi step) timed by events:
    # its rate over all the solve's column counts, up to 50
    L.check(lib.lace_timing_reset())
    L.check(lib.lace_timing_only(b"hh_mdot"))
    L.check(lib.lace_timing_enable(1))
    timed_solves(L, fp64, db, dx, zeros, 1)
    L.check(lib.lace_timing_enable(0))
    L.check(lib.lace_timing_only(None))
    md_ms, md_n = L.timing("hh_mdot")
    md_bytes = L.timing_work("hh_mdot")[1]
    hh_mdot = {"launches": md_n, "avg_us": round(md_ms * 1e3 / max(md_n, 1), 2),
               "GBps": round(md_bytes / max(md_ms * 1e-3, 1e-12) / 1e9, 1),
               "avg_columns": round(md_bytes / max(md_n, 1) / (8.0 * n) - 1.0, 1)}
    split = {"mixed": breakdown(L, mixed, db, dx, zeros), "fp64": breakdown(L, fp64, db, dx, zeros)}
    res_fp64n = timed_solves(L, fp64n, db, dx, zeros, args.steps)
    x_fp64n = dx.download()
    res_m16 = timed_solves(L, mixed16, db, dx, zeros, args.steps)
    x_m16 = dx.download()
    fp64x = L.Solver(H, make_opts(L, **FP64_EXACT))
    timed_solves(L, fp64x, db, dx, zeros, 1)
    res_fp64x = timed_solves(L, fp64x, db, dx, zeros, EXACT_STEPS)
    x_fp64x = dx.download()
    del fp64x
    # the same mixed steps without events (the events' cost), and each
    # started with nothing of its data in the caches
    res_plain = timed_solves(L, mixed, db, dx, zeros, args.steps)
    flush = make_flush(L)
    res_cold = timed_solves(L, mixed, db, dx, zeros, args.steps, flush=flush)
    t_cold = float(np.mean([t for t, _ in res_cold]))
    t_plain = float(np.mean([t for t, _ in res_plain]))

    x_mixed = dx.download()
    t_mixed = float(np.mean([t for t, _ in res_mixed]))
    t_fp64 = float(np.mean([t for t, _ in res_fp64]))
    t_fp64n = float(np.mean([t for t, _ in res_fp64n]))
    t_m16 = float(np.mean([t for t, _ in res_m16]))
    t_fp64x = float(np.mean([t for t, _ in res_fp64x]))
    st_m, st_f = res_mixed[-1][1], res_fp64[-1][1]
    st_fn, st_m16, st_fx = res_fp64n[-1][1], res_m16[-1][1], res_fp64x[-1][1]
    ok = all(s.converged == 1 and s.rel_true_residual <= 1e-8
             for _, s in res_mixed + res_fp64 + res_plain + res_cold + res_fp64n + res_m16 + res_fp64x)
    # FP64 oracle error of each solution: against an FP64 solve to 1e-12
    ref_o = make_opts(L, **FP64)
    ref_o.rtol = 1e-12
    ref = L.Solver(H, ref_o)
    timed_solves(L, ref, db, dx, zeros, 1)
    x_ref = dx.download()
    del ref
    err = lambda x: float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref))

    # per-kernel rates from the live timings, algorithmic bytes per launch
    traffic = pmc_traffic()
    kernels = {}
    for name, regions in KERNEL_CLASSES.items():
        ms = sum(live[r][0] for r in regions)
        cnt = sum(live[r][1] for r in regions)
        # algorithmic bytes of the timed launches, as each launch recorded
        # them for the storage form it read
        by = sum(work[r][1] for r in regions)
        if cnt == 0:
            continue
        avg_s = ms * 1e-3 / cnt
        ach = by / (ms * 1e-3) / 1e9
        # all launches of the class in the timed steps, at the sampled
        # launches' average duration per region
        est_ms = sum(live[r][0] / live[r][1] * calls[r] for r in regions if live[r][1])
        kernels[name] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                         "traffic": (traffic or {}).get(name),
                         "bytes_per_launch": round(by / cnt), "avg_launch_us": round(avg_s * 1e6, 2),
                         "launches_timed": cnt, "launches": sum(calls[r] for r in regions),
                         "total_ms_in_timed_steps": round(est_ms, 3)}
    dominant = (max(kernels, key=lambda k: kernels[k]["total_ms_in_timed_steps"])
                if kernels else None)
    sw = kernels.get(SWEEP_KERNEL, {})
    achieved = sw.get("achieved", float("nan"))
    P64 = L.ParILU(H, L.LACE_F64)
    L.check(lib.lace_timing_reset())
    L.check(lib.lace_timing_enable(1))
    P64.sweeps(20)
    L.check(lib.lace_device_synchronize())
    L.check(lib.lace_timing_enable(0))
    s64_ms, s64_cnt = L.timing("parilu_sweep")
    ach64 = P64.sweep_bytes() / (s64_ms / s64_cnt * 1e-3) / 1e9
    L.check(lib.lace_timing_only(None))
    copy64 = copy_rate(L, P64.sweep_bytes())
    del P64
    cache = cache_rates(L, H, n, flush)

    # configs 3 and 5 in compact form: solve times, iterations, residuals
    # and the roofline of each one's dominant kernel
    del mixed, fp64, fp64n, mixed16
    extra = {}
    if not args.no_configs:
        c3 = bench_bcsr(args, compact=True)
        extra["config3"] = {k: c3[k] for k in ("workload", "n", "blocks", "solve_f32", "solve_bf16", "solve_f64",
                                                "bilu_sweep_f32", "bilu_sweep_bf16", "bilu_sweep_f64",
                                                "bcsr_spmv_f64", "roofline", "all_converged")}
        c5 = bench_block(args)
        extra["config5"] = {k: c5[k] for k in ("metric", "value", "unit", "config", "solves_per_s", "setup_s",
                                                "forward_s", "adjoint_and_gradient_s", "forward", "adjoint",
                                                "err_vs_fp64_solution", "roofline", "all_converged")}
        ok = ok and extra["config3"]["all_converged"] and extra["config5"]["all_converged"]

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, host_cpus()["usable_logical_cpus"])
        cpu = cpu_baseline(row, col, val, b, threads)

    out = {
        "metric": METRIC,
        "value": round(t_mixed, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 preconditioner / f64 outer FGMRES",
        "data": "synthetic (generated 3-D convection-diffusion operator, N(0,1) right-hand side)",
        "config": {
            "workload": f"config 2: 3-D 7-point upwind convection-diffusion {N}^3, FGMRES to 1e-8",
            "n": n, "nnz": nnz, "parilu_sweeps": 3,
            "mixed": "outer FP64 Householder FGMRES(10) x inner FP32 GMRES(10), FP32 ParILU, Jacobi-5",
            "fp64": "FP64 Householder FGMRES(50), FP64 ParILU, Jacobi-5",
            "fp64_exact": "FP64 Householder FGMRES(50), FP64 ParILU, exact triangular solves",
        },
        "fp64_time_s": round(t_fp64, 5),
        "speedup_mixed_vs_fp64": round(t_fp64 / t_mixed, 3),
        # the same nested structure in FP64 (outer FGMRES(10) x inner FP64
        # GMRES(10), FP64 factors): the share of the speed-up the precision
        # accounts for
        "fp64_nested_time_s": round(t_fp64n, 5),
        "speedup_mixed_vs_fp64_nested": round(t_fp64n / t_mixed, 3),
        "fp64_nested_outer_iterations": st_fn.iterations,
        "fp64_nested_inner_iterations": st_fn.inner_iterations,
        # the mixed path with BF16 factor storage (per-row scales)
        "mixed_bf16_time_s": round(t_m16, 5),
        "mixed_bf16_outer_iterations": st_m16.iterations,
        "mixed_bf16_rel_true_residual": st_m16.rel_true_residual,
        # the reference's method (FP64 FGMRES(50), FP64 factors, exact
        # triangular solves) on the GPU, EXACT_STEPS solves
        "fp64_exact_time_s": round(t_fp64x, 5),
        "fp64_exact_iterations": st_fx.iterations,
        "speedup_mixed_vs_fp64_exact": round(t_fp64x / t_mixed, 3),
        "err_vs_fp64_solution": {"mixed": err(x_mixed), "mixed_bf16": err(x_m16), "fp64": err(x_fp64),
                                 "fp64_nested": err(x_fp64n), "fp64_exact": err(x_fp64x),
                                 "reference": "FP64 FGMRES(50) to 1e-12"},
        "mixed_outer_iterations": st_m.iterations,
        "mixed_inner_iterations": st_m.inner_iterations,
        "fp64_iterations": st_f.iterations,
        "mixed_rel_true_residual": st_m.rel_true_residual,
        "fp64_rel_true_residual": st_f.rel_true_residual,
        "all_converged": ok,
        "parilu_sweep_GBps": {"f32": round(achieved, 1), "f64": round(ach64, 1),
                              "f64_copy_same_bytes": copy64},
        "fp64_hh_mdot": hh_mdot,
        # one more solve of each path with its category regions timed
        "breakdown": split,
        # the kernel with the most time in the timed mixed steps
        "roofline": (dict(kernel=dominant, **kernels[dominant],
                          copy_same_bytes_GBps=copy_rate(L, kernels[dominant]["bytes_per_launch"]))
                     if dominant else None),
        "kernels": kernels,
        # value and ms_per_step: steps run back to back (each finds what the
        # previous one left in the 256 MB Infinity Cache); cold: each step
        # and each isolated launch preceded by a copy touching 1 GiB
        "cache": {"flush_bytes": FLUSH_BYTES, "ms_per_step_warm": round(t_mixed * 1e3, 3),
                  "ms_per_step_cold": round(t_cold * 1e3, 3), "isolated_launches": cache},
        # the timed steps carry events around every TIMING_SAMPLE-th hot
        # launch; the same steps without any events:
        "timing_sample": TIMING_SAMPLE,
        "options": dict(kv.partition("=")[::2] for kv in args.set),
        "ms_per_step_without_events": round(t_plain * 1e3, 3),
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
