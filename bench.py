"""Benchmark of the tile Cholesky path on one MI355X.

Workload (the configuration the headline metric is quoted on): FP64
dplasma_dpotrf, uplo = Lower, N = 32768, NB = 512, A = dplgsy(bump = N,
seed 3872) generated in HBM, tiles resident (no host staging inside the timed
region).  One step = one factorization of a freshly generated matrix; the
generation happens between steps, outside the timed region.  Each step is
bracketed by device synchronization; value = LAWN-41 flops / step time.

Also reported: the FP64 solve of the same system (dpotrs, NRHS = 1, b =
dplrnt(seed 3873)) with the reference's residual ratios computed on the GPU,
the mixed-precision dposv at N = 32768, NRHS = 16 on A = Q diag(lambda) Q^T
for kappa in {1e2, 1e5, 1e8} against the FP64 GPU dpotrf + dpotrs of the same
system (time, refinement steps, backward error, error vs the FP64 answer),
the adjoint of the factorization (df/dA from a given df/dL) at the same
size, the batched uneven Cholesky + solve (4096 matrices of orders 32..1024,
FP64 and FP32, one hipGraph-replayed persistent launch against a launch per
matrix), the differentiated Gaussian-process likelihood (N = 8192, FP64 and
FP32, plus 64 GPs of 1024 points), the roofline of the dominant kernel (the
trailing update, timed with HIP events on its own stream in a separate
profiled step) and the CPU baseline:
the oracle's tile-DAG executor (a restatement of src/zpotrf_L.jdf with
OpenBLAS tile kernels) on the host cores, on a bounded sample.  When that
sample is the full N = 32768 workload, its factor is compared element by
element with the GPU factor of the same matrix (vs_cpu_oracle).  The
reference's own CPU-runnable case (N = 4096, NB = 256) is timed the same two
ways, with its two factors compared (dpotrf_N4096_NB256).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TF = 78.6          # MI355X dense FP64 matrix peak, TFLOP/s
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, GB/s
UPDATE_KERNEL = "dpotrf_update"   # trailing SYRK/GEMM update, the dominant kernel
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "dpotrf_update_N32768_NB512_traffic.json")


def flops_dpotrf(n):
    n = float(n)
    return n * (((1. / 6.) * n + 0.5) * n + (1. / 3.)) + n * (((1. / 6.) * n) * n - (1. / 6.))


def host_cpu():
    """(threads to use, threads this process may run on, nproc, CPU model).
    One worker per host core the process may use, at most 16: the host-core
    share of one GPU on the MI355X boxes this runs on."""
    try:
        avail = len(os.sched_getaffinity(0))
    except Exception:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, min(avail, 16)), avail, os.cpu_count() or 1, model


def cpu_baseline(seconds_hint=True):
    """Oracle tile-DAG dpotrf on the host cores (test infrastructure, timed
    only as the reported CPU baseline).  The host tile size is tuned first
    (NB in 256/384/512 on an N=16384 sample, the fastest kept), then the full
    N=32768 workload runs with it.  Returns the baseline line and the
    oracle's factor at NB=512 (packed lower tiles, for the comparison with
    the GPU factor; a separate run when the host's best NB differs)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    blas = po.load_openblas()
    cores, avail, nproc, model = host_cpu()

    def run(N, NB=512):
        T = po.dplgsy_symtiles(N, NB, float(N), 3872)
        t0 = time.time()
        info = po.dpotrf_tiles(T, N, NB, nthreads=cores, use_blas=blas is not None)
        return time.time() - t0, info, T

    probe = {}
    for nb in (256, 384, 512):
        probe[nb] = run(16384, nb)[0]
    nb_host = min(probe, key=probe.get)
    N = 32768
    if probe[nb_host] > 6.0:                   # a slow host: keep the 1/8-flop sample
        N = 16384
        dt, info, T = probe[nb_host], 0, None
    else:
        dt, info, T = run(N, nb_host)
    T512 = T if nb_host == 512 else (run(N, 512)[2] if N == 32768 else None)
    what = "the full N=32768 workload" if N == 32768 else "1/8 of the N=32768 flops"
    return (T512 if N == 32768 and info == 0 else None), {
        "value": round(flops_dpotrf(N) / dt / 1e12, 4),
        "unit": "TFLOP/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"dpotrf Lower N={N} NB={nb_host} ({what}), dplgsy seed 3872, "
                   f"tile-DAG of zpotrf_L.jdf with {os.path.basename(blas) if blas else 'plain C'} "
                   f"tile kernels, {dt:.2f} s, info={info}; {cores} workers of {avail} usable threads "
                   f"(nproc {nproc}), CPU {model}; host NB probe at N=16384: "
                   + ", ".join(f"NB={k} {v:.2f} s" for k, v in probe.items())),
        "cpu_model": model,
        "nproc": nproc,
        "host_nb": nb_host,
    }


def small_case_leg(dp, N=4096, NB=256, runs=5):
    """The reference's own CPU-runnable dpotrf case (Lower, N=4096, NB=256,
    dplgsy seed 3872), timed on the GPU (event-timed, resident tiles) and
    with the oracle's tile DAG on the host cores, and the two factors
    compared element by element."""
    import ctypes
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    blas = po.load_openblas()
    cores = host_cpu()[0]
    M, _ = dp.libs()
    MT = (N + NB - 1) // NB
    nt = MT * (MT + 1) // 2 * NB * NB
    gpu_ms = []
    with dp.Device(0) as dev:
        dA = dev.alloc(nt * 8)
        info = ctypes.c_int()
        ms = ctypes.c_double()
        for _ in range(runs + 1):                  # the first run warms up
            dev.check(M.mi_dplgsy_tiles(dev.h, float(N), N, NB, dA, 3872), "dplgsy")
            dev.check(M.mi_dpotrf_tiles(dev.h, N, NB, dA, ctypes.byref(info)), "dpotrf")
            M.mi_last_span_ms(dev.h, ctypes.byref(ms))
            gpu_ms.append(ms.value)
        G = dev.download(dA, nt)
        dev.free(dA)
    cpu_s = []
    for _ in range(3):
        T = po.dplgsy_symtiles(N, NB, float(N), 3872)
        t0 = time.time()
        cinfo = po.dpotrf_tiles(T, N, NB, nthreads=cores, use_blas=blas is not None)
        cpu_s.append(time.time() - t0)
    g_ms = float(np.median(gpu_ms[1:]))
    c_ms = float(np.median(cpu_s)) * 1e3
    cmp = compare_lower_tiles(G, T, N, NB)
    cmp["what"] = f"GPU dpotrf factor vs the CPU oracle's factor of the same matrix (N={N}, NB={NB})"
    return {"N": N, "NB": NB, "gpu_ms": round(g_ms, 3), "gpu_tflops": round(flops_dpotrf(N) / g_ms / 1e9, 3),
            "gpu_info": info.value, "cpu_ms": round(c_ms, 2), "cpu_tflops": round(flops_dpotrf(N) / c_ms / 1e9, 4),
            "cpu_cores": cores, "cpu_info": cinfo,
            "cpu_kernels": os.path.basename(blas) if blas else "plain C", "factor_vs_cpu_oracle": cmp}


def compare_lower_tiles(G, C, N, NB):
    """Largest difference between two packed lower-tile factors over the
    referenced elements (off-diagonal tiles whole, diagonal tiles on and
    below the diagonal; N a multiple of NB)."""
    import numpy as np
    MT = N // NB
    g = G[:MT * (MT + 1) // 2 * NB * NB].reshape(-1, NB, NB)
    c = C[:MT * (MT + 1) // 2 * NB * NB].reshape(-1, NB, NB)
    diag = np.array([n * MT - n * (n - 1) // 2 for n in range(MT)])
    off = np.setdiff1d(np.arange(g.shape[0]), diag)
    d_off = max((np.abs(g[i] - c[i]).max() for i in np.array_split(off, 64) if len(i)), default=0.0)
    low = np.triu(np.ones((NB, NB), dtype=bool))      # [col][row]: row >= col
    d_diag = max(np.abs(g[i] - c[i])[low].max() for i in diag)
    scale = max(np.abs(c[off[:64]]).max(), max(np.abs(c[i])[low].max() for i in diag))
    d = max(d_off, d_diag)
    return {"max_abs_diff": float(d), "max_abs_L": float(scale), "rel": float(d / scale),
            "what": "GPU dpotrf factor vs the CPU oracle's factor of the same matrix (N=32768, NB=512)"}


def isolated_update(dp, N, NB):
    """The trailing update of step 0 over the whole trailing matrix (every
    tile below the first panel: one launch of the dominant kernel with the
    GPU to itself), event-timed: the kernel's rate without the panel
    stream beside it."""
    import ctypes
    M, _ = dp.libs()
    MT = (N + NB - 1) // NB
    nt = MT * (MT + 1) // 2 * NB * NB
    with dp.Device(0) as dev:
        dA = dev.alloc(nt * 8)
        dev.check(M.mi_dplgsy_tiles(dev.h, float(N), N, NB, dA, 3872), "dplgsy")
        info = ctypes.c_int()
        dev.check(M.mi_dpotrf_step_potrf(dev.h, N, NB, dA, 0, ctypes.byref(info)), "potrf")
        dev.check(M.mi_dpotrf_step_trsm(dev.h, N, NB, dA, 0), "trsm")
        out = []
        for _ in range(3):
            M.mi_prof_reset(dev.h)
            M.mi_prof_enable(dev.h, 1)
            dev.check(M.mi_dpotrf_step_update(dev.h, N, NB, dA, 0, 1, MT), "update")
            M.mi_prof_enable(dev.h, 0)
            n, ms, fl, by = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            M.mi_prof_get(dev.h, UPDATE_KERNEL.encode(), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(fl),
                          ctypes.byref(by))
            out.append((ms.value, fl.value))
        dev.free(dA)
    ms, fl = sorted(out)[1]
    return {"what": f"step-0 trailing update of N={N}, NB={NB} alone ({MT - 1} tile columns, one launch)",
            "ms": round(ms, 3), "tflops": round(fl / ms / 1e9, 3), "frac": round(fl / ms / 1e9 / FP64_PEAK_TF, 4)}


def staged_leg(dp, N, NB, runs=2):
    """dplasma_dpotrf on a descriptor that is not resident: host tiles
    (page-locked) -> HBM -> factor -> host, as every timed GPU run of the
    reference tester does (tests/testing_zpotrf.c:61 drops device copies
    before each run).  Host-clock wall time and its parts."""
    ctx = dp.Context(0)
    A = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    best = None
    for _ in range(runs):
        assert ctx.dplgsy(float(N), dp.LOWER, A, 3872) == 0     # into host memory
        t0 = time.perf_counter()
        assert ctx.dpotrf(dp.LOWER, A) == 0
        wall = (time.perf_counter() - t0) * 1e3
        t = ctx.last_timing()
        if best is None or wall < best[0]:
            best = (wall, t)
    wall, t = best
    nbytes = A.desc.super.nb_local_tiles * NB * NB * 8
    A.free()
    ctx.close()
    return {"what": "host tiles -> HBM -> dpotrf -> host (page-locked host memory), host clock",
            "wall_ms": round(wall, 3), "stage_in_ms": round(t["stage_in_ms"], 3),
            "factor_ms": round(t["total_ms"], 3), "stage_out_ms": round(t["stage_out_ms"], 3),
            "tflops_incl_staging": round(flops_dpotrf(N) / wall / 1e9, 3),
            "bytes_each_way": nbytes,
            "h2d_GBps": round(nbytes / t["stage_in_ms"] / 1e6, 2) if t["stage_in_ms"] > 0 else None,
            "d2h_GBps": round(nbytes / t["stage_out_ms"] / 1e6, 2) if t["stage_out_ms"] > 0 else None}


def mixed_leg(dp, N, NB, nrhs=16, kappas=(1e2, 1e5, 1e8)):
    """Mixed-precision dposv at the configuration's shape: A = Q diag(lambda)
    Q^T (Q Haar from the Householder QR of a Gaussian matrix, seed 7; lambda
    geometric from 1 to 1/kappa), B = dplrnt(N x 16, seed 3873).  For each
    kappa: the FP64 GPU path (dpotrf + dpotrs) and the mixed solve with the
    FP32 factor and with the BF16-split factor, each timed as one blocking
    call on tiles already in HBM (A copied fresh before each solve, outside
    the timed region), then checked on the GPU against the original A."""
    import ctypes
    import numpy as np
    mi, _ = dp.libs()
    eps = 2.0 ** -53
    out = {"N": N, "NB": NB, "nrhs": nrhs, "cases": []}
    with dp.Device(0) as dev:
        MT = (N + NB - 1) // NB
        nt = MT * (MT + 1) // 2 * NB * NB
        t0 = time.perf_counter()
        dQ = dev.alloc(N * N * 8)
        dev.check(mi.mi_dorth_haar(dev.h, N, 7, dQ), "orth")
        out["orth_s"] = round(time.perf_counter() - t0, 3)
        dA, dW = dev.alloc(nt * 8), dev.alloc(nt * 8)
        dB, dX, dX64 = dev.alloc(N * nrhs * 8), dev.alloc(N * nrhs * 8), dev.alloc(N * nrhs * 8)
        dev.check(mi.mi_dplrnt(dev.h, N, nrhs, dB, N, N, 0, 0, 3873), "B")

        def check(dx):
            a, b, x, r = (ctypes.c_double() for _ in range(4))
            dev.check(mi.mi_dcheck_axmb(dev.h, N, NB, dA, nrhs, dB, N, dx, N, ctypes.byref(a), ctypes.byref(b),
                                        ctypes.byref(x), ctypes.byref(r)), "check")
            berr = r.value / (a.value * x.value + b.value)
            return berr, berr / (N * eps)

        info, it = ctypes.c_int(), ctypes.c_int()
        warmed = set()
        for kappa in kappas:
            t0 = time.perf_counter()
            dev.check(mi.mi_dspd_from_orth_tiles(dev.h, N, NB, dQ, kappa, dA), "spd")
            case = {"kappa": kappa, "gen_s": round(time.perf_counter() - t0, 3)}
            dev.check(mi.mi_memcpy_d2d(dev.h, dW, dA, nt * 8), "copy")
            dev.check(mi.mi_memcpy_d2d(dev.h, dX64, dB, N * nrhs * 8), "copy")
            t0 = time.perf_counter()
            dev.check(mi.mi_dpotrf_tiles(dev.h, N, NB, dW, ctypes.byref(info)), "dpotrf")
            dev.check(mi.mi_dpotrs_tiles(dev.h, N, NB, dW, nrhs, dX64, N), "dpotrs")
            t64 = time.perf_counter() - t0
            berr, ratio = check(dX64)
            case["fp64"] = {"s": round(t64, 4), "info": info.value, "backward_error": berr, "check_daxmb": ratio}
            x64 = dev.download(dX64, (N * nrhs,))
            for prec, name in ((0, "fp32"), (1, "bf16x3")):
                if prec not in warmed:           # first call sizes the solver's workspaces
                    dev.check(mi.mi_memcpy_d2d(dev.h, dW, dA, nt * 8), "copy")
                    dev.check(mi.mi_dsposv_tiles(dev.h, N, NB, dW, nrhs, dB, N, dX, N, prec, ctypes.byref(it),
                                                 ctypes.byref(info)), "dsposv")
                    warmed.add(prec)
                dev.check(mi.mi_memcpy_d2d(dev.h, dW, dA, nt * 8), "copy")
                t0 = time.perf_counter()
                dev.check(mi.mi_dsposv_tiles(dev.h, N, NB, dW, nrhs, dB, N, dX, N, prec, ctypes.byref(it),
                                             ctypes.byref(info)), "dsposv")
                tm = time.perf_counter() - t0
                berr, ratio = check(dX)
                x = dev.download(dX, (N * nrhs,))
                case[name] = {"s": round(tm, 4), "iter": it.value, "info": info.value,
                              "backward_error": berr, "check_daxmb": ratio,
                              "err_vs_fp64": float(np.abs(x - x64).max() / np.abs(x64).max()),
                              "speedup_vs_fp64": round(t64 / tm, 3)}
            out["cases"].append(case)
            print(f"[mixed] kappa={kappa:g}: fp64 {t64*1e3:.1f} ms, fp32 {case['fp32']['s']*1e3:.1f} ms "
                  f"(iter {case['fp32']['iter']}), bf16x3 {case['bf16x3']['s']*1e3:.1f} ms "
                  f"(iter {case['bf16x3']['iter']})", file=sys.stderr, flush=True)
        for p in (dQ, dA, dW, dB, dX, dX64):
            dev.free(p)
    return out


def flops_adjoint(N, NB):
    """Flops of the blocked adjoint as its kernels run them: S1 2 M^2 NB,
    S2 + S3 4 M NB^2 per step (M = rows below the diagonal tile), and the
    three NB^3 products of the diagonal tile (6 NB^3) per step."""
    MT = (N + NB - 1) // NB
    f = 0.0
    for k in range(MT):
        M = max(0, N - (k + 1) * NB)
        f += 2.0 * M * M * NB + 4.0 * M * NB * NB + 6.0 * min(NB, N - k * NB) ** 3
    return f


def gp_leg(dp, N=8192, NB=512, batch=64, nb_small=1024):
    """Differentiated Gaussian-process NLL: N points uniform in [0,1]^3
    (seed 5), SE kernel (lengthscale 0.2, variance 1), noise 1e-2, y = L z
    ~ N(0, K).  Forward: tile Cholesky, two triangular solves, logdet;
    backward: the blocked adjoint (torch autograd over the C ABI).  The
    gradient dNLL/dK is compared with the analytic (K^-1 - alpha alpha^T)/2
    (torch's own FP64 Cholesky inverse) and dNLL/dl with central differences
    (h = 1e-6).  Also FP32 (the kernel matrix rounded to FP32) and a batch
    of 64 GPs of 1024 points each (one after another)."""
    import ctypes
    import torch
    import dplasma_amd.autograd as ag
    dev = torch.device("cuda")
    f64 = torch.float64

    def last_ms():
        d = ag._Lib.get()
        ms = ctypes.c_double()
        d.M.mi_last_span_ms(d.h, ctypes.byref(ms))
        return ms.value

    def problem(n, seed):
        g = torch.Generator().manual_seed(seed)
        X = torch.rand(n, 3, generator=g, dtype=f64).to(dev)
        z = torch.randn(n, generator=g, dtype=f64).to(dev)
        with torch.no_grad():
            y = ag.cholesky(ag.se_kernel(X, 0.2), NB if n == N else 256) @ z
        return X, y

    def run(X, y, nb, dtype=f64):
        ell = torch.tensor(0.2, dtype=f64, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        K = ag.se_kernel(X, ell)
        K.retain_grad()
        Kd = K if dtype == f64 else K.to(dtype)
        if dtype != f64:
            Kd.retain_grad()
        nll = ag.gp_nll(Kd, y.to(dtype), nb)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        potrf_ms = last_ms()
        nll.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return nll, ell, (Kd.grad if dtype != f64 else K.grad), K.detach(), t1 - t0, t2 - t1, potrf_ms, last_ms()

    def analytic(K, y):
        Lt = torch.linalg.cholesky(K)
        Ki = torch.cholesky_inverse(Lt)
        al = Ki @ y
        return 0.5 * (Ki - torch.outer(al, al))

    X, y = problem(N, 5)
    run(X, y, NB)                                  # warm-up (workspaces, code objects)
    nll, ell, G, K, tf, tb, pm, am = run(X, y, NB)
    Ga = analytic(K, y)
    rel = ((G - Ga).abs().max() / Ga.abs().max()).item()
    h = 1e-6
    with torch.no_grad():
        fd = (ag.gp_nll(ag.se_kernel(X, 0.2 + h), y, NB) - ag.gp_nll(ag.se_kernel(X, 0.2 - h), y, NB)).item() / (2 * h)
    out = {"N": N, "NB": NB, "fp64": {
        "nll": nll.item(), "forward_ms": round(tf * 1e3, 3), "backward_ms": round(tb * 1e3, 3),
        "dpotrf_ms": round(pm, 3), "adjoint_ms": round(am, 3),
        "adjoint_tflops": round(flops_adjoint(N, NB) / am / 1e9, 2),
        "grad_K_rel_err_vs_analytic": rel,
        "dnll_dlengthscale": ell.grad.item(), "fd_dnll_dlengthscale": fd,
        "lengthscale_rel_err_vs_fd": abs(ell.grad.item() - fd) / abs(fd)}}
    del G, Ga
    run(X, y, NB, torch.float32)
    nll32, ell32, G32, _, tf, tb, pm, am = run(X, y, NB, torch.float32)
    Ga = analytic(K, y)
    out["fp32"] = {"nll": nll32.item(), "forward_ms": round(tf * 1e3, 3), "backward_ms": round(tb * 1e3, 3),
                   "spotrf_ms": round(pm, 3), "adjoint_ms": round(am, 3),
                   "grad_K_rel_err_vs_fp64_analytic": ((G32.double() - Ga).abs().max() / Ga.abs().max()).item(),
                   "lengthscale_rel_err_vs_fp64_fd": abs(ell32.grad.item() - fd) / abs(fd)}
    del G32, Ga, K
    print(f"[gp] N={N}: fp64 fwd {out['fp64']['forward_ms']} ms bwd {out['fp64']['backward_ms']} ms "
          f"(adjoint {out['fp64']['adjoint_ms']} ms), grad rel err {rel:.2e}", file=sys.stderr, flush=True)
    # a batch of GPs of nb_small points, one after another
    worst = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b in range(batch):
        Xb, yb = problem(nb_small, 1000 + b)
        _, _, Gb, Kb, _, _, _, _ = run(Xb, yb, 256)
        if b < 4:
            worst = max(worst, ((Gb - analytic(Kb, yb)).abs().max() / analytic(Kb, yb).abs().max()).item())
    torch.cuda.synchronize()
    out["batch"] = {"gps": batch, "N": nb_small, "NB": 256, "total_s": round(time.perf_counter() - t0, 3),
                    "grad_K_rel_err_vs_analytic_first4": worst}
    return out


def batched_leg(dp, count=4096, lo=32, hi=1024, seed=11):
    """Batched uneven Cholesky + solve: `count` SPD matrices with orders
    uniform in [lo, hi] (splitmix64, seed 11), A_i = G_i^T G_i / n_i + I with
    G_i Gaussian, b uniform in [-0.5, 0.5], NRHS = 1.  FP64 and FP32 (the FP64
    batch rounded), each as one persistent launch replayed from a hipGraph
    with the tile shape the autotuner picked, against the same kernel
    launched once per matrix (FP64).  Flops: LAWN-41 dpotrf + dpotrs counts
    summed over the batch.  Backward errors of 8 matrices checked on the host."""
    import ctypes
    import numpy as np
    mi, _ = dp.libs()
    ns = (ctypes.c_int * count)()
    mi.mi_batch_sizes(seed, count, lo, hi, ns)
    nl = np.array(list(ns), dtype=np.float64)
    flops = float(np.sum(nl ** 3 / 3 + nl ** 2 / 2 + nl / 6 + 2 * nl ** 2))
    out = {"count": count, "n_range": [lo, hi], "flops": flops}
    with dp.Device(0) as dev:
        p = ctypes.c_void_p()
        dev.check(mi.mi_batch_create(dev.h, count, ns, ctypes.byref(p)), "batch_create")
        ae, be = ctypes.c_int64(), ctypes.c_int64()
        mi.mi_batch_extent(p, ctypes.byref(ae), ctypes.byref(be))
        A0, A, B0, B = dev.alloc(ae.value * 8), dev.alloc(ae.value * 8), dev.alloc(be.value * 8), dev.alloc(be.value * 8)
        dev.check(mi.mi_dbatch_spd(dev.h, count, ns, seed, A0), "batch_spd")
        dev.check(mi.mi_dplrnt(dev.h, be.value, 1, B0, be.value, be.value, 0, 0, seed), "b")
        ms = ctypes.c_double()
        sh_n, sh_w = ctypes.c_int(), ctypes.c_int()

        def timed(fn, a0, a, b0, b, nbytes_a, nbytes_b, reps=3):
            best = []
            for _ in range(reps):
                dev.check(mi.mi_memcpy_d2d(dev.h, a, a0, nbytes_a), "copy")
                dev.check(mi.mi_memcpy_d2d(dev.h, b, b0, nbytes_b), "copy")
                dev.check(fn(), "batched")
                mi.mi_batch_last_ms(p, ctypes.byref(ms))
                best.append(ms.value)
            return sorted(best)[len(best) // 2]

        def check(a0, x, b0, to64):
            ao = np.cumsum([0] + [n * n for n in ns])
            bo = np.cumsum([0] + list(ns))
            worst = 0.0
            for i in list(range(0, count, max(1, count // 8)))[:8]:
                n = ns[i]
                Ai = to64(a0, ao[i], n * n).reshape(n, n, order="F")
                Ai = np.tril(Ai) + np.tril(Ai, -1).T
                xi, bi = to64(x, bo[i], n), to64(b0, bo[i], n)
                r = np.abs(bi - Ai @ xi).max() / (np.abs(Ai).sum(1).max() * np.abs(xi).max() + np.abs(bi).max())
                worst = max(worst, r)
            return worst

        def dl64(ptr, off, k):
            buf = np.empty(k)
            dev.check(mi.mi_memcpy_d2h(dev.h, buf.ctypes.data, int(ptr) + int(off) * 8, int(k) * 8), "d2h")
            return buf

        ch = ctypes.c_int()
        dev.check(mi.mi_dbatch_autotune(p, A0, B0, A, B, ctypes.byref(ch)), "autotune")
        mi.mi_batch_get_shape(p, ctypes.byref(sh_n), ctypes.byref(sh_w))
        t64 = timed(lambda: mi.mi_dposv_batched(p, A, B, 1), A0, A, B0, B, ae.value * 8, be.value * 8)
        info = (ctypes.c_int * count)()
        mi.mi_batch_info(p, info)
        out["fp64"] = {"ms": round(t64, 3), "matrices_per_s": round(count / t64 * 1e3, 1),
                       "tflops": round(flops / t64 / 1e9, 3), "shape": {"nsplit": sh_n.value, "wg_per_cu": sh_w.value},
                       "failed": sum(1 for v in info if v),
                       "max_backward_error_sample": check(A0, B, B0, dl64)}
        tl = timed(lambda: mi.mi_dposv_batched_loop(p, A, B), A0, A, B0, B, ae.value * 8, be.value * 8, reps=1)
        out["fp64_launch_loop"] = {"ms": round(tl, 3), "matrices_per_s": round(count / tl * 1e3, 1),
                                   "tflops": round(flops / tl / 1e9, 4)}
        out["fp64"]["speedup_vs_launch_loop"] = round(tl / t64, 2)
        print(f"[batched] fp64 {t64:.2f} ms ({count / t64 * 1e3:.0f} matrices/s), launch loop {tl:.1f} ms",
              file=sys.stderr, flush=True)
        # FP32: the FP64 batch rounded
        A32, A32w, B32, B32w = (dev.alloc(ae.value * 4), dev.alloc(ae.value * 4), dev.alloc(be.value * 4),
                                dev.alloc(be.value * 4))
        ovf = ctypes.c_int()
        dev.check(mi.mi_dlag2s_tiles(dev.h, ae.value, A0, A32, ctypes.byref(ovf)), "d2s")
        dev.check(mi.mi_dlag2s_tiles(dev.h, be.value, B0, B32, ctypes.byref(ovf)), "d2s")
        dev.check(mi.mi_sbatch_autotune(p, A32, B32, A32w, B32w, ctypes.byref(ch)), "autotune32")
        mi.mi_batch_get_shape(p, ctypes.byref(sh_n), ctypes.byref(sh_w))
        t32 = timed(lambda: mi.mi_sposv_batched(p, A32w, B32w, 1), A32, A32w, B32, B32w, ae.value * 4, be.value * 4)

        def dl32(ptr, off, k):
            buf = np.empty(k, dtype=np.float32)
            dev.check(mi.mi_memcpy_d2h(dev.h, buf.ctypes.data, int(ptr) + int(off) * 4, int(k) * 4), "d2h")
            return buf.astype(np.float64)

        out["fp32"] = {"ms": round(t32, 3), "matrices_per_s": round(count / t32 * 1e3, 1),
                       "tflops": round(flops / t32 / 1e9, 3), "shape": {"nsplit": sh_n.value, "wg_per_cu": sh_w.value},
                       "max_backward_error_sample": check(A32, B32w, B32, dl32)}
        for q in (A0, A, B0, B, A32, A32w, B32, B32w):
            dev.free(q)
        mi.mi_batch_destroy(p)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nb", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision dposv leg")
    ap.add_argument("--mixed-only", action="store_true", help="run only the mixed-precision dposv leg")
    ap.add_argument("--kappas", default="1e2,1e5,1e8", help="condition numbers of the mixed leg")
    ap.add_argument("--no-gp", action="store_true", help="skip the Gaussian-process adjoint leg")
    ap.add_argument("--gp-only", action="store_true", help="run only the Gaussian-process adjoint leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched uneven Cholesky leg")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched uneven Cholesky leg")
    ap.add_argument("--profile-only", action="store_true",
                    help="one untimed factorization, for rocprofv3 runs")
    ap.add_argument("--profile-kind", default="dpotrf",
                    choices=["dpotrf", "spotrf", "spotrf_bf16x3", "dpotrs", "dsposv_bf16x3", "adjoint", "batched",
                             "batched_fp32"],
                    help="what --profile-only runs: the FP64 factorization, the FP32 factorization (FP32 "
                         "MFMA or BF16-split updates) of the same matrix rounded, the FP64 solve (16 RHS), "
                         "the mixed-precision solve with the BF16-split factor (kappa = 1e2, 16 RHS; two "
                         "calls, the first sizes the workspaces), the adjoint of the FP64 factorization, or "
                         "one launch of the batched solve of 4096 uneven matrices (FP64 or FP32) at the tuned "
                         "shape")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run the plain copy kernel on 1 GiB (scales HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); "
              "a block-cyclic tile map over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    import dplasma_amd as dp
    if args.copy_calibration:
        with dp.Device(0) as dev:
            nbytes = 1 << 30
            s, d = dev.alloc(nbytes), dev.alloc(nbytes)
            for _ in range(3):
                dev.check(dev.M.mi_dcopy(dev.h, nbytes // 8, s, d), "mi_dcopy")
            dev.free(s)
            dev.free(d)
        print(json.dumps({"copy_calibration_bytes_per_launch": {"read": 1 << 30, "write": 1 << 30}}))
        return
    N, NB = args.n, args.nb
    if args.gp_only:
        print(json.dumps({"gp_nll": gp_leg(dp)}))
        return
    if args.batched_only:
        print(json.dumps({"batched_posv": batched_leg(dp)}))
        return
    if args.mixed_only:
        kap = tuple(float(k) for k in args.kappas.split(","))
        print(json.dumps({"mixed_dposv": mixed_leg(dp, N, NB, kappas=kap)}))
        return
    if args.profile_only and args.profile_kind.startswith("batched"):
        import ctypes
        mi, _ = dp.libs()
        count, seed = 4096, 11
        ns = (ctypes.c_int * count)()
        mi.mi_batch_sizes(seed, count, 32, 1024, ns)
        fp32 = args.profile_kind == "batched_fp32"
        with dp.Device(0) as dev:
            p = ctypes.c_void_p()
            dev.check(mi.mi_batch_create(dev.h, count, ns, ctypes.byref(p)), "batch_create")
            ae, be = ctypes.c_int64(), ctypes.c_int64()
            mi.mi_batch_extent(p, ctypes.byref(ae), ctypes.byref(be))
            A, B = dev.alloc(ae.value * 8), dev.alloc(be.value * 8)
            dev.check(mi.mi_dbatch_spd(dev.h, count, ns, seed, A), "batch_spd")
            dev.check(mi.mi_dplrnt(dev.h, be.value, 1, B, be.value, be.value, 0, 0, seed), "b")
            dev.check(mi.mi_batch_set_shape(p, 1024, 2), "shape")     # the shape the tuner picks here
            if fp32:
                A32, B32 = dev.alloc(ae.value * 4), dev.alloc(be.value * 4)
                ovf = ctypes.c_int()
                dev.check(mi.mi_dlag2s_tiles(dev.h, ae.value, A, A32, ctypes.byref(ovf)), "A32")
                dev.check(mi.mi_dlag2s_tiles(dev.h, be.value, B, B32, ctypes.byref(ovf)), "B32")
                dev.check(mi.mi_sposv_batched(p, A32, B32, 0), "sposv_batched")
            else:
                dev.check(mi.mi_dposv_batched(p, A, B, 0), "dposv_batched")
            mi.mi_batch_destroy(p)
        print("profile step done")
        return
    if args.profile_only and args.profile_kind == "dsposv_bf16x3":
        import ctypes
        mi, _ = dp.libs()
        with dp.Device(0) as dev:
            MT = (N + NB - 1) // NB
            nt = MT * (MT + 1) // 2 * NB * NB
            dQ = dev.alloc(N * N * 8)
            dev.check(mi.mi_dorth_haar(dev.h, N, 7, dQ), "orth")
            dA, dW = dev.alloc(nt * 8), dev.alloc(nt * 8)
            dev.check(mi.mi_dspd_from_orth_tiles(dev.h, N, NB, dQ, 1e2, dA), "spd")
            dev.free(dQ)
            dB, dX = dev.alloc(N * 16 * 8), dev.alloc(N * 16 * 8)
            dev.check(mi.mi_dplrnt(dev.h, N, 16, dB, N, N, 0, 0, 3873), "B")
            info, it = ctypes.c_int(), ctypes.c_int()
            for _ in range(2):
                dev.check(mi.mi_memcpy_d2d(dev.h, dW, dA, nt * 8), "copy")
                dev.check(mi.mi_dsposv_tiles(dev.h, N, NB, dW, 16, dB, N, dX, N, 1, ctypes.byref(it),
                                             ctypes.byref(info)), "dsposv")
                assert info.value == 0 and it.value > 0, (info.value, it.value)
        print("profile step done")
        return
    if args.profile_only and args.profile_kind == "adjoint":
        import ctypes
        mi, _ = dp.libs()
        with dp.Device(0) as dev:
            MT = (N + NB - 1) // NB
            nt = MT * (MT + 1) // 2 * NB * NB
            dA, dG = dev.alloc(nt * 8), dev.alloc(nt * 8)
            dev.check(mi.mi_dplgsy_tiles(dev.h, float(N), N, NB, dA, 3872), "dplgsy")
            info = ctypes.c_int()
            dev.check(mi.mi_dpotrf_tiles(dev.h, N, NB, dA, ctypes.byref(info)), "dpotrf")
            assert info.value == 0
            dev.check(mi.mi_dplgsy_tiles(dev.h, 1.0, N, NB, dG, 11), "Lbar")
            dev.check(mi.mi_dpotrf_adjoint_tiles(dev.h, N, NB, dA, dG), "adjoint")
        print("profile step done")
        return
    if args.profile_only and args.profile_kind != "dpotrf":
        import ctypes
        mi, _ = dp.libs()
        with dp.Device(0) as dev:
            MT = (N + NB - 1) // NB
            nt = MT * (MT + 1) // 2 * NB * NB
            dA = dev.alloc(nt * 8)
            dev.check(mi.mi_dplgsy_tiles(dev.h, float(N), N, NB, dA, 3872), "dplgsy")
            info = ctypes.c_int()
            if args.profile_kind == "dpotrs":
                dev.check(mi.mi_dpotrf_tiles(dev.h, N, NB, dA, ctypes.byref(info)), "dpotrf")
                dB = dev.alloc(N * 16 * 8)
                dev.check(mi.mi_dplrnt(dev.h, N, 16, dB, N, N, 0, 0, 3873), "B")
                dev.check(mi.mi_dpotrs_tiles(dev.h, N, NB, dA, 16, dB, N), "dpotrs")
            else:
                d32 = dev.alloc(nt * 4)
                ovf = ctypes.c_int()
                dev.check(mi.mi_dlag2s_tiles(dev.h, nt, dA, d32, ctypes.byref(ovf)), "dlag2s")
                mi.mi_set_spotrf_update(1 if args.profile_kind == "spotrf_bf16x3" else 0)
                dev.check(mi.mi_spotrf_tiles(dev.h, N, NB, d32, ctypes.byref(info)), "spotrf")
                mi.mi_set_spotrf_update(0)
            assert info.value == 0
        print("profile step done")
        return
    if args.profile_only:
        ctx = dp.Context(0)
        A = dp.SymMatrix(ctx, N, NB, dp.LOWER)
        assert ctx.advise(dp.LOWER, A) == 0
        assert ctx.dplgsy(float(N), dp.LOWER, A, 3872) == 0
        assert ctx.dpotrf(dp.LOWER, A) == 0
        A.free()
        ctx.close()
        print("profile step done")
        return
    ctx = dp.Context(0)
    A = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    assert ctx.advise(dp.LOWER, A) == 0           # tiles live in HBM from here on

    def regen():
        assert ctx.dplgsy(float(N), dp.LOWER, A, 3872) == 0

    for _ in range(args.warmup):
        regen()
        assert ctx.dpotrf(dp.LOWER, A) == 0
    times = []
    for _ in range(args.steps):
        regen()
        t0 = time.perf_counter()              # dpotrf is blocking: synchronized on both sides
        info = ctx.dpotrf(dp.LOWER, A)
        times.append(time.perf_counter() - t0)
        assert info == 0, info
    total = sum(times)
    fl = flops_dpotrf(N)
    value = fl * args.steps / total / 1e12

    # profiled step (not timed above): per-kernel durations from HIP events
    regen()
    ctx.prof_reset()
    ctx.prof_enable(True)
    assert ctx.dpotrf(dp.LOWER, A) == 0
    prof = ctx.prof()
    ctx.prof_enable(False)
    # the dominant kernel: every trailing-update launch (the main stream's
    # and the panel stream's look-ahead columns), algorithmic flops over
    # the union of their launch intervals, so that the overlap of the two
    # streams counts once; the main-stream launches alone are reported beside
    upd = prof.get(UPDATE_KERNEL, (0, 0.0, 0.0, 0.0))
    upd_la = prof.get(UPDATE_KERNEL + "_la", (0, 0.0, 0.0, 0.0))
    busy_main = ctx.prof_busy_ms(UPDATE_KERNEL)
    busy_all = ctx.prof_busy_ms(UPDATE_KERNEL + "*")
    all_upd_flops = upd[2] + upd_la[2]
    all_upd_bytes = upd[3] + upd_la[3]
    n_all = upd[0] + upd_la[0]
    achieved = all_upd_flops / busy_all / 1e9 if busy_all > 0 else 0.0
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    # correctness of the timed path: reference residual checks on the GPU
    A0 = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    assert ctx.advise(dp.LOWER, A0) == 0
    assert ctx.dplgsy(float(N), dp.LOWER, A0, 3872) == 0
    chk_f = ctx.check_dpotrf(0, dp.LOWER, A, A0)
    B = dp.GeMatrix(ctx, N, 1, NB, NB)
    X = dp.GeMatrix(ctx, N, 1, NB, NB)
    ctx.dplrnt(0, B, 3873)
    ctx.dplrnt(0, X, 3873)
    assert ctx.advise(dp.UPPERLOWER, X) == 0
    # untimed solve of a scratch right-hand side: the solve's workspaces are
    # allocated on first use
    Wm = dp.GeMatrix(ctx, N, 1, NB, NB)
    ctx.dplrnt(0, Wm, 3874)
    assert ctx.advise(dp.UPPERLOWER, Wm) == 0
    assert ctx.dpotrs(dp.LOWER, A, Wm) == 0
    t0 = time.perf_counter()
    assert ctx.dpotrs(dp.LOWER, A, X) == 0
    t_potrs = time.perf_counter() - t0
    ctx.sync_host(X)
    chk_s = ctx.check_daxmb(0, dp.LOWER, A0, B, X)

    # the GPU factor on the host, for the comparison with the CPU oracle's
    L_gpu = None
    if not args.no_cpu and N % NB == 0:
        assert ctx.sync_host(A) == 0
        L_gpu = A.array.copy()

    # adjoint of the same factorization: Lbar = a dplgsy-pattern matrix (seed 11), in HBM
    Gm = A0                                      # reuse: A0 is no longer needed
    assert ctx.dplgsy(1.0, dp.LOWER, Gm, 11) == 0
    t0 = time.perf_counter()
    assert ctx.dpotrf_adjoint(dp.LOWER, A, Gm) == 0
    t_adj = time.perf_counter() - t0
    adj_ms = ctx.last_timing()["total_ms"]

    iso = isolated_update(dp, N, NB)
    staged = staged_leg(dp, N, NB)

    line = {
        "metric": "dposv s to FP64 backward err @N=32k: mixed vs FP64 GPU; dpotrf TFLOP/s",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: dplgsy(bump=N, seed 3872) generated in HBM",
        "config": {"workload": f"dpotrf Lower FP64 N={N} NB={NB}, tiles resident in HBM",
                   "N": N, "NB": NB, "uplo": "Lower"},
        "fraction_of_fp64_peak": round(value / FP64_PEAK_TF, 4),
        "checks": {"check_dpotrf_pass": chk_f == 0, "check_daxmb_pass": chk_s == 0},
        "fp64_dposv_s": round(total / args.steps + t_potrs, 4),
        "dpotrs_ms": round(t_potrs * 1e3, 3),

        "kernels_ms": {k: round(v[1], 3) for k, v in prof.items()},
        "adjoint": {"ms": round(adj_ms, 3), "host_ms": round(t_adj * 1e3, 3),
                    "flops": flops_adjoint(N, NB), "tflops": round(flops_adjoint(N, NB) / adj_ms / 1e9, 3),
                    "vs_dpotrf_time": round(adj_ms / (total / args.steps * 1e3), 3)},
        "roofline": {
            "bound": "mfma",
            "kernel": UPDATE_KERNEL,
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TF,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TF, 4),
            "traffic": traffic,
            "what": "all trailing-update launches (main stream and look-ahead): algorithmic flops over the "
                    "union of their launch intervals (overlaps counted once); traffic = measured HBM bytes "
                    "per launch averaged over the same launches",
            "launches": n_all,
            "flops_per_launch": round(all_upd_flops / n_all) if n_all else None,
            "algorithmic_bytes_per_launch": round(all_upd_bytes / n_all) if n_all else None,
            "busy_ms": round(busy_all, 3),
            "main_stream": {
                "what": "the bulk updates on the main stream alone (one launch at a time)",
                "launches": upd[0],
                "busy_ms": round(busy_main, 3),
                "achieved": round(upd[2] / busy_main / 1e9, 3) if busy_main > 0 else None,
            },
            "isolated": iso,
        },
        "staged_dpotrf": staged,
    }
    for M in (A, A0, B, X):
        M.free()
    ctx.close()
    if not args.no_mixed:
        line["mixed_dposv"] = mixed_leg(dp, N, NB, kappas=tuple(float(k) for k in args.kappas.split(",")))
    if not args.no_batched:
        line["batched_posv"] = batched_leg(dp)
    if not args.no_gp:
        line["gp_nll"] = gp_leg(dp)
    if not args.no_cpu:
        T_cpu, line["cpu_baseline"] = cpu_baseline()
        if T_cpu is not None and L_gpu is not None and N == 32768:
            line["vs_cpu_oracle"] = compare_lower_tiles(L_gpu, T_cpu, N, NB)
        line["dpotrf_N4096_NB256"] = small_case_leg(dp)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
