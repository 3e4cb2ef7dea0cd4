#!/usr/bin/env python3
"""L2C transforms/s at N = 2^20 on one MI355X (plain FP64 FMM, maxs = 64, M = 18).

A step is one Legendre->Chebyshev transform of one synthetic coefficient
vector (U(-1,1), SplitMix64 seed 1) that is already resident in HBM.  Warmup
steps are untimed; the timed region is bracketed by device synchronisation.
Every execute kernel is timed live with HIP events on its own dispatch; the
dominant kernel's average duration gives the roofline figure.  The result is
checked against the oracle (CPU restatement of the reference's execute) and
against extended-precision direct sums on sampled rows.  The oracle, run on
the host cores, is the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Other workloads, each one JSON line: --workload batch (4096 transforms of
uneven length), --workload field (a 256^3 field along every axis),
--workload dct (the length-N DCT-II, the reference's FFTW yardstick).
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "legendre-to-chebyshev_amd")
HBM_PEAK_GBS = 8000.0
METRIC = "L2C transforms/s at N=2^20 and batched N<=8192; max rel err vs FP64 oracle"
# The reference's published L2C figure at N = 2^20: 1.6530e-02 s per transform
# on one MacBook M3 Pro core, maxs = 32 (s = 32), M = 18
# (src/results/figure4.sh:54).
PUBLISHED = {"value": 60.5, "unit": "transforms/s", "source": "src/results/figure4.sh:54",
             "hardware": "MacBook M3 Pro, 1 core", "config": "L2C N=2^20 maxs=32 (s=32) M=18"}


def load_pkg():
    if "l2c_amd" in sys.modules:
        return sys.modules["l2c_amd"]
    spec = importlib.util.spec_from_file_location(
        "l2c_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["l2c_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_bytes(name, info, N):
    """Algorithmic HBM bytes of one launch of an execute kernel.

    k_m2l: the M2L matrices (Nc*M^2 doubles), two multipole vectors read and
    two M2L vectors written per block (Nb blocks, 2M doubles each).
    k_down: input, output, the Lambda table the band reads (a for L2C, an for
    C2L; the 2s-entry head is negligible) and one M2L vector per chunk slot.
    k_up: input, one multipole vector per chunk slot.
    k_direct (single transforms of N <= 8192 on an FMM plan): input, output
    and the Lambda table a (C2L reads an and dn, 1.5 N; counted as N).
    """
    M, L, Nb, Nc = info["M"], info["L"], info["Nb"], info["Nc"]
    slots = (4 << L) - 4
    if name == "k_m2l":
        return 8 * (Nc * M * M + 4 * Nb * 2 * M)
    if name == "k_down":
        return 8 * (3 * N + slots * 2 * M)
    if name == "k_up":
        return 8 * (N + slots * 2 * M)
    if name == "k_direct":
        return 8 * 3 * N
    return 8 * 2 * N


def workload_name(N, maxs, M, d):
    """Name of the bench workload, with the leaf size s of create_fmm's
    geometry (src/C/leg2cheb.c:860-871)."""
    L = math.ceil(math.log2(N / maxs)) - 2
    s = math.ceil(N / 2 ** (L + 2)) if L >= 1 else N
    return f"{'l2c' if d == 0 else 'c2l'}_fp64_N{N}_s{s}_M{M}"


def pmc_path(workload):
    return os.path.join(ROOT, "profiles", f"execute_kernels_{workload}_pmc.json")


def pmc_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary of this workload (written by --pmc-summary), or None."""
    path = pmc_path(workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f).get("kernels", {}).get(kernel)
    return None if k is None else k["hbm_bytes_per_launch"]


def _counter_means(csv_dir, counter, copy_kib):
    """Mean per-dispatch value of `counter` for each execute kernel found in
    the rocprofv3 counter_collection.csv files under csv_dir, and for the
    calibration copy (copy kernels reading at least half of copy_kib; smaller
    copies are the input uploads)."""
    import csv
    import glob
    acc = {}
    for path in glob.glob(os.path.join(csv_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                name = r["Kernel_Name"]
                if name.startswith("void "):
                    name = name[5:]
                if name.startswith("l2c::"):
                    name = name[5:]
                # profile slots: k_m2l covers k_m2l_t and k_m2l, k_down the
                # downward kernels (k_down8r, k_down8, k_down)
                for tag, slot in (("k_up", "k_up"), ("k_m2l_t", "k_m2l"), ("k_m2l", "k_m2l"),
                                  ("k_down8r", "k_down"),
                                  ("k_down8", "k_down"), ("k_down", "k_down"),
                                  ("k_direct", "k_direct")):
                    if name.startswith(tag + "<") or name.startswith(tag + "(") or name == tag:
                        name = slot
                        break
                else:
                    if "copy" not in name.lower() or float(r["Counter_Value"]) < 0.5 * copy_kib:
                        continue
                    name = "copy"
                s = acc.setdefault(name, [0.0, 0])
                s[0] += float(r["Counter_Value"])
                s[1] += 1
    return {k: v[0] / v[1] for k, v in acc.items()}


def pmc_summary(args):
    """Turn the FETCH_SIZE and WRITE_SIZE passes (KiB per dispatch) into HBM
    bytes per launch, each scaled by what the same counter reads for the
    plain copy of --copy-mb MiB run in the same process."""
    nbytes = args.copy_mb * 1024 * 1024
    fetch = _counter_means(args.pmc_summary[0], "FETCH_SIZE", nbytes / 1024.0)
    write = _counter_means(args.pmc_summary[1], "WRITE_SIZE", nbytes / 1024.0)
    f_scale = nbytes / (fetch["copy"] * 1024.0)
    w_scale = nbytes / (write["copy"] * 1024.0)
    out = {"copy_bytes": nbytes, "fetch_scale": f_scale, "write_scale": w_scale,
           "copy_fetch_kib": fetch["copy"], "copy_write_kib": write["copy"], "kernels": {}}
    for k in sorted(set(fetch) & set(write)):
        if k == "copy":
            continue
        fb, wb = fetch[k] * 1024.0 * f_scale, write[k] * 1024.0 * w_scale
        out["kernels"][k] = {"fetch_kib": fetch[k], "write_kib": write[k], "fetch_bytes": round(fb),
                             "write_bytes": round(wb), "hbm_bytes_per_launch": round(fb + wb)}
    return out


def copy_calibration(amd, mb):
    """Plain device-to-device copies of `mb` MiB (the runtime's copy kernel),
    run so the --pmc passes see a kernel of known traffic in the same process."""
    n = mb * 1024 * 1024 // 8
    a, b = amd.DeviceBuffer(n), amd.DeviceBuffer(n)  # contents irrelevant
    for _ in range(3):
        if amd.lib().l2c_memcpy_d2d(b.ptr, a.ptr, n * 8) != 0:
            raise RuntimeError(amd.last_error())
    del a, b


def per_transform_times(plan, din, dout, d, n):
    """Device time of each of n back-to-back transforms: HIP events recorded
    before and after every l2c_execute_device call on a stream of its own
    (the HIP runtime's event API through ctypes).  Returns (median, min) in
    microseconds."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    st = C.c_void_p()
    if hip.hipStreamCreate(C.byref(st)) != 0:
        raise RuntimeError("hipStreamCreate failed")
    evs = []
    try:
        for _ in range(2 * n):
            e = C.c_void_p()
            if hip.hipEventCreate(C.byref(e)) != 0:
                raise RuntimeError("hipEventCreate failed")
            evs.append(e)
        for k in range(n):
            hip.hipEventRecord(evs[2 * k], st)
            plan.execute_device(din.ptr, dout.ptr, d, stream=st.value)
            hip.hipEventRecord(evs[2 * k + 1], st)
        if hip.hipEventSynchronize(evs[-1]) != 0:
            raise RuntimeError("hipEventSynchronize failed")
        us = []
        for k in range(n):
            ms = C.c_float()
            hip.hipEventElapsedTime(C.byref(ms), evs[2 * k], evs[2 * k + 1])
            us.append(1e3 * ms.value)
    finally:
        for e in evs:
            hip.hipEventDestroy(e)
        hip.hipStreamDestroy(st)
    us.sort()
    return us[len(us) // 2], us[0]


def bench_batch(amd, orc, args):
    """The batched half of the metric: `--batch` transforms of uneven length
    n = floor(exp(U(ln 64, ln 8193))) (SplitMix64 seed 2024), inputs
    U(-1,1)/(k+1) packed back to back and resident in HBM, L2C, one tuned
    batch plan replayed as one HIP graph per step.  Beside it: the same
    transforms as sequential single-transform calls (plans made beforehand,
    timed on a sample and scaled) and the oracle on the host cores."""
    import tempfile
    lengths = amd.log_uniform_lengths(args.batch, seed=2024)
    u = amd.packed_coefficients(lengths, seed=77)
    # the tuner with a fresh cache file: timed choices on the first plan,
    # read back by the second (the plan the bench runs)
    cache = os.path.join(tempfile.mkdtemp(prefix="l2c_bench_"), "tune.txt")
    amd.tune_cache(cache)
    t0 = time.perf_counter()
    bp_cold = amd.BatchPlan(lengths, args.M, amd.L2C, tune=True)
    plan_s = time.perf_counter() - t0
    del bp_cold
    t0 = time.perf_counter()
    bp = amd.BatchPlan(lengths, args.M, amd.L2C, tune=True)
    plan_warm_s = time.perf_counter() - t0
    amd.tune_cache(None)
    din, dout = amd.DeviceBuffer(bp.total), amd.DeviceBuffer(bp.total)
    din.upload(u)
    bp.graph(True)
    for _ in range(max(3, args.warmup // 4)):
        bp.execute_device(din.ptr, dout.ptr, 0)
    bp.synchronize()
    steps = max(10, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(steps):
        bp.execute_device(din.ptr, dout.ptr, 0)
    bp.synchronize()
    el = time.perf_counter() - t0
    value = steps * args.batch / el
    g = dout.download()
    rng = np.random.default_rng(5)
    sample = sorted(rng.choice(args.batch, min(args.batch, 256), replace=False).tolist())
    err = 0.0
    for t in sample:
        n, o = int(lengths[t]), int(bp.offsets[t])
        ref = orc.FMMPlan(n, 64, args.M, 0).execute(u[o:o + n], 0)
        err = max(err, float(np.max(np.abs(g[o:o + n] - ref)) / np.max(np.abs(ref))))
    # sequential single-transform calls on the same device data
    seq_sample = sample[:128]
    plans = {int(lengths[t]): amd.FMMPlan(int(lengths[t]), 64, args.M, 0) for t in seq_sample}
    for t in seq_sample[:8]:
        plans[int(lengths[t])].execute_device(din.ptr + 8 * int(bp.offsets[t]),
                                              dout.ptr + 8 * int(bp.offsets[t]), 0)
    for p in plans.values():
        p.synchronize()
    t0 = time.perf_counter()
    for t in seq_sample:
        p = plans[int(lengths[t])]
        p.execute_device(din.ptr + 8 * int(bp.offsets[t]), dout.ptr + 8 * int(bp.offsets[t]), 0)
        p.synchronize()
    seq = len(seq_sample) / (time.perf_counter() - t0)
    cpu = None
    if not args.no_cpu:
        nt = orc.num_threads()
        reps, t0 = 0, time.perf_counter()
        oplans = {}
        while time.perf_counter() - t0 < args.cpu_seconds / 2:
            t = reps % args.batch
            n, o = int(lengths[t]), int(bp.offsets[t])
            op = oplans.get(n)
            if op is None:
                op = oplans[n] = orc.FMMPlan(n, 64, args.M, 0)
            op.execute(u[o:o + n], 0)
            reps += 1
        el_c = time.perf_counter() - t0
        cpu = {"value": round(reps / el_c, 1), "unit": "transforms/s", "cores": nt, "kind": "port",
               "sample": f"oracle single transforms over the first {reps} of the batch in "
                         f"{el_c:.1f} s (plans included), OpenMP {nt} threads"}
    # whole-batch roofline: every input and output once, plus each class
    # plan's tables once (FMM: the M2L matrices, 3 x total_blocks(L) x M^2;
    # direct: the two parity blocks of the dense product, n^2 / 2 doubles)
    classes = bp.classes()
    tab_bytes = 0
    for c in classes:
        if c["method"] == "fmm":
            tab_bytes += 8 * 3 * ((2 << c["L"]) - (c["L"] + 2)) * args.M * args.M
        else:
            tab_bytes += 8 * c["n_max"] * c["n_max"] // 2
    alg = 16 * int(bp.total) + tab_bytes
    t_batch = el / steps
    roof = {"bound": "hbm", "achieved": round(alg / t_batch / 1e9, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(alg / t_batch / 1e9 / HBM_PEAK_GBS, 4),
            "algorithmic_bytes": alg, "traffic": None,
            "note": "whole batch per graph replay; its kernels are bound by FP64 issue and "
                    "latency, not by these bytes"}
    out = {"value": round(value, 1), "unit": "transforms/s", "count": args.batch,
           "sum_n": int(bp.total), "ms_per_batch": round(1e3 * el / steps, 4),
           "roofline": roof,
           "graph": True, "plan_seconds": round(plan_s, 3),
           "plan_seconds_warm_tune_cache": round(plan_warm_s, 4),
           "max_rel_err_vs_oracle": err, "checked": len(sample),
           "sequential_calls_transforms_s": round(seq, 1),
           "vs_sequential": round(value / seq, 1), "classes": classes,
           "cpu_baseline": cpu}
    log(f"[bench] batch {args.batch}: {out}")
    return out


def bench_field(amd, args, steps=None):
    """A 256^3 FP64 field c_ijk = U(-1,1)/((i+1)(j+1)(k+1)) resident in HBM:
    L2C along axes 2, 1, 0 (65,536 strided lines of N = 256 per axis, one
    batched call each), then C2L back along 0, 1, 2; 64 vectors of N = 2^18
    (L2C) run at the same time on the batch plan's own streams.  One step is
    the six field passes plus the 64 vectors.  The six passes are also timed
    alone, and their dense line kernel with HIP events for its roofline.
    Returns the figures as a dict."""
    from oracle import orc
    n = args.field_n
    shape = (n, n, n)
    rng_u = amd.splitmix_uniform(n ** 3, 3)
    i, j, k = np.meshgrid(*[np.arange(n, dtype=np.float64)] * 3, indexing="ij")
    x = (rng_u.reshape(shape) / ((i + 1) * (j + 1) * (k + 1))).ravel()
    del i, j, k
    fp = amd.FMMPlan(n, 64, args.M, amd.BOTH)
    a, b = amd.DeviceBuffer(x.size), amd.DeviceBuffer(x.size)
    a.upload(x)
    # line maps (n0, ls0, ls1, es) of the three axes of a C-ordered n^3 array
    maps = {2: (n * n, n, 0, 1), 1: (n, 1, n * n, n), 0: (n * n, 1, 0, n * n)}
    nv, nvec = args.field_vectors, 1 << 18
    bp = amd.BatchPlan([nvec] * nv, args.M, amd.L2C, tune=False) if nv else None
    uv = amd.packed_coefficients([nvec] * nv, seed=4) if nv else None
    if nv:
        vin, vout = amd.DeviceBuffer(nv * nvec), amd.DeviceBuffer(nv * nvec)
        vin.upload(uv)

    def step():
        if nv:
            bp.execute_device(vin.ptr, vout.ptr, 0)
        src, dst = a, b
        for d, axes in ((0, (2, 1, 0)), (1, (0, 1, 2))):
            for ax in axes:
                n0, ls0, ls1, es = maps[ax]
                fp.execute_lines_device(src.ptr, dst.ptr, d, n * n, n0, ls0, ls1, es)
                src, dst = dst, src
        fp.synchronize()
        if nv:
            bp.synchronize()

    step()
    # correctness: the first L2C pass (axis 2) on sampled lines, the round trip,
    # two of the 2^18 vectors
    fp.execute_lines_device(a.ptr, b.ptr, 0, n * n, *maps[2])
    fp.synchronize()
    g = b.download().reshape(n * n, n)
    op = orc.FMMPlan(n, 64, args.M, orc.BOTH)
    xl = x.reshape(n * n, n)
    rows = np.random.default_rng(1).choice(n * n, 64, replace=False)
    err_lines = max(float(np.max(np.abs(g[r] - op.execute(xl[r], 0))) / np.max(np.abs(op.execute(xl[r], 0))))
                    for r in rows)
    step()  # a -> ... -> a after six passes
    rt = float(np.max(np.abs(a.download() - x)))
    a.upload(x)
    err_vec = 0.0
    if nv:
        gv = vout.download()
        ov = orc.FMMPlan(nvec, 64, args.M, 0)
        for t in (0, nv - 1):
            ref = ov.execute(uv[t * nvec:(t + 1) * nvec], 0)
            err_vec = max(err_vec, float(np.max(np.abs(gv[t * nvec:(t + 1) * nvec] - ref)) / np.max(np.abs(ref))))
    for _ in range(args.warmup // 10 + 2):
        step()
    steps = steps or max(5, args.steps // 20)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    el = (time.perf_counter() - t0) / steps
    # the six passes alone, then once more with HIP events around each pass's
    # kernel (the dense line product, one launch per pass)
    nv_saved, nv = nv, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    el_alone = (time.perf_counter() - t0) / steps
    fp.profile(1)
    for _ in range(max(2, steps // 2)):
        step()
    prof = fp.profile_read()
    fp.profile(False)
    nv = nv_saved
    kname, (kms, kcnt) = next(iter(prof.items()))
    pass_us = 1e3 * kms / kcnt
    # algorithmic work of one pass: every line read and written once; the
    # dense product per line and parity is the upper triangle of K_p at
    # 16 x 16 tile granularity (NPP = 128: 36 tiles of 16 x 16 x 2 flops x 16)
    npp = 128 if n == 256 else ((n + 1) // 2 + 15) // 16 * 16
    tiles = (npp // 16) * (npp // 16 + 1) // 2
    pass_bytes = 16 * n * n * n
    pass_flops = n * n * 2 * tiles * 256 * 2
    lines = 6 * n * n
    out = {"lines_per_s": round(lines / el, 1), "transforms_per_s": round((lines + nv) / el, 1),
           "ms_per_step": round(1e3 * el, 4), "steps": steps,
           "alone": {"lines_per_s": round(lines / el_alone, 1),
                     "ms_per_6_passes": round(1e3 * el_alone, 4)},
           "dense_kernel": {"kernel": "k_lines_dense", "avg_pass_us": round(pass_us, 2),
                            "algorithmic_bytes": pass_bytes,
                            "achieved_GBps": round(pass_bytes / (pass_us * 1e-6) / 1e9, 1),
                            "hbm_frac": round(pass_bytes / (pass_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                            "flops": pass_flops,
                            "achieved_TFLOPs": round(pass_flops / (pass_us * 1e-6) / 1e12, 2),
                            "fp64_mfma_frac": round(pass_flops / (pass_us * 1e-6) / 78.6e12, 4)},
           "config": {"workload": f"field{n}_x6_plus_{nv}x2^18", "field": list(shape),
                      "vectors": nv, "vector_n": nvec, "M": args.M},
           "max_rel_err": {"field_lines_vs_oracle": err_lines, "vectors_vs_oracle": err_vec},
           "round_trip_max_abs_err": rt,
           "tol": {"field_lines": 1e-14 * math.log2(n), "vectors": 1e-14 * 18, "round_trip": 1e-10}}
    out["ok"] = bool(err_lines <= 1e-14 * math.log2(n) and err_vec <= 1e-14 * 18 and rt < 1e-10)
    log(f"[bench] field: {out}")
    return out


FFTW_REDFT10_2_20 = {"value": round(1 / 6.263459e-03, 2), "unit": "transforms/s",
                     "source": "src/results/figure4.sh:137 (FFTW REDFT10, N=2^20, "
                               "6.263459e-03 s, one core)"}


def bench_dct(amd, orc, args, steps):
    """Length-N DCT-II (FFTW REDFT10, the reference's yardstick,
    src/bin/l2c.c:414-463 on fun[i] = i^2) of one vector resident in HBM,
    and the same transform batched over many short vectors.  Algorithmic HBM
    bytes: 16 per element (read and write once); the multi-pass form moves
    16 more per element per pass boundary through its workspace."""
    N = args.dct_n
    x = np.arange(N, dtype=np.float64) ** 2
    plan = amd.DCTPlan(N)
    din, dout = amd.DeviceBuffer(N), amd.DeviceBuffer(N)
    din.upload(x)
    for _ in range(max(5, steps // 10)):
        plan.execute_device(din.ptr, dout.ptr, 2)
    plan.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute_device(din.ptr, dout.ptr, 2)
    plan.synchronize()
    el = (time.perf_counter() - t0) / steps
    ref = orc.redft(x, 2)
    err = float(np.max(np.abs(dout.download() - ref)) / np.max(np.abs(ref)))
    logm = int(math.log2(N)) - 1
    passes = 1 if logm <= 11 else (2 if logm < 23 else 3)
    out = {"N": N, "kind": "REDFT10", "us": round(1e6 * el, 3),
           "value": round(1 / el, 1), "unit": "transforms/s",
           "vs_fftw_published": round(1 / el / FFTW_REDFT10_2_20["value"], 1) if N == 1 << 20 else None,
           "algorithmic_GBps": round(16 * N / el / 1e9, 1),
           "hbm_passes": passes, "traffic_bytes": 16 * N * passes,
           "max_rel_err_vs_oracle": err}
    # many short vectors in one call: N = 1024, 8192 of them (64 MB each way)
    n_b, c_b = 1024, 8192
    pb = amd.DCTPlan(n_b, c_b)
    xb = amd.splitmix_uniform(n_b * c_b, 11)
    bi, bo = amd.DeviceBuffer(n_b * c_b), amd.DeviceBuffer(n_b * c_b)
    bi.upload(xb)
    for _ in range(5):
        pb.execute_device(bi.ptr, bo.ptr, 2)
    pb.synchronize()
    sb = max(10, steps // 4)
    t0 = time.perf_counter()
    for _ in range(sb):
        pb.execute_device(bi.ptr, bo.ptr, 2)
    pb.synchronize()
    eb = (time.perf_counter() - t0) / sb
    yb = bo.download().reshape(c_b, n_b)
    rows = [0, 1, c_b // 2, c_b - 1]
    errb = max(float(np.max(np.abs(yb[r] - orc.redft(xb.reshape(c_b, n_b)[r], 2)))
                     / np.max(np.abs(orc.redft(xb.reshape(c_b, n_b)[r], 2)))) for r in rows)
    out["batched"] = {"N": n_b, "count": c_b, "us": round(1e6 * eb, 2),
                      "transforms_s": round(c_b / eb, 1),
                      "algorithmic_GBps": round(16 * n_b * c_b / eb / 1e9, 1),
                      "max_rel_err_vs_oracle_rows": errb}
    # a 256^3 field: DCT-II along axes 2, 1, 0 (65,536 lines each, one launch per axis)
    n = 256
    fx = amd.splitmix_uniform(n ** 3, 12)
    pf = amd.DCTPlan(n)
    fa, fb = amd.DeviceBuffer(n ** 3), amd.DeviceBuffer(n ** 3)
    fa.upload(fx)
    maps = {2: (n * n, n * n, n, 0, 1), 1: (n * n, n, 1, n * n, n), 0: (n * n, n * n, 1, 0, n * n)}

    def field_step():
        src, dst = fa, fb
        for ax in (2, 1, 0):
            pf.execute_lines_device(src.ptr, dst.ptr, 2, *maps[ax])
            src, dst = dst, src

    # spot check of the axis-2 pass on two lines
    pf.execute_lines_device(fa.ptr, fb.ptr, 2, *maps[2])
    pf.synchronize()
    a2 = fb.download().reshape(n, n, n)
    errf = max(relerr_(a2[i, j], orc.redft(fx.reshape(n, n, n)[i, j], 2)) for i, j in ((0, 0), (n - 1, 7)))
    for _ in range(3):
        field_step()
    pf.synchronize()
    sf = max(5, steps // 20)
    t0 = time.perf_counter()
    for _ in range(sf):
        field_step()
    pf.synchronize()
    ef = (time.perf_counter() - t0) / sf
    out["field256"] = {"passes": 3, "ms_per_3_passes": round(1e3 * ef, 4),
                       "lines_s": round(3 * n * n / ef, 1),
                       "algorithmic_GBps": round(3 * 16 * n ** 3 / ef / 1e9, 1),
                       "max_rel_err_vs_oracle_lines": errf}
    out["tol"] = 1e-15 * math.log2(2 * N)
    log(f"[bench] dct: {out}")
    return out


def host_cpu():
    """The host CPU the baseline runs on: model name, physical cores and
    logical CPUs from /proc/cpuinfo, and the CPUs this process may use."""
    model, cores, logical = None, set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            phys = core = None
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"model": model, "physical_cores": len(cores) or None, "logical_cpus": logical or None,
            "cpus_available_to_process": avail}


def relerr_(x, ref):
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("fmm", "batch", "field", "dct"), default="fmm",
                    help="fmm: the N=2^20 transform and the uneven batch (default); batch: the "
                         "uneven batch alone; field: the 256^3 strided field with 2^18 vectors "
                         "mixed in")
    ap.add_argument("--dct-n", type=int, default=1 << 20,
                    help="length of the DCT-II of the dct workload and of the default line's dct entry")
    ap.add_argument("--dct-steps", type=int, default=200,
                    help="timed DCT-II calls in the default line (0 = skip)")
    ap.add_argument("--field-n", type=int, default=256)
    ap.add_argument("--field-steps", type=int, default=5,
                    help="timed steps of the 256^3 field (with its 2^18 vectors) in the default "
                         "line (0 = skip)")
    ap.add_argument("--field-vectors", type=int, default=64)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--N", type=int, default=1 << 20)
    ap.add_argument("--maxs", type=int, default=64)
    ap.add_argument("--M", type=int, default=18)
    ap.add_argument("--direction", type=int, default=0, choices=(0, 1))
    ap.add_argument("--decay", type=float, default=0.0,
                    help="input family U(-1,1)*(n+1)^-decay (0: plain U(-1,1))")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="bounded CPU-baseline sample (seconds of oracle work)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--check-rows", type=int, default=4352,
                    help="rows checked against the double-double direct sum")
    ap.add_argument("--per-transform", type=int, default=200,
                    help="transforms timed one by one with HIP events for the median and min "
                         "device time (0 = skip)")
    ap.add_argument("--prof-every", type=int, default=1,
                    help="time the kernels with HIP events in a separate pass after the timed "
                         "steps (0 = skip)")
    ap.add_argument("--settle", type=float, default=0.25,
                    help="seconds of untimed transforms before the warmup steps")
    ap.add_argument("--graph", action="store_true",
                    help="replay a HIP graph of each step's launches (measured slower than "
                         "launching them directly)")
    ap.add_argument("--compare-maxs", type=int, default=32,
                    help="also time this maxs (the published figure's leaf size); 0 = skip")
    ap.add_argument("--pcie-steps", type=int, default=20,
                    help="host-pointer execute() calls timed for the PCIe-inclusive rate; 0 = skip")
    ap.add_argument("--lines", type=int, default=8,
                    help="also time this many vectors of length N per step in one batched call; "
                         "0 or 1 = skip (this and --lines-more)")
    ap.add_argument("--lines-more", type=int, default=32,
                    help="and this many (reported as lines_per_step_more); 0 or 1 = skip")
    ap.add_argument("--cold-steps", type=int, default=40,
                    help="transforms timed after a 256 MiB device copy has evicted the "
                         "last-level cache (cold-cache kernel times); 0 = skip")
    ap.add_argument("--split-steps", type=int, default=200,
                    help="steps timed with precision form 1 (split M2L matrices); 0 = skip")
    ap.add_argument("--batch", type=int, default=4096,
                    help="transforms in the uneven batch (log-uniform n in [64, 8192]); 0 = skip")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run a plain copy of this many MiB first (--pmc calibration)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this "
                         "command into profiles/execute_kernels_<workload>_pmc.json and exit")
    args = ap.parse_args()

    if args.pmc_summary:
        wl = workload_name(args.N, args.maxs, args.M, args.direction)
        out = pmc_summary(args)
        out["workload"] = wl
        out["command"] = ("rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) -- python3 bench.py "
                          + " ".join(a for a in sys.argv[1:] if a not in args.pmc_summary
                                     and a != "--pmc-summary"))
        os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
        with open(pmc_path(wl), "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(out, indent=1))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 "
              "(no multi-GPU path is launched).", file=sys.stderr)
        sys.exit(2)

    amd = load_pkg()
    if args.workload == "field":
        f = bench_field(amd, args)
        line = {"metric": "field lines/s: 256^3 FP64 field, L2C along 3 axes then C2L back, "
                          "with 64 x 2^18 L2C transforms mixed in",
                "value": f["lines_per_s"], "unit": "lines/s", "n_gpus": 1, "steps": f["steps"],
                "ms_per_step": f["ms_per_step"], "higher_is_better": True, "dtype": "f64",
                "data": "synthetic (SplitMix64)", "config": f["config"],
                "roofline": {"bound": "hbm", "kernel": "k_lines_dense",
                             "achieved": f["dense_kernel"]["achieved_GBps"], "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": f["dense_kernel"]["hbm_frac"],
                             "traffic": None},
                "field": f}
        print(json.dumps(line), flush=True)
        if not f["ok"]:
            log("[bench] FAILED: field error above tolerance")
            sys.exit(1)
        return
    if args.workload == "dct":
        sys.path.insert(0, ROOT)
        from oracle import orc
        if args.copy_mb:
            log(f"[bench] calibration copy {args.copy_mb} MiB")
            copy_calibration(amd, args.copy_mb)
        dc = bench_dct(amd, orc, args, args.steps)
        cpu = None
        if not args.no_cpu:
            x = np.arange(args.dct_n, dtype=np.float64) ** 2
            reps, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < args.cpu_seconds:
                orc.redft(x, 2)
                reps += 1
            el_c = time.perf_counter() - t0
            cpu = {"value": round(reps / el_c, 3), "unit": "transforms/s", "cores": 1, "kind": "port",
                   "sample": f"oracle REDFT10 (long-double length-4N FFT), N={args.dct_n}, "
                             f"{reps} repeats in {el_c:.1f} s"}
            try:
                import scipy.fft as sfft
                reps, t0 = 0, time.perf_counter()
                while time.perf_counter() - t0 < args.cpu_seconds / 2:
                    sfft.dct(x, type=2, workers=1)
                    reps += 1
                cpu["scipy_pocketfft_one_core"] = round(reps / (time.perf_counter() - t0), 1)
            except ImportError:
                pass
        line = {"metric": f"DCT-II (FFTW REDFT10) transforms/s at N={args.dct_n}, fun[i] = i^2",
                "value": dc["value"], "unit": "transforms/s", "n_gpus": 1, "steps": args.steps,
                "warmup": max(5, args.steps // 10), "ms_per_step": round(dc["us"] / 1e3, 5),
                "higher_is_better": True, "scaling": "weak",
                "vs_baseline": dc["vs_fftw_published"], "baseline": FFTW_REDFT10_2_20,
                "dtype": "f64", "data": "fun[i] = i^2 as in the reference's test_dct0",
                "config": {"workload": f"dct2_N{args.dct_n}", "N": args.dct_n, "kind": "REDFT10",
                           "inputs_resident": "HBM"},
                "roofline": {"bound": "hbm", "achieved": dc["algorithmic_GBps"], "peak": 8000.0,
                             "unit": "GB/s", "frac": round(dc["algorithmic_GBps"] / 8000.0, 4),
                             "traffic": dc["traffic_bytes"]},
                "dct": dc, "cpu_baseline": cpu}
        print(json.dumps(line), flush=True)
        if not (dc["max_rel_err_vs_oracle"] <= dc["tol"]
                and dc["batched"]["max_rel_err_vs_oracle_rows"] <= 1e-15 * math.log2(2048)
                and dc["field256"]["max_rel_err_vs_oracle_lines"] <= 1e-15 * math.log2(512)):
            log("[bench] FAILED: DCT error above tolerance")
            sys.exit(1)
        return
    if args.workload == "batch":
        sys.path.insert(0, ROOT)
        from oracle import orc
        b = bench_batch(amd, orc, args)
        line = {"metric": "batched L2C transforms/s, 4096 transforms of log-uniform N in [64, 8192]",
                "value": b["value"], "unit": "transforms/s", "n_gpus": 1,
                "steps": max(10, args.steps // 4), "ms_per_step": b["ms_per_batch"],
                "higher_is_better": True, "dtype": "f64",
                "data": "synthetic (SplitMix64 lengths seed 2024, inputs seed 77, U(-1,1)/(k+1))",
                "config": {"workload": f"batch{args.batch}_loguniform_64_8192", "M": args.M},
                "roofline": b["roofline"], "batched": b}
        print(json.dumps(line), flush=True)
        if not b["max_rel_err_vs_oracle"] <= 1e-14 * math.log2(8192):
            sys.exit(1)
        return
    N, maxs, M, d = args.N, args.maxs, args.M, args.direction
    u = amd.coefficients(N, seed=1, decay=args.decay)

    log(f"[bench] plan N={N} maxs={maxs} M={M} direction={d}")
    t0 = time.perf_counter()
    plan = amd.FMMPlan(N, maxs, M, d)
    plan_s = time.perf_counter() - t0
    info = plan.info()
    din, dout = amd.DeviceBuffer(N), amd.DeviceBuffer(N)
    din.upload(u)

    if args.copy_mb:
        log(f"[bench] calibration copy {args.copy_mb} MiB")
        copy_calibration(amd, args.copy_mb)

    # --graph: steps replay a HIP graph of the launches (l2c_plan_graph)
    if args.graph:
        plan.graph(True)
    # bring the GPU to its steady clocks before the warmup steps: untimed
    # transforms for --settle seconds (a few warmup steps alone leave the
    # first timed steps slower than the rest)
    t0 = time.perf_counter()
    nsettle = 0
    while time.perf_counter() - t0 < args.settle:
        for _ in range(50):
            plan.execute_device(din.ptr, dout.ptr, d)
        plan.synchronize()
        nsettle += 50
    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        plan.execute_device(din.ptr, dout.ptr, d)
    plan.synchronize()

    log(f"[bench] timed {args.steps} steps")
    plan.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute_device(din.ptr, dout.ptr, d)
    plan.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    # kernel durations: a separate pass of the same steps in which every
    # kernel launch carries dispatch start/stop HIP events (kept out of the
    # timed loop, whose throughput the event handling would lower)
    prof_steps = max(20, args.steps // 4) if args.prof_every > 0 else 0
    kernels = {}
    if prof_steps:
        plan.profile(1)
        for _ in range(prof_steps):
            plan.execute_device(din.ptr, dout.ptr, d)
        plan.synchronize()
        prof = plan.profile_read()
        plan.profile(False)
        kernels = {k: {"avg_us": 1e3 * ms / max(c, 1), "launches": c} for k, (ms, c) in prof.items()}
    # per-transform device times (median and min over separately bracketed
    # transforms, outside the timed loop)
    per_tx = None
    if args.per_transform > 0:
        med, mn = per_transform_times(plan, din, dout, d, args.per_transform)
        per_tx = {"median_us": round(med, 3), "min_us": round(mn, 3), "samples": args.per_transform,
                  "timing": "HIP events before and after each l2c_execute_device call"}
    log(f"[bench] {value:.1f} transforms/s, {ms_per_step * 1e3:.2f} us/step; kernels {kernels}; "
        f"per transform {per_tx}")

    # correctness against the oracle
    g = dout.download()
    sys.path.insert(0, ROOT)
    from oracle import orc
    o = orc.FMMPlan(N, maxs, M, d).execute(u, d)
    err_fmm = float(np.max(np.abs(g - o)) / np.max(np.abs(o)))
    rng = np.random.default_rng(0)
    nrows = min(args.check_rows, N)
    rows = np.unique(np.concatenate([np.arange(min(256, nrows)),
                                     rng.integers(0, N, max(0, nrows - 256))]))
    ref = orc.direct_dd_rows(u, rows, d)
    scale = float(np.max(np.abs(o)))
    err_direct = float(np.max(np.abs(g[rows] - ref)) / scale)
    tol = 1e-14 * math.log2(N)
    log(f"[bench] rel err vs oracle FMM {err_fmm:.3e}, vs double-double direct (rows) {err_direct:.3e}")

    # roofline of the dominant (longest) kernel (none when the timing pass
    # is off, as in the counter runs)
    workload = workload_name(N, maxs, M, d)
    assert f"_s{info['s']}_" in workload
    roofline, dom, alg = None, None, None
    if kernels:
        dom = max(kernels, key=lambda k: kernels[k]["avg_us"])
        avg_us = kernels[dom]["avg_us"]
        alg = kernel_bytes(dom, info, N)
        achieved = alg / (avg_us * 1e-6) / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": pmc_traffic(workload, dom), "algorithmic_bytes": alg,
                    "avg_launch_us": round(avg_us, 3)}

    # cold cache: before each transform a 256 MiB device copy (512 MiB of
    # traffic, twice the 256 MB last-level cache) evicts the plan's tables,
    # the input and the previous output; HIP events time every kernel
    cold = None
    if args.cold_steps > 0:
        nf = 256 * 1024 * 1024 // 8
        fa, fb = amd.DeviceBuffer(nf), amd.DeviceBuffer(nf)
        plan.profile(1)
        for _ in range(args.cold_steps):
            if amd.lib().l2c_memcpy_d2d(fb.ptr, fa.ptr, nf * 8) != 0:
                raise RuntimeError(amd.last_error())
            plan.execute_device(din.ptr, dout.ptr, d)
            plan.synchronize()
        cprof = plan.profile_read()
        plan.profile(False)
        del fa, fb
        ck = {k: round(1e3 * ms / max(c, 1), 3) for k, (ms, c) in cprof.items()}
        csum = sum(ck.values())
        wsum = sum(v["avg_us"] for v in kernels.values())
        cold = {"kernels_us": ck, "sum_us": round(csum, 3), "warm_sum_us": round(wsum, 3),
                "transforms_s_from_kernel_sum": round(1e6 / csum, 1), "steps": args.cold_steps,
                "flush": "256 MiB device-to-device copy before each transform"}
        cd = cold["kernels_us"].get(dom) if dom else None
        if cd:
            cold["dominant_achieved_GBps"] = round(alg / (cd * 1e-6) / 1e9, 1)
        gc = dout.download()
        cold["max_rel_err_vs_oracle"] = float(np.max(np.abs(gc - o)) / np.max(np.abs(o)))
        log(f"[bench] cold cache: {cold}")

    # the same transform on B independent vectors per step, one batched
    # launch sequence (l2c_execute_lines_device): what the GPU sustains when
    # transforms of one size arrive together (reported beside the value)
    def time_lines(B):
        ub = np.concatenate([amd.coefficients(N, seed=1 + b, decay=args.decay) for b in range(B)])
        bi, bo = amd.DeviceBuffer(B * N), amd.DeviceBuffer(B * N)
        bi.upload(ub)
        for _ in range(max(2, args.warmup // 8)):
            plan.execute_lines_device(bi.ptr, bo.ptr, d, B, B, N, 0, 1)
        plan.synchronize()
        ls = max(5, args.steps // 8)
        t0 = time.perf_counter()
        for _ in range(ls):
            plan.execute_lines_device(bi.ptr, bo.ptr, d, B, B, N, 0, 1)
        plan.synchronize()
        el_l = time.perf_counter() - t0
        gb = bo.download()
        e_l = float(np.max(np.abs(gb[:N] - o)) / np.max(np.abs(o)))
        r = {"vectors_per_step": B, "value": round(B * ls / el_l, 2), "unit": "transforms/s",
             "ms_per_step": round(1e3 * el_l / ls, 4), "max_rel_err_vs_oracle_first": e_l}
        log(f"[bench] {B} vectors per step: {r}")
        return r

    lines = time_lines(args.lines) if args.lines > 1 else None
    lines_more = time_lines(args.lines_more) if args.lines > 1 and args.lines_more > 1 else None

    # precision form 1 (split M2L matrices) on the same plan and input,
    # reported beside the FP64 value
    split = None
    if args.split_steps > 0:
        plan.set_precision(1)
        for _ in range(args.warmup):
            plan.execute_device(din.ptr, dout.ptr, d)
        plan.profile(args.prof_every)
        plan.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.split_steps):
            plan.execute_device(din.ptr, dout.ptr, d)
        plan.synchronize()
        el_s = time.perf_counter() - t0
        prof_s = plan.profile_read()
        plan.profile(False)
        gs = dout.download()
        plan.set_precision(0)
        e_s = float(np.max(np.abs(gs - o)) / np.max(np.abs(o)))
        e_sd = float(np.max(np.abs(gs[rows] - ref)) / scale)
        split = {"form": "split M2L matrices (per-position FP64/FP32/BF16/dropped), FP64 arithmetic",
                 "value": round(args.split_steps / el_s, 2), "unit": "transforms/s",
                 "ms_per_step": round(1e3 * el_s / args.split_steps, 5),
                 "kernels_us": {k: round(1e3 * ms / max(c, 1), 3) for k, (ms, c) in prof_s.items()},
                 "max_rel_err": {"vs_oracle_fmm": e_s, "vs_dd_direct_rows": e_sd}}
        log(f"[bench] split form: {split}")
        if not (e_s <= tol and e_sd <= tol):
            log("[bench] split form error above tolerance")
            split["failed"] = True

    # precision form 2 (the M2L products below the top levels as BF16 slices
    # on MFMA), single transforms and 8 lines per step, beside form 0
    form2 = None
    if args.split_steps > 0:
        plan.set_precision(2)
        for _ in range(args.warmup):
            plan.execute_device(din.ptr, dout.ptr, d)
        plan.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.split_steps):
            plan.execute_device(din.ptr, dout.ptr, d)
        plan.synchronize()
        el2 = time.perf_counter() - t0
        plan.profile(1)
        for _ in range(20):
            plan.execute_device(din.ptr, dout.ptr, d)
        plan.synchronize()
        prof2 = plan.profile_read()
        plan.profile(False)
        g2 = dout.download()
        e2o = float(np.max(np.abs(g2 - o)) / np.max(np.abs(o)))
        e2d = float(np.max(np.abs(g2[rows] - ref)) / scale)
        form2 = {"form": "M2L below the top levels as 6 x 6 BF16 slices on MFMA, exact FP32 "
                         "anti-diagonal sums, FP64 recombination",
                 "value": round(args.split_steps / el2, 2), "unit": "transforms/s",
                 "ms_per_step": round(1e3 * el2 / args.split_steps, 5),
                 "vs_form0": round((args.split_steps / el2) / value, 3),
                 "kernels_us": {k: round(1e3 * ms / max(c, 1), 3) for k, (ms, c) in prof2.items()},
                 "max_rel_err": {"vs_oracle_fmm": e2o, "vs_dd_direct_rows": e2d, "tol": 1e-13}}
        if args.lines > 1:
            B = args.lines
            ub = np.concatenate([amd.coefficients(N, seed=1 + b, decay=args.decay) for b in range(B)])
            bi, bo = amd.DeviceBuffer(B * N), amd.DeviceBuffer(B * N)
            bi.upload(ub)
            for _ in range(3):
                plan.execute_lines_device(bi.ptr, bo.ptr, d, B, B, N, 0, 1)
            plan.synchronize()
            ls = max(5, args.steps // 8)
            t0 = time.perf_counter()
            for _ in range(ls):
                plan.execute_lines_device(bi.ptr, bo.ptr, d, B, B, N, 0, 1)
            plan.synchronize()
            el_l2 = time.perf_counter() - t0
            gb = bo.download()
            form2["lines"] = {"vectors_per_step": B, "value": round(B * ls / el_l2, 2),
                              "ms_per_step": round(1e3 * el_l2 / ls, 4),
                              "max_rel_err_vs_oracle_first": float(np.max(np.abs(gb[:N] - o)) / np.max(np.abs(o)))}
            del bi, bo
        plan.set_precision(0)
        log(f"[bench] form 2: {form2}")
        if not (e2o <= 1e-13 and e2d <= 1e-13):
            form2["failed"] = True

    # the drop-in execute() on host arrays: staged through pinned buffers,
    # PCIe included (reported beside the value, never as the value)
    pcie = None
    if args.pcie_steps > 0:
        out_h = np.zeros(N)
        plan.execute(u, d, out=out_h)
        t0 = time.perf_counter()
        for _ in range(args.pcie_steps):
            plan.execute(u, d, out=out_h)
        pcie = round(args.pcie_steps / (time.perf_counter() - t0), 2)
        log(f"[bench] host-pointer execute (PCIe inclusive): {pcie} transforms/s")

    # the same transform at the published figure's leaf size
    cmp_line = None
    if args.compare_maxs and args.compare_maxs != maxs:
        p2 = amd.FMMPlan(N, args.compare_maxs, M, d)
        for _ in range(args.warmup):
            p2.execute_device(din.ptr, dout.ptr, d)
        p2.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            p2.execute_device(din.ptr, dout.ptr, d)
        p2.synchronize()
        v2 = args.steps / (time.perf_counter() - t0)
        g2 = dout.download()
        o2 = orc.FMMPlan(N, args.compare_maxs, M, d).execute(u, d)
        e2 = float(np.max(np.abs(g2 - o2)) / np.max(np.abs(o2)))
        cmp_line = {"maxs": args.compare_maxs, "s": p2.info()["s"], "value": round(v2, 2),
                    "vs_published": round(v2 / PUBLISHED["value"], 1), "max_rel_err_vs_oracle": e2}
        if not args.no_cpu:
            # the oracle at the same leaf size on the host cores, all threads
            # and one core (the published figure is one core)
            op2 = orc.FMMPlan(N, args.compare_maxs, M, d)
            out2 = np.zeros(N)
            op2.execute(u, d, out=out2)
            cpu2 = {}
            nt = orc.num_threads()
            for label, thr, secs in (("all_threads", nt, args.cpu_seconds / 4),
                                     ("one_core", 1, args.cpu_seconds / 4)):
                orc.set_num_threads(thr)
                reps2, t0 = 0, time.perf_counter()
                while time.perf_counter() - t0 < secs:
                    op2.execute(u, d, out=out2)
                    reps2 += 1
                cpu2[label] = {"value": round(reps2 / (time.perf_counter() - t0), 3), "cores": thr,
                               "sample": f"{reps2} repeats in {time.perf_counter() - t0:.1f} s"}
            orc.set_num_threads(nt)
            cmp_line["cpu_baseline"] = dict(cpu2, kind="port", unit="transforms/s")
            del op2
        log(f"[bench] maxs={args.compare_maxs}: {cmp_line}")
        del p2

    batched = bench_batch(amd, orc, args) if args.batch > 0 else None
    dct = bench_dct(amd, orc, args, args.dct_steps) if args.dct_steps > 0 else None
    field = bench_field(amd, args, steps=5) if args.field_steps > 0 else None

    cpu = None
    if not args.no_cpu:
        nt = orc.num_threads()
        op = orc.FMMPlan(N, maxs, M, d)
        out = np.zeros(N)
        op.execute(u, d, out=out)
        # each call timed on its own as well, for the min and average of the
        # reference's own protocol (src/bin/l2c.c:181-212)
        reps, t0, per = 0, time.perf_counter(), []
        while True:
            tc = time.perf_counter()
            op.execute(u, d, out=out)
            per.append(time.perf_counter() - tc)
            reps += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        cpu_t = time.perf_counter() - t0
        cpu = {"value": round(reps / cpu_t, 3), "unit": "transforms/s", "cores": nt, "kind": "port",
               "sample": f"oracle FMM execute, N={N} L2C maxs={maxs} M={M}, {reps} repeats "
                         f"in {cpu_t:.1f} s, OpenMP {nt} threads, own loops (no BLAS)",
               "min_s": min(per), "avg_s": sum(per) / len(per), "host": host_cpu()}
        log(f"[bench] cpu baseline {cpu}")
        # the same restatement on one core (the reference's published figures
        # are single-core)
        orc.set_num_threads(1)
        reps1, t0, per1 = 0, time.perf_counter(), []
        while True:
            tc = time.perf_counter()
            op.execute(u, d, out=out)
            per1.append(time.perf_counter() - tc)
            reps1 += 1
            if time.perf_counter() - t0 > args.cpu_seconds / 2:
                break
        t1 = time.perf_counter() - t0
        orc.set_num_threads(nt)
        cpu["one_core"] = {"value": round(reps1 / t1, 3), "unit": "transforms/s", "cores": 1,
                           "sample": f"{reps1} repeats in {t1:.1f} s",
                           "min_s": min(per1), "avg_s": sum(per1) / len(per1)}
        log(f"[bench] cpu baseline, one core: {cpu['one_core']}")
        # the same restatement built at -Ofast, all threads
        r, n, el, mn, av = orc.ofast_rate(N, maxs, M, d, u, args.cpu_seconds / 2, threads=nt)
        cpu["ofast"] = {"value": round(r, 3), "unit": "transforms/s", "cores": nt,
                        "sample": f"{n} repeats in {el:.1f} s, gcc -Ofast", "min_s": mn, "avg_s": av}
        log(f"[bench] cpu baseline, -Ofast: {cpu['ofast']}")

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "transforms/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": round(value / PUBLISHED["value"], 1), "baseline": PUBLISHED, "dtype": "f64",
        "data": "synthetic (SplitMix64 U(-1,1) coefficients, seed 1"
                + (f", times (n+1)^-{args.decay:g})" if args.decay else ")"),
        "config": {"workload": workload, "N": N, "maxs": maxs, "s": info["s"], "L": info["L"],
                   "M": M, "direction": "L2C" if d == 0 else "C2L", "form": "plain-fp64",
                   "launch": "hip-graph replay" if args.graph else "direct",
                   "settle_s": args.settle,
                   "inputs_resident": "HBM", "plan_seconds": round(plan_s, 4)},
        "max_rel_err": {"vs_oracle_fmm": err_fmm, "vs_dd_direct_rows": err_direct,
                        "rows_checked": int(rows.size), "tol": tol},
        "kernels_us": {k: round(v["avg_us"], 3) for k, v in kernels.items()},
        "per_transform": per_tx,
        "roofline": roofline,
        "cold_cache": cold,
        "cpu_baseline": cpu,
        "host_pointer_transforms_s": pcie,
        "split_form": split,
        "form2_split_products": form2,
        "lines_per_step": lines,
        "lines_per_step_more": lines_more,
        "batched": batched,
        "dct2": dct,
        "field": field,
        "at_published_leaf_size": cmp_line,
    }
    print(json.dumps(line), flush=True)
    bad = not (err_fmm <= tol and err_direct <= tol)
    if cold is not None and not cold["max_rel_err_vs_oracle"] <= tol:
        bad = True
    if cmp_line is not None and not cmp_line["max_rel_err_vs_oracle"] <= tol:
        bad = True
    if split is not None and split.get("failed"):
        bad = True
    if form2 is not None and form2.get("failed"):
        bad = True
    for ln in (lines, lines_more):
        if ln is not None and not ln["max_rel_err_vs_oracle_first"] <= tol:
            bad = True
    if batched is not None and not batched["max_rel_err_vs_oracle"] <= 1e-14 * math.log2(8192):
        bad = True
    if field is not None and not field["ok"]:
        bad = True
    if dct is not None and not (dct["max_rel_err_vs_oracle"] <= dct["tol"]
                                and dct["batched"]["max_rel_err_vs_oracle_rows"] <= 1e-15 * math.log2(2048)
                                and dct["field256"]["max_rel_err_vs_oracle_lines"] <= 1e-15 * math.log2(512)):
        bad = True
    if bad:
        log(f"[bench] FAILED: error above {tol:.2e}")
        sys.exit(1)


if __name__ == "__main__":
    main()
