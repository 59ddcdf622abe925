"""Benchmark of the headline workload: a 512^3 Float32 RFFT + normalised
BRFFT pair on one MI355X.

A step is one pair, as in the reference's benchmark_rfft
(benchmarks/benchmarks.jl:236-245): mul!(v, plan, u) followed by
ldiv!(u, plan, v), inputs resident in HBM.  Prints one JSON line.

Beside the headline line's required fields, "extras" holds the other
workloads of the benchmark set, each timed the same way (HIP events around
K pairs on the launch stream):
  * 512^3 ComplexF32 FFT + BFFT pair, 256^3 Float64 RFFT pair;
  * (RFFT, FFT, R2R REDFT10) on a Float32 384x384x256 field;
  * the precision form: 1024^3 ComplexF32 FFT pair with plain FP32 radix
    stages, split-FP16 MFMA and BF16x3 MFMA radix-16 stages, each with its
    relative L2 error against the Float64 plan of the same input;
  * 12 uneven Float32 vector fields, loss + gradient in one hipGraph replay;
  * the 64^3 Float64 round trip (forward and backward timed apart).

The headline line also carries per-pair mean/std/min/max, the error of the
forward transform against the Float64 plan, a measured device-copy ceiling
and the compulsory (fused 3D) bytes.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X datasheet
METRIC = "512³ r2c+c2r FFT: ms/pair, % HBM peak; split-FP16 MFMA vs FP32 at 1e-6 rel"
ELT_BYTES = {0: 4, 1: 8, 2: 8, 3: 16}    # F32, F64, C32, C64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dims", type=int, default=512)
    ap.add_argument("--transform", default="RFFT", choices=["RFFT", "FFT", "REDFT10"],
                    help="RFFT (the headline), FFT (complex data) or REDFT10 (a Chebyshev "
                         "plan of real data, R2R(REDFT10) along every dimension)")
    ap.add_argument("--real", default="Float32", choices=["Float32", "Float64"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="headline workload only")
    ap.add_argument("--extra", action="append", default=None,
                    choices=["plain", "small_64_f64", "precision_1024", "batched_graph", "r2r",
                             "generic"],
                    help="run only the named extra workload(s)")
    ap.add_argument("--no-checks", action="store_true",
                    help="skip the Float64-plan error and the copy-ceiling measurement "
                         "(profiled runs: only the pair's own kernels)")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--flags", type=int, default=0)
    ap.add_argument("--export-wisdom", metavar="PATH",
                    help="write the tune data of this run to PATH (pfft_tune_export)")
    ap.add_argument("--import-wisdom", metavar="PATH",
                    help="plan from the tune data in PATH: nothing is timed at plan time "
                         "(profiled runs then trace only the pairs' own launches)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with two 1 GiB device copies: the known-size reference that "
                         "scales HBM byte counters (rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def host_workers():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(dims, transform, real, reps):
    """The oracle pipeline (per-stage permuted copy + 1D pocketfft transform,
    a restatement of the reference's execution layer) timed on the host
    cores, in the workload's precision, on `reps` full pairs."""
    import numpy as np

    from oracle import pipeline as OP
    from oracle import plan as OM

    workers = host_workers()
    meta = OM.PlanMeta(dims, (transform,), "Float64")
    rng = np.random.default_rng(42)
    if transform == "RFFT":
        x = rng.uniform(-1, 1, size=dims).astype(np.float32 if real == "Float32" else np.float64)
    else:
        x = (rng.standard_normal(dims) + 1j * rng.standard_normal(dims)).astype(
            np.complex64 if real == "Float32" else np.complex128)
    xm = np.asfortranarray(x)
    single = real == "Float32"
    # warm-up on a small case to load pocketfft plans
    OP.forward(OM.PlanMeta((16, 16, 16), (transform,)), np.zeros((16, 16, 16)))

    def median_ms(fn, n, warm=True):
        if warm:
            fn()                  # warm-up at the full size
        ts = []
        for _ in range(n):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        ts.sort()
        return round(1e3 * ts[len(ts) // 2], 3)

    def pipeline(w):
        v = OP.forward(meta, xm, method="fast", workers=w, single=single)
        OP.backward(meta, v, scaled=True, method="fast", workers=w, single=single)

    # the same pair as one fused n-dimensional scipy.fft call per direction
    # (no pencil chain): a second CPU reference point
    import scipy.fft as sfft

    def fused(w):
        if transform == "RFFT":
            sfft.irfftn(sfft.rfftn(x, workers=w), s=dims, workers=w)
        else:
            sfft.ifftn(sfft.fftn(x, workers=w), workers=w)

    ms = median_ms(lambda: pipeline(workers), reps)
    ms_fused = median_ms(lambda: fused(workers), reps)
    # one thread (the pocketfft plans are warm by now)
    ms1 = median_ms(lambda: pipeline(1), reps, warm=False)
    ms1_fused = median_ms(lambda: fused(1), reps, warm=False)
    # the reference's CPU-runnable case, forward and backward timed apart
    m64 = OM.PlanMeta((64, 64, 64), ("RFFT",), "Float64")
    x64 = np.asfortranarray(np.random.default_rng(42).uniform(0.0, 1.0, size=(64, 64, 64)))
    tf, tb = [], []
    for _ in range(20):
        t0 = time.perf_counter()
        v64 = OP.forward(m64, x64, method="fast", workers=1)
        t1 = time.perf_counter()
        OP.backward(m64, v64, scaled=True, method="fast", workers=1)
        tb.append(time.perf_counter() - t1)
        tf.append(t1 - t0)
    cpu = "unknown"
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if info:
                        break          # the first processor's block is enough
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping", "cpu MHz",
                         "cache size", "cpu cores", "siblings"):
                    info[k] = v
        cpu = "%s (family %s, model %s, stepping %s; %s CPUs visible, %d used)" % (
            info.get("model name", "unknown"), info.get("cpu family", "?"),
            info.get("model", "?"), info.get("stepping", "?"), os.cpu_count(), workers)
    except OSError:
        pass
    return {"value": ms, "unit": "ms/pair", "cores": workers, "kind": "port",
            "sample": "full %s^3 %s %s pairs (oracle pipeline: permuted copy + scipy.fft "
                      "(pocketfft; FFTW is absent from the image) per stage, %s, "
                      "workers=%d), median of %d after a warm-up pair; host CPU: %s"
                      % (dims[0], real, transform, "float32" if single else "float64",
                         workers, reps, cpu),
            "cpu_model": cpu, "library": "scipy %s pocketfft" % _scipy_version(),
            "fused_nd_ms_per_pair": ms_fused,
            "workers_1": {"ms_per_pair": ms1, "fused_nd_ms_per_pair": ms1_fused, "cores": 1,
                          "sample": "the same pipeline and fused pair on one thread, "
                                    "median of %d" % reps},
            "small_64_f64": {"ms_forward": round(1e3 * sorted(tf)[len(tf) // 2], 4),
                             "ms_backward": round(1e3 * sorted(tb)[len(tb) // 2], 4),
                             "cores": 1,
                             "sample": "64^3 Float64 RFFT / ldiv!, median of 20"}}


def _scipy_version():
    import scipy
    return scipy.__version__


# ---------------------------------------------------------------- helpers

def pass_bytes(pf, plan):
    """Algorithmic HBM bytes of each forward pass (read its input pencil once,
    write its output pencil once), in launch order; the backward passes move
    the same bytes in reverse order."""
    T = pf.Transforms
    ext = math.prod(plan.extra_dims) if plan.extra_dims else 1
    et = plan.eltype_input
    out = []
    for n in range(1, plan.N + 1):
        _, di, do, _ = plan.stage(n)
        et_out = T.eltype_output(plan.transforms[n - 1], et)
        out.append(ext * (math.prod(di) * ELT_BYTES[et] + math.prod(do) * ELT_BYTES[et_out]))
        et = et_out
    return out


def fill(torch, a, gen, kind="uniform"):
    p = a.parent
    if p.is_complex():
        rt = torch.float32 if p.dtype == torch.complex64 else torch.float64
        p.copy_(torch.complex(torch.randn(p.shape, generator=gen, device=p.device, dtype=rt),
                              torch.randn(p.shape, generator=gen, device=p.device, dtype=rt)))
    elif kind == "uniform":
        p.uniform_(-1, 1, generator=gen)
    else:
        p.normal_(generator=gen)


def time_pairs(torch, pf, plan, u, v, steps, warmup):
    """ms per pair (mul! then ldiv!), HIP events on the launch stream."""
    for _ in range(warmup):
        pf.mul_(v, plan, u)
        pf.ldiv_(u, plan, v)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        pf.mul_(v, plan, u)
        pf.ldiv_(u, plan, v)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def pair_stats(torch, pf, plan, u, v, n):
    """mean / std / min / max ms of n pairs timed one by one (HIP events on
    the launch stream), as the reference's benchmark reports them."""
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(n + 1)]
    ev[0].record(stream)
    for i in range(n):
        pf.mul_(v, plan, u)
        pf.ldiv_(u, plan, v)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    t = [ev[i].elapsed_time(ev[i + 1]) for i in range(n)]
    mean = sum(t) / n
    std = math.sqrt(sum((x - mean) ** 2 for x in t) / max(1, n - 1))
    return {"n": n, "mean": round(mean, 4), "std": round(std, 4), "min": round(min(t), 4),
            "max": round(max(t), 4)}


def bandwidth_lib():
    """The streaming-copy kernel (pencilffts.jl_amd/csrc/bandwidth.hip)."""
    import ctypes as C
    path = os.path.join(ROOT, "pencilffts.jl_amd", "lib", "libpfft_bandwidth.so")
    lib = C.CDLL(path)
    lib.pfft_bandwidth_copy.restype = C.c_int
    lib.pfft_bandwidth_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_int,
                                        C.c_int, C.c_void_p]
    return lib


def stream_copy(torch, lib, dst, src, per_cu, unroll=8, nt_loads=0):
    st = torch.cuda.current_stream()
    r = lib.pfft_bandwidth_copy(dst.data_ptr(), src.data_ptr(), src.numel() * 4, per_cu,
                                unroll, nt_loads, st.cuda_stream)
    if r:
        raise RuntimeError("pfft_bandwidth_copy failed: hipError %d" % r)


def copy_ceiling(torch, nbytes=2 << 30, reps=10):
    """Measured HBM ceiling: read + write bytes per second of the
    hand-written streaming copy (16-byte loads, nontemporal stores) over a
    2 GiB buffer, the best over its grid shapes (resident workgroups per CU,
    or one per chunk), vectors in flight per lane and load hints; torch's
    device copy of the same buffer beside it."""
    lib = bandwidth_lib()
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return 2 * nbytes / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9

    sweep = {}
    for per_cu in (0, 1, 2, 4, 8):
        for unroll in (4, 8, 16):
            for ntl in (0, 1):
                key = "%s,unroll%d%s" % ("grid=chunks" if per_cu == 0 else "%d/CU" % per_cu,
                                         unroll, ",ntload" if ntl else "")
                sweep[key] = round(timed(lambda: stream_copy(torch, lib, dst, src, per_cu,
                                                             unroll, ntl)), 1)
    ok = bool(torch.equal(dst, src))
    torch_gbs = timed(lambda: dst.copy_(src))
    del src, dst
    torch.cuda.empty_cache()
    best = max(sweep, key=sweep.get)
    top = dict(sorted(sweep.items(), key=lambda kv: -kv[1])[:6])
    return {"GBs": sweep[best], "kernel": "k_stream_copy (16-byte loads, nontemporal stores; "
                                          "%s)" % best,
            "bytes": 2 * nbytes, "sweep_top_GBs": top, "sweep_min_GBs": min(sweep.values()),
            "copy_correct": ok, "torch_copy_GBs": round(torch_gbs, 1)}


def error_vs_f64(torch, pf, dims, tr, u32):
    """Relative L2 error of the Float32 forward transform against the Float64
    plan on the same (rounded) input."""
    T = pf.Transforms
    p32 = pf.PencilFFTPlan(dims, tr, (1, 1), None, T.Float32)
    v32 = pf.allocate_output(p32)
    pf.mul_(v32, p32, u32)
    p64 = pf.PencilFFTPlan(dims, tr, (1, 1), None, T.Float64)
    u64 = pf.allocate_input(p64)
    u64.parent.copy_(u32.parent)
    v64 = pf.allocate_output(p64)
    pf.mul_(v64, p64, u64)
    torch.cuda.synchronize()
    err = rel_l2_chunked(torch, v32.parent, v64.parent)
    del p32, p64, u64, v64, v32
    torch.cuda.empty_cache()
    return err


def pass_times(pf, plan, u, v, nrep):
    """Per-pass kernel ms of a forward and a backward (library HIP events)."""
    plan.set_timing(True)
    fwd = [0.0] * plan.N
    bwd = [0.0] * plan.N
    for _ in range(nrep):
        pf.mul_(v, plan, u)
        for i, t in enumerate(plan.stage_times()):
            fwd[i] += t / nrep
        pf.ldiv_(u, plan, v)
        for i, t in enumerate(plan.stage_times()):
            bwd[i] += t / nrep
    plan.set_timing(False)
    return fwd, bwd


def pair_entry(torch, pf, plan, u, v, steps, warmup, label):
    ms = time_pairs(torch, pf, plan, u, v, steps, warmup)
    pb = pass_bytes(pf, plan)
    pair = 2 * sum(pb)
    fwd, bwd = pass_times(pf, plan, u, v, 5)
    return {"workload": label, "ms_per_pair": round(ms, 4),
            "hbm_frac": round(pair / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "pair_algorithmic_bytes": pair,
            "passes_ms": {"fwd": [round(t, 4) for t in fwd], "bwd": [round(t, 4) for t in bwd]},
            "kernels": plan.kernels(0) + plan.kernels(2)}


def rel_l2_chunked(torch, a, b, chunks=16):
    """||a - b|| / ||b|| in float64 on the GPU, over the leading dim in chunks."""
    num = 0.0
    den = 0.0
    n = a.shape[0]
    step = max(1, (n + chunks - 1) // chunks)
    for i in range(0, n, step):
        x = a[i:i + step].to(b.dtype)
        y = b[i:i + step]
        num += float(torch.sum(torch.abs(x - y) ** 2))
        den += float(torch.sum(torch.abs(y) ** 2))
    return math.sqrt(num / den)


# ----------------------------------------------------------------- extras

def extra_plain(torch, pf, steps, warmup):
    T = pf.Transforms
    res = {}
    g = torch.Generator(device="cuda").manual_seed(7)
    for key, dims, tr, rt, label in (
            ("fft_512_c32", (512,) * 3, T.FFT(), T.Float32, "512^3 ComplexF32 FFT+BFFT pair"),
            ("rfft_256_f64", (256,) * 3, T.RFFT(), T.Float64, "256^3 Float64 RFFT+BRFFT pair"),
            ("mixed_384x384x256_f32", (384, 384, 256),
             (T.RFFT(), T.FFT(), T.R2R(T.REDFT10)), T.Float32,
             "384x384x256 Float32 (RFFT, FFT, REDFT10) pair"),
            ("rfft_1024_f64", (1024,) * 3, T.RFFT(), T.Float64,
             "1024^3 Float64 RFFT+BRFFT pair")):
        plan = pf.PencilFFTPlan(dims, tr, (1, 1), None, rt)
        u = pf.allocate_input(plan)
        v = pf.allocate_output(plan)
        if key.startswith("mixed"):
            # the mixed workload's data: 32 Fourier-Chebyshev modes + 1e-3 noise
            u.logical().copy_(pf.workloads.fourier_chebyshev_field(dims, n_modes=32, noise=1e-3,
                                                                   seed=7))
            label += ", 32 Fourier-Chebyshev modes + 1e-3 noise"
        else:
            fill(torch, u, g)
        res[key] = pair_entry(torch, pf, plan, u, v, steps, warmup, label)
        del plan, u, v
        torch.cuda.empty_cache()
    return res


def extra_r2r(torch, pf, steps, warmup):
    """R2R workloads beyond the DCT-II pair: a channel-flow layout with a
    Chebyshev (DCT-I, Gauss-Lobatto) wall-normal dimension, and the DCT-IV
    (self-inverse) on every dimension."""
    T = pf.Transforms
    res = {}
    g = torch.Generator(device="cuda").manual_seed(11)
    for key, dims, tr, label in (
            ("chan_512x512x513_f32", (512, 512, 513),
             (T.RFFT(), T.FFT(), T.R2R(T.REDFT00)),
             "512x512x513 Float32 (RFFT, FFT, REDFT00) pair"),
            ("redft11_512_f32", (512,) * 3, T.R2R(T.REDFT11),
             "512^3 Float32 REDFT11 x3 pair")):
        plan = pf.PencilFFTPlan(dims, tr, (1, 1), None, T.Float32)
        u = pf.allocate_input(plan)
        v = pf.allocate_output(plan)
        fill(torch, u, g)
        res[key] = pair_entry(torch, pf, plan, u, v, steps, warmup, label)
        del plan, u, v
        torch.cuda.empty_cache()
    return res


def extra_generic(torch, pf, steps, warmup):
    """Lengths without a specialised kernel: a prime line in the Bluestein
    form, and lines with radix-7/11/13 direct stages."""
    T = pf.Transforms
    res = {}
    g = torch.Generator(device="cuda").manual_seed(13)
    for key, dims, label in (
            ("bluestein_1031x128x128_c32", (1031, 128, 128),
             "1031x128x128 ComplexF32 FFT+BFFT pair (1031-point lines: Bluestein)"),
            ("primes_448x352x208_c32", (448, 352, 208),
             "448x352x208 ComplexF32 FFT+BFFT pair (radices 7, 11, 13)")):
        plan = pf.PencilFFTPlan(dims, T.FFT(), (1, 1), None, T.Float32)
        u = pf.allocate_input(plan)
        v = pf.allocate_output(plan)
        fill(torch, u, g)
        res[key] = pair_entry(torch, pf, plan, u, v, steps, warmup, label)
        del plan, u, v
        torch.cuda.empty_cache()
    return res


def extra_small_f64(torch, pf, steps, warmup):
    """The reference's CPU-runnable case: 64^3 Float64 field, uniform [0,1)
    with seed 42, RFFT then normalised BRFFT (ldiv!), forward and backward
    timed separately, and the round-trip error."""
    import numpy as np
    T = pf.Transforms
    dims = (64, 64, 64)
    plan = pf.PencilFFTPlan(dims, T.RFFT(), (1, 1), None, T.Float64)
    u = pf.allocate_input(plan)
    v = pf.allocate_output(plan)
    w = pf.allocate_input(plan)
    x = np.random.default_rng(42).uniform(0.0, 1.0, size=dims)
    u.logical().copy_(torch.from_numpy(x).to("cuda"))
    steps = max(steps, 100)
    for _ in range(warmup):
        pf.mul_(v, plan, u)
        pf.ldiv_(w, plan, v)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(stream)
    for _ in range(steps):
        pf.mul_(v, plan, u)
    ev[1].record(stream)
    for _ in range(steps):
        pf.ldiv_(w, plan, v)
    ev[2].record(stream)
    torch.cuda.synchronize()
    fwd = ev[0].elapsed_time(ev[1]) / steps
    bwd = ev[1].elapsed_time(ev[2]) / steps
    err = rel_l2_chunked(torch, w.logical().reshape(-1), u.logical().reshape(-1), 1)
    out = {"workload": "64^3 Float64 RFFT + ldiv! round trip, uniform[0,1) seed 42",
           "ms_forward": round(fwd, 5), "ms_backward": round(bwd, 5),
           "round_trip_rel_l2": err, "kernels": plan.kernels(0) + plan.kernels(2)}
    del plan, u, v, w
    torch.cuda.empty_cache()
    return out


def extra_precision(torch, pf, steps, warmup, edge=1024):
    """1024^3 ComplexF32 FFT pair: plain FP32 radix stages vs radix-16 MFMA
    stages (split-FP16 hi+lo, BF16x3), error vs the Float64 plan."""
    T = pf.Transforms
    dims = (edge,) * 3
    g = torch.Generator(device="cuda").manual_seed(11)
    p64 = pf.PencilFFTPlan(dims, T.FFT(), (1, 1), None, T.Float64)
    u64 = pf.allocate_input(p64)
    fill(torch, u64, g)
    # the FP32 input is the rounded FP64 input; the reference spectrum is the
    # FP64 transform of exactly that rounded input
    u32_data = u64.parent.to(torch.complex64)
    u64.parent.copy_(u32_data)
    v64 = pf.allocate_output(p64)
    pf.mul_(v64, p64, u64)
    torch.cuda.synchronize()
    del p64
    ref = v64.parent
    res = {"workload": "%d^3 ComplexF32 FFT+BFFT pair" % edge, "input": "re/im iid N(0,1)",
           "error_reference": "Float64 plan on the same (rounded) input"}
    v = u = None
    for name, flags in (("fp32", 0), ("split16", 4), ("bf16x3", 8)):
        plan = pf.PencilFFTPlan(dims, T.FFT(), (1, 1), None, T.Float32, flags=flags)
        if u is None:
            u = pf.allocate_input(plan)
            v = pf.allocate_output(plan)
        u.parent.copy_(u32_data)
        pf.mul_(v, plan, u)
        torch.cuda.synchronize()
        err_f = rel_l2_chunked(torch, v.parent, ref)
        pf.ldiv_(u, plan, v)
        torch.cuda.synchronize()
        err_rt = rel_l2_chunked(torch, u.parent, u64.parent)
        ent = pair_entry(torch, pf, plan, u, v, steps, warmup, name)
        ent.pop("workload")
        ent["rel_l2_forward_vs_f64"] = float("%.3e" % err_f)
        ent["rel_l2_roundtrip"] = float("%.3e" % err_rt)
        res[name] = ent
        del plan
        torch.cuda.empty_cache()
    f, s = res["fp32"]["ms_per_pair"], res["split16"]["ms_per_pair"]
    res["split16_speedup_vs_fp32"] = round(f / s, 3)
    del u, v, u64, v64, ref, u32_data
    torch.cuda.empty_cache()
    return res


def extra_batched(torch, pf, steps, warmup):
    """12 three-component Float32 fields of uneven cube edges; loss terms and
    gradient of all of them in one hipGraph replay.  The default path hands
    all groups to the library in one pfft_execute_batch call (groups on
    forked streams); the grouped path runs group after group on one stream."""
    from importlib import import_module
    B = import_module(pf.__name__ + ".batched")
    edges = B.draw_edges(12, seed=42)
    m = B.GradientEnergy(edges, 0)
    B.taylor_green_fields(m, seed=0)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    res = {}
    for batch in (True, False):
        m.batch = batch
        m.capture()
        res[batch] = (timed(m.replay), timed(m.run), float(m.loss()))
    nbytes = m.algorithmic_bytes()
    ms_graph, ms_eager, loss = res[True]
    out = {"workload": "12 Float32 3-component fields, edges %s: loss + gradient" % edges,
           "ms_per_graph_replay": round(ms_graph, 4), "ms_eager": round(ms_eager, 4),
           "plan_bytes_per_replay": nbytes,
           "plan_hbm_frac": round(nbytes / (ms_graph * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "loss": loss,
           "grouped_one_stream": {"ms_per_graph_replay": round(res[False][0], 4),
                                  "ms_eager": round(res[False][1], 4), "loss": res[False][2]}}
    del m
    torch.cuda.empty_cache()
    return out


def run_extras(torch, pf, steps, warmup, only=None):
    res = {}
    for key, fn in (("plain", extra_plain), ("small_64_f64", extra_small_f64),
                    ("precision_1024", extra_precision), ("batched_graph", extra_batched),
                    ("r2r", extra_r2r), ("generic", extra_generic)):
        if only and key not in only:
            continue
        try:
            r = fn(torch, pf, steps, warmup)
        except Exception as e:     # a failed extra is reported, the headline stands
            r = {"error": "%s: %s" % (type(e).__name__, e)}
        if key == "plain":
            res.update(r)
        else:
            res[key] = r
        print("extra %s done" % key, file=sys.stderr, flush=True)
    return res


# ------------------------------------------------------------------- main

def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch

    import __graft_entry__ as ge
    pf = ge.load_package()
    T = pf.Transforms
    dims = (args.dims,) * 3
    real_type = T.Float32 if args.real == "Float32" else T.Float64
    tr = {"RFFT": T.RFFT(), "FFT": T.FFT(), "REDFT10": T.R2R(T.REDFT10)}[args.transform]
    if args.import_wisdom:
        pf.import_wisdom(args.import_wisdom)
    plan = pf.PencilFFTPlan(dims, tr, (1, 1), None, real_type, flags=args.flags)
    u = pf.allocate_input(plan)
    v = pf.allocate_output(plan)
    g = torch.Generator(device="cuda").manual_seed(42)
    fill(torch, u, g)
    # tuning happens here (or nothing is timed, with imported tune data)
    t0 = time.perf_counter()
    plan.prepare()
    tune_s = time.perf_counter() - t0

    t0 = time.perf_counter()
    ms = time_pairs(torch, pf, plan, u, v, args.steps, args.warmup)
    wall = time.perf_counter() - t0

    # per-pass kernel durations from HIP events recorded by the library on the
    # launch stream (pfft_plan_set_timing), averaged over a few pairs
    fwd_ms, bwd_ms = pass_times(pf, plan, u, v, max(3, min(args.steps, 20)))
    stats = pair_stats(torch, pf, plan, u, v, max(5, min(args.steps, 100)))
    fwd_bytes = pass_bytes(pf, plan)
    bwd_bytes = list(reversed(fwd_bytes))
    X = dims[0]
    pair_bytes = sum(fwd_bytes) + sum(bwd_bytes)
    launches = [(fwd_ms[i], fwd_bytes[i], "fwd%d" % (i + 1)) for i in range(3)]
    launches += [(bwd_ms[i], bwd_bytes[i], "bwd%d" % (3 - i)) for i in range(3)]
    # the dominant kernel: all six passes run the same kernel family here, so
    # the roofline is quoted for the family over the pair (bytes / time summed)
    kern_ms = sum(t for t, _, _ in launches) / len(launches)
    kern_bytes = sum(b for _, b, _ in launches) / len(launches)
    achieved = kern_bytes / (kern_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s_%d.json" % (args.transform.lower(), X))
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    kernels = plan.kernels(0) + plan.kernels(2)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "pass kernels (6 launches per pair)",
                "bytes_per_launch": kern_bytes, "launch_ms": round(kern_ms, 4),
                "passes_ms": {name: round(t, 4) for t, _, name in launches},
                "passes_kernels": kernels}
    out = {
        "metric": METRIC,
        "value": round(ms, 4), "unit": "ms/pair", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if real_type == T.Float32 else "f64", "data": "synthetic",
        "config": {"workload": "%d^3 %s %s+%s pair (mul! + ldiv!)"
                               % (X, args.real, args.transform,
                                  {"RFFT": "BRFFT", "FFT": "BFFT",
                                   "REDFT10": "REDFT01"}[args.transform]),
                   "dims": list(dims), "transforms": args.transform, "real_type": args.real,
                   "input": "uniform[-1,1) seed 42" if args.transform == "RFFT"
                            else "re/im iid N(0,1) seed 42",
                   "parallelism": "single GPU"},
        "pair_hbm_frac": round(pair_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "pair_algorithmic_bytes": pair_bytes,
        "pair_ms_stats": stats,
        "pairs_per_s": round(1e3 / ms, 2),
        "wall_s": round(wall, 3),
        "plan_prepare_s": round(tune_s, 3),
        "roofline": roofline,
    }
    if not args.no_checks and not args.flags:
        # the same pair on the first candidate of every pass (no timing at
        # plan time: the choice of a plan made with FFTW.ESTIMATE and no
        # tune data)
        p0 = pf.PencilFFTPlan(dims, tr, (1, 1), None, real_type,
                              flags=pf._lib.FLAG_NO_AUTOTUNE)
        ms0 = time_pairs(torch, pf, p0, u, v, max(5, args.steps // 2), max(2, args.warmup // 2))
        out["untuned"] = {"ms_per_pair": round(ms0, 4),
                          "hbm_frac": round(pair_bytes / (ms0 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "kernels": p0.kernels(0) + p0.kernels(2)}
        del p0
    # compulsory bytes of a fused 3D transform (read the input, write the
    # output once per direction): the lower bound beside the 3-pass figure
    comp = 2 * (u.parent.numel() * u.parent.element_size() +
                v.parent.numel() * v.parent.element_size())
    out["compulsory_bytes"] = comp
    out["compulsory_hbm_frac"] = round(comp / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    if real_type == T.Float32 and not args.flags and not args.no_checks:
        u.parent.uniform_(-1, 1, generator=torch.Generator(device="cuda").manual_seed(42))
        out["rel_l2_forward_vs_f64_plan"] = float("%.3e" % error_vs_f64(torch, pf, dims, tr, u))
    del plan, u, v
    torch.cuda.empty_cache()
    if not args.no_checks:
        cc = copy_ceiling(torch)
        out["copy_ceiling"] = cc
        out["copy_ceiling_GBs"] = cc["GBs"]
        out["pair_frac_of_copy_ceiling"] = round(pair_bytes / (ms * 1e-3) / 1e9 / cc["GBs"], 4)
        out["roofline"]["frac_of_copy_ceiling"] = round(achieved / cc["GBs"], 4)
    if args.copy_calibration:
        # two 1 GiB streaming copies: the known-size launches that scale the
        # HBM byte counters of a rocprofv3 --pmc run
        lib = bandwidth_lib()
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            stream_copy(torch, lib, dst, src, 0, 8, 0)
        torch.cuda.synchronize()
        del src, dst
    if not args.no_extras:
        out["extras"] = run_extras(torch, pf, max(5, args.steps // 5), max(2, args.warmup // 5),
                                   args.extra)
    if args.export_wisdom:
        # the choices of every plan this run tuned (headline and extras)
        pf.export_wisdom(args.export_wisdom)
    if not args.no_cpu and args.transform in ("RFFT", "FFT"):
        out["cpu_baseline"] = cpu_baseline(dims, args.transform, args.real, args.cpu_reps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
