"""Benchmark: FP64 ray-surface intersections/s of the fused whole-optic trace
on one MI355X, with its HBM roofline and the reference CPU path beside it.

Workload (one step): the 12-interface TMA + 3-lens telescope
(batoid_amd.workloads.tma) traced over a 9,998,544-ray hexapolar pupil at
one (field, wavelength); successive steps cycle through 6 pre-generated
(field, wavelength) batches that are resident in HBM before timing starts.
Each step reads its batch and writes a separate output buffer, so inputs stay
pristine from step to step (the same per-ray traffic as the in-place
Optic.trace).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

METRIC = "ray-surface intersections/s (FP64, ≤1e-12 m) + Huygens PSF TFLOP/s vs plain"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0
# Bytes the fused trace moves per ray: reads x, y, z, vx, vy, vz, t,
# wavelength (8 B each) + 2 flag bytes; writes x..t + 2 flag bytes.
BYTES_PER_RAY = 8 * 8 + 2 + 7 * 8 + 2
# HBM bytes per ray of the same kernel on this workload, measured with
# rocprofv3 FETCH_SIZE and WRITE_SIZE (each scaled by its reading for a
# 1 GiB device copy: x2.000 and x1.000), from
# profiles/trace_plan_tma_1e7_counters.md.
MEASURED_TRAFFIC_PER_RAY = 124.06
# FP64 flops per ray of the same kernel on this workload: 64 lanes x
# (ADD + MUL + 2 FMA) FP64 wave instructions (SQ_INSTS_VALU_*_F64), same file
# (current build's figure; issued lanes, an upper bound on useful flops).
FP64_FLOP_PER_RAY = 6220.1
# The same two figures for the C3 kernel (perturbed TMA, reference start),
# from profiles/trace_plan_c3_perturbed_tma_1e7.md.
C3_FP64_FLOP_PER_RAY = 18232.3
C3_TRAFFIC_PER_RAY = 124.85
# Ray generation writes 9 FP64 columns and 2 flag bytes per ray.
RAYGEN_BYTES_PER_RAY = 9 * 8 + 2


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cpu():
    """The host CPU the baselines run on: threads used (every CPU this
    process may run on, what `nproc` reports), the model name, the logical
    CPU count and the cgroup CPU quota where one is set."""
    threads = len(os.sched_getaffinity(0))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()
            if q != "max":
                quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    return {"threads": threads, "model": model, "logical_cpus": os.cpu_count(),
            "cgroup_cpu_quota": quota, "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def median_of_runs(run, runs=5):
    """One warm-up call of run(), then `runs` timed calls; returns the median
    and all times in seconds."""
    run()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), times


def cpu_trace(optic, rv, what, run_s=2.0, runs=5):
    """The reference's CPU trace (its own C++ core, oracle/_ref, one call per
    Interface exactly as batoid's Optic.trace issues them; the oracle's C
    restatement where _ref is absent) with every host CPU, on a strided
    sample of `rv` sized so one run takes about run_s: the median of `runs`
    runs after a warm-up, in ray-surface intersections/s."""
    import oracle
    from oracle.host import HostRays, host_records, _HostTable
    from batoid_amd.plan import PlanSpec, collect_steps
    cpu = host_cpu()
    threads = cpu["threads"]
    use_ref = oracle.ref_available()
    lib = oracle.load_ref() if use_ref else oracle.load()
    spec = PlanSpec(collect_steps(optic), rv.coordSys)
    tabs = [host_records(i.surface) for i in spec.interfaces]
    obscs = [host_records(i.obscuration) if i.obscuration is not None else None
             for i in spec.interfaces]
    meds = {st.medium: _HostTable([spec.med_recs[st.medium]], [spec.med_arrs[st.medium]])
            for st in spec.steps if st.medium >= 0}
    full = HostRays.from_rv(rv)
    pre = "ref_" if use_ref else "or_"

    def trace(h):
        for st, tab, ob in zip(spec.steps, tabs, obscs):
            dr = (ctypes.c_double * 3)(*st.dr)
            drot = (ctypes.c_double * 9)(*st.drot)
            if st.action == 2:
                getattr(lib, pre + "refract")(tab.ptr(), dr, drot, meds[st.medium].ptr(),
                                              h.struct, 5, threads)
            else:
                name = "intersect" if st.action == 0 else "reflect"
                getattr(lib, pre + name)(tab.ptr(), dr, drot, h.struct, 5, threads)
            if ob is not None:
                getattr(lib, pre + "obscure")(ob.ptr(), ob.count, h.struct, threads)

    def make(n, reps=1):
        """A run: `reps` traces of a strided n-ray sample, each from fresh
        input columns (restored outside the timed calls); returns the time
        spent in the trace calls."""
        idx = np.linspace(0, len(full) - 1, n).astype(np.int64)
        base = {k: v[idx] for k, v in full.cols.items()}
        h = HostRays(**base)

        def run():
            total = 0.0
            for _ in range(reps):
                for k, v in base.items():
                    np.copyto(h.cols[k], v)
                t0 = time.perf_counter()
                trace(h)
                total += time.perf_counter() - t0
            return total
        return run

    def median(run, count):
        run()                                     # warm-up
        times = [run() for _ in range(count)]
        return float(np.median(times)), times

    # size one run to about run_s: a strided sample of the batch, or the
    # whole batch traced `reps` times over when it takes less than that
    n = min(len(full), 20000)
    t_small, _ = median(make(n), 1)
    n = int(min(len(full), max(n, n * run_s / max(t_small, 1e-6))))
    reps = 1
    if n == len(full):
        t_full, _ = median(make(n), 1)
        reps = max(1, int(round(run_s / max(t_full, 1e-6))))
    t, times = median(make(n, reps), runs)
    value = reps * n * len(spec.steps) / t
    log(f"cpu {what}: {value:.3e} ray-surfaces/s ({n} rays x {reps}, median of {runs}: "
        f"{t:.2f} s, {threads} threads)")
    return {"value": value, "unit": "ray-surface intersections/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": f"{what}: {n} of {len(full)} rays (strided)"
                      + (f", traced {reps} times per run" if reps > 1 else "")
                      + f", {len(spec.steps)} interfaces, median of {runs} runs after a warm-up "
                      f"({', '.join(f'{x:.2f}' for x in times)} s), "
                      f"{'reference C++ core (oracle/_ref)' if use_ref else 'C restatement (oracle)'}"
                      f" with {threads} OpenMP threads",
           "host": cpu}


def cpu_huygens(rays, point, prim, npix, time_, sub=16, runs=5):
    """The reference's huygensPSF image loop (analysis.py:302-303: one
    sumAmplitude per pixel, rayVector.cpp:169-222, OpenMP over rays) through
    its own C++ core, on a sub x sub block of the npix x npix lattice around
    the centre; the time scales with pixels x rays, so s per PSF =
    block time x (npix / sub)^2."""
    import oracle
    from oracle.host import HostRays
    cpu = host_cpu()
    threads = cpu["threads"]
    use_ref = oracle.ref_available()
    lib = oracle.load_ref() if use_ref else oracle.load()
    fn = lib.ref_sum_amplitude if use_ref else lib.or_sum_amplitude
    h = HostRays.from_rv(rays)
    out = np.empty(2)
    pts = [np.asarray(point[:2], float) + (i - sub // 2) * prim[0] + (j - sub // 2) * prim[1]
           for j in range(sub) for i in range(sub)]

    def run():
        for p in pts:
            fn(h.struct, float(p[0]), float(p[1]), 0.0, time_, 1, out.ctypes.data, threads)

    t, times = median_of_runs(run, runs=runs)
    per_psf = t * (npix / sub) ** 2
    nsum = int((~rays.vignetted & ~rays.failed).sum())
    log(f"cpu huygens: {per_psf:.1f} s per {npix}^2 PSF ({sub}^2 block, {threads} threads)")
    return {"value": per_psf, "unit": "s per PSF", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "tflops": 8.0 * npix * npix * nsum / per_psf / 1e12,
            "sample": f"{sub}x{sub} of the {npix}x{npix} pixels (one sumAmplitude per pixel "
                      f"over {nsum} summed rays, as the reference's huygensPSF loop), median "
                      f"of {runs} runs after a warm-up ({', '.join(f'{x:.2f}' for x in times)}"
                      f" s), scaled by ({npix}/{sub})^2",
            "host": cpu}


def bench_huygens(optic, reps=10, nx=512, cpu=False):
    """C4: the TMA pupil on an nx x nx grid (one field, 620 nm), traced, then
    a 256 x 256 Huygens PSF at 0.5 um pixels by the MFMA GEMM (FP32 and
    split-BF16) and by the FP64 direct sum.  TFLOP/s counts 8 flops per
    complex multiply-add: 8 * pixels * (unvignetted rays)."""
    import torch
    import batoid_amd as bt
    from batoid_amd import _lib, _device, workloads
    thx, thy = workloads.tma_fields()
    dirCos = bt.utils.fieldToDirCos(thx[0], thy[0])
    rays = bt.RayVector.asGrid(optic=optic, wavelength=620e-9, dirCos=dirCos, nx=nx)
    optic.trace(rays)
    nvalid = int((~rays.vignetted & ~rays.failed).sum())
    cr = (nx // 2) * nx + nx // 2
    point = rays.r[cr]
    time_ = float(rays.t[0])
    npix = 256
    prim = np.array([[0.5e-6, 0.0], [0.0, 0.5e-6]])
    lib = _lib.load()
    ws_bytes = int(lib.bt_huygens_workspace_bytes(len(rays), npix, npix))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=_device.device())
    out = torch.empty(2 * npix * npix, dtype=torch.float64, device=_device.device())
    args = (rays._rays, _lib.dvec(point[:2]), _lib.dvec(prim[0]), _lib.dvec(prim[1]),
            npix, npix, time_, 1)
    stream = _device.stream()

    def run(kind):
        if kind == "direct":
            return lib.bt_huygens_direct(*args, _device.ptr(out), _device.ptr(ws), ws_bytes, stream)
        prec = _lib.PSF_FP32 if kind == "fp32" else _lib.PSF_BF16X3
        return lib.bt_huygens_gemm(*args, prec, _device.ptr(out), _device.ptr(ws), ws_bytes,
                                   stream)

    res = {"rays_traced": len(rays), "rays_summed": nvalid, "pixels": npix * npix,
           "pixel_um": 0.5, "wavelength_nm": 620,
           "flop_per_psf": 8.0 * npix * npix * nvalid}
    images = {}
    for kind, n_rep in (("direct", 1), ("fp32", reps), ("bf16x3", reps)):
        _lib.check(run(kind), kind)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n_rep):
            run(kind)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / n_rep
        amp = out.cpu().numpy().view(np.complex128).reshape(npix, npix)
        images[kind] = np.abs(amp) ** 2
        res[kind] = {"ms_per_psf": ms, "tflops": res["flop_per_psf"] / (ms * 1e-3) / 1e12}
        log(f"huygens {kind}: {ms:.3f} ms/PSF, {res[kind]['tflops']:.2f} TFLOP/s")
    ref = images["direct"]
    for kind in ("fp32", "bf16x3"):
        res[kind]["rel_l2_vs_fp64"] = float(np.sqrt(np.sum((images[kind] - ref) ** 2) /
                                                     np.sum(ref ** 2)))
    res["fp32"]["frac_of_fp32_mfma_peak"] = res["fp32"]["tflops"] / FP32_MFMA_PEAK_TFLOPS
    # split-BF16 issues three bf16 products per complex MAC's real products
    res["bf16x3"]["frac_of_bf16_mfma_peak"] = 3 * res["bf16x3"]["tflops"] / BF16_MFMA_PEAK_TFLOPS
    res["speedup_fp32_vs_fp64_direct"] = res["direct"]["ms_per_psf"] / res["fp32"]["ms_per_psf"]
    res["speedup_bf16x3_vs_fp64_direct"] = res["direct"]["ms_per_psf"] / res["bf16x3"]["ms_per_psf"]
    if cpu:
        res["cpu_baseline"] = cpu_huygens(rays, point, prim, npix, time_)
        res["speedup_fp32_vs_cpu_reference"] = (res["cpu_baseline"]["value"] * 1e3
                                                / res["fp32"]["ms_per_psf"])
    return res


def _out_structs(batches, dev):
    """One output buffer for a list of input batches of equal size; each
    step reads its batch and writes the buffer (wavelength and flux are
    shared with the input).  Returns the columns and (in, out) bt_rays."""
    import torch
    from batoid_amd import _lib
    n = len(batches[0])
    cols = {k: torch.empty(n, dtype=torch.float64, device=dev)
            for k in ("x", "y", "z", "vx", "vy", "vz", "t")}
    cols.update({k: torch.empty(n, dtype=torch.uint8, device=dev)
                 for k in ("vignetted", "failed")})
    structs = []
    for rv in batches:
        o = _lib.BtRays()
        for k, v in cols.items():
            setattr(o, k, v.data_ptr())
        o.wavelength = rv.torch("wavelength").data_ptr()
        o.flux = rv.torch("flux").data_ptr()
        o.n = n
        structs.append((rv._rays, o))
    return cols, structs


def _time_plan(plan, structs, steps, warmup):
    """Wall time of `steps` bt_trace_plan launches (cycling through the
    batches) and the mean launch time by HIP events on the launch stream."""
    import torch
    from batoid_amd import _lib, _device
    from batoid_amd.global_vars import _batoid_niter
    lib = _lib.load()
    stream = _device.stream()

    def step(i):
        rin, rout = structs[i % len(structs)]
        rc = lib.bt_trace_plan(plan.struct, rin, rout, _batoid_niter, stream)
        if rc:
            _lib.check(rc, "bt_trace_plan")

    for i in range(warmup):
        step(i)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record()
        step(i)
        ends[i].record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(starts, ends)]))
    return elapsed, kernel_ms


def _plan_kernel_name(plan):
    from batoid_amd import _lib
    f = plan.struct.flags
    mode = 2 if f & _lib.PLAN_MLP_START else (1 if f & _lib.PLAN_FAST_START else 0)
    kinds = ("kKindsAnalytic" if f & _lib.PLAN_ANALYTIC
             else "kKindsSumAPB" if f & _lib.PLAN_SUM_APB else "kKindsAll")
    return f"k_plan<{mode}, {kinds}> (bt_trace_plan)"


def bench_c1(steps=50, warmup=3, cpu=True):
    """C1: the Newtonian telescope (Paraboloid mirror + flat detector at its
    focus), 1e6 seeded rays from a uniform pupil at 0.1 deg, 500 nm: the
    fused trace (reflect + obscure, intersect) from an input batch into an
    output buffer.  HBM: reads x..t + 2 flags (58 B), writes the same
    (58 B): 116 B per ray (no refraction, so no wavelength)."""
    from batoid_amd import workloads, _device
    optic = workloads.newtonian()
    rv = workloads.newtonian_rays(optic)
    n = len(rv)
    plan = optic._plan(rv.coordSys, False)
    _, structs = _out_structs([rv], _device.device())
    elapsed, kernel_ms = _time_plan(plan, structs, steps, warmup)
    nstep = plan.struct.nstep
    gbs = 116.0 * n / (kernel_ms * 1e-3) / 1e9
    res = {"workload": "Newtonian: Paraboloid(R=10 m) mirror, 0.2-1.0 m clear annulus, flat "
                       "detector at z = 5 m; asPolar nrandom=1e6, seed 57721, 0.1 deg, 500 nm",
           "rays": n, "interfaces": nstep, "kernel": _plan_kernel_name(plan),
           "ms_per_trace": elapsed * 1e3 / steps, "kernel_ms": kernel_ms,
           "value": n * nstep * steps / elapsed, "unit": "ray-surface intersections/s",
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS, "bytes_per_ray": 116}}
    log(f"C1: {elapsed * 1e3 / steps:.3f} ms per 1e6-ray trace (kernel {kernel_ms:.3f} ms), "
        f"{res['value']:.3e} ray-surfaces/s")
    if cpu:
        res["cpu_baseline"] = cpu_trace(optic, rv, "C1 Newtonian")
        res["speedup_vs_cpu_reference"] = res["value"] / res["cpu_baseline"]["value"]
    return res


def bench_c3(batches, steps=20, warmup=2, cpu=True):
    """C3: the perturbed TMA (every mirror Sum(Asphere, Zernike j<=28,
    Bicubic 204^2)) over the same resident pupils as C2, reference Newton
    start (bit-identical to the reference), one (field, wavelength) batch
    per step into an output buffer.  HBM bytes as C2 (124 B per ray); the
    kernel is FP64 bound."""
    from batoid_amd import workloads, _device
    optic = workloads.perturbed_tma()
    n = len(batches[0])
    plan = optic._plan(batches[0].coordSys, False)
    _, structs = _out_structs(batches, _device.device())
    elapsed, kernel_ms = _time_plan(plan, structs, steps, warmup)
    nstep = plan.struct.nstep
    value = n * nstep * steps / elapsed
    gbs = BYTES_PER_RAY * n / (kernel_ms * 1e-3) / 1e9
    res = {"workload": "perturbed TMA: C2 with each mirror Sum(Asphere, Zernike j<=28 "
                       "N(0, 50 nm), Bicubic 204^2 10 nm rms); the C2 pupils and fields",
           "rays": n, "interfaces": nstep, "batches": len(batches),
           "ms_per_step": elapsed * 1e3 / steps, "kernel_ms": kernel_ms, "value": value,
           "unit": "ray-surface intersections/s", "start": "reference (bit-identical)"}
    roof = {"kernel": _plan_kernel_name(plan), "bound": "fp64_valu"}
    hbm = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
           "bytes_per_ray": BYTES_PER_RAY}
    if C3_FP64_FLOP_PER_RAY:
        tf = C3_FP64_FLOP_PER_RAY * n / (kernel_ms * 1e-3) / 1e12
        roof.update({"achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tf / FP64_PEAK_TFLOPS, "flop_per_ray": C3_FP64_FLOP_PER_RAY})
    if C3_TRAFFIC_PER_RAY:
        roof["traffic"] = C3_TRAFFIC_PER_RAY * n
        hbm["measured_bytes_per_ray"] = C3_TRAFFIC_PER_RAY
    roof["hbm"] = hbm
    res["roofline"] = roof
    log(f"C3: {elapsed * 1e3 / steps:.3f} ms per step (kernel {kernel_ms:.3f} ms), "
        f"{value:.3e} ray-surfaces/s")
    if cpu:
        res["cpu_baseline"] = cpu_trace(optic, batches[0], "C3 perturbed TMA, plain start")
        res["speedup_vs_cpu_reference"] = value / res["cpu_baseline"]["value"]
    return res


def bench_raygen(optic, reps=20, cpu=True):
    """Ray generation: the C2 pupil, asPolar(outer=4.18, inner=2.558,
    nrad=1120, naz=11080) at one field, written by one bt_generate_rays
    pass (hexapolar ring lookup, sin/cos, stop sag and frame change,
    finishParallel, constant columns).  HBM bound: 74 B written per ray.
    Beside it, the reference's host path (numpy hexapolar + cos/sin, then the
    serial sag / transform / finishParallel loops, here the oracle's C
    restatement of them, one thread) on the same pupil."""
    import torch
    import batoid_amd as bt
    from batoid_amd import _lib, _device, workloads
    from batoid_amd.rayVector import RayVector
    thx, thy = workloads.tma_fields()
    wl = 630e-9
    dirCos = bt.utils.fieldToDirCos(thx[0], thy[0])
    # end to end through the public generator (host ring table + launch)
    e2e = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rv = workloads.tma_rays(optic, thx[0], thy[0], wl)
        torch.cuda.synchronize()
        e2e.append(time.perf_counter() - t0)
        if len(e2e) < 3:
            del rv
    n = len(rv)
    gen = _lib.BtRaygen()
    npts, tables = RayVector._polar_tables(gen, 4.18, 2.558, 1120, 11080)
    RayVector._generate(gen, npts, optic.stopSurface, optic.coordSys, wl, optic.inMedium,
                        optic.backDist, None, dirCos, 1.0, out=rv)
    lib = _lib.load()
    stream = _device.stream()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        lib.bt_generate_rays(ctypes.byref(gen), rv._rays, stream)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = RAYGEN_BYTES_PER_RAY * n
    gbs = nbytes / (ms * 1e-3) / 1e9
    res = {"workload": "C2 pupil asPolar(outer=4.18, inner=2.558, nrad=1120, naz=11080), "
                       "field 0, 630 nm",
           "rays": n, "kernel": "k_generate (bt_generate_rays)", "ms_per_generation": ms,
           "rays_per_s": n / (ms * 1e-3), "end_to_end_ms": min(e2e) * 1e3,
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS, "bytes_per_ray": RAYGEN_BYTES_PER_RAY}}
    log(f"raygen: {ms:.3f} ms per {n} rays ({gbs:.0f} GB/s, {gbs / HBM_PEAK_GBS:.1%} of HBM), "
        f"end to end {min(e2e) * 1e3:.1f} ms")
    # the same pass with the lattice pattern (no ring search, no sin/cos):
    # the write-stream rate the hexapolar pass is compared with
    ng = 3162
    grid = RayVector.asGrid(optic=optic, wavelength=wl, dirCos=dirCos, nx=ng,
                            lx=optic.pupilSize)
    ggen = _lib.BtRaygen()
    ggen.pattern = _lib.GEN_GRID
    ggen.nx = ggen.ny = ng
    ggen.fx = ggen.fy = 1.0 / (ng * ((ng - 2) / ng))
    ggen.lx[:] = [optic.pupilSize, 0.0]
    ggen.ly[:] = [0.0, optic.pupilSize]
    RayVector._generate(ggen, ng * ng, optic.stopSurface, optic.coordSys, wl, optic.inMedium,
                        optic.backDist, None, dirCos, 1.0, out=grid)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        lib.bt_generate_rays(ctypes.byref(ggen), grid._rays, stream)
    e1.record()
    torch.cuda.synchronize()
    gms = e0.elapsed_time(e1) / reps
    res["lattice_same_size"] = {"rays": ng * ng, "ms_per_generation": gms,
                                "GBps": RAYGEN_BYTES_PER_RAY * ng * ng / (gms * 1e-3) / 1e9}
    log(f"raygen lattice {ng}x{ng}: {gms:.3f} ms "
        f"({res['lattice_same_size']['GBps']:.0f} GB/s)")
    del grid
    if cpu:
        import oracle
        from oracle import raygen as G
        from oracle.host import host_records
        tab = host_records(optic.inMedium)
        w = np.array([wl])
        nidx = np.empty(1)
        oracle.load().or_get_n_n(tab.ptr(), w.ctypes.data, nidx.ctypes.data, 1, 1)
        ct = bt.CoordTransform(optic.stopSurface.coordSys, optic.coordSys)
        t0 = time.perf_counter()
        rr, th = G.hexapolar(4.18, 2.558, 1120, 11080)
        x, y = G.polar_xy(rr, th)
        G.finish(x, y, optic.stopSurface.surface, (ct.dr, ct.drot), dirCos, nidx[0],
                 optic.backDist, wl)
        t = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": n / t, "unit": "rays/s", "cores": 1, "kind": "port",
                               "sample": f"the whole {n}-ray pupil, {t:.2f} s: numpy as the "
                                         "reference + the C oracle's sag, transform and "
                                         "finishParallel loops"}
        log(f"raygen cpu: {t:.2f} s ({n / t:.3e} rays/s)")
    del tables
    return res


def bench_zernike(optic, nx=512, jmax=37, reps=20, cpu=True):
    """Zernike decomposition of the C2 wavefront: the TMA at field 0,
    620 nm, nx x nx pupil grid, Noll j <= jmax.  Times the whole
    batoid.zernike call (mean of 10) and the bt_zernike_gram pass alone
    (HIP events),
    and the reference's host fit (closed-form basis + numpy lstsq, as the
    oracle restates it) on the same rays."""
    import torch
    import batoid_amd as bt
    from batoid_amd import _lib, _device, workloads
    from batoid_amd.analysis import _sphere_rays, zernike_table
    thx, thy = workloads.tma_fields()
    wl = 620e-9
    bt.zernike(optic, thx[0], thy[0], wl, nx=nx, jmax=jmax)          # warm-up
    torch.cuda.synchronize()
    t0w = time.perf_counter()
    for _ in range(10):
        coefs = bt.zernike(optic, thx[0], thy[0], wl, nx=nx, jmax=jmax)
    torch.cuda.synchronize()
    e2e = (time.perf_counter() - t0w) / 10
    # the Gram pass alone, on the same rays
    dirCos = bt.utils.fieldToDirCos(thx[0], thy[0])
    rays = bt.RayVector.asGrid(optic=optic, wavelength=wl, nx=nx, dirCos=dirCos)
    ep = rays.toCoordSys(optic.stopSurface.coordSys)
    optic.stopSurface.surface.intersect(ep)
    srays, t0 = _sphere_rays(optic, thx[0], thy[0], wl, 'postel', nx, None, 'chief')
    C, nmax = zernike_table(jmax)
    coef = _device.f64(C)
    lib = _lib.load()
    ws = torch.empty(int(lib.bt_zernike_workspace_bytes(jmax)), dtype=torch.uint8,
                     device=_device.device())
    gram = torch.empty((jmax + 1) * (jmax + 2) // 2, dtype=torch.float64, device=ws.device)
    stream = _device.stream()
    args = (_device.ptr(ep.torch("x")), _device.ptr(ep.torch("y")), _device.ptr(srays.torch("t")),
            None, _device.ptr(srays.torch("vignetted")), nx * nx, float(t0), wl,
            optic.pupilSize / 2, _device.ptr(coef), jmax, nmax, _device.ptr(gram),
            _device.ptr(ws), ws.numel(), stream)
    _lib.check(lib.bt_zernike_gram(*args), "bt_zernike_gram")
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        lib.bt_zernike_gram(*args)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    M = C.shape[1]
    flop_per_ray = 2.0 * (jmax * M + (jmax + 1) * (jmax + 2) // 2)
    tflops = flop_per_ray * nx * nx / (ms * 1e-3) / 1e12
    res = {"workload": f"TMA field 0, 620 nm, asGrid nx={nx}, Noll j<={jmax}",
           "rays": nx * nx, "unvignetted": int((srays.torch("vignetted") == 0).sum().item()),
           "end_to_end_ms": e2e * 1e3, "gram_kernel_ms": ms,
           "fp64": {"kernel": "k_zernike_gram (FP64 MFMA)", "achieved": tflops,
                    "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": tflops / FP64_PEAK_TFLOPS, "flop_per_ray": flop_per_ray},
           "z4_z11_waves": [float(coefs[4]), float(coefs[11])]}
    log(f"zernike nx={nx} j<={jmax}: end to end {e2e * 1e3:.1f} ms, gram pass {ms:.3f} ms "
        f"({tflops:.2f} TFLOP/s FP64)")
    if cpu:
        from oracle import zfit
        x = ep.x
        y = ep.y
        w = (t0 - srays.t) / wl
        keep = ~srays.vignetted
        t1 = time.perf_counter()
        basis = zfit.zernike_basis(jmax, x[keep], y[keep], R_outer=optic.pupilSize / 2)
        np.linalg.lstsq(basis.T, w[keep], rcond=None)
        t = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": t * 1e3, "unit": "ms per fit", "cores": 1,
                               "kind": "port",
                               "sample": "the same rays: closed-form basis + numpy lstsq "
                                         "(the reference's galsim + lstsq step)"}
        log(f"zernike cpu fit: {t * 1e3:.1f} ms")
    return res


def bench_trace_split(optic_factory, nrad=354, naz=3503, reps=10, cpu=True):
    """Ghost splitting: the TMA with every refractive interface and the
    detector 2 % reflective (the reference's traceSplit test setup),
    traceSplit of a ~1e6-ray hexapolar pupil (minFlux 1e-4), and the
    bt_rsplit kernel alone on the 9,998,544-ray C2 pupil at L1's entrance
    (206 B of HBM traffic per ray: read 74, write 58 + 74)."""
    import torch
    import batoid_amd as bt
    from batoid_amd import _device, workloads
    optic = optic_factory()
    for item in optic.itemDict.values():
        if isinstance(item, bt.RefractiveInterface):
            item.forwardCoating = bt.SimpleCoating(0.02, 0.98)
            item.reverseCoating = bt.SimpleCoating(0.02, 0.98)
        if isinstance(item, bt.Detector):
            item.forwardCoating = bt.SimpleCoating(0.02, 0.98)
    thx, thy = workloads.tma_fields()
    rays = workloads.tma_rays(optic, thx[0], thy[0], 620e-9, nrad=nrad, naz=naz)
    optic.traceSplit(rays.copy(), minFlux=1e-4)               # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rF, rR = optic.traceSplit(rays.copy(), minFlux=1e-4)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    res = {"workload": f"TMA, 2 % reflective glass and detector, {len(rays)} rays, "
                       "minFlux 1e-4",
           "rays": len(rays), "ms_per_traceSplit": ms,
           "forward_paths": len(rF), "reverse_paths": len(rR),
           "rays_out": int(sum(len(r) for r in rF) + sum(len(r) for r in rR))}
    log(f"traceSplit {len(rays)} rays: {ms:.1f} ms, {len(rF)} forward / {len(rR)} reverse "
        f"paths")
    # the rSplit kernel alone, C2 pupil at L1's entrance
    big = workloads.tma_rays(optic, thx[0], thy[0], 620e-9)
    l1 = optic['L1_entrance']
    coat = l1.reverseCoating
    work = big.copy()
    bt.rSplit(l1.surface, work, l1.inMedium, l1.outMedium, coat, coordSys=l1.coordSys)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(reps):
        work = big.copy()
        torch.cuda.synchronize()
        e0.record()
        bt.rSplit(l1.surface, work, l1.inMedium, l1.outMedium, coat, coordSys=l1.coordSys)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    kms = float(np.median(times))
    gbs = 206.0 * len(big) / (kms * 1e-3) / 1e9
    res["rsplit_kernel"] = {"rays": len(big), "ms": kms,
                            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS,
                                         "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                                         "bytes_per_ray": 206}}
    log(f"rSplit kernel {len(big)} rays: {kms:.3f} ms ({gbs:.0f} GB/s incl. allocation)")
    if cpu:
        from oracle import split as OS
        from oracle.host import HostRays
        small = workloads.tma_rays(optic, thx[0], thy[0], 620e-9, nrad=112, naz=1108)
        h = OS.Rays(HostRays.from_rv(small), small.coordSys)
        t1 = time.perf_counter()
        OS.trace_split(optic, h, 1e-4)
        t = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": len(small) / t, "unit": "input rays/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{len(small)} rays through the oracle's traceSplit "
                                         f"restatement, {t:.2f} s"}
        log(f"traceSplit cpu: {len(small)} rays in {t:.2f} s")
    return res


def bench_newton_start(nrad, naz, optic_name="perturbed_tma", reps=5, train_steps=3000):
    """An optic with Newton mirrors traced with the reference start, the
    analytic base start ('conic': the mirror's conic root, or the ray origin
    for polynomial mirrors) and learned starts.  perturbed_tma: every mirror
    Sum(Asphere, Zernike j<=28, Bicubic 204^2); freeform_tma: the same
    mirrors as Sum(Zernike fit, Bicubic), no conic base.  The networks are
    trained first on 12 other fields (recorded with their wall time); each
    mode is then timed on one 1e7-ray batch and checked against the
    reference-start trace (max |dx|, flags)."""
    import torch
    from batoid_amd import workloads
    from batoid_amd.newton_start import train_newton_starts
    optic = getattr(workloads, optic_name)()
    thx, thy = workloads.tma_fields()
    # training set: 12 other fields x 838,554 hexapolar rays = 10,062,648
    # tracer-generated rays (the 1e7 of the learned-start configuration)
    t0 = time.perf_counter()
    train_rvs = [workloads.tma_rays(optic, thx[k], thy[k], 620e-9, nrad=325, naz=3205)
                 for k in range(6, 18)]
    train_rays = sum(len(r) for r in train_rvs)
    nets = train_newton_starts(optic, train_rvs, steps=train_steps, batch=8192)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del train_rvs
    log(f"newton-start networks trained in {train_s:.1f} s: "
        + ", ".join(f"{k} loss {v.history[-1][1]:.2e}" for k, v in nets.items()))
    rv0 = workloads.tma_rays(optic, thx[0], thy[0], 620e-9, nrad=nrad, naz=naz)
    n = len(rv0)
    nsteps = optic._plan(rv0.coordSys, False).struct.nstep
    res = {"workload": {"perturbed_tma": "perturbed TMA (Sum(Asphere, Zernike j<=28 "
                                          "N(0,50nm), Bicubic 204^2 10nm rms) per mirror)",
                        "freeform_tma": "freeform TMA (Sum(Zernike j<=45 fit of the asphere "
                                         "+ perturbation, Bicubic) per mirror)"}[optic_name]
                       + ", field 0, 620 nm",
           "rays": n, "interfaces": nsteps, "train_seconds": train_s,
           "train_rays": train_rays,
           "train_samples_per_surface": {k: int(v.samples) for k, v in nets.items()},
           "train_steps": train_steps}
    ref_cols = None
    for mode in ("reference", "conic", "mlp"):
        start = nets if mode == "mlp" else mode
        rv = rv0.copy()
        optic.trace(rv, start=start)            # plan build + warm-up
        torch.cuda.synchronize()
        cols = {k: rv.torch(k).clone() for k in ("x", "y", "z", "vx", "vy", "vz", "t",
                                                "vignetted", "failed")}
        work = [rv0.copy() for _ in range(reps)]
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for w in work:
            optic.trace(w, start=start)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        entry = {"ms_per_trace": ms, "ray_surfaces_per_s": n * nsteps / (ms * 1e-3)}
        if ref_cols is None:
            ref_cols = cols
        else:
            ok = ref_cols["failed"] == 0
            entry["max_abs_dpos_m"] = max(float((cols[k] - ref_cols[k])[ok].abs().max())
                                          for k in ("x", "y", "z"))
            entry["max_abs_ddir"] = max(float((cols[k] - ref_cols[k])[ok].abs().max())
                                        for k in ("vx", "vy", "vz"))
            entry["flag_mismatches"] = int((cols["failed"] != ref_cols["failed"]).sum() +
                                           (cols["vignetted"] != ref_cols["vignetted"]).sum())
            entry["speedup_vs_reference_start"] = res["reference"]["ms_per_trace"] / ms
        res[mode] = entry
        log(f"newton start {mode}: {ms:.3f} ms/trace, "
            f"{entry['ray_surfaces_per_s']:.3e} ray-surfaces/s"
            + (f", max |dpos| {entry['max_abs_dpos_m']:.2e} m, flags differ "
               f"{entry['flag_mismatches']}" if "flag_mismatches" in entry else ""))
        del work
    res["_nets"] = nets                         # reused by the batch leg, not reported
    return res


def bench_batch(reps=5, nets=None):
    """C5: the perturbed TMA, 64 fields (default_rng(1)) x 6 wavelengths =
    384 (field, wavelength) segments with N per segment log-uniform in
    [1e4, 2e5] (default_rng(2)) as an asGrid pupil of nx = round(sqrt(N)),
    each traced and turned into a 128 x 128 Huygens PSF at 0.5 um pixels
    (FP32 MFMA) -- the whole batch as one HIP graph replay.  The work-item
    size and persistent workgroup count are autotuned on the batch first."""
    import torch
    from batoid_amd import workloads
    from batoid_amd.batch import PSFBatch
    from batoid_amd.analysis import huygens_amplitudes, _chief_index
    optic = workloads.perturbed_tma()
    thx, thy = workloads.tma_fields(n=64, seed=1)
    rng = np.random.default_rng(2)
    nseg = 64 * len(workloads.WAVELENGTHS)
    nrays = np.exp(rng.uniform(np.log(1e4), np.log(2e5), nseg))
    nx = np.maximum(np.round(np.sqrt(nrays)).astype(int), 2)
    fx = np.repeat(thx, len(workloads.WAVELENGTHS))
    fy = np.repeat(thy, len(workloads.WAVELENGTHS))
    wl = np.tile(workloads.WAVELENGTHS, 64)
    t0 = time.perf_counter()
    b = PSFBatch(optic, fx, fy, wl, nx, nxOut=128, dx=0.5e-6, reference='chief',
                 precision='fp32')
    torch.cuda.synchronize()
    log(f"batch: {nseg} segments, {b.n} rays, set up in {time.perf_counter() - t0:.1f} s")
    timings = b.autotune(reps=2)
    best = min(timings, key=timings.get)
    log(f"batch autotune: best chunk {best[0]}, {best[1]} workgroups, {timings[best]:.2f} ms")
    eager_ms = b.time(reps)
    b.capture()
    graph_ms = b.time(reps)
    nvalid = int(b.valid_rays().sum())
    nsteps = b.plan.struct.nstep
    # the trace alone, for the split
    from batoid_amd.global_vars import _batoid_niter
    from batoid_amd import _lib, _device
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        _lib.call("bt_trace_plan", b.plan.struct, b.rays_in._rays, b.rays._rays,
                  _batoid_niter, _device.stream())
    e1.record()
    torch.cuda.synchronize()
    trace_ms = e0.elapsed_time(e1) / reps
    b.run()
    amp = b.amplitudes()
    errs = []
    for s in (0, nseg // 2, nseg - 1):
        rays = b.rays[int(b.bounds[s]):int(b.bounds[s + 1])]
        point = rays.r[_chief_index(int(nx[s]))]
        ref = huygens_amplitudes(rays, point, b.lattices[s], 128, 128, float(rays.t[0]),
                                 method='direct')
        errs.append(float(np.sqrt(np.sum(np.abs(amp[s] - ref) ** 2) / np.sum(np.abs(ref) ** 2))))
    flop = 8.0 * 128 * 128 * nvalid
    res = {"workload": "perturbed TMA, 64 fields x 6 wavelengths, uneven asGrid pupils "
                       "(N log-uniform in [1e4, 2e5]), 128x128 PSF at 0.5 um per segment",
           "segments": nseg, "rays": b.n, "rays_summed": nvalid, "interfaces": nsteps,
           "autotuned": {"chunk": best[0], "workgroups": best[1]},
           "ms_per_batch_graph": graph_ms, "ms_per_batch_eager": eager_ms,
           "trace_ms": trace_ms, "psf_ms": graph_ms - trace_ms,
           "rays_per_s": b.n / (graph_ms * 1e-3),
           "ray_surfaces_per_s": b.n * nsteps / (graph_ms * 1e-3),
           "psf_latency_ms": graph_ms / nseg,
           "psf_tflops": flop / ((graph_ms - trace_ms) * 1e-3) / 1e12,
           "rel_l2_vs_fp64_direct": errs}
    log(f"batch: {graph_ms:.2f} ms per batch (graph), {eager_ms:.2f} ms eager, trace "
        f"{trace_ms:.2f} ms, PSF {res['psf_tflops']:.1f} TFLOP/s, rel L2 {max(errs):.1e}")
    # the same batch with the conic Newton start (positions within 1e-12 m of
    # the reference start, flags identical)
    ref_x = b.rays.torch("x").clone()
    ref_fail = b.rays.torch("failed").clone()
    b.set_start("conic")
    b.capture()
    conic_ms = b.time(reps)
    torch.cuda.synchronize()
    ok = ref_fail == 0
    res["conic_start"] = {
        "ms_per_batch_graph": conic_ms,
        "rays_per_s": b.n / (conic_ms * 1e-3),
        "ray_surfaces_per_s": b.n * nsteps / (conic_ms * 1e-3),
        "max_abs_dx_m": float((b.rays.torch("x") - ref_x)[ok].abs().max()),
        "flag_mismatches": int((b.rays.torch("failed") != ref_fail).sum()),
    }
    log(f"batch, conic start: {conic_ms:.2f} ms per batch (graph)")
    if nets is not None:
        # the learned start of the perturbed TMA's networks (trained on 12
        # fields at 620 nm by the newton-start leg), over all 384 segments
        b.set_start(nets)
        b.capture()
        mlp_ms = b.time(reps)
        torch.cuda.synchronize()
        res["mlp_start"] = {
            "ms_per_batch_graph": mlp_ms,
            "rays_per_s": b.n / (mlp_ms * 1e-3),
            "ray_surfaces_per_s": b.n * nsteps / (mlp_ms * 1e-3),
            "max_abs_dx_m": float((b.rays.torch("x") - ref_x)[ok].abs().max()),
            "flag_mismatches": int((b.rays.torch("failed") != ref_fail).sum()),
        }
        log(f"batch, learned start: {mlp_ms:.2f} ms per batch (graph), max |dx| "
            f"{res['mlp_start']['max_abs_dx_m']:.1e} m, flags differ "
            f"{res['mlp_start']['flag_mismatches']}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batches", type=int, default=6)
    ap.add_argument("--nrad", type=int, default=1120)
    ap.add_argument("--naz", type=int, default=11080)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-huygens", action="store_true")
    ap.add_argument("--no-newton-start", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-raygen", action="store_true")
    ap.add_argument("--no-zernike", action="store_true")
    ap.add_argument("--no-split", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run a plain device copy of this many MiB (calibrates "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    import batoid_amd as bt
    from batoid_amd import _lib, _device, workloads
    from batoid_amd.global_vars import _batoid_niter

    dev = _device.device()
    optic = workloads.tma()
    thx, thy = workloads.tma_fields()
    batches = []
    for b in range(args.batches):
        wl = workloads.WAVELENGTHS[b % len(workloads.WAVELENGTHS)]
        rv = workloads.tma_rays(optic, thx[b], thy[b], wl, nrad=args.nrad, naz=args.naz)
        batches.append(rv)
        log(f"batch {b}: field ({np.rad2deg(thx[b]):.3f}, {np.rad2deg(thy[b]):.3f}) deg, "
            f"{wl*1e9:.0f} nm, {len(rv)} rays")
    n = len(batches[0])
    plan = optic._plan(batches[0].coordSys, False)
    nsteps = plan.struct.nstep
    # one output buffer; wavelength and flux are shared with each input
    out_cols = {k: torch.empty(n, dtype=torch.float64, device=dev)
                for k in ("x", "y", "z", "vx", "vy", "vz", "t")}
    out_cols.update({k: torch.empty(n, dtype=torch.uint8, device=dev)
                     for k in ("vignetted", "failed")})
    structs = []
    for rv in batches:
        o = _lib.BtRays()
        for k, v in out_cols.items():
            setattr(o, k, v.data_ptr())
        o.wavelength = rv.torch("wavelength").data_ptr()
        o.flux = rv.torch("flux").data_ptr()
        o.n = n
        structs.append((rv._rays, o))
    lib = _lib.load()
    stream = _device.stream()

    def step(i):
        rin, rout = structs[i % len(structs)]
        rc = lib.bt_trace_plan(plan.struct, rin, rout, _batoid_niter, stream)
        if rc:
            _lib.check(rc, "bt_trace_plan")

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")
    if args.calib_copy_mb:
        nb = args.calib_copy_mb << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {nb} bytes read, {nb} bytes written")
        del src, dst

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record()
        step(i)
        ends[i].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    value = n * nsteps * args.steps / elapsed
    log(f"timed {args.steps} steps: {elapsed*1e3/args.steps:.3f} ms/step, "
        f"kernel {kernel_ms:.3f} ms, {value:.3e} ray-surfaces/s")

    # sanity: the traced output is a real trace (finite positions for the
    # rays that reached the detector)
    vig = out_cols["vignetted"].cpu().numpy().astype(bool)
    frac_unvig = float(1.0 - vig.mean())

    # the same steps with the conic Newton start (tolerance exit instead of
    # the reference's fixed iteration count): positions within 1e-12 m,
    # flags identical -- reported beside the bit-exact value
    step(len(structs) * args.steps - 1)            # reference output of the last batch
    torch.cuda.synchronize()
    ref_last = {k: out_cols[k].clone() for k in ("x", "y", "z", "failed", "vignetted")}
    fplan = optic._plan(batches[0].coordSys, False, "conic")

    def fstep(i):
        rin, rout = structs[i % len(structs)]
        rc = lib.bt_trace_plan(fplan.struct, rin, rout, _batoid_niter, stream)
        if rc:
            _lib.check(rc, "bt_trace_plan")

    fstep(len(structs) * args.steps - 1)
    torch.cuda.synchronize()
    ok = ref_last["failed"] == 0
    fast_dpos = max(float((out_cols[k] - ref_last[k])[ok].abs().max()) for k in ("x", "y", "z"))
    fast_flags = int((out_cols["failed"] != ref_last["failed"]).sum()
                     + (out_cols["vignetted"] != ref_last["vignetted"]).sum())
    for i in range(args.warmup):
        fstep(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        fstep(i)
    torch.cuda.synchronize()
    fast_elapsed = time.perf_counter() - t0
    fast_value = n * nsteps * args.steps / fast_elapsed
    log(f"conic start: {fast_elapsed*1e3/args.steps:.3f} ms/step, {fast_value:.3e} "
        f"ray-surfaces/s, max |dpos| {fast_dpos:.1e} m, flags differ {fast_flags}")

    bytes_per_launch = BYTES_PER_RAY * n
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    fp64_tflops = FP64_FLOP_PER_RAY * n / (kernel_ms * 1e-3) / 1e12
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "ray-surface intersections/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "TMA + 3 fused-silica lenses + filter + detector (12 interfaces), "
                        "hexapolar pupil, one (field, wavelength) per step",
            "rays": n, "interfaces": nsteps, "batches": len(batches),
            "wavelengths_nm": [round(w * 1e9) for w in workloads.WAVELENGTHS[:len(batches)]],
            "fields": "first 6 of 189 uniform in a 1.75 deg disk",
            "niter": _batoid_niter, "unvignetted_fraction": round(frac_unvig, 4),
        },
        # the kernel keeps each ray in registers across all 12 interfaces:
        # 124 B of HBM per ray against ~7k FP64 flops, so FP64 VALU issue
        # binds it, not HBM (the HBM figures are beside it)
        "roofline": {
            "kernel": _plan_kernel_name(plan),
            "bound": "fp64_valu",
            "achieved": fp64_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": fp64_tflops / FP64_PEAK_TFLOPS,
            "flop_per_ray": FP64_FLOP_PER_RAY,
            "traffic": MEASURED_TRAFFIC_PER_RAY * n,
            "kernel_ms": kernel_ms,
            "hbm": {
                "achieved": achieved,
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "bytes_per_launch": bytes_per_launch,
                "measured_bytes_per_ray": MEASURED_TRAFFIC_PER_RAY,
            },
        },
    }
    result["conic_start"] = {
        "ms_per_step": fast_elapsed * 1e3 / args.steps, "value": fast_value,
        "max_abs_dpos_m": fast_dpos, "flag_mismatches": fast_flags,
        "note": "same workload, Newton surfaces seeded by their conic root and "
                "converged to a 1e-14 m sag gap (not bit-identical to the reference)"}
    if not args.no_c3:
        result["c3"] = bench_c3(batches, cpu=not args.no_cpu)
    if not args.no_c1:
        result["c1"] = bench_c1(cpu=not args.no_cpu)
    if not args.no_huygens:
        result["huygens"] = bench_huygens(optic, cpu=not args.no_cpu)
        # the same PSF from a ~2e5-ray pupil (asGrid nx=640 over the annulus)
        result["huygens_nx640"] = bench_huygens(optic, nx=640)
    if not args.no_raygen:
        result["raygen"] = bench_raygen(optic, cpu=not args.no_cpu)
    if not args.no_zernike:
        result["zernike"] = bench_zernike(optic, cpu=not args.no_cpu)
    if not args.no_split:
        result["trace_split"] = bench_trace_split(workloads.tma, cpu=not args.no_cpu)
    nets = None
    if not args.no_newton_start:
        result["newton_start"] = {
            name: bench_newton_start(args.nrad, args.naz, name)
            for name in ("perturbed_tma", "freeform_tma")}
        nets = result["newton_start"]["perturbed_tma"].pop("_nets")
        result["newton_start"]["freeform_tma"].pop("_nets")
    if not args.no_batch:
        result["batch"] = bench_batch(nets=nets)
    if not args.no_cpu:
        log("timing the CPU baseline ...")
        result["cpu_baseline"] = cpu_trace(optic, batches[0], "C2 TMA, first batch")
        result["speedup_vs_cpu_reference"] = value / result["cpu_baseline"]["value"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
