#!/usr/bin/env python3
"""Throughput of the Fourier-gridding backprojection (SPIDER `BP 32F`) on one GPU.

A step = one whole `BP 32F` job over HBM-resident synthetic projections:
reset the two half-set accumulator pairs, insert every projection into the
half the RANDLIST split gives it (pad + 2D FFT + KB scatter), finalise the
overall volume and the two sample volumes (SYMPLANE0 + normalise + 3D inverse
FFT + window, bp32f.f:275-410).  Default workload: 20,000 projections of
128x128 fp32 into the 256^3 padded grid (the recipe: 60-blob phantom, uniform
orientations, SNR 0.1); the same line carries a `geom_f64` record (the job
with FP64 sample geometry), a `host_images` record (the job with the images
handed over in host memory: the PCIe-inclusive rate, never `value`) and a
`config5` record (100,000 projections of 256x256 stored as BF16 into the
512^3 grid with compensated weights).

Prints one JSON line.  `roofline` is for the insertion kernel: algorithmic
bytes per projection B = 4 NX^2 + 24 U(NX) (image read + read/write of the
complex64 value and fp32 weight of each distinct voxel touched) over the
kernel's HIP-event-timed launch durations.  `cpu_baseline` times the CPU
oracle's restatement of the reference's OpenMP insertion loop on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "projections/s into 256³ volume (plain vs batched/BF16); rel-L2 vs FP64 oracle"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def distinct_voxels(bpmod, nx, ang, n=16):
    """U(NX): mean count of distinct half-grid voxels with nonzero weight from
    one projection, over n orientations."""
    import torch
    counts = []
    img = torch.ones((1, nx, nx), device="cuda")
    with bpmod.Backprojector(nx) as b:
        for i in range(n):
            b.reset()
            b.insert(img, ang[i:i + 1])
            _, W = b.export_raw(0)
            counts.append(int((W > 0).sum()))
    return sum(counts) / len(counts)


def cpu_baseline(imgs_host, ang, seconds):
    import numpy as np
    import oracle
    nx = imgs_host.shape[-1]
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    dms = oracle.rot_from_doc(ang, "f32")

    def run(m):
        X, W = oracle.empty_grid(nx)
        t = time.perf_counter()
        oracle.insert_omp_f32(imgs_host[:m], dms[:m], X, W, threads)
        return time.perf_counter() - t

    m0 = min(8, imgs_host.shape[0])
    t0 = run(m0)
    m = int(max(m0, min(imgs_host.shape[0], seconds / max(t0 / m0, 1e-6))))
    t = run(m)
    return {"value": m / t, "unit": "projections/s", "cores": threads, "kind": "port",
            "sample": f"{m} of the workload's projections ({nx}x{nx} fp32), insertion only "
                      f"(pad + FFT + ONELINE, OpenMP rows stride 6 as bp32f.f:873-888), "
                      f"{t:.1f} s on {threads} threads; oracle FFT is its own double-precision FFT"}


def measured_traffic(kernel, nx, per_launch, images_dtype="f32", twosum=False):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE measurement (profiles/traffic.json) for the same NX, image
    storage and weight compensation, scaled to this launch size; None when no
    measurement matches."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    for rec in t.get("records", []):
        if (rec.get("kernel") == kernel and rec.get("nx") == nx
                and rec.get("images_dtype", "f32") == images_dtype and bool(rec.get("twosum", False)) == twosum):
            return round(rec["bytes_per_unit"] * per_launch)
    return None


def measured_counters(kernel, nx, images_dtype="f32", twosum=False):
    """The committed rocprofv3 counter record (profiles/atomics.json) of
    `kernel` for the same NX, image storage and weight compensation: LDS
    wave-instructions and global atomic requests per projection; None when no
    measurement matches."""
    path = os.path.join(ROOT, "profiles", "atomics.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    for rec in t.get("records", []):
        if (rec.get("kernel") == kernel and rec.get("nx") == nx
                and rec.get("images_dtype", "f32") == images_dtype and bool(rec.get("twosum", False)) == twosum):
            return rec
    return None


def copy_rates():
    """Device-to-device copy rates, read + write bytes per second (GB/s): one
    2 GiB buffer copied 4 times (beyond the 256 MB last-level cache: HBM) and
    one 64 MiB buffer copied 64 times (cache-resident).  Timed with events on
    torch's current stream, the stream the copies run on."""
    import torch
    out = {}
    for name, mib, reps in (("hbm", 2048, 4), ("cache", 64, 64)):
        src = torch.empty(mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            dst.copy_(src)
        e1.record()
        e1.synchronize()
        out[name] = round(2.0 * (mib << 20) * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9, 1)
        del src, dst
    return out


def atomic_roofline(kname, nx, per_launch, avg_launch_s, U, images_dtype, twosum):
    """Atomic throughput of the insertion against measured peaks.  LDS: the
    kernel's LDS wave-instructions per launch (rocprofv3 SQ_INSTS_LDS, from
    the committed counter record) over the live launch time, against the
    library's calibration of 64-bit LDS atomic adds.  Global: the algorithmic
    atomic count per projection, 3 U (complex64 value and fp32 weight of each
    distinct voxel; 4 U with compensated weights), against non-returning fp32
    global atomics spread over 1 GiB; the kernel issues fewer (one flush per
    brick per work item), as the counter record's L2 atomic count shows."""
    from spider_amd import _lib
    import ctypes as C
    L = _lib.lib()
    peaks = {}
    for name, what, nbytes in (("lds", _lib.CAL_LDS_U64, 0), ("global", _lib.CAL_GLOBAL_F32, 1 << 30)):
        r = C.c_double()
        _lib.check(L.spbp_calibrate(what, nbytes, C.byref(r), None))
        peaks[name] = r.value
    rec = measured_counters(kname, nx, images_dtype, twosum)
    a_unit = (4.0 if twosum else 3.0) * U
    g_ach = a_unit * per_launch / avg_launch_s
    out = {"global": {"unit": "atomics/s", "algorithmic_per_unit": round(a_unit), "achieved": round(g_ach, -6),
                      "peak": round(peaks["global"], -6), "frac": round(g_ach / peaks["global"], 4),
                      "peak_source": "spbp_calibrate: fp32 non-returning atomics over 1 GiB",
                      "measured_l2_atomics_per_unit": None if rec is None else rec.get("l2_atomics_per_unit"),
                      "measured_l2_atomics_per_s": None if rec is None else round(
                          rec["l2_atomics_per_unit"] * per_launch / avg_launch_s, -6)},
           "lds": {"unit": "wave-instructions/s", "peak": round(peaks["lds"], -6),
                   "peak_source": "spbp_calibrate: ds_add_u64, 64 active lanes, no bank conflicts",
                   "achieved": None, "frac": None, "per_unit": None}}
    if rec is not None and rec.get("lds_instr_per_unit"):
        l_ach = rec["lds_instr_per_unit"] * per_launch / avg_launch_s
        out["lds"].update(achieved=round(l_ach, -6), frac=round(l_ach / peaks["lds"], 4),
                          per_unit=rec["lds_instr_per_unit"])
    return out


def parity_subset(bpmod, imgs, ang, m, half=None, **kw):
    """rel-L2 of the library's volumes on the first m projections against the
    FP64 and the FP32-faithful oracles: `BP 3F` (half None: the full volume)
    or `BP 32F` (half given: the worst of the full and the two sample volumes)."""
    import numpy as np
    import oracle
    sub = imgs[:m].contiguous()
    if half is None:
        vols = {"full": bpmod.bp3f(sub, ang[:m], **kw)}
        hp = None
    else:
        hp = np.asarray(half[:m], np.int8)
        f, h1, h2 = bpmod.bp32f(sub, ang[:m], half=hp, **kw)
        vols = {"full": f, "h1": h1, "h2": h2}
    vols = {k: v.cpu().numpy().astype(np.float64) for k, v in vols.items()}
    host = sub.float().cpu().numpy()   # BF16 images widen exactly
    o64 = oracle.bp(host, ang[:m], half=hp, prec="f64")
    o32 = oracle.bp(host, ang[:m], half=hp, prec="f32")
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    r64 = max(rel(v, o64[k]) for k, v in vols.items())
    r32 = max(rel(v, o32[k].astype(np.float64)) for k, v in vols.items())
    return r64, r32


def run_throughput(args, config, nproj, steps, warmup, op="bp32f", bf16=False, twosum=False, geom_f64=False,
                   parity_proj=100, cpu=True, calib=True, tune=False, extras=True, cpu_seconds=None):
    """One workload's throughput record.  A step is one whole job over the
    HBM-resident synthetic projections: reset the accumulators, insert every
    projection (pad + 2D FFT + lists + KB scatter; `BP 32F`: each into the
    half RANDLIST gives it), finalise (`BP 32F`: the overall volume and the
    two sample volumes)."""
    import numpy as np
    import torch
    from spider_amd import _lib, synth
    from spider_amd import bp as bpmod

    rec = synth.CONFIGS[config]
    nx = rec["nx"]
    nproj = nproj or rec["nproj"]
    log(f"bench: generating config {config}: {nproj} projections {nx}x{nx}")
    ang, ph, imgs = synth.make_dataset(config, nproj=nproj, device="cuda")
    if bf16:
        imgs = imgs.to(torch.bfloat16)
    halves = op == "bp32f"
    half = bpmod.random_halves(nproj, seed=synth.config_seed(config)) if halves else None
    torch.cuda.synchronize()

    if args.copy_calib:
        src = torch.empty(args.copy_calib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"bench: calibration copy of {args.copy_calib} MiB done")

    batch = args.batch if args.batch else (nproj if args.one_launch else 0)
    b = bpmod.Backprojector(nx, two_halves=halves, timing=True, batch=batch, strategy=args.strategy,
                            brick_proj=args.brick_proj, twosum=twosum, geom_f64=geom_f64)
    t_tune = None
    if args.shape is not None:
        b.set_plan("direct" if args.strategy == "direct" else "brick", args.shape, args.brick_proj)
    if tune:
        t0 = time.perf_counter()
        b.tune(imgs[:1024], ang[:1024])
        torch.cuda.synchronize()
        t_tune = time.perf_counter() - t0
    plan = b.plan()
    log(f"bench: plan {plan['strategy']} brick {plan['brick']} C={plan['brick_proj']}")

    def step():
        b.reset()
        b.insert(imgs, ang, half)
        return b.finalize(full=True, halves=halves)

    for w in range(warmup):
        step()
        torch.cuda.synchronize()
        log(f"bench: warmup step {w + 1} done")
    b.reset_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(f"bench: {steps} timed steps done")
    dt = t1 - t0
    value = nproj * steps / dt
    ins_ms, launches, units = b.kernel_stats(_lib.K_INSERT)
    pad_ms, _, _ = b.kernel_stats(_lib.K_PAD)
    fft_ms, _, _ = b.kernel_stats(_lib.K_FFT2)
    fin_ms, nfin, _ = b.kernel_stats(_lib.K_FINALIZE)
    bin_ms, _, _ = b.kernel_stats(_lib.K_BIN)
    res_ms, nres, _ = b.kernel_stats(_lib.K_RESOLVE)
    thresh = b.threshold_stats()
    b.close()
    opname = "BP 32F" if halves else "BP 3F"
    out = {"value": round(value, 2), "unit": "projections/s", "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": round(dt / steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16 images, f32 accumulation" if bf16 else "f32",
           "data": "synthetic (seeded blob phantom, analytic projections, white noise)",
           "config": {"workload": f"{opname} plain, {nx}x{nx} {'bf16' if bf16 else 'fp32'} images, {nproj} projections, "
                                  f"{2 * nx}^3 padded Fourier grid{', compensated weights' if twosum else ''}"
                                  f"{', FP64 sample geometry' if geom_f64 else ''}"
                                  f"{' (RANDLIST halves; overall + 2 sample volumes)' if halves else ''}"
                                  f" (config {config})",
                      "operation": opname, "nx": nx, "nproj": nproj, "grid": 2 * nx,
                      "images_dtype": "bf16" if bf16 else "f32", "twosum_weights": twosum, "geom_f64": bool(geom_f64),
                      "insert_launches_per_step": round(launches / steps, 2)},
           "breakdown_ms_per_step": {"pad": round(pad_ms / steps, 3), "fft2": round(fft_ms / steps, 3),
                                     "insert": round(ins_ms / steps, 3), "bin": round(bin_ms / steps, 3),
                                     "finalize": round(fin_ms / steps, 3),
                                     "nrmw2_resum": round(res_ms / steps, 3)},
           "nrmw2_threshold": thresh,
           "plan": {"strategy": plan["strategy"], "brick": list(plan["brick"]), "brick_proj": plan["brick_proj"],
                    "owner_computes": plan.get("owner_computes", False),
                    "tune_seconds": None if t_tune is None else round(t_tune, 3)}}
    if not extras:
        return out

    U = distinct_voxels(bpmod, nx, synth.uniform_angles(16, np.random.default_rng(5)))
    # bytes per projection: the image once, and the read + write of each
    # distinct voxel's complex64 value and fp32 weight (+ fp32 compensation)
    b_alg = (2.0 if bf16 else 4.0) * nx * nx + (32.0 if twosum else 24.0) * U
    avg_launch_s = ins_ms / 1e3 / launches
    per_launch = units / launches
    achieved = b_alg * per_launch / avg_launch_s / 1e9
    kname = ("insert_direct_kernel" if plan["strategy"] == "direct"
             else "insert_brick_kernel<%d,%d,%d%s>" % (*plan["brick"], ",own" if plan.get("owner_computes") else ""))
    idt = "bf16" if bf16 else "f32"
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(kname, nx, per_launch, idt, twosum),
            "kernel": kname, "bytes_per_unit": round(b_alg), "unit_of_work": "projection",
            "units_per_launch": round(per_launch, 1), "avg_launch_ms": round(avg_launch_s * 1e3, 4),
            "U_distinct_voxels": round(U, 1)}
    if calib:
        cr = copy_rates()
        roof.update(copy_GBps=cr, frac_of_copy_hbm=round(achieved / cr["hbm"], 5),
                    frac_of_copy_cache=round(achieved / cr["cache"], 5))
        if plan["strategy"] == "brick":
            roof["atomics"] = atomic_roofline(kname, nx, per_launch, avg_launch_s, U, idt, twosum)
    out["roofline"] = roof
    if parity_proj:
        log("bench: parity subset vs oracle")
        m = min(parity_proj, nproj)
        r64, r32 = parity_subset(bpmod, imgs, ang, m, half=half, twosum=twosum)
        out["rel_l2_vs_fp64"] = r64
        out["rel_l2_vs_ref_f32"] = r32
        g64, _ = parity_subset(bpmod, imgs, ang, m, half=half, twosum=twosum, geom_f64=True)
        out["rel_l2_vs_fp64_with_geom_f64"] = g64
        out["parity_subset_proj"] = m
    if cpu:
        log("bench: cpu baseline")
        m = min(nproj, 20000 if nx <= 128 else 2000)
        out["cpu_baseline"] = cpu_baseline(imgs[:m].float().cpu().numpy(), ang[:m], cpu_seconds or args.cpu_seconds)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    del imgs
    return out


def run_host_images(args, steps=3, warmup=1):
    """The config-2 BP 32F job with the images handed over in host memory
    (pageable, fp32) through spbp_insert_host: each step uploads them over
    PCIe, in sub-chunks that overlap the insertion.  Reported beside `value`,
    never as it."""
    import numpy as np
    import torch
    from spider_amd import synth
    from spider_amd import bp as bpmod

    nproj = args.nproj or synth.CONFIGS[2]["nproj"]
    ang, _, imgs = synth.make_dataset(2, nproj=nproj, device="cuda")
    host = np.ascontiguousarray(imgs.cpu().numpy())
    del imgs
    half = bpmod.random_halves(nproj, seed=synth.config_seed(2))
    nx = host.shape[-1]

    def timed(src):
        with bpmod.Backprojector(nx, two_halves=True) as b:
            def step():
                b.reset()
                b.insert_host(src, ang, half)
                return b.finalize(full=True, halves=True)
            for _ in range(warmup):
                step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                step()
            torch.cuda.synchronize()
            return time.perf_counter() - t0

    dt = timed(host)
    pinned = torch.from_numpy(host).pin_memory()   # a caller's page-locked buffer: asynchronous DMA
    dtp = timed(pinned)
    del pinned
    return {"value": round(nproj * steps / dt, 2), "unit": "projections/s", "ms_per_step": round(dt / steps * 1e3, 3),
            "steps": steps, "bytes_uploaded_per_step": int(host.nbytes),
            "pinned": {"value": round(nproj * steps / dtp, 2), "ms_per_step": round(dtp / steps * 1e3, 3)},
            "note": ("the config-2 BP 32F job with the images in host memory (spbp_insert_host), pageable "
                     "(`value`) and page-locked (`pinned`): PCIe-inclusive, uploads overlapped with insertion "
                     "in ~256 MiB sub-chunks; not the line's `value`")}


def run_batched(args):
    """Twelve BP 32F reconstructions with box sizes from {64, 96, 128, 192} and
    U[500, 30000] projections each (seeded blob phantoms, SNR 0.1).  Plain
    form: one context after another, fp32 images, default plan, insert +
    finalise.  Batched form: every context tuned on its own images, BF16
    image storage, all insertions captured once in one HIP graph (one
    persistent scatter launch per round of chunks) and replayed, then one
    batched finalisation of all 36 volumes (spbp_finalize_batch).  Both forms
    are timed over the same work; the FSC between their volumes and against
    the phantoms shows they agree.  The finalisations are also timed alone,
    one context after another and batched."""
    import numpy as np
    import torch
    from spider_amd import bp as bpmod
    from spider_amd import fsc as F
    from spider_amd import synth

    rng = np.random.Generator(np.random.PCG64(synth.config_seed(3)))
    nvol = args.volumes
    sizes = rng.choice([64, 96, 128, 192], nvol)
    counts = rng.integers(500, 30001, nvol)
    if args.nproj:
        counts = np.minimum(counts, args.nproj)
    vols = []
    for v in range(nvol):
        nx, m = int(sizes[v]), int(counts[v])
        r = np.random.Generator(np.random.PCG64(synth.config_seed(3, v + 1)))
        ph = synth.blob_phantom(nx, max(10, nx * 60 // 128), nx / 2 - 6, r)
        ang = synth.uniform_angles(m, r)
        imgs = synth.project(ph, synth.rotations(ang), device="cuda")
        synth.add_noise(imgs, 0.1, synth.config_seed(3, v + 1) + 7)
        vols.append(dict(nx=nx, m=m, ph=ph, ang=ang, imgs=imgs, imgs16=imgs.to(torch.bfloat16),
                         half=bpmod.random_halves(m, seed=v)))
        log(f"batched: volume {v}: {nx}^3 from {m} projections")
    total = int(sum(counts))
    torch.cuda.synchronize()

    def finalize_all(ctxs):
        return [c.finalize(full=True, halves=True) for c in ctxs]

    # plain: one context after another
    plain = [bpmod.Backprojector(d["nx"], two_halves=True) for d in vols]

    def run_plain():
        for c, d in zip(plain, vols):
            c.reset()
            c.insert(d["imgs"], d["ang"], d["half"])
        return finalize_all(plain)

    run_plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out_plain = run_plain()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / args.steps
    log(f"batched: plain form {t_plain:.3f} s per pass")

    # batched: tuned plans, BF16 storage, one captured graph
    fast = [bpmod.Backprojector(d["nx"], two_halves=True) for d in vols]
    t0 = time.perf_counter()
    plans = [c.tune(d["imgs16"], d["ang"]) for c, d in zip(fast, vols)]
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = bpmod.BatchGraph([(c, d["imgs16"], d["ang"], d["half"]) for c, d in zip(fast, vols)], reset_first=True)
    torch.cuda.synchronize()
    t_capture = time.perf_counter() - t0

    def run_fast():
        g.launch()
        return bpmod.finalize_batch(fast, full=True, halves=True)

    run_fast()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out_fast = run_fast()
    torch.cuda.synchronize()
    t_fast = (time.perf_counter() - t0) / args.steps
    log(f"batched: batched form {t_fast:.3f} s per pass")

    # the finalisations alone: one context after another vs one batched call
    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    t_fin_plain = timed(lambda: finalize_all(fast))
    t_fin_batch = timed(lambda: bpmod.finalize_batch(fast, full=True, halves=True))
    log(f"batched: finalisation {t_fin_plain:.4f} s one by one, {t_fin_batch:.4f} s batched")

    per = []
    for d, a, b in zip(vols, out_plain, out_fast):
        truth = torch.tensor(d["ph"].volume(), dtype=torch.float32, device="cuda")
        fsc_ab = F.fsc(a["full"], b["full"])
        per.append({"nx": d["nx"], "nproj": d["m"],
                    "fsc05_plain_vs_phantom": round(F.resolution(F.fsc(a["full"], truth)), 4),
                    "fsc05_batched_vs_phantom": round(F.resolution(F.fsc(b["full"], truth)), 4),
                    "min_fsc_plain_vs_batched": round(float(np.min(fsc_ab["fsc"][fsc_ab["voxels"] > 0])), 6)})
    g.close()
    for c in plain + fast:
        c.close()
    out = {"metric": "projections/s, 12 uneven BP 32F volumes: batched/BF16/graph vs plain",
           "value": round(total / t_fast, 1), "unit": "projections/s", "higher_is_better": True,
           "n_gpus": 1, "steps": args.steps, "dtype": "bf16 images, f32 accumulation",
           "data": "synthetic (seeded blob phantoms, analytic projections, SNR 0.1)",
           "config": {"workload": f"{nvol} BP 32F volumes, boxes {sorted(set(int(x) for x in sizes))}, "
                                  f"{total} projections in all (config 3)",
                      "sizes": [int(x) for x in sizes], "counts": [int(x) for x in counts]},
           "plain_proj_per_s": round(total / t_plain, 1), "speedup_vs_plain": round(t_plain / t_fast, 3),
           "plain_seconds": round(t_plain, 4), "batched_seconds": round(t_fast, 4),
           "tune_seconds": round(t_tune, 3), "capture_seconds": round(t_capture, 3),
           "finalize_seconds_one_by_one": round(t_fin_plain, 4), "finalize_seconds_batched": round(t_fin_batch, 4),
           "plans": [{"strategy": p["strategy"], "brick": p["brick"], "brick_proj": p["brick_proj"],
                      "best_brick_ms_per_proj": round(float(np.min(p["ms_per_proj"])), 6),
                      "direct_ms_per_proj": round(p["direct_ms_per_proj"], 6),
                      "gather_ms_per_proj": round(p["gather_ms_per_proj"], 6)} for p in plans],
           "volumes": per}
    print(json.dumps(out), flush=True)


def run_recon(args):
    """128^3 from 20,000 noisy projections: plain BP 32F against CGLS with the
    KB-gather projector (A) and its adjoint (the insertion, A^T), both scored
    by FSC against the phantom; also one autograd step (loss on a batch of
    256 projections + backward) and the gather kernel's own rate."""
    import numpy as np
    import torch
    from spider_amd import _lib, synth
    from spider_amd import bp as bpmod
    from spider_amd import fsc as F
    from spider_amd import project

    rec = synth.CONFIGS[4]
    nx = rec["nx"]
    nproj = args.nproj or rec["nproj"]
    log(f"recon: generating {nproj} projections {nx}x{nx}, SNR {rec['snr']}")
    ang, ph, imgs = synth.make_dataset(4, nproj=nproj, device="cuda")
    truth = torch.tensor(ph.volume(), dtype=torch.float32, device="cuda")
    half = bpmod.random_halves(nproj, seed=synth.config_seed(4))
    torch.cuda.synchronize()

    def bp32f():
        with bpmod.Backprojector(nx, two_halves=True) as b:
            b.insert(imgs, ang, half)
            v = b.finalize(full=True, halves=True)
        torch.cuda.synchronize()
        return v

    bp32f()   # warm-up (plans, allocations)
    t0 = time.perf_counter()
    vols = bp32f()
    t_bp = time.perf_counter() - t0
    c_bp = F.fsc(vols["full"], truth)
    res_bp = F.resolution(c_bp)
    res_halves = F.resolution(F.fsc(vols["h1"], vols["h2"]), 0.143)
    log(f"recon: BP 32F {t_bp:.3f} s, FSC(0.5) vs phantom at {res_bp:.4f}")

    P = project.Projector(nx, angles=ang, batch=args.proj_batch, timing=True)
    curve = []

    def cb(it, x):
        curve.append(F.resolution(F.fsc(x, truth)))
        log(f"recon: CG iteration {it}: FSC(0.5) vs phantom at {curve[-1]:.4f}")

    x, hist = project.cgls(P, imgs, iters=args.cg_iters, callback=cb)
    g_ms, g_launch, g_units = P.ctx.kernel_stats(_lib.K_GATHER)
    s_ms, s_launch, s_units = P.ctx.kernel_stats(_lib.K_INSERT)
    it_s = [h["seconds"] for h in hist]
    eq = next((h["iter"] for h, r in zip(hist, curve) if r >= res_bp), None)

    # CG on the normal equations preconditioned by the sampling density
    # (its first step is a gridding reconstruction)
    t0 = time.perf_counter()
    M = project.DensityPreconditioner(P)
    t_pre = time.perf_counter() - t0
    pcurve = []
    _, phist = project.pcg(P, imgs, iters=args.cg_iters, precond=M,
                           callback=lambda it, x: pcurve.append(F.resolution(F.fsc(x, truth))))
    p_eq = next((h["iter"] for h, r in zip(phist, pcurve) if r >= res_bp), None)
    log(f"recon: PCG FSC(0.5) first {pcurve[0]:.4f}, best {max(pcurve):.4f}")

    # one autograd step on a batch of 256 projections
    sel = slice(0, min(256, nproj))
    v = x.clone().requires_grad_(True)
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = 0.5 * torch.sum((P(v, sel) - imgs[sel]) ** 2)
        loss.backward()
        torch.cuda.synchronize()
        t_ad = time.perf_counter() - t0
        v.grad = None
    P.close()
    out = {"metric": "CG reconstruction with the KB-gather projector vs plain BP 32F (FSC vs phantom)",
           "value": round(float(np.median(it_s)), 4), "unit": "s per CG iteration",
           "higher_is_better": False, "n_gpus": 1, "dtype": "f32",
           "data": "synthetic (seeded blob phantom, analytic projections, white noise)",
           "config": {"workload": f"{nx}^3 from {nproj} projections {nx}x{nx} SNR {rec['snr']}, "
                                  f"CGLS batches of 256 (config 4)", "nx": nx, "nproj": nproj},
           "bp32f_seconds": round(t_bp, 4), "bp32f_fsc05_vs_phantom": round(res_bp, 4),
           "bp32f_halves_fsc0143": round(res_halves, 4),
           "cg_iters": len(hist), "cg_seconds_total": round(sum(it_s), 3),
           "cg_fsc05_vs_phantom": [round(r, 4) for r in curve],
           "cg_iter_reaching_bp32f_fsc": eq,
           "cg_residual": [round(h["residual"], 3) for h in hist],
           "pcg_density_fsc05_vs_phantom": [round(r, 4) for r in pcurve],
           "pcg_iter_reaching_bp32f_fsc": p_eq, "pcg_best_fsc05": round(max(pcurve), 4),
           "pcg_seconds_per_iter": round(float(np.median([h["seconds"] for h in phist])), 4),
           "pcg_precond_setup_seconds": round(t_pre, 3),
           "gather_proj_per_s": round(g_units / (g_ms / 1e3), 1) if g_ms > 0 else None,
           "scatter_proj_per_s": round(s_units / (s_ms / 1e3), 1) if s_ms > 0 else None,
           "autograd_step_256_ms": round(t_ad * 1e3, 3)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5],
                    help="2: plain BP 32F throughput (default; with an FP64-geometry record and a config-5 record: "
                         "256^2 BF16 x 100,000 into 512^3, compensated weights); 1 / 5: that workload alone; 3: 12 uneven BP 32F volumes, "
                         "plain one-by-one vs batched (tuned plans, BF16 images, one captured graph); "
                         "4: CG reconstruction with the gather projector beside BP 32F")
    ap.add_argument("--volumes", type=int, default=12, help="config 3: number of volumes")
    ap.add_argument("--cg-iters", type=int, default=30)
    ap.add_argument("--proj-batch", type=int, default=2048,
                    help="config 4: projections per gather / scatter call of the explicit A and A^T")
    ap.add_argument("--nproj", type=int, default=None, help="override the projection count")
    ap.add_argument("--batch", type=int, default=0, help="projections per internal chunk (0 = auto)")
    ap.add_argument("--strategy", default="auto", choices=["auto", "direct", "brick"])
    ap.add_argument("--brick-proj", type=int, default=0, help="projections per brick work item (0 = auto)")
    ap.add_argument("--shape", type=int, default=None, help="brick shape 0: 16x16x8, 1: 16x8x8, 2: 8x8x8")
    ap.add_argument("--tune", action="store_true", help="run the plan tuner on 1024 of the images first")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--parity-proj", type=int, default=100)
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="MiB: run one device-to-device copy of this size first (calibrates PMC byte counters)")
    ap.add_argument("--bf16", action="store_true", help="store the images as BF16 (default for config 5)")
    ap.add_argument("--twosum", action="store_true", help="compensated weights (default for config 5)")
    ap.add_argument("--geom-f64", action="store_true", help="FP64 sample geometry (SPBP_GEOM_F64) in the timed job")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-calib", action="store_true", help="skip the copy-rate and atomic-rate calibrations")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--bp3f", action="store_true", help="time BP 3F (one volume) instead of BP 32F")
    ap.add_argument("--one-launch", action="store_true",
                    help="one chunk per job: every projection's spectrum resident, one persistent scatter launch")
    ap.add_argument("--no-extra", action="store_true", help="config 2: skip the FP64-geometry and config-5 records")
    ap.add_argument("--no-config5", action="store_true", help="config 2: skip the config-5 record")
    ap.add_argument("--config5-nproj", type=int, default=None, help="config-5 record: projection count")
    ap.add_argument("--config5-steps", type=int, default=3)
    ap.add_argument("--config5-parity-proj", type=int, default=60)
    args = ap.parse_args()
    if args.no_parity:
        args.parity_proj = 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(projection sharding + one sum-reduction of the half grid is the design, not built).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from spider_amd import _lib, synth
    from spider_amd import bp as bpmod

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.config == 4:
        run_recon(args)
        return
    if args.config == 3:
        run_batched(args)
        return
    op = "bp3f" if args.bp3f else "bp32f"
    out = run_throughput(args, args.config, args.nproj, args.steps, args.warmup, op=op,
                         bf16=args.bf16 or args.config == 5, twosum=args.twosum or args.config == 5,
                         geom_f64=args.geom_f64, parity_proj=args.parity_proj, cpu=not args.no_cpu,
                         calib=not args.no_calib, tune=args.tune)
    out = {"metric": METRIC, **out}
    if args.config == 2 and not args.no_extra:
        # the same BP 32F job with FP64 sample geometry (the mode that meets the
        # FP64 bar), and the 256^2 BF16 -> 512^3 compensated-weights workload
        log("bench: FP64-geometry BP 32F")
        g = run_throughput(args, 2, args.nproj, max(2, min(args.steps, 5)), 1, op=op, bf16=False, twosum=False,
                           geom_f64=True, parity_proj=0, cpu=False, calib=False, tune=False, extras=False)
        out["geom_f64"] = {k: g[k] for k in ("value", "unit", "ms_per_step", "steps", "config")}
        out["geom_f64"]["rel_l2_vs_fp64"] = out.get("rel_l2_vs_fp64_with_geom_f64")
        log("bench: host images (PCIe-inclusive)")
        torch.cuda.empty_cache()
        out["host_images"] = run_host_images(args)
        if not args.no_config5:
            torch.cuda.empty_cache()
            log("bench: config 5")
            c5 = run_throughput(args, 5, args.config5_nproj, args.config5_steps, 1, op=op, bf16=True, twosum=True,
                                geom_f64=False, parity_proj=args.config5_parity_proj, cpu=not args.no_cpu,
                                calib=False, tune=False, cpu_seconds=min(args.cpu_seconds, 12.0))
            out["config5"] = c5
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
