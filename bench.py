"""Benchmark: W-projection gridding throughput on one MI355X.

Workload (default): the 31,395,840-visibility / 601-plane / support-72 /
16384^2 synthetic shape of the reference's SKA-LOW EL56-82 data set
(README:126-134), FP32, inputs resident in HBM.  A step is one pass of the
gridding pipeline over all visibilities: kernel-table expansion, selection,
bucketing by tile, tile gridding and the ordered merge, accumulating into
the device grid (the drop-in's semantics).

Prints one JSON line.  `roofline` prices the gridding kernel: its
algorithmic FP32 work, 8 flops per complex tap MAC summed over gridded
visibilities' (2s+1)^2 taps, divided by the kernel's average duration from
HIP events recorded on the launch stream inside the timed steps.  The work
is FP32 vector FMAs (no matrix instructions: the tap lookup is an indexed
gather); `peak` is MI355X's FP32 figure, 157.3 TFLOP/s, which needs packed
FMAs, and `frac_nonpacked` prices against the 78.6 TFLOP/s of one FMA per
lane per clock.  `traffic` is the kernel's HBM bytes per launch from the
FETCH_SIZE / WRITE_SIZE counter summary in profiles/, when present.
`accuracy` compares one pass's FP32 grid (plain and compensated) with the
FP64-accumulation truth and times the compensated form, the truth and the
FP64 gridder (FP64 selection and accumulation) on the same inputs.
`cpu_baseline` times the reference's serial oskar_grid_wproj_f (compiled
from its sources into oracle/_ref, else the oracle's restatement) on one
core on a prefix of the same visibilities.  `dropin` (with --dropin) times
the host-pointer drop-in mi355_grid_wproj_f, host-device copies included.
`driver` (with --driver) times main.cpp's flow over .vis / FITS files of the
2.15M-visibility shape.  `image` times the grid -> image transform (centred inverse 2-D FFT of the
gridded 16384^2 grid, real part / norm) beside numpy's FFT of a 4096^2
grid on one core.  `weights` times the uniform-weighting pre-pass on the same points (device
write of the 16384^2 weights grid, then the read of every point's weight),
beside the reference's serial oskar_grid_weights_write_f / _read_f over all
points on one core.

--copy-calibration MIB only runs plain device copies of MIB MiB (the
reference kernel for scaling the HBM byte counters) and exits.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gpu-gridding_amd"))

PEAK_FP32_TFLOPS = 157.3          # MI355X FP32, packed FMA
PEAK_FP32_NONPACKED_TFLOPS = 78.6  # one FMA per lane per clock
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def host_facts():
    """CPU model and core counts of the host the bench runs on."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return dict(cpu_model=model, nproc=os.cpu_count(), nproc_usable=usable)


def grid_errors(test, truth):
    """max |G - T|, max |G - T| / max |T| and ||G - T||_2 / ||T||_2 over the
    grids' real and imaginary parts (numpy or torch arrays, any precision)."""
    import numpy as np
    t = np.asarray(truth, np.float64)
    dd = np.asarray(test, np.float64) - t
    mt, lt = float(np.max(np.abs(t))), float(np.linalg.norm(t))
    ma = float(np.max(np.abs(dd)))
    return dict(max_abs=ma, max_rel=ma / mt if mt else ma,
                l2_rel=float(np.linalg.norm(dd)) / lt if lt else float(np.linalg.norm(dd)))


def cpu_baseline(case, seconds, d=None, ks=None, ctx=None):
    """Serial gridder on one core over a prefix of the workload (about
    `seconds` of CPU work), and on the same prefix the accuracy statement:
    the serial FP64 gridder (oskar_grid_wproj_d's arithmetic) and the FP64
    accumulation with FP32 selection as truths, the serial FP32 gridder's own
    error against them (the reference's FP32 error) and, given the device
    inputs d, the GPU's plain and compensated FP32 grids of the same prefix
    against them."""
    import numpy as np
    from oracle import oracle
    use_ref = oracle.ref() is not None
    kind = "reference" if use_ref else "port"
    G = case["grid_size"]
    grid = np.zeros((G, G, 2), np.float32)

    def sub(n):
        s = dict(case)
        for k in ("uu", "vv", "ww", "weight"):
            s[k] = case[k][:n]
        s["vis"] = case["vis"][:n]
        s["num_points"] = n
        return s

    def run(n):
        grid[...] = 0
        t0 = time.perf_counter()
        _, nrm, _ = oracle.grid_f(sub(n), use_ref=use_ref, grid=grid)
        return time.perf_counter() - t0, nrm

    n0 = 20_000
    t0, _ = run(n0)
    n = int(min(case["num_points"], max(n0, n0 * seconds / max(t0, 1e-6))))
    t, norm_f = run(n)
    facts = host_facts()
    out = dict(value=n / t / 1e6, unit="Mvis/s", cores=1, kind=kind,
               sample="first %d of %d visibilities (%.1f s), serial oskar_grid_wproj_f on one "
                      "core of %d (%s)" % (n, case["num_points"], t, facts["nproc"],
                                           facts["cpu_model"]),
               host=facts)
    # accuracy on the same prefix
    s = sub(n)
    g_d, norm_d, _ = oracle.grid_d(s)                 # FP64 selection and accumulation
    g_t, norm_t, _ = oracle.grid_fsel_d(s)            # FP32 selection, FP64 accumulation
    acc = dict(prefix_points=n,
               serial_f_vs_d=grid_errors(grid, g_d),
               serial_f_vs_fsel_d=grid_errors(grid, g_t),
               serial_norm_rel_vs_d=abs(norm_f - norm_d) / abs(norm_d))
    if d is not None:
        import torch
        from mi355_gridding import _lib, torch_ops as ops
        comp0 = ctx.get_option(_lib.OPT_COMPENSATED)
        for name, comp in (("gpu_plain", 0), ("gpu_compensated", 1)):
            gg = torch.zeros((G, G, 2), dtype=torch.float32, device="cuda")
            nn = torch.zeros(1, dtype=torch.float64, device="cuda")
            sk = torch.zeros(1, dtype=torch.int64, device="cuda")
            ctx.set_option(_lib.OPT_COMPENSATED, comp)
            ops.grid(ks, d["uu"][:n], d["vv"][:n], d["ww"][:n], d["vis"][:2 * n]
                     if d["vis"].dim() == 1 else d["vis"][:n], d["weight"][:n],
                     case["cell_size_rad"], case["w_scale"], G, gg, nn, sk, ctx=ctx)
            torch.cuda.synchronize()
            h = gg.cpu().numpy()
            del gg
            acc[name + "_vs_fsel_d"] = grid_errors(h, g_t)
            acc[name + "_vs_d"] = grid_errors(h, g_d)
            acc[name + "_norm_rel_vs_serial_f"] = abs(nn.item() - norm_f) / abs(norm_f)
        ctx.set_option(_lib.OPT_COMPENSATED, comp0)
    acc["note"] = ("errors over the real and imaginary parts of the grid of the prefix; d = the "
                   "serial FP64 gridder (FP64 selection, oskar_grid_wproj_d), fsel_d = FP32 "
                   "selection with FP64 accumulation (the truth with the FP32 gridder's own "
                   "cells and kernel indices)")
    out["accuracy"] = acc
    del g_d, g_t
    return out


def bench_batched(a):
    """The batched workload: 64 snapshot grids (4096^2) with log-uniform visibility
    counts, autotuned, gridded as one captured HIP graph per step; compared
    with the same pass launched eagerly and with 64 sequential single-grid
    calls."""
    import numpy as np
    import torch
    from mi355_gridding import synth, _lib, torch_ops as ops

    t0 = time.time()
    b = synth.make_batched()
    log("generated %d jobs, %d visibilities in %.1fs" % (
        len(b["jobs"]), sum(len(j["uu"]) for j in b["jobs"]), time.time() - t0))
    ks = ops.KernelStack(b["support"], b["oversample"], b["conv_size_half"], b["conv_func"])
    cat = {k: torch.from_numpy(np.ascontiguousarray(
               np.concatenate([j[k] for j in b["jobs"]]).reshape(-1))).cuda()
           for k in ("uu", "vv", "ww", "vis", "weight")}
    N = cat["uu"].numel()
    jobs, off = [], 0
    for j in b["jobs"]:
        G, n = j["grid_size"], len(j["uu"])
        jobs.append(dict(vis_offset=off, num_points=n, cell_size_rad=j["cell_size_rad"],
                         w_scale=j["w_scale"], grid_size=G,
                         grid=torch.zeros((G, G, 2), dtype=torch.float32, device="cuda"),
                         norm=torch.zeros(1, dtype=torch.float64, device="cuda"),
                         num_skipped=torch.zeros(1, dtype=torch.int64, device="cuda")))
        off += n
    args = (ks, jobs, cat["uu"], cat["vv"], cat["ww"], cat["vis"], cat["weight"])
    ctx = ops.context()
    if a.batched_leg != "all":
        # one leg alone, default options (for kernel traces of that leg)
        fn = ((lambda: ops.grid_batched_graph(*args, ctx=ctx)) if a.batched_leg == "graph"
              else (lambda: ops.grid_batched(*args, ctx=ctx)))
        for _ in range(max(a.warmup, 1)):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(a.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        log("%s batched pass: %.3f ms" % (a.batched_leg, e0.elapsed_time(e1) / a.steps))
        return
    best = ops.autotune_batched(*args, ctx=ctx)
    log("autotuned: %.3f ms per pass" % best)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    graph_ms = timed(lambda: ops.grid_batched_graph(*args, ctx=ctx), a.steps, max(a.warmup, 1))
    st = ctx.stats()                      # counts of the last graph launch
    eager_ms = timed(lambda: ops.grid_batched(*args, ctx=ctx), max(2, a.steps // 2), 1)
    singles = []
    for jb in jobs:
        o, n = jb["vis_offset"], jb["num_points"]
        singles.append(tuple(cat[k][o:o + n] for k in ("uu", "vv", "ww")) +
                       (cat["vis"][2 * o:2 * (o + n)], cat["weight"][o:o + n], jb))

    def sequential():
        for uu, vv, ww, vis, wt, jb in singles:
            ops.grid(ks, uu, vv, ww, vis, wt, jb["cell_size_rad"], jb["w_scale"], jb["grid_size"],
                     jb["grid"], jb["norm"], jb["num_skipped"], ctx=ctx)
    seq_ms = timed(sequential, 2, 1)
    flops = 8.0 * st["num_taps"]
    out = {
        "metric": "Mvis/s, 64 snapshot grids (4096^2, 339 planes, support 44) in one HIP graph",
        "value": round(N / (graph_ms * 1e-3) / 1e6, 3),
        "unit": "Mvis/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(graph_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded generator, batched-64x4096 shapes)",
        "config": {"workload": "batched-64x4096", "num_jobs": len(jobs), "num_points": N,
                   "grid_size": 4096, "num_w_planes": b["num_w_planes"],
                   "max_support": b["max_support"],
                   "autotuned": {"grid_kernel": ctx.get_option(_lib.OPT_GRID_KERNEL),
                                 "visits_per_item": ctx.get_option(_lib.OPT_VISITS_PER_ITEM),
                                 "min_stage_run": ctx.get_option(_lib.OPT_MIN_STAGE_RUN),
                                 "min_stage_run_per_class_w33_w65_w97_w111": [
                                     ctx.get_option(k) for k in (
                                         _lib.OPT_MIN_STAGE_RUN_W33, _lib.OPT_MIN_STAGE_RUN_W65,
                                         _lib.OPT_MIN_STAGE_RUN_W97,
                                         _lib.OPT_MIN_STAGE_RUN_W111)],
                                 "item_group": ctx.get_option(_lib.OPT_ITEM_GROUP)}},
        "batched": {"graph_ms": round(graph_ms, 3), "eager_batched_ms": round(eager_ms, 3),
                    "sequential_64_ms": round(seq_ms, 3),
                    "graph_vs_sequential": round(seq_ms / graph_ms, 3),
                    "autotune_best_ms": round(best, 3),
                    "pipeline_tflops": round(flops / (graph_ms * 1e-3) / 1e12, 3)},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_weights(case, d, ctx, no_cpu, reps=5):
    """Uniform-weighting pre-pass on the device (write, then read) over all
    points of the workload, and the reference's serial functions on the
    host over the same points."""
    import numpy as np
    import torch
    from mi355_gridding import torch_ops as ops
    N, G = case["num_points"], case["grid_size"]
    cell = float(case["cell_size_rad"])
    wgrid = torch.zeros((G, G), dtype=torch.float32, device="cuda")
    wout = torch.empty_like(d["weight"])
    sk = torch.zeros(1, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms_w = timed(lambda: ops.grid_weights_write(d["uu"], d["vv"], d["weight"], cell, G, wgrid,
                                                sk, ctx=ctx))
    wgrid.zero_()
    ops.grid_weights_write(d["uu"], d["vv"], d["weight"], cell, G, wgrid, sk, ctx=ctx)
    ms_r = timed(lambda: ops.grid_weights_read(d["uu"], d["vv"], d["weight"], wout, cell, G,
                                               wgrid, sk, ctx=ctx))
    out = dict(ms_write=round(ms_w, 4), ms_read=round(ms_r, 4),
               value=round(N / ((ms_w + ms_r) * 1e-3) / 1e6, 1), unit="Mpoints/s",
               num_skipped=int(sk.item()),
               note="device pointers; write = cell keys + stable radix sort + ordered "
                    "per-cell sums into a zeroed FP32 grid; read = one division per point")
    del wgrid, wout
    if not no_cpu:
        from oracle import oracle
        use_ref = oracle.ref() is not None
        hgrid = np.zeros((G, G), np.float32)
        hout = np.empty(N, np.float32)
        t0 = time.perf_counter()
        oracle.weights_write(case["uu"], case["vv"], case["weight"], np.float32(cell), G, hgrid,
                             use_ref=use_ref)
        t1 = time.perf_counter()
        oracle.weights_read(case["uu"], case["vv"], case["weight"], hout, np.float32(cell), G,
                            hgrid, use_ref=use_ref)
        t2 = time.perf_counter()
        out["cpu_baseline"] = dict(
            ms_write=round((t1 - t0) * 1e3, 1), ms_read=round((t2 - t1) * 1e3, 1),
            value=round(N / (t2 - t0) / 1e6, 2), unit="Mpoints/s", cores=1,
            kind="reference" if use_ref else "port",
            sample="all %d points, serial oskar_grid_weights_write_f + _read_f" % N)
        del hgrid, hout
    log("weights %s" % json.dumps(out))
    return out


def bench_differentiable(case, d, reps=3):
    """Forward (gridder) and backward (degridding adjoint) of the
    differentiable gridder, loss 0.5 ||grid - target||^2 with a seeded
    complex target, gradients for the amplitudes and the weights."""
    import numpy as np
    import torch
    from mi355_gridding import torch_ops as ops
    G = case["grid_size"]
    ks = ops.KernelStack(case["support"], case["oversample"], case["conv_size_half"],
                         case["conv_func"])
    gen = torch.Generator(device="cuda").manual_seed(1234)
    target = torch.randn((G, G, 2), generator=gen, device="cuda")
    vis = d["vis"].reshape(-1, 2).clone().requires_grad_(True)
    wt = d["weight"].clone().requires_grad_(True)
    args = (ks, d["uu"], d["vv"], d["ww"], case["cell_size_rad"], case["w_scale"], G)
    fw, bw = [], []
    for r in range(reps + 1):
        vis.grad = None
        wt.grad = None
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        g, _ = ops.grid_autograd(vis, wt, *args)
        loss = 0.5 * ((g - target) ** 2).sum()
        e1.record()
        loss.backward()
        e2.record()
        torch.cuda.synchronize()
        if r:
            fw.append(e0.elapsed_time(e1))
            bw.append(e1.elapsed_time(e2))
    N = case["num_points"]
    # the degridding call alone (selection, grid-tile ordering, k_degrid)
    y = ops.degrid(ks, d["uu"], d["vv"], d["ww"], d["weight"], case["cell_size_rad"],
                   case["w_scale"], G, target, apply_weight=False)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        ops.degrid(ks, d["uu"], d["vv"], d["ww"], d["weight"], case["cell_size_rad"],
                   case["w_scale"], G, target, apply_weight=False, out=y)
    e1.record()
    torch.cuda.synchronize()
    dg_ms = e0.elapsed_time(e1) / 5
    out = dict(forward_ms=round(float(np.mean(fw)), 3), backward_ms=round(float(np.mean(bw)), 3),
               degrid_call_ms=round(dg_ms, 3),
               value=round(N / ((np.mean(fw) + np.mean(bw)) * 1e-3) / 1e6, 3),
               unit="Mvis/s (forward + backward)",
               note="torch_ops.GridWProj: forward = gridder plus the loss; backward = "
                    "degridding kernel for d/dvis and d/dweight; degrid_call_ms = one "
                    "mi355_degrid_wproj_f_dev call (selection, grid-tile order, k_degrid)")
    log("differentiable %s" % json.dumps(out))
    return out


def bench_image(grid, G, norm, ctx, no_cpu, reps=5):
    """Grid -> image on the device over the gridded grid; numpy's FFT of a
    4096^2 grid on the host as the CPU figure."""
    import numpy as np
    import torch
    from mi355_gridding import torch_ops as ops
    img = torch.empty((G, G), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    ops.grid_to_image(grid, G, img, norm=norm, ctx=ctx)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        ops.grid_to_image(grid, G, img, norm=norm, ctx=ctx)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del img
    gb = (8.0 * G * G + 4.0 * G * G) / 1e9       # grid read once, image written once
    out = dict(ms=round(ms, 4), grid_size=G, value=round(G * G / (ms * 1e-3) / 1e9, 3),
               unit="Gcells/s", min_bytes_gb=round(gb, 3),
               eff_gbps=round(gb / (ms * 1e-3), 1),
               note="hipFFT backward C2C in place on a work copy, plus the centring-sign "
                    "and real-part/scale passes; eff_gbps counts only the grid read and the "
                    "image write")
    if not no_cpu:
        n = 4096
        rng = np.random.default_rng(0)
        h = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))).astype(np.complex64)
        t0 = time.perf_counter()
        np.fft.ifft2(h)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=round(n * n / t / 1e9, 4), unit="Gcells/s", cores=1,
                                   kind="numpy", sample="numpy.fft.ifft2 of one %d^2 complex64 "
                                   "grid (pocketfft, FP64 internally): %.2f s" % (n, t))
    log("image %s" % json.dumps(out))
    return out


def bench_driver():
    """main.cpp's flow on the GPU (mi355_grid_vis_file: read the blocks of a
    .vis file, uniform weights, grid every block, FP32) over files of the
    2.15M-visibility / 714-plane / support-36 / 8192^2 shape (208 stations x
    100 times, 10 times per block) written to a scratch directory; the
    coordinates are written in metres at a frequency equal to c, so they
    read back as the same wavelengths."""
    import shutil
    import tempfile
    import numpy as np
    from mi355_gridding import synth, vis_io
    c = synth.make(1)
    S, T, tpb = 208, 100, 10
    B = S * (S - 1) // 2
    assert B * T == c["num_points"]
    tmp = tempfile.mkdtemp(prefix="mi355_driver_")
    root = os.path.join(tmp, "synthetic")
    try:
        kw = (c["support"], c["oversample"], c["grid_size"], float(c["cell_size_rad"]),
              float(c["w_scale"]))
        vis_io.write_kernel(root + "_KERNELS_REAL.fits", c["conv_func"][..., 0], *kw)
        vis_io.write_kernel(root + "_KERNELS_IMAG.fits", c["conv_func"][..., 1], *kw)
        vis = (c["vis"][:, 0] + 1j * c["vis"][:, 1]).reshape(T, B)
        vis_io.write_vis(root + ".vis", c["uu"].reshape(T, B), c["vv"].reshape(T, B),
                         c["ww"].reshape(T, B), vis, S, tpb, 299792458.0)
        vis_io.grid_vis_file(root, want_grid=False)            # first call: device set-up
        runs = []
        for _ in range(2):
            t0 = time.perf_counter()
            _, res = vis_io.grid_vis_file(root, want_grid=False)
            runs.append((time.perf_counter() - t0, res))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    wall, res = min(runs, key=lambda r: r[0])
    # one block's weights read and gridding on resident inputs, for
    # comparison with the driver's per-block sum (blocks x this)
    import torch
    from mi355_gridding import torch_ops as ops
    n = B * tpb
    ks = ops.KernelStack(c["support"], c["oversample"], c["conv_size_half"], c["conv_func"])
    d = {k: torch.from_numpy(np.ascontiguousarray(c[k][:n])).cuda()
         for k in ("uu", "vv", "ww", "vis", "weight")}
    G = c["grid_size"]
    grid = torch.zeros((G, G, 2), dtype=torch.float32, device="cuda")
    nrm = torch.zeros(1, dtype=torch.float64, device="cuda")
    skp = torch.zeros(1, dtype=torch.int64, device="cuda")
    call = lambda: ops.grid(ks, d["uu"], d["vv"], d["ww"], d["vis"], d["weight"],
                            c["cell_size_rad"], c["w_scale"], G, grid, nrm, skp)
    for _ in range(3):
        call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        call()
    e1.record()
    torch.cuda.synchronize()
    block_ms = e0.elapsed_time(e1) / 10
    return dict(seconds_total=round(wall, 4), seconds_wall=round(res["seconds_wall"], 4),
                seconds_read=round(res["seconds_read"], 4),
                seconds_weights=round(res["seconds_weights"], 4),
                seconds_grid=round(res["seconds_grid"], 4), num_blocks=res["num_blocks"],
                num_points=res["num_points"],
                block_grid_ms=round(block_ms, 3),
                grid_vs_blocks_x_block=round(res["seconds_grid"] * 1e3 /
                                             (res["num_blocks"] * block_ms), 3),
                mvis_s_grid=round(res["num_points"] / res["seconds_grid"] / 1e6, 2),
                mvis_s_total=round(res["num_points"] / wall / 1e6, 2),
                note="mi355_grid_vis_file over a .vis file and FITS kernel cubes in a scratch "
                     "directory (file reads, weights and gridding of every block; best of 2); "
                     "seconds_grid sums each block's device time from its copies to its last "
                     "pass; block_grid_ms is one %d-visibility block gridded on resident "
                     "inputs (mean of 10)" % n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--batched-leg", choices=("all", "eager", "graph"), default="all",
                    help="config 3: time only the eager batched pass or only the graph "
                         "replays, with default options (no autotuning, no other legs)")
    ap.add_argument("--compensated", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--kernel", choices=("strips", "staged", "direct"), default="strips",
                    help="gridding kernel: 16-column strips on 128x128 regions, or 64x64 "
                         "tiles with tables staged in LDS / read directly")
    ap.add_argument("--visits-per-item", type=int, default=0,
                    help="tile visits per work item (0: the library's automatic choice)")
    ap.add_argument("--min-stage-run", type=int, default=0,
                    help="shortest staged run (0: the library's automatic choice)")
    ap.add_argument("--item-group", type=int, default=1,
                    help="strip kernel: consecutive work items dealt to one XCD")
    ap.add_argument("--dropin", action="store_true",
                    help="also time the host-pointer drop-in (PCIe included)")
    ap.add_argument("--driver", action="store_true",
                    help="also time main.cpp's flow over .vis / FITS files of the "
                         "2.15M-visibility shape")
    ap.add_argument("--no-weights", action="store_true",
                    help="skip the uniform-weighting pre-pass leg")
    ap.add_argument("--no-image", action="store_true",
                    help="skip the grid -> image transform leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run plain device copies of MIB MiB, then exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: measures one GPU only (the gridder has no multi-GPU "
              "path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from mi355_gridding import synth, _lib, torch_ops as ops

    torch.cuda.set_device(0)
    if a.config == 3:
        return bench_batched(a)
    if a.copy_calibration:
        n = a.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copied %d MiB 5 times" % a.copy_calibration)
        return
    t0 = time.time()
    case = synth.make(a.config)
    log("generated %s: N=%d G=%d P=%d smax=%d in %.1fs" % (
        case["name"], case["num_points"], case["grid_size"], case["num_w_planes"],
        case["max_support"], time.time() - t0))
    N, G = case["num_points"], case["grid_size"]
    ks = ops.KernelStack(case["support"], case["oversample"], case["conv_size_half"],
                         case["conv_func"])
    d = {k: torch.from_numpy(np.ascontiguousarray(case[k])).cuda()
         for k in ("uu", "vv", "ww", "vis", "weight")}
    grid = torch.zeros((G, G, 2), dtype=torch.float32, device="cuda")
    norm = torch.zeros(1, dtype=torch.float64, device="cuda")
    skipped = torch.zeros(1, dtype=torch.int64, device="cuda")
    ctx = ops.context()
    ctx.set_option(_lib.OPT_PHASE_TIMING, 1)
    ctx.set_option(_lib.OPT_COMPENSATED, 1 if a.compensated else 0)
    ctx.set_option(_lib.OPT_GRID_KERNEL, {"strips": 0, "staged": 2, "direct": 1}[a.kernel])
    ctx.set_option(_lib.OPT_VISITS_PER_ITEM, a.visits_per_item)
    ctx.set_option(_lib.OPT_MIN_STAGE_RUN, a.min_stage_run)
    ctx.set_option(_lib.OPT_ITEM_GROUP, a.item_group)

    def step():
        ops.grid(ks, d["uu"], d["vv"], d["ww"], d["vis"], d["weight"], case["cell_size_rad"],
                 case["w_scale"], G, grid, norm, skipped, ctx=ctx)

    for i in range(a.warmup):
        step()
        log("warmup %d done" % i)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    phases = []
    w0 = time.perf_counter()
    ev0.record(stream)
    for i in range(a.steps):
        step()
        phases.append(ctx.stats())
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_time(ev1) / a.steps
    log("timed %d steps: %.3f ms/step (wall %.3f ms/step)" % (a.steps, ms, wall * 1e3 / a.steps))
    st = phases[-1]
    grid_ms = float(np.mean([p["ms_grid"] for p in phases]))
    flops = 8.0 * st["num_taps"]
    achieved = flops / (grid_ms * 1e-3) / 1e12
    phase_ms = {k: round(float(np.mean([p[k] for p in phases])), 4)
                for k in ("ms_prepare", "ms_select", "ms_bucket", "ms_grid", "ms_merge",
                          "ms_total")}

    accuracy = None
    if not a.no_accuracy:
        # FP32 grid of one fresh pass vs the FP64 truth with the same selection.
        g32 = torch.zeros_like(grid)
        n32 = torch.zeros_like(norm)
        step_args = (ks, d["uu"], d["vv"], d["ww"], d["vis"], d["weight"], case["cell_size_rad"],
                     case["w_scale"], G)
        ops.grid(*step_args, g32, n32, skipped, ctx=ctx)
        g64 = torch.zeros((G, G, 2), dtype=torch.float64, device="cuda")
        n64 = torch.zeros_like(norm)
        torch.cuda.synchronize()
        f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        f0.record(stream)
        ops.grid_fsel_d(*step_args, g64, n64, skipped, ctx=ctx)
        f1.record(stream)
        torch.cuda.synchronize()
        fsel_d_ms = f0.elapsed_time(f1)
        ctx.set_option(_lib.OPT_COMPENSATED, 1)
        gc = torch.zeros_like(grid)
        ops.grid(*step_args, gc, torch.zeros_like(norm), skipped, ctx=ctx)
        # time the compensated form on the same inputs (3 passes)
        torch.cuda.synchronize()
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record(stream)
        for _ in range(3):
            ops.grid(*step_args, gc, torch.zeros_like(norm), skipped, ctx=ctx)
        c1.record(stream)
        torch.cuda.synchronize()
        comp_ms = c0.elapsed_time(c1) / 3
        gc.zero_()
        ops.grid(*step_args, gc, torch.zeros_like(norm), skipped, ctx=ctx)
        ctx.set_option(_lib.OPT_COMPENSATED, 1 if a.compensated else 0)
        torch.cuda.synchronize()
        m64 = g64.abs().max().item()
        l64 = torch.linalg.norm(g64).item()

        def err(g):
            dd = g.double() - g64
            ma = dd.abs().max().item()
            return dict(max_abs=ma, max_rel=ma / m64, l2_rel=torch.linalg.norm(dd).item() / l64)
        # the FP64 gridder (oskar_grid_wproj_d's arithmetic: FP64 selection
        # and accumulation) on the same inputs: one warm pass, one timed
        ks64 = ops.KernelStack(case["support"], case["oversample"], case["conv_size_half"],
                               case["conv_func"], dtype=torch.float64)
        args64 = (ks64,) + tuple(d[k].double() for k in ("uu", "vv", "ww", "vis", "weight")) + \
            (case["cell_size_rad"], case["w_scale"], G)
        gd = torch.zeros((G, G, 2), dtype=torch.float64, device="cuda")
        nd = torch.zeros_like(norm)
        ops.grid(*args64, gd, nd, skipped, ctx=ctx)
        torch.cuda.synchronize()
        f0.record(stream)
        ops.grid(*args64, gd, nd, skipped, ctx=ctx)
        f1.record(stream)
        torch.cuda.synchronize()
        fp64_ms = f0.elapsed_time(f1)
        del gd, args64, ks64
        accuracy = dict(plain_fp32=err(g32), compensated_fp32=err(gc),
                        compensated_ms_per_step=round(comp_ms, 3),
                        fp64_accumulation_ms=round(fsel_d_ms, 3),
                        fp64_gridder_ms=round(fp64_ms, 3),
                        fp64_gridder_mvis_s=round(N / fp64_ms / 1e3, 2),
                        norm_rel=abs(n32.item() - n64.item()) / abs(n64.item()),
                        truth="FP64 accumulation, FP32 selection (mi355_grid_wproj_fsel_d_dev, "
                              "checked against the serial oracle at small sizes and on tiles); "
                              "cpu_baseline.accuracy states the same errors against the serial "
                              "FP64 oracle on a prefix, beside the serial FP32 gridder's own")
        del g64, gc, g32
        log("accuracy %s" % json.dumps(accuracy))

    dropin = None
    if a.dropin:
        from mi355_gridding import wproj
        hgrid = np.zeros((G, G), np.complex64)
        times = []
        for _ in range(2):
            t0 = time.perf_counter()
            wproj.grid_wproj_f(case["num_w_planes"], case["support"], case["oversample"],
                               case["conv_size_half"], case["conv_func"], N, case["uu"],
                               case["vv"], case["ww"], case["vis"], case["weight"],
                               case["cell_size_rad"], case["w_scale"], G, 0.0, hgrid)
            times.append(time.perf_counter() - t0)
        del hgrid
        dropin = dict(ms_per_call=round(min(times) * 1e3, 3),
                      value=round(N / min(times) / 1e6, 3), unit="Mvis/s",
                      note="host arrays; inputs and the touched grid window copied "
                           "host<->device inside the call (best of 2)")
        log("drop-in %s" % json.dumps(dropin))

    driver = None
    if a.driver:
        driver = bench_driver()
        log("driver %s" % json.dumps(driver))

    differentiable = bench_differentiable(case, d) if a.config == 4 else None
    weights = None if a.no_weights else bench_weights(case, d, ctx, a.no_cpu_baseline)
    image = None if a.no_image else bench_image(grid, G, float(norm.item()), ctx,
                                                a.no_cpu_baseline)

    traffic = None
    kname = {"strips": "k_grid_strips", "staged": "k_grid_staged", "direct": "k_grid_tiles"}[a.kernel]
    if os.path.exists(TRAFFIC_FILE):
        tr = json.load(open(TRAFFIC_FILE)).get(kname)
        if tr and tr.get("workload") == case["name"]:
            traffic = tr["bytes_per_launch"]

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(case, a.cpu_seconds, d, ks, ctx)
        log("cpu baseline %s" % json.dumps(cpu))

    out = {
        "metric": "Mvis/s and % FP32 peak, 31.4M-vis/601-plane grid, at <=1e-5 rel err vs FP64",
        "value": round(N / (ms * 1e-3) / 1e6, 3),
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, shapes of the reference data set)",
        "config": {"workload": case["name"], "num_points": N, "grid_size": G,
                   "num_w_planes": case["num_w_planes"], "max_support": case["max_support"],
                   "oversample": case["oversample"],
                   "accumulation": "compensated" if a.compensated else "plain",
                   "tile_kernel": a.kernel,
                   "visits_per_item": int(phases[-1]["visits_per_item"]) if phases
                   else a.visits_per_item,
                   "min_stage_run": a.min_stage_run or "auto", "item_group": a.item_group},
        "roofline": {"bound": "mfma",
                     "bound_note": "the kernel's binding unit is the FP32 vector FMA pipe (no "
                                   "matrix instructions: the tap lookup is an indexed gather); "
                                   "the line's two classes are hbm and mfma, so the "
                                   "compute-bound class is named",
                     "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "frac_nonpacked": round(achieved / PEAK_FP32_NONPACKED_TFLOPS, 4),
                     "compute": "FP32 vector FMA (compute-bound class; no matrix instructions)",
                     "kernel": kname, "kernel_ms": round(grid_ms, 4),
                     "flops_per_launch": flops, "traffic": traffic},
        "phases_ms": phase_ms,
        "work": {"num_taps": st["num_taps"], "num_visits": st["num_visits"],
                 "num_items": st["num_items"], "num_skipped": st["num_skipped"]},
        "accuracy": accuracy,
        "cpu_baseline": cpu,
        "dropin": dropin,
        "driver": driver,
        "differentiable": differentiable,
        "weights": weights,
        "image": image,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
