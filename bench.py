#!/usr/bin/env python3
"""MBIR throughput benchmark (slices/s) on one MI355X.

Headline workload (default): Toeplitz-gradient MBIR on 64 slices of
2048 x 2048, 1501 angles uniform in [0, pi) x 2048 rays, centre 1024,
seeded 30-random-ellipse phantoms projected with the library's NUFFT
projector plus 0.5 % Gaussian noise, 100 iterations, qGGMRF p = 1.2, q = 2,
c = 0.001, sigma = 100.  One step = one full mbir2 call (normalise, FBP
initial guess, padding, P^T y, PSF, step size, 100 iterations, crop) on a
sinogram already resident in HBM.  `value` is the mixed-precision form's
slices/s; plain FP32 is timed beside it on the same input.

The JSON line also carries:
  - `roofline`: the dominant kernel class of the timed (mixed) run, its
    algorithmic bytes over its HIP-event time on the launch stream;
  - `accuracy`: both forms against the FP64 oracle on one 2048^2 slice of
    the workload's geometry at 100 iterations (a committed fixture), and
    mixed against FP32 on the whole timed output;
  - `cpu_baseline`: the FP32 CPU restatement of the same algorithm on a
    bounded sample (set-up once, one warm-up iteration, the median of three
    timed runs), extrapolated per slice-iteration;
  - `extra`: the other workload shapes (32 x 1024^2 FP32, the 200-stack
    uneven batch, the autograd projector step).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

WORKLOADS = {
    # the headline: 100-iteration MBIR at 2048^2, mixed precision against
    # plain FP32 (sigma = 100, the reference's default smoothness 0.01: the
    # iteration converges, so errors against FP64 are meaningful)
    "mbir-64x2048": dict(
        slices=64, npix=2048, nproj=1501, iterations=100, sigma=100.0, p=1.2,
        noise=0.005, seed_group=3, lipschitz_iters=20,
        label="Toeplitz MBIR, mixed precision (FP16 gradient storage, FP32 "
              "spectrum, FFTs, intermediates and state, FP64 line search, L "
              "from 20 power-iteration steps) against plain FP32: 64 slices "
              "2048x2048, 1501 angles x 2048 rays, 100 iterations, qGGMRF "
              "sigma=100 p=1.2"),
    "plain-fp32-32x1024": dict(
        slices=32, npix=1024, nproj=720, iterations=100, sigma=0.1, p=1.2,
        noise=0.01, seed_group=1,
        label="plain FP32 Toeplitz MBIR: 32 slices 1024x1024, 720 angles x "
              "1024 rays, 100 iterations, qGGMRF sigma=0.1 p=1.2"),
    "cpu-path-4x256": dict(
        slices=4, npix=256, nproj=180, iterations=50, sigma=0.1, p=1.2,
        noise=0.01, seed_group=0, kind="shepp",
        label="4 slices 256x256 Shepp-Logan, 180 angles, 50 iterations"),
    "uneven-200": dict(
        kind="uneven", stacks=200, sizes=(256, 512, 768, 1024, 1536),
        pmin=180, pmax=1500, smin=1, smax=16, iterations=50, sigma=0.1,
        p=1.2, noise=0.01, seed_group=4,
        label="uneven batching: 200 seeded stacks, image size in "
              "{256,512,768,1024,1536}, angles U{180..1500}, rays = size, "
              "1-16 slices each, 50 iterations, qGGMRF sigma=0.1 p=1.2"),
    # one forward + backward of the differentiable loss
    # 0.5 ||P x - y||^2 + beta qGGMRF(x) (projector with the hand-written
    # adjoint backward, free-boundary prior with its exact gradient kernel)
    "autograd-8x512": dict(
        kind="autograd", slices=8, npix=512, nproj=360, iterations=1,
        sigma=0.1, p=1.2, beta=1e-3, noise=0.01, seed_group=2,
        label="differentiation: forward + backward of 0.5||Px-y||^2 + "
              "1e-3 qGGMRF(x) on 8 slices 512x512, 360 angles x 512 rays"),
}
HEADLINE = "mbir-64x2048"

PEAK_HBM_GBS = 8000.0

# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# counter runs (calibrated on a copy kernel), per workload
TRAFFIC_FILES = {
    "plain-fp32-32x1024": "mbir_fp32_32x1024_hbm_traffic.json",
    "mbir-64x2048": "mbir_mixed_64x2048_hbm_traffic.json",
    # the same counter runs hold the plain FP32 classes of that workload
    "plain-fp32-64x2048": "mbir_mixed_64x2048_hbm_traffic.json",
}


PEAK_VALU_TFLOPS = 78.6  # FP32 vector FMA without packed math (2 flop/lane/clk)


def valu_roofline(kd, kernel, n, L, wt_bytes=8, spec_bytes=4):
    """The column pass against the FP32 vector-issue peak.  Its time goes to
    FFT arithmetic (VALU bound), so next to the HBM roofline it is also
    rated by algorithmic flops: 5 L log2(L) per length-L complex transform,
    two per column (forward and inverse), plus the 2L of the real spectrum
    multiply; H = L/2 + 1 columns per slice.  The slices per launch follow
    from the launch's algorithmic bytes (ns * 2 * wt_bytes * n * H +
    spec_bytes * H * L)."""
    import math
    H = L // 2 + 1
    ns = (kd["bytes_per_launch"] - spec_bytes * H * L) / (2.0 * wt_bytes * n * H)
    flops = ns * H * (2 * 5.0 * L * math.log2(L) + 2.0 * L)
    ach = flops / (kd["avg_ms"] * 1e-3) / 1e12
    return {"kernel": kernel, "bound": "valu", "achieved": ach,
            "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_VALU_TFLOPS, "flops_per_launch": flops,
            "slices_per_launch": ns, "fft_length": L}


def fft_length(n):
    import ctypes as C
    from tomocam_amd import _lib
    L = C.c_int()
    _lib.call("tc_toeplitz_fft_length", int(n), C.byref(L))
    return L.value


def measured_traffic(workload, kernel_class, bytes_per_launch):
    """HBM bytes per launch: the class's counter-measured traffic /
    algorithmic ratio (committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs)
    times this run's algorithmic bytes per launch, so that it holds for the
    slice-group size the autotuner picked here."""
    name = TRAFFIC_FILES.get(workload)
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    c = t["classes"].get(kernel_class)
    return None if c is None else c["ratio"] * bytes_per_launch


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_sinogram(torch, gpu, phantom, w):
    kind = w.get("kind", "ellipses")
    ang = phantom.angles(w["nproj"])
    if kind == "ellipses" and w["npix"] * w["npix"] * w["slices"] > (1 << 26):
        # large stacks: the same phantoms (phantom_stack's seeds) rasterised
        # on the GPU
        seeds = [1000 * w["seed_group"] + s for s in range(w["slices"])]
        img = _ellipse_stack_torch(torch, w["npix"], seeds)
    else:
        img = torch.from_numpy(phantom.phantom_stack(
            w["npix"], w["slices"], kind, w["seed_group"])).cuda()
    sino = gpu.project(img, ang)
    del img
    torch.cuda.synchronize()
    g = torch.Generator(device="cuda")
    g.manual_seed(1000 * w["seed_group"] + 7)
    sd = w["noise"] * float(sino.max())
    sino = sino + sd * torch.randn(sino.shape, generator=g, device="cuda")
    return sino.contiguous(), ang


def _ellipse_stack_torch(torch, n, seeds):
    """random_ellipses (same parameter draws) rasterised on the GPU."""
    t = (torch.arange(n, dtype=torch.float64, device="cuda") - n / 2.0) / (
        n / 2.0)
    yy, xx = torch.meshgrid(t, t, indexing="ij")
    out = torch.zeros((len(seeds), n, n), dtype=torch.float32, device="cuda")
    for i, seed in enumerate(seeds):
        rng = np.random.default_rng(seed)
        img = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        for _ in range(30):
            x0, y0 = rng.uniform(-0.5, 0.5, size=2)
            a, b = rng.uniform(0.02, 0.4, size=2)
            phi = rng.uniform(0.0, np.pi)
            rho = rng.uniform(0.0, 1.0)
            c, sn = np.cos(phi), np.sin(phi)
            dx, dy = xx - x0, yy - y0
            u = dx * c + dy * sn
            v = -dx * sn + dy * c
            img += rho * (((u / a) ** 2 + (v / b) ** 2) <= 1.0)
        out[i] = img.to(torch.float32)
    return out


def make_uneven(torch, gpu, phantom, w):
    """The uneven-batching workload: stack sizes drawn from a seeded rng;
    slice phantoms seeded 1000*config + running slice index."""
    rng = np.random.default_rng(1000 * w["seed_group"])
    stacks, sl = [], 0
    for k in range(w["stacks"]):
        n = int(rng.choice(w["sizes"]))
        P = int(rng.integers(w["pmin"], w["pmax"] + 1))
        S = int(rng.integers(w["smin"], w["smax"] + 1))
        seeds = [1000 * w["seed_group"] + sl + i for i in range(S)]
        sl += S
        img = _ellipse_stack_torch(torch, n, seeds)
        ang = phantom.angles(P)
        sino = gpu.project(img, ang)
        g = torch.Generator(device="cuda")
        g.manual_seed(7919 * k + 1)
        sd = w["noise"] * float(sino.max())
        sino = sino + sd * torch.randn(sino.shape, generator=g, device="cuda")
        stacks.append((sino.contiguous(), ang, n / 2.0))
    torch.cuda.synchronize()
    return stacks


def cpu_baseline(sino_host, ang, w, total_slices, sample_slices,
                 sample_iters):
    """The FP32 CPU restatement on `sample_slices` slices: set-up once, one
    warm-up iteration, the median of three runs of `sample_iters`
    iterations, extrapolated to the workload (per slice-iteration; the PSF
    once per stack)."""
    from oracle import cpu_port
    r = cpu_port.sample_rate(sino_host, ang, w["npix"] / 2.0, w["iterations"],
                             w["sigma"], sample_slices, sample_iters,
                             total_slices=total_slices)
    return {
        "value": r["slices_per_s"],
        "unit": "slices/s",
        "cores": r["threads"],
        "kind": "port",
        "sample": (f"FP32 CPU restatement of mbir2 (4 Toeplitz products per "
                   f"attempt, scipy.fft + C/OpenMP) on {r['sample_slices']} "
                   f"of the {total_slices} slices: set-up once "
                   f"({r['setup_s']:.1f} s per sample, PSF {r['psf_s']:.1f} s "
                   f"per stack), 1 warm-up iteration, median of 3 runs of "
                   f"{sample_iters} iterations ({r['iter_s']:.2f} s per "
                   f"iteration); extrapolated linearly to "
                   f"{w['iterations']} iterations per slice"),
    }


# The HIP-event kernel timers bracket one launch in KTIMER_STRIDE of each
# class during the timed steps: a timing event costs about 1 us of GPU time,
# which at one-slice Toeplitz groups (60 us launches, two streams) slowed the
# 64 x 2048^2 loop by 2.4 % when every launch was bracketed.  3 is coprime to
# the power-of-two group counts, so every group position is sampled in turn.
KTIMER_STRIDE = 3


def timed_steps(torch, gpu, step, steps, warmup, classes, what):
    """`warmup` untimed calls, then `steps` timed ones with the kernel
    classes' HIP-event timers on (one launch in KTIMER_STRIDE of each class);
    returns (seconds, last result, kernels)."""
    res = None
    def brief(r):
        return r[1] if isinstance(r[1], dict) else f"{len(r[1])} results"

    for i in range(warmup):
        res = step()
        log(f"{what}: warmup {i} done: {brief(res)}")
    gpu.ktimer_stride(KTIMER_STRIDE)
    for k in classes:
        gpu.ktimer_collect(k)  # discard earlier records
        gpu.ktimer_enable(k, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kernels = {}
    for k in classes:
        ms, cnt, nbytes = gpu.ktimer_collect(k)
        gpu.ktimer_enable(k, False)
        if cnt:
            kernels[gpu.KERNEL_CLASSES[k]] = dict(
                launches_timed=cnt, timer_stride=KTIMER_STRIDE,
                total_ms=ms, avg_ms=ms / cnt,
                bytes_per_launch=nbytes / cnt,
                gbs=nbytes / (ms * 1e-3) / 1e9)
    gpu.ktimer_stride(1)
    log(f"{what}: {steps} timed steps in {elapsed:.3f} s; info {brief(res)}")
    return elapsed, res, kernels


def roofline_of(kernels, workload):
    if not kernels:
        return None
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    kd = kernels[dom]
    return {
        "kernel": dom,
        "bound": "hbm",
        "achieved": kd["gbs"],
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": kd["gbs"] / PEAK_HBM_GBS,
        # HBM bytes per launch (rocprofv3 counters), or None
        "traffic": measured_traffic(workload, dom, kd["bytes_per_launch"]),
        "bytes_per_launch": kd["bytes_per_launch"],
        "avg_launch_ms": kd["avg_ms"],
    }


def step_roofline(kernels, elapsed, steps):
    """Whole-step rate: the algorithmic HBM bytes of every timed kernel class
    per step over the wall time of a step (set-up included).  Unlike the
    per-kernel figure it credits kernels that overlap on two streams.  The
    launches of a class are the timed ones times the timer stride (exact to
    within stride - 1 launches of thousands)."""
    if not kernels:
        return None
    nbytes = sum(k["bytes_per_launch"] * k["launches_timed"] * k["timer_stride"]
                 for k in kernels.values()) / steps
    ach = nbytes / (elapsed / steps) / 1e9
    return {"achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": ach / PEAK_HBM_GBS, "bytes_per_step": nbytes,
            "classes": sorted(kernels)}


def accuracy_vs_fp64(torch, gpu, phantom, w):
    """Both forms on one 2048^2 slice of the workload's geometry against the
    FP64 oracle's 100-iteration result (tests/golden/
    mbir2_2048_slice_100it.npz, made by tests/golden/make_mbir2_golden.py
    from the same phantom and noise seeds).  The input here is projected by
    the library's FP32 projector, the fixture's by the oracle's FP64 one;
    the line reports the relative difference of their checksums."""
    path = os.path.join(ROOT, "tests", "golden", "mbir2_2048_slice_100it.npz")
    if not os.path.exists(path):
        return None
    gold = np.load(path)
    N, P, seed, frac, nseed = 2048, 1501, 3000, 0.005, 3007
    ang = phantom.angles(P)
    img = torch.from_numpy(phantom.random_ellipses(N, seed)[None]).cuda()
    sino = gpu.project(img, ang).cpu().numpy().astype(np.float64)
    sino = sino + np.random.default_rng(nseed).normal(0, frac * sino.max(),
                                                      sino.shape)
    chk = np.array([sino.sum(), np.linalg.norm(sino)])
    out = {"case": "one 2048^2 slice, 1501 angles, 100 iterations, "
                   "sigma=100, against the FP64 oracle (sampled 64 x 64 "
                   "grid + whole-image norm)",
           "input_checksum_rel_diff": float(np.max(
               np.abs(chk - gold["sino_checksum"]) /
               np.abs(gold["sino_checksum"])))}
    st = torch.tensor(sino, dtype=torch.float32).cuda()
    k = int(gold["stride"])
    for prec in ("fp32", "mixed"):
        kw = dict(lipschitz_iters=w["lipschitz_iters"]) if prec == "mixed" \
            else {}
        rec, info = gpu.mbir(st, ang, N / 2, int(gold["iterations"]),
                             w["sigma"], precision=prec, **kw)
        im = rec.cpu().numpy()[0].astype(np.float64)
        samp = im[::k, ::k]
        d = samp - gold["sampled"]
        out[prec] = {
            "relL2_sampled": float(np.linalg.norm(d) /
                                   np.linalg.norm(gold["sampled"])),
            "rel_norm_diff": float(abs(np.linalg.norm(im) - gold["l2"]) /
                                   gold["l2"]),
            "step": info["step"],
        }
    out["reference_step"] = float(gold["step"])
    out["within_1e-4"] = all(out[p]["relL2_sampled"] <= 1e-4
                             for p in ("fp32", "mixed"))
    return out


def extra_workloads(torch, gpu, phantom):
    """The other shapes, one line each: 32 x 1024^2 plain FP32 (3 timed
    calls), the 200-stack uneven batch (1 timed call), the autograd step
    (3 timed calls)."""
    out = {}
    classes = sorted(gpu.KERNEL_CLASSES)
    w = dict(WORKLOADS["plain-fp32-32x1024"])
    sino, ang = make_sinogram(torch, gpu, phantom, w)
    el, res, ker = timed_steps(
        torch, gpu, lambda: gpu.mbir(sino, ang, w["npix"] / 2.0,
                                     w["iterations"], w["sigma"]),
        3, 1, classes, "32x1024")
    out["plain-fp32-32x1024"] = {
        "workload": w["label"], "slices_per_s": w["slices"] * 3 / el,
        "ms_per_call": el / 3 * 1e3,
        "roofline": roofline_of(ker, "plain-fp32-32x1024"),
        "step_roofline": step_roofline(ker, el, 3),
        "accepted_iterations": int(res[1]["iterations"])}
    del sino
    w = dict(WORKLOADS["uneven-200"])
    stacks = make_uneven(torch, gpu, phantom, w)
    nsl = sum(int(st[0].shape[0]) for st in stacks)
    el, res, ker = timed_steps(
        torch, gpu, lambda: gpu.mbir_batch(stacks, w["iterations"],
                                           w["sigma"]),
        1, 1, [], "uneven-200")
    recs, infos = res
    out["uneven-200"] = {
        "workload": w["label"], "stacks": len(stacks), "slices": nsl,
        "slices_per_s": nsl / el, "s_per_call": el,
        "all_finite": bool(all(bool(torch.isfinite(r).all()) for r in recs)),
        "accepted_iterations_min": int(min(i["iterations"] for i in infos))}
    del stacks, recs
    from tomocam_amd import autograd as AG
    w = dict(WORKLOADS["autograd-8x512"])
    img = torch.from_numpy(phantom.phantom_stack(
        w["npix"], w["slices"], "ellipses", w["seed_group"])).cuda()
    ang = phantom.angles(w["nproj"])
    sino = gpu.project(img, ang)
    x = torch.zeros_like(img, requires_grad=True)

    def step():
        x.grad = None
        loss = AG.mbir_loss(x, sino, ang, beta=w["beta"], sigma=w["sigma"],
                            p=w["p"])
        loss.backward()
        torch.cuda.synchronize()
        return None, {"loss": float(loss.detach())}

    # a step is ~1 ms: 100 timed steps keep the host's jitter out of it
    nsteps = 100
    el, res, _ = timed_steps(torch, gpu, step, nsteps, 5, [], "autograd-8x512")
    out["autograd-8x512"] = {"workload": w["label"],
                             "slices_per_s": w["slices"] * nsteps / el,
                             "ms_per_step": el / nsteps * 1e3,
                             "steps": nsteps}
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=HEADLINE, choices=sorted(WORKLOADS))
    ap.add_argument("--fp32-steps", type=int, default=5,
                    help="timed calls of the plain FP32 comparison "
                         "(headline workload; at most --steps)")
    ap.add_argument("--cpu-sample-slices", type=int, default=2)
    ap.add_argument("--cpu-sample-iters", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other workload shapes")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP64 fixture comparison")
    ap.add_argument("--no-ktimer", action="store_true",
                    help="do not bracket kernels with HIP events")
    ap.add_argument("--iterations", type=int, default=None,
                    help="override the workload's iteration count (for "
                         "short counter-collection runs)")
    ap.add_argument("--slices", type=int, default=None,
                    help="override the workload's slice count (for "
                         "counter-collection runs; stacks up to 2^26 "
                         "pixels rasterise their phantoms on the host)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device-to-device copy, whose "
                         "FETCH_SIZE / WRITE_SIZE calibrate HBM counters")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on exactly one GPU; multi-GPU "
              "(--gpus > 1) is not supported (slab decomposition is "
              "design-only).", file=sys.stderr)
        sys.exit(2)

    import torch
    from tomocam_amd import gpu, phantom, _lib
    _lib.lib()
    w = dict(WORKLOADS[args.workload])
    if args.iterations is not None:
        w["iterations"] = args.iterations
        w["label"] += f" (iterations overridden to {args.iterations})"
    if args.slices is not None:
        w["slices"] = args.slices
        w["label"] += f" (slices overridden to {args.slices})"
    log(f"workload {args.workload}: {w['label']}")
    if args.copy_calibration:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        c = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.add(a, 1.0, out=c)  # elementwise kernel: 1 GiB in, 1 GiB out
        torch.cuda.synchronize()
        s = float(c.sum())  # reduction: 1 GiB in
        log(f"copy calibration: 1 GiB read + 1 GiB written; 1 GiB read "
            f"({s:.0f})")
        del a, c
    classes = [] if args.no_ktimer else sorted(gpu.KERNEL_CLASSES)
    headline = args.workload == HEADLINE
    kind = w.get("kind", "ellipses")
    extra = {}
    if kind == "autograd":
        from tomocam_amd import autograd as AG
        img = torch.from_numpy(phantom.phantom_stack(
            w["npix"], w["slices"], "ellipses", w["seed_group"])).cuda()
        ang = phantom.angles(w["nproj"])
        sino = gpu.project(img, ang)
        x = torch.zeros_like(img, requires_grad=True)

        def step():
            x.grad = None
            loss = AG.mbir_loss(x, sino, ang, beta=w["beta"],
                                sigma=w["sigma"], p=w["p"])
            loss.backward()
            torch.cuda.synchronize()
            return None, {"loss": float(loss.detach())}
    elif kind == "uneven":
        stacks = make_uneven(torch, gpu, phantom, w)
        w["slices"] = sum(int(st[0].shape[0]) for st in stacks)
        log(f"{len(stacks)} stacks, {w['slices']} slices")

        def step():
            recs, infos = gpu.mbir_batch(stacks, w["iterations"], w["sigma"])
            return recs, infos[-1]
    else:
        sino, ang = make_sinogram(torch, gpu, phantom, w)
        center = w["npix"] / 2.0
        prec = "mixed" if headline else "fp32"
        kw = dict(lipschitz_iters=w["lipschitz_iters"]) if headline else {}

        def step():
            return gpu.mbir(sino, ang, center, w["iterations"], w["sigma"],
                            precision=prec, **kw)

    elapsed, res, kernels = timed_steps(torch, gpu, step, args.steps,
                                        args.warmup, classes, args.workload)
    info = res[1]
    value = w["slices"] * args.steps / elapsed
    roofline = roofline_of(kernels, args.workload)
    valu = None
    cols = "toeplitz_cols_f16" if headline else "toeplitz_cols"
    if cols in kernels and info and info.get("padded"):
        n = int(info["padded"])
        # both forms: FP32 intermediates and spectrum
        valu = valu_roofline(kernels[cols], cols, n, fft_length(n), 8, 4)

    fp32 = None
    if headline:
        out_mixed = res[0]
        fsteps = max(1, min(args.fp32_steps, args.steps))

        def fstep():
            return gpu.mbir(sino, ang, center, w["iterations"], w["sigma"],
                            precision="fp32")

        fel, fres, fker = timed_steps(torch, gpu, fstep, fsteps, 1, classes,
                                      "fp32")
        d = (out_mixed.double() - fres[0].double()).norm()
        fp32 = {
            "slices_per_s": w["slices"] * fsteps / fel,
            "steps": fsteps, "warmup": 1,
            "ms_per_step": fel / fsteps * 1e3,
            "speedup_mixed_vs_fp32": value / (w["slices"] * fsteps / fel),
            "relL2_mixed_vs_fp32": float(d / fres[0].double().norm()),
            "step": fres[1]["step"], "lipschitz": fres[1]["lipschitz"],
            "roofline": roofline_of(fker, "plain-fp32-64x2048"),
            "step_roofline": step_roofline(fker, fel, fsteps),
            "kernels": fker,
        }
        del out_mixed, fres

    accuracy = None
    if headline and not args.no_accuracy:
        log("accuracy against the FP64 fixture ...")
        accuracy = accuracy_vs_fp64(torch, gpu, phantom, w)
        log(f"accuracy: {accuracy}")

    cpu = None
    if (not args.no_cpu_baseline and kind == "ellipses"
            and int(os.environ.get("RANK", "0")) == 0):
        log("cpu baseline sample ...")
        ns = max(1, min(args.cpu_sample_slices, w["slices"]))
        cpu = cpu_baseline(sino[:ns].cpu().numpy(), ang, w, w["slices"], ns,
                           args.cpu_sample_iters)
        log(f"cpu baseline {cpu['value']:.4g} slices/s on {cpu['cores']} "
            f"threads")

    if headline and not args.no_extra:
        del sino
        torch.cuda.empty_cache()
        extra = extra_workloads(torch, gpu, phantom)

    mixed_dtype = "f16 storage (gradient) / f32 compute / f64 line search"
    out = {
        "metric": "MBIR slices/s (mixed-prec vs plain FP32) at relL2≤1e-4 "
                  "vs FP64; HBM GB/s",
        "value": value,
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": mixed_dtype if headline else "f32",
        "data": "synthetic: seeded random-ellipse phantoms projected by the "
                "library's NUFFT projector + Gaussian noise",
        "config": ({
            "workload": w["label"], "stacks": w["stacks"],
            "slices": w["slices"], "iterations": w["iterations"],
            "sigma": w["sigma"],
        } if kind == "uneven" else {
            "workload": w["label"], "slices": w["slices"],
            "npix": w["npix"], "nproj": w["nproj"], "beta": w["beta"],
        } if kind == "autograd" else {
            "workload": w["label"],
            "slices": w["slices"], "npix": w["npix"], "nproj": w["nproj"],
            "iterations": w["iterations"], "sigma": w["sigma"],
            "padded_width": int(info["padded"]),
            "accepted_iterations": int(info["iterations"]),
            "step": info["step"],
        }),
        "roofline": roofline,
        "valu_roofline": valu,
        "step_roofline": step_roofline(kernels, elapsed, args.steps),
        "kernels": kernels,
        "fp32": fp32,
        "accuracy": accuracy,
        "cpu_baseline": cpu,
        "extra": extra or None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
