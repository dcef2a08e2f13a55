#!/usr/bin/env python3
"""MBIR throughput benchmark (slices/s) on one MI355X.

Workload (default): plain FP32 Toeplitz-gradient MBIR on 32 slices of
1024 x 1024, 720 angles uniform in [0, pi) x 1024 rays, centre 512, seeded
30-random-ellipse phantoms projected with the library's NUFFT projector plus
1% Gaussian noise, 100 iterations, qGGMRF p = 1.2, q = 2, c = 0.001,
sigma = 0.1.  One step = one full mbir2 call (normalise, FBP initial guess,
padding, P^T y, PSF, step size, 100 iterations, crop) on a sinogram already
resident in HBM.

Prints one JSON line with the throughput, the roofline of the dominant
kernel (HIP-event timed inside the timed region), a CPU baseline (the FP32
CPU restatement of the same algorithm on a bounded sample) and the
mixed-precision form against plain FP32 on 64 x 2048^2 (one timed call each,
sigma = 100).
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
    "mixed-size-64x2048": dict(
        slices=64, npix=2048, nproj=1501, iterations=100, sigma=0.1, p=1.2,
        noise=0.005, seed_group=3,
        label="64 slices 2048x2048, 1501 angles, 100 iterations (plain FP32)"),
    # one forward + backward of the differentiable loss
    # 0.5 ||P x - y||^2 + beta qGGMRF(x) (projector with the hand-written
    # adjoint backward, free-boundary prior with its exact gradient kernel)
    "autograd-8x512": dict(
        kind="autograd", slices=8, npix=512, nproj=360, iterations=1,
        sigma=0.1, p=1.2, beta=1e-3, noise=0.01, seed_group=2,
        label="differentiation: forward + backward of 0.5||Px-y||^2 + "
              "1e-3 qGGMRF(x) on 8 slices 512x512, 360 angles x 512 rays"),
    # the mixed-precision comparison: sigma = 100 (the reference's default
    # smoothness 0.01), where the iteration converges and relative errors
    # between precisions are meaningful
    "mixed-prec-64x2048": dict(
        slices=64, npix=2048, nproj=1501, iterations=100, sigma=100.0, p=1.2,
        noise=0.005, seed_group=3, precision="mixed",
        label="mixed precision (FP16 gradient and Toeplitz intermediates, "
              "FP32 FFTs and state, FP64 dots, FP32 refresh every 5): 64 "
              "slices 2048x2048, 1501 angles x 2048 rays, 100 iterations, "
              "qGGMRF sigma=100 p=1.2"),
}

PEAK_HBM_GBS = 8000.0

# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# counter runs (calibrated on a copy kernel) for the default workload
TRAFFIC_FILE = os.path.join(ROOT, "profiles",
                            "mbir_fp32_32x1024_hbm_traffic.json")


PEAK_VALU_TFLOPS = 78.6  # FP32 vector FMA without packed math (2 flop/lane/clk)


def valu_roofline(kd, n, L, half=False):
    """The column pass against the FP32 vector-issue peak.  Its time goes to
    FFT arithmetic (VALU bound), so next to the HBM roofline it is also
    rated by algorithmic flops: 5 L log2(L) per length-L complex transform,
    two per column (forward and inverse), plus the 2L of the real spectrum
    multiply; H = L/2 + 1 columns per slice.  The slices per launch follow
    from the launch's algorithmic bytes (ns * 2 * 8nH + 4HL in FP32,
    ns * 2 * 4nH + 2HL with FP16 intermediates)."""
    import math
    H = L // 2 + 1
    if half:
        ns = (kd["bytes_per_launch"] - 2.0 * H * L) / (8.0 * n * H)
    else:
        ns = (kd["bytes_per_launch"] - 4.0 * H * L) / (16.0 * n * H)
    flops = ns * H * (2 * 5.0 * L * math.log2(L) + 2.0 * L)
    ach = flops / (kd["avg_ms"] * 1e-3) / 1e12
    return {"kernel": "toeplitz_cols", "bound": "valu", "achieved": ach,
            "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_VALU_TFLOPS, "flops_per_launch": flops,
            "slices_per_launch": ns, "fft_length": L}


def measured_traffic(workload, kernel_class):
    if workload != "plain-fp32-32x1024" or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        t = json.load(f)
    c = t["classes"].get(kernel_class)
    return None if c is None else c["traffic_bytes"]


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


def cpu_baseline(sino_host, ang, w, sample_slices, sample_iters):
    from oracle import cpu_port
    r = cpu_port.sample_rate(sino_host, ang, w["npix"] / 2.0, w["iterations"],
                             w["sigma"], sample_slices, sample_iters)
    return {
        "value": r["slices_per_s"],
        "unit": "slices/s",
        "cores": r["threads"],
        "kind": "port",
        "sample": (f"FP32 CPU restatement of mbir2 (4 Toeplitz products per "
                   f"attempt, scipy.fft + C/OpenMP): full setup and "
                   f"{sample_iters} iterations on {r['sample_slices']} of the "
                   f"{w['slices']} slices ({r['setup_s']:.1f} s setup, "
                   f"{r['iter_s']:.1f} s iterations), extrapolated linearly to "
                   f"{w['iterations']} iterations per slice"),
    }


def mixed_compare(torch, gpu, phantom, warmup):
    """Mixed precision against plain FP32 on the 64 x 2048^2 workload: one
    timed call of each (after `warmup` short calls priming the plans), their
    slices/s and the relative L2 difference of the reconstructions."""
    w = dict(WORKLOADS["mixed-prec-64x2048"])
    sino, ang = make_sinogram(torch, gpu, phantom, w)
    center = w["npix"] / 2.0
    res, recs = {}, {}
    for prec in ("fp32", "mixed"):
        for _ in range(warmup):
            gpu.mbir(sino, ang, center, 2, w["sigma"], precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rec, info = gpu.mbir(sino, ang, center, w["iterations"], w["sigma"],
                             precision=prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[prec] = (w["slices"] / dt, info)
        recs[prec] = rec
        log(f"{prec}: {w['slices'] / dt:.3f} slices/s ({dt:.2f} s)")
    d = (recs["mixed"].double() - recs["fp32"].double()).norm()
    rel = float(d / recs["fp32"].double().norm())
    return {
        "workload": w["label"],
        "fp32_slices_per_s": res["fp32"][0],
        "mixed_slices_per_s": res["mixed"][0],
        "speedup": res["mixed"][0] / res["fp32"][0],
        "relL2_mixed_vs_fp32": rel,
        "saturated_slice_attempts": int(res["mixed"][1]["saturated"]),
        "accepted_iterations": int(res["mixed"][1]["iterations"]),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plain-fp32-32x1024",
                    choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample-slices", type=int, default=8)
    ap.add_argument("--cpu-sample-iters", type=int, default=12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision vs FP32 comparison on the "
                         "64 x 2048^2 workload")
    ap.add_argument("--no-ktimer", action="store_true",
                    help="do not bracket kernels with HIP events")
    ap.add_argument("--iterations", type=int, default=None,
                    help="override the workload's iteration count (for "
                         "short counter-collection runs)")
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
    log(f"workload {args.workload}: {w['label']}")
    if args.copy_calibration:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        c = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.add(a, 1.0, out=c)  # elementwise kernel: 1 GiB in, 1 GiB out
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del a, c
    uneven = w.get("kind") == "uneven"
    autograd = w.get("kind") == "autograd"
    if autograd:
        from tomocam_amd import autograd as AG
        img = torch.from_numpy(phantom.phantom_stack(
            w["npix"], w["slices"], "ellipses", w["seed_group"])).cuda()
        ang = phantom.angles(w["nproj"])
        sino = gpu.project(img, ang)
        x = torch.zeros_like(img, requires_grad=True)
        center = w["npix"] / 2.0

        def step():
            x.grad = None
            loss = AG.mbir_loss(x, sino, ang, beta=w["beta"],
                                sigma=w["sigma"], p=w["p"])
            loss.backward()
            torch.cuda.synchronize()
            return None, {"loss": float(loss)}
    elif uneven:
        stacks = make_uneven(torch, gpu, phantom, w)
        w["slices"] = sum(int(st[0].shape[0]) for st in stacks)
        log(f"{len(stacks)} stacks, {w['slices']} slices")

        def step():
            recs, infos = gpu.mbir_batch(stacks, w["iterations"], w["sigma"])
            return recs, infos[-1]
    else:
        sino, ang = make_sinogram(torch, gpu, phantom, w)
        center = w["npix"] / 2.0

        prec = w.get("precision", "fp32")

        def step():
            return gpu.mbir(sino, ang, center, w["iterations"], w["sigma"],
                            verbose=False, precision=prec)

    for i in range(args.warmup):
        _, info = step()
        log(f"warmup {i} done: {info}")
    classes = sorted(gpu.KERNEL_CLASSES)
    if not args.no_ktimer:
        for k in classes:
            gpu.ktimer_collect(k)  # discard warm-up records
            gpu.ktimer_enable(k, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, info = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    log(f"timed {args.steps} steps in {elapsed:.3f} s; info {info}")

    kernels = {}
    if not args.no_ktimer:
        for k in classes:
            ms, cnt, nbytes = gpu.ktimer_collect(k)
            gpu.ktimer_enable(k, False)
            if cnt:
                kernels[gpu.KERNEL_CLASSES[k]] = dict(
                    launches=cnt, total_ms=ms, avg_ms=ms / cnt,
                    bytes_per_launch=nbytes / cnt,
                    gbs=nbytes / (ms * 1e-3) / 1e9)
    roofline = None
    if kernels:
        dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
        kd = kernels[dom]
        roofline = {
            "kernel": dom,
            "bound": "hbm",
            "achieved": kd["gbs"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": kd["gbs"] / PEAK_HBM_GBS,
            # HBM bytes per launch (rocprofv3 counters), or None
            "traffic": measured_traffic(args.workload, dom),
            "bytes_per_launch": kd["bytes_per_launch"],
            "avg_launch_ms": kd["avg_ms"],
        }
    valu = None
    if ("toeplitz_cols" in kernels and not uneven and not autograd
            and info and info.get("padded")):
        import ctypes as C
        n = int(info["padded"])
        L = C.c_int()
        _lib.call("tc_toeplitz_fft_length", n, C.byref(L))
        valu = valu_roofline(kernels["toeplitz_cols"], n, L.value,
                             half=w.get("precision") == "mixed")

    cpu = None
    if (not args.no_cpu_baseline and not uneven and not autograd
            and int(os.environ.get("RANK", "0")) == 0):
        log("cpu baseline sample ...")
        cpu = cpu_baseline(sino.cpu().numpy(), ang, w, args.cpu_sample_slices,
                           args.cpu_sample_iters)
        log(f"cpu baseline {cpu['value']:.4g} slices/s on {cpu['cores']} threads")

    mixed = None
    if not args.no_mixed and args.workload == "plain-fp32-32x1024":
        if not uneven:
            del sino
        torch.cuda.empty_cache()
        log("mixed precision vs FP32 on 64 x 2048^2 ...")
        mixed = mixed_compare(torch, gpu, phantom, max(args.warmup, 1))

    value = w["slices"] * args.steps / elapsed
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
        "dtype": "f16 storage / f32 compute" if w.get("precision") == "mixed"
                 else "f32",
        "data": "synthetic: seeded random-ellipse phantoms projected by the "
                "library's NUFFT projector + Gaussian noise",
        "config": ({
            "workload": w["label"], "stacks": w["stacks"],
            "slices": w["slices"], "iterations": w["iterations"],
            "sigma": w["sigma"],
        } if uneven else {
            "workload": w["label"], "slices": w["slices"],
            "npix": w["npix"], "nproj": w["nproj"], "beta": w["beta"],
        } if autograd else {
            "workload": w["label"],
            "slices": w["slices"], "npix": w["npix"], "nproj": w["nproj"],
            "iterations": w["iterations"], "sigma": w["sigma"],
            "padded_width": int(info["padded"]),
            "accepted_iterations": int(info["iterations"]),
        }),
        "roofline": roofline,
        "valu_roofline": valu,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
