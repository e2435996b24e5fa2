"""Benchmark of the image-model hot path on one MI355X.

Workload (N=1): one 512 x 512 image, Sersic bulge + exponential disk, 51 x 51
Moffat PSF (beta = 3, FWHM = 3 px), threshold-selected adaptive supersampling
of the centre (AstroPhot's defaults), PSF convolution.  One step = one
``model.sample()`` of the whole group (render + supersample + direct
convolution of both components).  value = rendered output Mpix/s.

Also reported: Jacobian evaluations/s of the same model (13 columns), a plain
LM fit of it, the dominant kernel's roofline (live HIP-event timing on the
launch stream), and the float64 oracle timed on the host cores (baseline).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LM fits/s, batched+learned-init vs plain, at matched chi2; render Mpix/s"
PEAK_FP32 = 157.3e12      # MI355X dense FP32 (TFLOP/s published figure)
PEAK_HBM = 8.0e12         # MI355X HBM3E bytes/s
SEED = 20261202
SIZE = 512
PSF_SIDE = 51
BETA = 3.0
FWHM = 3.0


def c1_params(seed=SEED):
    """Seeded bulge + disk parameters (numpy PCG64)."""
    rng = np.random.default_rng(seed)
    cx, cy = SIZE / 2 + rng.uniform(-0.5, 0.5, 2)
    pa = rng.uniform(0, math.pi)
    bulge = {"center": [cx, cy], "q": rng.uniform(0.3, 1.0), "PA": pa, "n": rng.uniform(2, 5),
             "Re": rng.uniform(5, 20), "Ie": 1.0}
    disk = {"center": [cx, cy], "q": rng.uniform(0.3, 1.0), "PA": pa, "Re": rng.uniform(15, 60), "Ie": 0.5}
    return bulge, disk


def moffat_rd(beta=BETA, fwhm=FWHM):
    return fwhm / (2 * math.sqrt(2 ** (1 / beta) - 1))


def build_gpu_model(psf, data=None, var=None):
    import astrophot_amd as ap

    bulge, disk = c1_params()
    tgt = ap.image.Target_Image(data=np.zeros((SIZE, SIZE)) if data is None else data, pixelscale=1.0,
                                variance=var, psf=ap.image.PSF_Image(data=psf, pixelscale=1.0))
    b = ap.models.AstroPhot_Model(name="bulge", model_type="sersic galaxy model", target=tgt, parameters=bulge,
                                  psf_mode="full")
    d = ap.models.AstroPhot_Model(name="disk", model_type="exponential galaxy model", target=tgt,
                                  parameters=disk, psf_mode="full")
    return ap.models.AstroPhot_Model(name="galaxy", model_type="group model", models=[b, d], target=tgt)


def cpu_baseline(psf, budget_s=20.0):
    """float64 oracle (the reference algorithm restated, FFT convolution) on the
    host cores: forward renders of the same model until ~budget_s."""
    from oracle import hotpath as H

    ncores = len(os.sched_getaffinity(0))
    torch.set_num_threads(ncores)
    bulge, disk = c1_params()
    g = H.Grid(SIZE, SIZE, 1.0)
    pt = torch.tensor(psf, dtype=torch.float64)
    m = H.Model([H.Component("sersic", g.copy(), bulge, psf=pt, psf_mode="full"),
                 H.Component("exponential", g.copy(), disk, psf=pt, psf_mode="full")], window=g.copy())
    rep = m.vector_representation()
    m.forward(rep)  # warm-up
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s or n < 2:
        m.forward(rep)
        n += 1
    dt = (time.perf_counter() - t0) / n
    return {"value": SIZE * SIZE / dt / 1e6, "unit": "Mpix/s", "cores": ncores, "kind": "port",
            "sample": f"{n} float64 forward renders of the 512x512 bulge+disk model (oracle, torch CPU, FFT conv), "
                      f"{dt:.3f} s each"}


def main():
    ap_ = argparse.ArgumentParser()
    ap_.add_argument("--gpus", type=int, default=1)
    ap_.add_argument("--steps", type=int, default=50)
    ap_.add_argument("--warmup", type=int, default=5)
    ap_.add_argument("--no-cpu", action="store_true")
    ap_.add_argument("--no-fit", action="store_true")
    args = ap_.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the fits of this workload are independent; see DESIGN.md for the sharding plan)",
              file=sys.stderr)
        return 2

    import astrophot_amd as ap
    from astrophot_amd import _engine
    from astrophot_amd.utils import moffat_psf

    psf = moffat_psf(BETA, moffat_rd(), PSF_SIDE, 1.0)
    model = build_gpu_model(psf)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        model()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(args.steps):
        model()
    e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = s.elapsed_time(e) / args.steps
    mpix = SIZE * SIZE / (ms * 1e-3) / 1e6

    # dominant kernel: live event timing of every launch over a second timed region
    _engine.TIMING = {}
    for _ in range(args.steps):
        model()
    torch.cuda.synchronize()
    timing, _engine.TIMING = _engine.TIMING, None
    kern = {}
    for name, evs in timing.items():
        t = [a.elapsed_time(b) for a, b, _ in evs]
        kern[name] = {"launches": len(evs), "avg_us": 1e3 * float(np.mean(t)),
                      "work_per_launch": float(np.mean([w for _, _, w in evs]))}
    conv = kern["convolve"]
    achieved = conv["work_per_launch"] / (conv["avg_us"] * 1e-6) / 1e12
    roof = {"bound": "valu", "kernel": "convolve_kernel", "achieved": achieved, "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s", "frac": achieved * 1e12 / PEAK_FP32, "traffic": None,
            "algorithmic_per_launch": conv["work_per_launch"], "avg_launch_us": conv["avg_us"]}

    # Jacobian evaluations/s (13 columns, representation units)
    rep = model.parameters.vector_representation()
    for _ in range(2):
        model.jacobian(parameters=rep, as_representation=True)
    torch.cuda.synchronize()
    nj = max(3, args.steps // 5)
    s.record()
    for _ in range(nj):
        model.jacobian(parameters=rep, as_representation=True)
    e.record()
    torch.cuda.synchronize()
    jac_ms = s.elapsed_time(e) / nj

    fit = None
    if not args.no_fit:
        from oracle import hotpath as H  # truth image for the fit only
        truth = model().data.double().cpu().numpy()
        rng = np.random.default_rng(SEED + 1)
        var = truth / 5.0 + 0.01
        data = truth + rng.normal(size=truth.shape) * np.sqrt(var)
        fm = build_gpu_model(psf, data=data, var=var)
        start = fm.parameters.vector_representation().clone()
        start = start * (1 + torch.as_tensor(rng.uniform(-0.02, 0.02, len(start))))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = ap.fit.LM(fm, initial_state=start).fit()
        torch.cuda.synchronize()
        fit = {"fit_s": time.perf_counter() - t1, "iterations": len(res.loss_history) - 1,
               "chi2_start": res.loss_history[0], "chi2_end": res.loss_history[-1], "message": res.message,
               "forwards": res.n_forward, "jacobians": res.n_jacobian}

    line = {
        "metric": METRIC, "value": mpix, "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "C1: 512x512 Sersic bulge + exponential disk, 51x51 Moffat PSF (beta=3, FWHM=3), "
                               "threshold-adaptive supersampling, one model.sample() per step",
                   "image": [SIZE, SIZE], "psf": PSF_SIDE, "seed": SEED, "parallelism": "single GPU"},
        "roofline": roof,
        "kernels": kern,
        "jacobian_evals_per_s": 1e3 / jac_ms,
        "jacobian_ms": jac_ms,
        "lm_fit": fit,
        "wall_s_timed_region": wall,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(psf)
        line["gpu_vs_cpu"] = mpix / line["cpu_baseline"]["value"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
