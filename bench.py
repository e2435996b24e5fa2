"""Benchmark of the image-model hot path on one MI355X.

Workload (N=1): 4096 independent Sersic cutouts (side log-uniform 32-256 px,
n ~ U(0.5, 6), Re log-uniform 2-40 px, q ~ U(0.2, 1), PA ~ U(0, pi), S/N
log-uniform 10-1000, Gaussian PSF sigma ~ U(1, 3) in a 4-sigma box), each
fitted by AstroPhot's Levenberg-Marquardt (psf_mode "full", threshold
supersampling, defaults of fit/lm.py) from a +-10 % perturbed start.  One step
= fitting all 4096 cutouts to convergence with the batched device-resident LM
(HIP-graph ticks).  value = LM fits/s.

Beside it: the plain single-galaxy GPU path (astrophot_amd.fit.LM, one cutout
at a time) on a subset, with the batched results checked against it (matched
chi^2 and parameters); the batched convolution kernel's roofline measured live
with HIP events; the float64 oracle LM on the host cores as the CPU baseline;
and the 512x512 bulge+disk render (render Mpix/s, Jacobian evals/s).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--ngal 4096] [--no-cpu]
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
PEAK_FP32 = 157.3e12      # MI355X dense FP32, FLOP/s
SEED = 20261202

# C1: 512x512 bulge + disk, 51x51 Moffat PSF
C1_SIZE, C1_PSF, BETA, FWHM = 512, 51, 3.0, 3.0


def moffat_rd(beta=BETA, fwhm=FWHM):
    return fwhm / (2 * math.sqrt(2 ** (1 / beta) - 1))


def c1_params(seed=SEED):
    rng = np.random.default_rng(seed)
    cx, cy = C1_SIZE / 2 + rng.uniform(-0.5, 0.5, 2)
    pa = rng.uniform(0, math.pi)
    bulge = {"center": [cx, cy], "q": rng.uniform(0.3, 1.0), "PA": pa, "n": rng.uniform(2, 5),
             "Re": rng.uniform(5, 20), "Ie": 1.0}
    disk = {"center": [cx, cy], "q": rng.uniform(0.3, 1.0), "PA": pa, "Re": rng.uniform(15, 60), "Ie": 0.5}
    return bulge, disk


def c1_model(psf):
    import astrophot_amd as ap

    bulge, disk = c1_params()
    tgt = ap.image.Target_Image(data=np.zeros((C1_SIZE, C1_SIZE)), pixelscale=1.0,
                                psf=ap.image.PSF_Image(data=psf, pixelscale=1.0))
    b = ap.models.AstroPhot_Model(name="bulge", model_type="sersic galaxy model", target=tgt, parameters=bulge,
                                  psf_mode="full")
    d = ap.models.AstroPhot_Model(name="disk", model_type="exponential galaxy model", target=tgt, parameters=disk,
                                  psf_mode="full")
    return ap.models.AstroPhot_Model(name="galaxy", model_type="group model", models=[b, d], target=tgt)


def c1_render_leg(steps=50):
    """render Mpix/s and Jacobian evals/s of the 512x512 bulge+disk model."""
    from astrophot_amd.utils import moffat_psf

    model = c1_model(moffat_psf(BETA, moffat_rd(), C1_PSF, 1.0))
    for _ in range(5):
        model()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        model()
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / steps
    rep = model.parameters.vector_representation()
    for _ in range(2):
        model.jacobian(parameters=rep, as_representation=True)
    torch.cuda.synchronize()
    nj = max(5, steps // 5)
    s.record()
    for _ in range(nj):
        model.jacobian(parameters=rep, as_representation=True)
    e.record()
    torch.cuda.synchronize()
    jms = s.elapsed_time(e) / nj
    return {"workload": "512x512 Sersic bulge + exponential disk, 51x51 Moffat PSF", "render_ms": ms,
            "render_mpix_per_s": C1_SIZE * C1_SIZE / (ms * 1e-3) / 1e6, "jacobian_ms": jms,
            "jacobian_evals_per_s": 1e3 / jms, "jacobian_columns": int(rep.numel())}


def c2_problem(ngal, seed=SEED):
    """Seeded C2 population rendered by the batch renderer, plus noise."""
    from astrophot_amd.batch import BatchFit, Cutout, values_to_rep
    from astrophot_amd.population import perturbed_start, sample_population

    pop = sample_population(ngal, seed)
    cuts = [Cutout(np.broadcast_to(0.0, (s, s)), None, p) for s, p in zip(pop.sides, pop.psfs)]
    dev = "cuda"
    npix = int(sum(int(s) * int(s) for s in pop.sides))
    zeros = torch.zeros(npix, dtype=torch.float32, device=dev)
    bf0 = BatchFit(cuts, data=zeros, variance=torch.ones_like(zeros))
    truth = bf0.render(values_to_rep(pop.truth))
    sigma = torch.repeat_interleave(torch.tensor(pop.sky_sigma, dtype=torch.float32, device=dev),
                                    torch.tensor(pop.sides.astype(np.int64) ** 2, device=dev))
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 1)
    data = truth + sigma * torch.randn(npix, generator=gen, device=dev)
    var = sigma.double() ** 2
    del bf0
    start = perturbed_start(pop.truth, seed + 2)
    return pop, cuts, data, var, start


def plain_fits(pop, cuts, data, var, start, idx):
    """The single-galaxy GPU path (astrophot_amd.fit.LM) on cutouts ``idx``."""
    import astrophot_amd as ap

    offs = np.concatenate([[0], np.cumsum(pop.sides.astype(np.int64) ** 2)])
    out = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in idx:
        s = int(pop.sides[k])
        d = data[offs[k]:offs[k + 1]].view(s, s)
        v = var[offs[k]:offs[k + 1]].view(s, s)
        tgt = ap.image.Target_Image(data=d, pixelscale=1.0, variance=v,
                                    psf=ap.image.PSF_Image(data=cuts[k].psf, pixelscale=1.0))
        p0 = start[k]
        m = ap.models.AstroPhot_Model(name=f"c{k}", model_type="sersic galaxy model", target=tgt, psf_mode="full",
                                      parameters={"center": [p0[0], p0[1]], "q": p0[2], "PA": p0[3], "n": p0[4],
                                                  "Re": p0[5], "Ie": p0[6]})
        lm = ap.fit.LM(m, initial_state=m.parameters.vector_representation()).fit()
        out.append((m.parameters.vector_values().numpy(), min(lm.loss_history), lm.message))
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def cpu_baseline(pop, data, var, start, budget_s=20.0):
    """float64 oracle LM (the reference algorithm restated, FFT convolution)
    on the host cores, cutout after cutout until ~budget_s."""
    from oracle import hotpath as H

    ncores = len(os.sched_getaffinity(0))
    torch.set_num_threads(ncores)
    offs = np.concatenate([[0], np.cumsum(pop.sides.astype(np.int64) ** 2)])
    dh, vh = data.double().cpu(), var.double().cpu()
    n, t0, npx = 0, time.perf_counter(), 0
    while time.perf_counter() - t0 < budget_s and n < len(pop.sides):
        k = n
        s = int(pop.sides[k])
        p0 = start[k]
        p = {"center": [p0[0], p0[1]], "q": p0[2], "PA": p0[3], "n": p0[4], "Re": p0[5], "Ie": p0[6]}
        om = H.Model([H.Component("sersic", H.Grid(s, s), p, psf=torch.tensor(pop.psfs[k]), psf_mode="full")])
        H.LM(om, dh[offs[k]:offs[k + 1]], weight=1.0 / vh[offs[k]:offs[k + 1]]).fit()
        n += 1
        npx += s * s
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "fits/s", "cores": ncores, "kind": "port",
            "sample": f"the first {n} cutouts of the seeded population fitted one after another by the float64 "
                      f"oracle LM (torch CPU, FFT conv, {npx} px in total) in {dt:.1f} s; the population's mean "
                      f"cutout is {float(np.mean(pop.sides.astype(float) ** 2)):.0f} px"}


def main():
    ap_ = argparse.ArgumentParser()
    ap_.add_argument("--gpus", type=int, default=1)
    ap_.add_argument("--steps", type=int, default=3)
    ap_.add_argument("--warmup", type=int, default=1)
    ap_.add_argument("--ngal", type=int, default=4096)
    ap_.add_argument("--plain", type=int, default=24, help="cutouts fitted by the plain single-galaxy path")
    ap_.add_argument("--no-cpu", action="store_true")
    ap_.add_argument("--no-c1", action="store_true")
    args = ap_.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the fits are independent "
              f"and would shard by cutout, see DESIGN.md)", file=sys.stderr)
        return 2

    from astrophot_amd.batch import BatchFit, values_to_rep

    pop, cuts, data, var, start = c2_problem(args.ngal)
    bf = BatchFit(cuts, data=data, variance=var)
    rep0 = torch.tensor(values_to_rep(start), dtype=torch.float64, device="cuda")
    for _ in range(args.warmup):
        res = bf.fit(rep0)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    for _ in range(args.steps):
        res = bf.fit(rep0)
    e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = s.elapsed_time(e) / args.steps
    fits_per_s = args.ngal / (ms * 1e-3)
    status = {int(k): int(v) for k, v in zip(*np.unique(res["status"], return_counts=True))}

    # dominant kernel: the Jacobian-stage convolution (value with K, dK/dcx,
    # dK/dcy and the five partial planes with K: 8 kernel passes per output
    # pixel) of every cutout, timed live with HIP events on the launch stream;
    # the value convolution alone is reported beside it
    truth_rep = torch.tensor(values_to_rep(pop.truth), dtype=torch.float64, device="cuda")
    out = torch.zeros(bf.npix, dtype=torch.float32, device="cuda")

    def timed(fn, n=10, skip=2):
        evs = []
        for _ in range(n):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            evs.append((a, b))
        torch.cuda.synchronize()
        return 1e3 * float(np.mean([a.elapsed_time(b) for a, b in evs[skip:]]))

    bf.jacobian(truth_rep, out=out, parts=1)
    jac_us = timed(lambda: bf.jacobian(truth_rep, out=out, parts=2))
    bf.render(truth_rep, out=out, parts=1)
    val_us = timed(lambda: bf.render(truth_rep, out=out, parts=2))
    vflops = bf.conv_flops()
    jflops = 8.0 * vflops
    achieved = jflops / (jac_us * 1e-6)
    roof = {"bound": "valu", "kernel": "conv_kernel<1> (batched Jacobian-stage convolution, all 4096 cutouts)",
            "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32, "traffic": None, "algorithmic_per_launch": jflops,
            "avg_launch_us": jac_us,
            "value_conv": {"kernel": "conv_kernel<0>", "algorithmic_per_launch": vflops, "avg_launch_us": val_us,
                           "achieved": vflops / (val_us * 1e-6) / 1e12}}

    # plain single-galaxy GPU path on a subset, and agreement with the batch
    idx = list(range(min(args.plain, args.ngal)))
    plain, t_plain = plain_fits(pop, cuts, data, var, start, idx)
    plain_fps = len(idx) / t_plain
    agree = {"fits": len(idx), "matched": 0, "param_rtol": 1e-4, "chi2_rtol": 1e-6,
             "max_param_rel_matched": 0.0, "max_chi2_rel_matched": 0.0, "unmatched": []}
    for k, (v, c2, msg) in zip(idx, plain):
        perr = float(np.max(np.abs(res["values"][k] - v) / np.maximum(np.abs(v), 1.0)))
        cerr = (float(res["chi2"][k]) - c2) / c2
        if perr <= agree["param_rtol"] and abs(cerr) <= agree["chi2_rtol"]:
            agree["matched"] += 1
            agree["max_param_rel_matched"] = max(agree["max_param_rel_matched"], perr)
            agree["max_chi2_rel_matched"] = max(agree["max_chi2_rel_matched"], abs(cerr))
        else:
            # signed: negative means the batched fit ended at the lower chi^2
            agree["unmatched"].append({"cutout": k, "param_rel": perr, "chi2_rel_batched_minus_plain": cerr,
                                       "q_fit": float(v[2]), "plain_message": msg[:40]})

    line = {
        "metric": METRIC, "value": fits_per_s, "unit": "fits/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "C2: 4096 Sersic cutouts (side 32-256 px log-uniform), per-cutout Gaussian PSF, "
                               "batched device-resident LM from a +-10% start; one step = all cutouts fitted",
                   "ngal": args.ngal, "total_pixels": int(bf.npix), "seed": SEED, "init": "perturbed truth +-10%",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "batched": {"fits_per_s": fits_per_s, "ticks": int(res["ticks"]), "status_counts": status,
                    "mean_iterations": float(np.mean(res["iterations"]))},
        "plain_gpu": {"fits_per_s": plain_fps, "cutouts": len(idx), "seconds": t_plain},
        "batched_vs_plain": fits_per_s / plain_fps,
        "batched_vs_plain_agreement": agree,
        "wall_s_timed_region": wall,
    }
    if not args.no_c1:
        line["c1"] = c1_render_leg()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(pop, data, var, start)
        line["gpu_vs_cpu"] = fits_per_s / line["cpu_baseline"]["value"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
