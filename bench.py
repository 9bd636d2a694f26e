#!/usr/bin/env python3
"""Burst-merge benchmark on one MI355X.

Metric: burst frames*MPix/s merged = N * (W*H / 1e6) / t, where t is the time of one full merge (global
pre-alignment, tracking images, tile alignment + ShiftMinimizer, kernel parameters, Lucas-Kanade flow,
robustness, accumulation, final weighting) with the raw u16 frames already resident in HBM. The headline
workload is the full-size FP32 plain configuration: 20-frame 6000x4000 RGGB 14-bit synthetic burst with global
shifts U(-64, 64) px, 4-level pyramid (levels 1/2/4/8, tiles 64/32/32/16, max shift 2/4/4/4), merge at 1x.

Prints ONE JSON line. Extra diagnostics go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

C1_LEVELS = [(1, 64, 2), (2, 32, 4), (4, 32, 4), (8, 16, 4)]
CONFIGS = {
    # full-size FP32 plain configuration (the metric's workload); global shifts up to +-64 px per frame, so up to
    # 128 px between two frames: the pre-alignment searches +-128 px
    "full_fp32": dict(w=6000, h=4000, n=20, levels=C1_LEVELS, gmax=64.0, flow=1.5, occ=0.02, prealign=128),
    # small CPU-runnable configuration
    "small": dict(w=1024, h=1024, n=8, levels=[(1, 32, 2), (2, 32, 4), (4, 32, 4)], gmax=16.0, flow=1.5, occ=0.0,
                  prealign=0),
    # precision variants of the tile search at their own size
    "c2": dict(w=4032, h=3024, n=30, levels=[(1, 64, 2), (2, 32, 4), (4, 32, 4), (8, 16, 4)], gmax=16.0, flow=1.5,
               occ=0.02, prealign=0),
    # combined largest: 50 frames, global shifts up to +-128 px, local flow up to 3 px, 5 % occluders
    "c4": dict(w=6000, h=4000, n=50, levels=C1_LEVELS, gmax=128.0, flow=3.0, occ=0.05, prealign=256),
}

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_MFMA_TFLOPS = {"bf16x3": 2500.0, "fp16x3": 2500.0, "fp32": 157.3}
XC_CODES = {"bf16x3": 1, "fp16x3": 2, "fp32": 3}
KERNEL_TRAFFIC = ROOT / "profiles" / "kernel_traffic_full_fp32.json"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def track_flops(cfg, level_idx, pairs, products=1):
    """Algorithmic flops of one tile-search launch: the cross terms sum_x r(x) t(x+s) (one multiply-add = 2 flops
    per interior pixel and lag), per tile and pair, times the products per term (3 for the split forms)."""
    l, t, m = cfg["levels"][level_idx]
    lw, lh = cfg["w"] // l, cfg["h"] // l
    tiles = ((lw - 2 * m) // t) * ((lh - 2 * m) // t)
    return 2.0 * t * t * (2 * m + 1) ** 2 * tiles * pairs * products


def host_info():
    model = platform.processor() or "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)), "cpu_model": model}


def cpu_baseline(cfg):
    """The oracle (reference semantics, FP32) on a crop of the same burst, timed on this host with every usable
    core."""
    import numpy as np

    import imagestackalignator_amd as isa
    from oracle import oracle as O

    info = host_info()
    # every core this process may use, within the host's thread budget for one GPU (OMP_NUM_THREADS when set)
    threads = min(info["usable_cores"], int(os.environ.get("OMP_NUM_THREADS", info["usable_cores"])))
    O.set_threads(threads)
    cw, ch = 2048, 1536
    scfg = isa.default_synth(seed=7, width=cw, height=ch, frames=cfg["n"], global_shift_max=cfg["gmax"] / 8,
                             flow_max=cfg["flow"], occluded_fraction=cfg["occ"])
    raw_t, _ = isa.synth_burst(scfg)
    import torch

    torch.cuda.synchronize()
    raw = raw_t.cpu().numpy().view(np.uint16)
    params = isa.default_params(levels=cfg["levels"])
    t0 = time.perf_counter()
    O.merge(params, raw)
    dt = time.perf_counter() - t0
    return {
        "value": round(cfg["n"] * (cw * ch / 1e6) / dt, 4),
        "unit": "burst frames*MPix/s",
        "cores": threads,
        "kind": "port",
        "nproc": info["nproc"],
        "cpu_model": info["cpu_model"],
        "sample": f"{cfg['n']}-frame {cw}x{ch} crop of the same workload (same levels/tiles, shifts /8 so no "
                  f"pre-alignment is needed), oracle FP32 pipeline with std::thread over {threads} threads "
                  f"(usable cores {info['usable_cores']}, thread budget OMP_NUM_THREADS), {dt:.1f} s",
    }


# Compulsory HBM bytes per pixel of each merger stage (each input read once, each output written once;
# u16 raw, f32 intermediates): per frame (f), per non-reference frame (g) and once per merge (o).
STAGE_BYTES = {
    "tracking_images": dict(f=30.0),  # debayer + grey 6, high-pass 8, two Gaussians 8 + 8
    "prepare_accumulation": dict(o=16.0),  # structure tensor, smoothing and kernel parameters
    "lucas_kanade": dict(g=80.0),  # flow from tiles 8, three LK iterations 24 each
    "robustness": dict(f=31.0),  # half-res debayer 5, mask + erosion 26
    "accumulate": dict(f=14.0, o=36.0),
    "final_weighting": dict(o=48.0),
}


def stage_rooflines(stages, n, w, h, levels):
    """Achieved rate of every stage against the MI355X peak that bounds it: HBM bytes for the image stages,
    FP64 flops of the tile cross terms for track_patches (FP64 VALU), time only for the latency-bound rest."""
    P = w * h
    out = {}
    for k, ms in stages.items():
        if k in STAGE_BYTES:
            c = STAGE_BYTES[k]
            b = P * (c.get("f", 0.0) * n + c.get("g", 0.0) * (n - 1) + c.get("o", 0.0))
            gbs = b / 1e6 / ms if ms > 0 else 0.0
            out[k] = {"ms": round(ms, 3), "bound": "hbm", "algorithmic_gb": round(b / 1e9, 3),
                      "achieved_gbs": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4)}
        elif k == "track_patches":
            gf = sum(l["track_gflop"] for l in levels)
            tf = gf / ms if ms > 0 else 0.0
            out[k] = {"ms": round(ms, 3), "bound": "fp64-valu", "algorithmic_gflop": round(gf, 1),
                      "achieved_tflops": round(tf, 2), "frac": round(tf / PEAK_FP64_TFLOPS, 4)}
        else:
            out[k] = {"ms": round(ms, 3), "bound": "latency"}
    return out


def timed_runs(torch, fn, steps, warmup):
    """Warm-up, then `steps` runs each bracketed by device synchronisation: (mean ms, median ms, all)."""
    for i in range(warmup):
        fn()
        torch.cuda.synchronize()
    times = []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1000.0)
    total = (time.perf_counter() - t_all) * 1000.0
    return total / max(1, steps), statistics.median(times), times


def psnr_t(torch, x, y, b=4):
    import numpy as np

    d = (x - y)[b:-b, b:-b].double()
    fin = torch.isfinite(d)
    mse = float(torch.mean(d[fin] ** 2))
    return 99.0 if mse == 0 else round(10 * np.log10(1.0 / mse), 3)


def color_export_bench(isa, torch, img, steps):
    """The step after the merge on the merged image: isa_color_pipeline to 16-bit RGB (the TIFF's pixels) with a
    256-level tone curve, timed with HIP events on the current stream; 12 bytes read and 6 written per pixel."""
    from imagestackalignator_amd import export

    h, w = int(img.shape[0]), int(img.shape[1])
    lx = torch.arange(256, dtype=torch.float32, device="cuda") / 255.0
    ly = (lx + 0.1 * torch.sin(2 * torch.pi * lx) * lx * (1 - lx)).contiguous()
    m1 = [[0.71, 0.21, 0.05, 0.0], [0.12, 0.88, -0.03, 0.0], [0.01, -0.09, 1.07, 0.0]]
    m2 = [[1.82, -0.51, -0.21, 0.0], [-0.19, 1.22, -0.03, 0.0], [0.01, -0.13, 1.12, 0.0]]
    out16 = torch.empty((h, w, 3), dtype=torch.int16, device="cuda")
    run = lambda: export.color_pipeline(img, [1.0, 1.0, 1.0], m1, lx, ly, m2, 0.0, out=out16)
    run()
    times = []
    for _ in range(max(3, steps)):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        run()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    ms = statistics.median(times)
    gb = 18.0 * w * h / 1e9
    return {"workload": f"{w}x{h} merged image -> 16-bit RGB (threshold, 2 colour twists, 256-level cubic tone "
                        f"curve, exposure, sRGB gamma, round to nearest even)",
            "ms": round(ms, 4), "algorithmic_gb": round(gb, 3), "achieved_gbs": round(gb / ms * 1000.0, 1),
            "frac": round(gb / ms * 1000.0 / PEAK_HBM_GBS, 4)}


def shifts_vs_truth(torch, merger, truth, b=64):
    """Recovered per-frame global motion against the synthesiser's truth: the pre-alignment's integer initial
    shifts (g_f - g_0) and the median over the interior of each frame's final flow (g_ref - g_f), taken on every
    4th pixel of every 4th row."""
    import numpy as np

    truth = np.asarray(truth, np.float64)
    n = len(truth)
    out = {}
    if merger.params.prealign_max_shift > 0:
        pre = merger.buffer_tensor(10, 0, 4).cpu().numpy().reshape(n, 4)[:, :2].astype(np.float64)
        err = pre - (truth - truth[0])
        out["prealign_max_err_px"] = round(float(np.abs(err).max()), 3)
    r = merger.reference_index
    errs = []
    for f in range(n):
        if f == r:
            continue
        fl = merger.buffer_tensor(3, f, 2)[b:-b:4, b:-b:4].reshape(-1, 2)
        med = fl.median(dim=0).values.cpu().numpy().astype(np.float64)
        errs.append(np.abs(med - (truth[r] - truth[f])))
    errs = np.array(errs)
    out["median_flow_max_err_px"] = round(float(errs.max()), 4)
    out["median_flow_rms_err_px"] = round(float(np.sqrt(np.mean(errs ** 2))), 4)
    out["truth_max_abs_px"] = round(float(np.abs(truth).max()), 2)
    return out


def kernel_times(isa, merger, n_levels):
    res = {}
    for name in merger.KERNELS:
        levels = range(n_levels) if name in ("track", "exact_queue") else [-1]
        res[name] = [merger.kernel_ms(name, lv) for lv in levels]
    return res


def precision_vs_fp64(isa, torch):
    """Errors of the GPU paths against the FP64 oracle on the 8-frame 1024x1024 configuration: finest-level
    one-to-one shifts (max, RMS px), flow (RMS px over the interior), eroded mask (max abs), merged image (PSNR
    and max relative error over the interior) and the PSNR loss against the noiseless ground truth. The FP32
    oracle rows are the reference's arithmetic (lk_ill_ratio 0) and the same with the FP64 solve of
    ill-conditioned Lucas-Kanade windows the GPU paths use (lk_ill_ratio 0.01)."""
    import numpy as np

    from oracle import oracle as O

    cfg = CONFIGS["small"]
    w, h, n = cfg["w"], cfg["h"], cfg["n"]
    scfg = isa.default_synth(seed=20240602, width=w, height=h, frames=n, global_shift_max=cfg["gmax"],
                             flow_max=cfg["flow"], occluded_fraction=0.02)
    raw_t, _ = isa.synth_burst(scfg)
    torch.cuda.synchronize()
    raw = raw_t.cpu().numpy().view(np.uint16)
    O.set_threads(len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    params = isa.default_params(levels=cfg["levels"])
    ref = O.merge(params, raw, dt=np.float64, want=("one_to_one", "flows", "masks"))
    t_oracle = time.perf_counter() - t0
    gt = isa.ground_truth(scfg, ref["ref"], w, h).cpu().numpy().astype(np.float64)
    b = 8  # interior: the reference's borders (Q4-Q6) and the LK window are excluded
    sl = (slice(b, -b), slice(b, -b))

    def psnr(x, y):
        d = (x - y)[sl]
        d = d[np.isfinite(d)]
        mse = float(np.mean(d ** 2))
        return 99.0 if mse == 0 else 10 * np.log10(1.0 / mse)

    p_gt_ref = psnr(ref["out"], gt)
    res = {"workload": f"{n}-frame {w}x{h} RGGB 14-bit burst, levels "
                       + "/".join(str(l[0]) for l in cfg["levels"]) + ", 1x merge",
           "oracle_s": round(t_oracle, 2), "psnr_vs_gt_fp64_db": round(p_gt_ref, 3)}
    paths = [("oracle_fp32_reference", None, 0.0), ("oracle_fp32", None, 0.01), ("plain", 0, 0.01),
             ("bf16x3", 1, 0.01), ("fp16x3", 2, 0.01), ("fp32", 3, 0.01)]
    for name, code, ratio in paths:
        if code is None:
            r32 = O.merge(isa.default_params(levels=cfg["levels"], lk_ill_ratio=ratio), raw, dt=np.float32,
                          want=("one_to_one", "flows", "masks"))
            out, o2o, ref_index = r32["out"].astype(np.float64), r32["one_to_one"], r32["ref"]
            flow_of = lambda f, r=r32: r["flows"][f].astype(np.float64)  # noqa: E731
            mask_of = lambda f, r=r32: r["masks"][f].astype(np.float64)[..., :3]  # noqa: E731
        else:
            m = isa.BurstMerger(w, h, n, isa.default_params(levels=cfg["levels"], xc_precision=code,
                                                            lk_ill_ratio=ratio))
            out = m.run(raw_t).cpu().numpy().astype(np.float64)
            ref_index = m.reference_index
            o2o = m.buffer_tensor(2, 0, 2).cpu().numpy().reshape(ref["one_to_one"].shape)
            flow_of = lambda f, m=m: m.buffer_tensor(3, f, 2).cpu().numpy().astype(np.float64)  # noqa: E731
            mask_of = lambda f, m=m: m.buffer_tensor(4, f, 4).cpu().numpy().astype(np.float64)[..., :3]  # noqa: E731
        if ref_index != ref["ref"]:
            res[name] = {"error": "reference frame differs from the oracle's"}
            continue
        ds = np.abs(o2o - ref["one_to_one"])
        fl_err, mk_err = [], 0.0
        for f in range(n):
            if f == ref["ref"]:
                continue
            d = (flow_of(f) - ref["flows"][f])[sl]
            fl_err.append(d[np.isfinite(d)] ** 2)
            dm = np.abs(mask_of(f) - ref["masks"][f][..., :3])[b // 2:-b // 2, b // 2:-b // 2]
            mk_err = max(mk_err, float(np.nanmax(dm)))
        fl_err = np.concatenate(fl_err)
        rel = np.abs(out - ref["out"])[sl] / np.maximum(np.abs(ref["out"][sl]), 1e-3)
        res[name] = {"lk_ill_ratio": ratio,
                     "shift_max_px": float(f"{ds.max():.3g}"), "shift_rms_px": float(f"{np.sqrt(np.mean(ds ** 2)):.3g}"),
                     "flow_rms_px": float(f"{np.sqrt(np.mean(fl_err)):.3g}"), "mask_max_abs": float(f"{mk_err:.3g}"),
                     "psnr_vs_fp64_db": round(psnr(out, ref["out"]), 2),
                     "max_rel_vs_fp64": float(f"{np.nanmax(rel):.3g}"),
                     "flow_frac_over_0.01px": round(float(np.mean(fl_err > 1e-4)), 5),
                     "psnr_loss_vs_gt_db": round(p_gt_ref - psnr(out, gt), 4)}
    return res


def adjoint_bench(isa, torch, steps: int, warmup: int):
    """Differentiable merge on a 16-frame 2048x2048 burst at 2x output: forward (relaxed robustness +
    accumulation, loss against the noiseless 2x rendering) and backward (flow and kernel parameters)
    timed with HIP events on the launch stream, a central-difference gradcheck (eps 1e-3 px) at 1,000
    random pixels, and 5 refinement steps through isa_merge_refine_step."""
    import numpy as np

    from imagestackalignator_amd.merge_diff import MergeDiff

    w = h = 2048
    n = 16
    scfg = isa.default_synth(seed=20240603, width=w, height=h, frames=n, global_shift_max=16.0, flow_max=1.5,
                             occluded_fraction=0.02)
    raw, _ = isa.synth_burst(scfg)
    params = isa.default_params(levels=[(1, 32, 2), (2, 32, 4), (4, 32, 4)])
    merger = isa.BurstMerger(w, h, n, params)
    merger.run(raw)
    md, flow = MergeDiff.from_merger(merger, raw)
    ref = merger.reference_index
    gt = isa.ground_truth(scfg, ref, 2 * w, 2 * h, 0.5, -0.25, -0.25)
    theta = md.theta_of(params)
    out = torch.empty((2 * h, 2 * w, 3), dtype=torch.float32, device="cuda")
    dflow = torch.empty_like(flow)

    def timed(fn):
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

    # the loss read-back synchronises; time the device work without it, then take one loss
    fwd_ms = timed(lambda: md.forward(flow, theta, None, out))
    _, loss0 = md.forward(flow, theta, gt, out)
    bwd_ms = timed(lambda: md.backward(flow, theta, gt, dflow))
    g, dth = md.backward(flow, theta, gt, dflow)
    g = g.clone()
    # gradcheck: 4 passes of 250 pixels on a 4-px lattice (their 4x4 output blocks are disjoint), one
    # (frame, component) per pass; the loss difference of each pixel is summed over its own block
    rng = np.random.default_rng(5)
    eps = 1e-3
    others = [f for f in range(n) if f != ref]
    rel = []
    for k in range(4):
        f, comp = int(rng.choice(others)), k % 2
        ys = rng.choice(np.arange(8, h // 4 - 8), 250) * 4 + int(rng.integers(0, 4))
        xs = rng.choice(np.arange(8, w // 4 - 8), 250) * 4 + int(rng.integers(0, 4))
        keep = np.unique(np.stack([ys, xs], 1), axis=0)
        ys_t, xs_t = torch.as_tensor(keep[:, 0], device="cuda"), torch.as_tensor(keep[:, 1], device="cuda")
        errs = []
        for sgn in (1.0, -1.0):
            fl = flow.clone()
            fl[f, ys_t, xs_t, comp] += sgn * eps
            o, _ = md.forward(fl, theta, None)
            e = ((o.double() - gt.double()) ** 2).sum(-1)
            e[:4], e[-4:], e[:, :4], e[:, -4:] = 0, 0, 0, 0
            errs.append(e.reshape(h // 2, 4, w // 2, 4).sum((1, 3)))
        hy, hx = ys_t // 2, xs_t // 2
        fd = (errs[0][hy, hx] - errs[1][hy, hx]) / (2 * eps)
        an = g[f, ys_t, xs_t, comp].double()
        rel.append((fd - an).abs() / torch.clamp(an.abs(), min=1e-3 * float(g.abs().max())))
    rel = torch.cat(rel).cpu().numpy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, theta_r, losses = md.refine(flow, theta, gt, steps=5, lr_flow=2.0)
    torch.cuda.synchronize()
    ref_ms = (time.perf_counter() - t0) * 1000.0
    return {
        "workload": f"{n}-frame {w}x{h} RGGB 14-bit burst, relaxed merge at 2x ({2 * w}x{2 * h}), loss vs the "
                    f"noiseless 2x rendering; flow init from the plain merge's Lucas-Kanade",
        "forward_ms": round(fwd_ms, 3),
        "backward_ms": round(bwd_ms, 3),
        "value": round(n * (w * h / 1e6) / ((fwd_ms + bwd_ms) / 1000.0), 3),
        "unit": "frames*MPix/s (forward + backward)",
        "loss0": round(loss0, 6),
        "dtheta": [float(f"{v:.6g}") for v in dth],
        "gradcheck": {"pixels": int(rel.size), "eps_px": eps, "median_rel_err": float(f"{np.median(rel):.3g}"),
                      "p99_rel_err": float(f"{np.percentile(rel, 99):.3g}"),
                      "frac_rel_err_below_1e-2": round(float(np.mean(rel < 1e-2)), 4)},
        "refine": {"steps": 5, "losses": [round(v, 6) for v in losses], "ms": round(ref_ms, 1)},
    }


def mixed_section(isa, torch, cfg, raw, out_exact, gt, psnr_gt_exact, steps, warmup, names, m_pairs):
    """The tile search's cross terms on the matrix cores, same workload: merge rate, MFMA launch rate against the
    dense matrix-core peak (3 products per term for the split forms), exact-search fractions, PSNR."""
    w, h, n = cfg["w"], cfg["h"], cfg["n"]
    res = {}
    for name in names:
        pm = isa.default_params(levels=cfg["levels"], xc_precision=XC_CODES[name],
                                prealign_max_shift=cfg["prealign"])
        mm = isa.BurstMerger(w, h, n, pm)
        o2 = torch.empty_like(out_exact)
        mean_ms, med_ms, _ = timed_runs(torch, lambda: mm.run(raw, o2), steps, warmup)
        st2 = mm.stage_ms()
        kt = kernel_times(isa, mm, len(cfg["levels"]))
        prods = 1 if name == "fp32" else 3
        levels = []
        for i, (ms_launch, _) in enumerate(kt["track"]):
            if kt["exact_queue"][i][1] == 0:  # xc_levels (default: 32-pixel tiles) left this level to the exact search
                levels.append({"level": cfg["levels"][i][0], "search": "exact", "exact_ms": round(ms_launch, 3)})
                continue
            fl = track_flops(cfg, i, m_pairs, prods)
            tf = fl / 1e9 / ms_launch if ms_launch > 0 else 0.0
            levels.append({"level": cfg["levels"][i][0], "search": "mfma", "mfma_ms": round(ms_launch, 3),
                           "exact_queue_ms": round(kt["exact_queue"][i][0], 3),
                           "mfma_tflops": round(tf, 1), "mfma_util": round(tf / PEAK_MFMA_TFLOPS[name], 4)})
        ex_items = mm.exact_items()
        p_gt = psnr_t(torch, o2, gt)
        res[name] = {"value": round(n * (w * h / 1e6) / (med_ms / 1000.0), 3), "ms_per_step": round(mean_ms, 3),
                     "median_ms": round(med_ms, 3), "track_patches_ms": round(st2["track_patches"], 3),
                     "psnr_vs_exact_db": psnr_t(torch, o2, out_exact), "psnr_vs_gt_db": p_gt,
                     "psnr_loss_vs_exact_db": round(psnr_gt_exact - p_gt, 4),
                     "exact_search_fraction": [round(a / max(1, b), 5) for a, b in ex_items],
                     "levels": levels,
                     "mfma_peak_tflops": PEAK_MFMA_TFLOPS[name]}
        log(f"[bench] xc {name}: {med_ms:.2f} ms/merge (median), track {st2['track_patches']:.2f} ms, "
            f"levels {levels}")
        del mm
    return res


def c2_section(isa, torch, steps, warmup):
    """The precision workload at its size (30 x 4032x3024): the exact FP64 search and the three matrix-core forms,
    timed; their merged images against the exact path's and the ground truth. (Errors against the FP64 oracle are
    measured on the small configuration, where the oracle runs in seconds.)"""
    cfg = CONFIGS["c2"]
    w, h, n = cfg["w"], cfg["h"], cfg["n"]
    scfg = isa.default_synth(seed=20240604, width=w, height=h, frames=n, global_shift_max=cfg["gmax"],
                             flow_max=cfg["flow"], occluded_fraction=cfg["occ"])
    raw, _ = isa.synth_burst(scfg)
    m = isa.BurstMerger(w, h, n, isa.default_params(levels=cfg["levels"]))
    out = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
    mean_ms, med_ms, _ = timed_runs(torch, lambda: m.run(raw, out), steps, warmup)
    gt = isa.ground_truth(scfg, m.reference_index, w, h)
    p_gt = psnr_t(torch, out, gt)
    m_pairs = isa.lib().isa_shift_tables(n, m.reference_index, 0, 5, None, None, None)
    res = {"workload": f"c2: {n}-frame {w}x{h} RGGB 14-bit burst, levels 1/2/4/8, tiles 64/32/32/16, 1x merge",
           "exact": {"value": round(n * (w * h / 1e6) / (med_ms / 1000.0), 3), "median_ms": round(med_ms, 3),
                     "track_patches_ms": round(m.stage_ms()["track_patches"], 3), "psnr_vs_gt_db": p_gt}}
    del m
    res.update(mixed_section(isa, torch, cfg, raw, out, gt, p_gt, steps, warmup, list(XC_CODES), m_pairs))
    return res


def c4_section(isa, torch, refine_steps=3):
    """Combined largest: 50 x 6000x4000 burst (global shifts U(-128, 128), local flow up to 3 px, 5 % occluders).
    Pre-alignment, bf16x3 matrix-core tile search, full-frame 2x merge; then the differentiable merge on the
    merger's flows with `refine_steps` refinement steps against the noiseless 2x rendering (isa_merge_refine_step).
    The same merge with the exact FP64 tile search is the plain reference for PSNR."""
    from imagestackalignator_amd.merge_diff import MergeDiff

    cfg = CONFIGS["c4"]
    w, h, n = cfg["w"], cfg["h"], cfg["n"]
    scfg = isa.default_synth(seed=20240605, width=w, height=h, frames=n, global_shift_max=cfg["gmax"],
                             flow_max=cfg["flow"], occluded_fraction=cfg["occ"])
    raw, truth = isa.synth_burst(scfg)
    res = {"workload": f"c4: {n}-frame {w}x{h} RGGB 14-bit burst, shifts U(-128,128), flow <= 3 px, 5 % occluders; "
                       f"pre-alignment, bf16x3 tile search, 2x merge ({2 * w}x{2 * h}), {refine_steps} refinement steps"}
    outs = {}
    for name, code in (("plain", 0), ("bf16x3", 1)):
        p = isa.default_params(levels=cfg["levels"], xc_precision=code, super_resolution=2,
                               prealign_max_shift=cfg["prealign"])
        m = isa.BurstMerger(w, h, n, p)
        out = torch.empty((2 * h, 2 * w, 3), dtype=torch.float32, device="cuda")
        mean_ms, med_ms, _ = timed_runs(torch, lambda: m.run(raw, out), 2, 1)
        st = m.stage_ms()
        outs[name] = out
        gt = isa.ground_truth(scfg, m.reference_index, 2 * w, 2 * h, 0.5, -0.25, -0.25)
        res[name] = {"merge_ms": round(med_ms, 2), "value": round(n * (w * h / 1e6) / (med_ms / 1000.0), 3),
                     "stages_ms": {k: round(v, 3) for k, v in st.items()},
                     "psnr_vs_gt_db": psnr_t(torch, out, gt), "shifts_vs_truth": shifts_vs_truth(torch, m, truth)}
        log(f"[bench] c4 {name}: merge {med_ms:.1f} ms, stages {res[name]['stages_ms']}")
        if name == "bf16x3":
            res[name]["psnr_vs_plain_db"] = psnr_t(torch, out, outs["plain"])
            res[name]["exact_search_fraction"] = [round(a / max(1, b), 5) for a, b in m.exact_items()]
            # refinement on the merger's flows
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            md, flow = MergeDiff.from_merger(m, raw)
            theta = md.theta_of(p)
            flow_r, theta_r, losses = md.refine(flow, theta, gt, steps=refine_steps, lr_flow=2.0)
            rel_out, _ = md.forward(flow_r, theta_r, None)
            torch.cuda.synchronize()
            ref_ms = (time.perf_counter() - t0) * 1000.0
            res["refine"] = {"steps": refine_steps, "ms": round(ref_ms, 1), "losses": [round(v, 3) for v in losses],
                             "psnr_vs_gt_db": psnr_t(torch, rel_out, gt)}
            total = med_ms + ref_ms
            res["value"] = round(n * (w * h / 1e6) / (total / 1000.0), 3)
            res["total_ms"] = round(total, 1)
            res["unit"] = "frames*MPix/s (bf16x3 merge at 2x + refinement)"
            log(f"[bench] c4 refinement: {ref_ms:.1f} ms, losses {res['refine']['losses']}")
            del md, flow, flow_r, rel_out
        del m
    return res


def ingest_section(isa, torch, raw, steps, warmup, cpu=True, distinct=4, frames=20):
    """Raw ingest of a burst: `frames` PEF files (Pentax Huffman) and `frames` DNG files (256 x 256 lossless-JPEG
    tiles, two components, predictor 1) holding `distinct` frames of the synthetic burst, cycled. Timed: one
    isa_raw_decode call over all files, file bytes resident in HBM, outputs and workspace allocated beforehand
    (median of `steps`). The host parse (isa_raw_open) is timed separately. Decoded frames are checked against
    the originals. cpu_baseline: the oracle's decoders (the reference's loops), one thread per file."""
    import concurrent.futures as cf

    import numpy as np

    sys.path.insert(0, str(ROOT / "tests"))
    import raw_writer as W
    from imagestackalignator_amd import raw as P

    n, h, w = raw.shape
    frames_np = [raw[i].cpu().numpy().view(np.uint16).copy() for i in range(min(distinct, n))]
    out = {}
    for kind in ("pef", "dng"):
        t0 = time.perf_counter()
        if kind == "pef":
            blobs = [W.write_pef(f) for f in frames_np]
        else:
            blobs = [W.write_dng(f, tile=(256, 256), comps=2) for f in frames_np]
        enc_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        files = [P.RawFile(blobs[i % len(blobs)]) for i in range(frames)]
        parse_ms = (time.perf_counter() - t0) * 1000.0 / frames
        bufs = []
        for f in files:
            b = torch.empty(((len(f.data) + 3) // 4 * 4,), dtype=torch.uint8, device="cuda")
            b[: len(f.data)].copy_(torch.frombuffer(bytearray(f.data), dtype=torch.uint8))
            bufs.append(b)
        outs = [torch.empty((h, w), dtype=torch.int16, device="cuda") for _ in files]
        need = __import__("ctypes").c_size_t()
        hs = (__import__("ctypes").c_void_p * frames)(*[f._h.value for f in files])
        isa.call("isa_raw_decode_workspace", hs, frames, __import__("ctypes").byref(need))
        ws = torch.empty((need.value,), dtype=torch.uint8, device="cuda")
        run = lambda: P.decode_device(files, bufs, outs, workspace=ws)
        mean_ms, med_ms, _ = timed_runs(torch, run, steps, warmup)
        ok = all(np.array_equal(outs[i].cpu().numpy().view(np.uint16), frames_np[i]) for i in range(len(frames_np)))
        walks, redec = __import__("ctypes").c_int(), __import__("ctypes").c_int()
        isa.call("isa_raw_decode_stats", __import__("ctypes").byref(walks), __import__("ctypes").byref(redec))
        comp_bytes = sum(len(f.data) for f in files)
        algo = comp_bytes + 2 * frames * w * h
        rec = {"workload": f"{frames} {kind.upper()} files of {w}x{h} 14-bit ({len(frames_np)} distinct frames of the "
                           f"burst, cycled)" + (", Pentax Huffman strip" if kind == "pef" else
                                                ", 256x256 lossless-JPEG tiles, 2 components, predictor 1"),
               "value": round(frames * (w * h / 1e6) / (med_ms / 1000.0), 1), "unit": "frames*MPix/s decoded",
               "median_ms": round(med_ms, 3), "compressed_mb": round(comp_bytes / 1e6, 1),
               "host_parse_ms_per_file": round(parse_ms, 3), "bit_exact": ok,
               "algorithmic_gb": round(algo / 1e9, 3), "achieved_gbs": round(algo / med_ms / 1e6, 1),
               "frac": round(algo / med_ms / 1e6 / PEAK_HBM_GBS, 4), "encode_s": round(enc_s, 1),
               "sync_walks_max": walks.value, "chunk_redecodes": redec.value}
        if cpu:
            from oracle import raw as R

            threads = max(1, min(16, len(os.sched_getaffinity(0))))
            t0 = time.perf_counter()
            with cf.ThreadPoolExecutor(threads) as ex:
                list(ex.map(lambda i: R.RawFile(blobs[i % len(blobs)]), range(frames)))
            cpu_s = time.perf_counter() - t0
            rec["cpu_baseline"] = {"value": round(frames * (w * h / 1e6) / cpu_s, 2), "unit": "frames*MPix/s decoded",
                                   "cores": threads, "kind": "port",
                                   "sample": f"the same {frames} files, the oracle's decoders (the reference's "
                                             f"symbol-by-symbol loops), one thread per file, {cpu_s:.2f} s"}
        out[kind] = rec
        log(f"[bench] ingest {kind}: {json.dumps(rec)}")
        del bufs, outs, ws, files
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="full_fp32", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--xc", default="bf16x3,fp16x3,fp32",
                    help="mixed-precision (MFMA) tile-search variants timed beside the exact path; '' for none")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-merge section (c3)")
    ap.add_argument("--no-precision", action="store_true", help="skip the errors against the FP64 oracle")
    ap.add_argument("--no-extra", action="store_true", help="skip the 2x, c2 and c4 sections")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the differentiable-merge section (prints its JSON object)")
    ap.add_argument("--no-ingest", action="store_true", help="skip the raw-ingest section (PEF / DNG decode)")
    ap.add_argument("--ingest-only", action="store_true", help="only the raw-ingest section (prints its JSON object)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(the multi-GPU sharding is a design note, not an implementation)"}))
        sys.exit(2)

    import torch

    import imagestackalignator_amd as isa

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.calibrate_copy:
        n_bytes = args.calibrate_copy << 20
        a = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n_bytes, "copy_bytes_written": n_bytes, "copies": 3}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_bench(isa, torch, args.steps, args.warmup)}), flush=True)
        return
    cfg = CONFIGS[args.config]
    w, h, n = cfg["w"], cfg["h"], cfg["n"]
    scfg = isa.default_synth(seed=20240601, width=w, height=h, frames=n, global_shift_max=cfg["gmax"],
                             flow_max=cfg["flow"], occluded_fraction=cfg["occ"])
    log(f"[bench] synthesising {n} x {w}x{h} burst")
    raw, truth = isa.synth_burst(scfg)
    if args.ingest_only:
        print(json.dumps({"ingest": ingest_section(isa, torch, raw, args.steps, args.warmup,
                                                   cpu=not args.no_cpu_baseline)}), flush=True)
        return
    params = isa.default_params(levels=cfg["levels"], prealign_max_shift=cfg["prealign"])
    merger = isa.BurstMerger(w, h, n, params)
    out = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
    mean_ms, med_ms, all_ms = timed_runs(torch, lambda: merger.run(raw, out), args.steps, args.warmup)
    value = n * (w * h / 1e6) / (mean_ms / 1000.0)
    stages = merger.stage_ms()
    kt = kernel_times(isa, merger, len(cfg["levels"]))
    import ctypes as C

    m_pairs = isa.lib().isa_shift_tables(n, merger.reference_index, params.tracking_strategy,
                                         params.tracking_block_size, None, None, None)
    levels = []
    for i in range(len(cfg["levels"])):
        buf = (C.c_float * 3)()
        isa.call("isa_merger_level_ms", merger._h, i, buf)
        l, t, m = cfg["levels"][i]
        tiles = ((w // l - 2 * m) // t) * ((h // l - 2 * m) // t)
        levels.append({"level": l, "tiles": tiles, "prep_ms": round(buf[0], 4), "track_ms": round(buf[1], 4),
                       "minimize_ms": round(buf[2], 4), "track_gflop": round(track_flops(cfg, i, m_pairs) / 1e9, 2)})
    log(f"[bench] {mean_ms:.2f} ms/merge (median {med_ms:.2f}), stages: "
        + ", ".join(f"{k}={v:.2f}" for k, v in stages.items()))
    # H2D-inclusive: the raw frames start in pinned host memory and are uploaded on the merger's stream
    raw_host = raw.cpu().pin_memory()
    raw_dev2 = torch.empty_like(raw)

    def with_upload():
        raw_dev2.copy_(raw_host, non_blocking=True)
        merger.run(raw_dev2, out)

    h2d_mean, h2d_med, _ = timed_runs(torch, with_upload, args.steps, 1)
    del raw_dev2, raw_host
    color = color_export_bench(isa, torch, out, args.steps)
    # per-kernel rooflines from the live HIP-event timings of the last run
    traffic = {}
    if KERNEL_TRAFFIC.exists():
        try:
            traffic = json.loads(KERNEL_TRAFFIC.read_text()).get("bytes_per_launch", {})
        except (OSError, ValueError):
            traffic = {}
    P = w * h
    lk_ms, lk_n = kt["lucas_kanade"][0]
    lk_avg = lk_ms / max(1, lk_n)
    lk_bytes = 24.0 * P  # source, target, flow in (8 B), flow out (8 B) per pixel and iteration
    lk_gbs = lk_bytes / 1e6 / lk_avg if lk_avg > 0 else 0.0
    roofline = {
        "kernel": f"k_lk_fused<5>: one Lucas-Kanade iteration of one frame ({w}x{h}), warp + derivatives + 11x11 "
                  f"window sums + solve; {lk_n} launches per merge",
        "bound": "hbm",
        "achieved": round(lk_gbs, 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(lk_gbs / PEAK_HBM_GBS, 4),
        "traffic": traffic.get("k_lk_fused<5>"),
        "traffic_source": str(KERNEL_TRAFFIC.relative_to(ROOT)) if "k_lk_fused<5>" in traffic else None,
        "avg_launch_ms": round(lk_avg, 4),
        "algorithmic_bytes": lk_bytes,
        "note": "the dominant kernel by time; algorithmic bytes 24 per pixel and iteration, timed live with HIP "
                "events on the merger's stream around every frame's iterations; in practice bound by VALU issue",
    }
    kernels = [roofline]
    for i, (ms_l, cnt) in enumerate(kt["track"]):
        fl = track_flops(cfg, i, m_pairs)
        tf = fl / 1e9 / ms_l if ms_l > 0 else 0.0
        l, t, mm = cfg["levels"][i]
        key = f"k_track_cols<{mm}>:level{l}"
        kernels.append({"kernel": f"k_track_cols<{mm}> level {l} (tile {t}, max shift {mm}): {m_pairs} pairs x "
                                  f"{levels[i]['tiles']} tiles in one launch",
                        "bound": "fp64-valu", "achieved": round(tf, 2), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tf / PEAK_FP64_TFLOPS, 4), "avg_launch_ms": round(ms_l, 4),
                        "traffic": traffic.get(key)})
    acc_ms = kt["accumulate"][0][0]
    acc_bytes = P * (14.0 * n + 36.0)
    kernels.append({"kernel": f"k_accumulate_frames_wide: all {n} frames in one pass", "bound": "hbm",
                    "achieved": round(acc_bytes / 1e6 / acc_ms, 1) if acc_ms > 0 else 0.0, "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": round(acc_bytes / 1e6 / acc_ms / PEAK_HBM_GBS, 4) if acc_ms > 0 else 0.0,
                    "avg_launch_ms": round(acc_ms, 4), "traffic": traffic.get("k_accumulate_frames_wide")})
    result = {
        "metric": "burst frames·MPix/s merged (FP32 plain vs MFMA mixed) at stated PSNR vs FP64",
        "value": round(value, 3),
        "unit": "frames·MPix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mean_ms, 3),
        "median_ms": round(med_ms, 3),
        "value_median": round(n * (w * h / 1e6) / (med_ms / 1000.0), 3),
        "step_ms": [round(v, 3) for v in all_ms],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {n}-frame {w}x{h} RGGB 14-bit burst, global shifts U(-{cfg['gmax']:g}, "
                               f"{cfg['gmax']:g}) px with pre-alignment (+-{cfg['prealign']} px), levels "
                               + "/".join(str(l[0]) for l in cfg["levels"]) + ", tiles "
                               + "/".join(str(l[1]) for l in cfg["levels"]) + ", 1x merge",
                   "frames": n, "width": w, "height": h, "pairs": m_pairs},
        "h2d_inclusive": {"value": round(n * (w * h / 1e6) / (h2d_med / 1000.0), 3), "median_ms": round(h2d_med, 3),
                          "upload_bytes": int(raw.numel() * 2),
                          "note": "raw frames copied from pinned host memory on the merge stream, then merged"},
        "stages_ms": {k: round(v, 3) for k, v in stages.items()},
        "color_export": color,
        "stage_roofline": stage_rooflines(stages, n, w, h, levels),
        "levels": levels,
        "roofline": roofline,
        "roofline_kernels": kernels,
    }
    gt = isa.ground_truth(scfg, merger.reference_index, w, h)
    torch.cuda.synchronize()
    result["psnr_vs_gt_db"] = psnr_t(torch, out, gt)
    result["shifts_vs_truth"] = shifts_vs_truth(torch, merger, truth)
    log(f"[bench] shifts vs truth: {result['shifts_vs_truth']}")
    names = [v for v in args.xc.split(",") if v]
    if names:
        result["mixed"] = mixed_section(isa, torch, cfg, raw, out, gt, result["psnr_vs_gt_db"], args.steps,
                                        args.warmup, names, m_pairs)
    del merger
    if not args.no_ingest:
        log("[bench] raw ingest")
        result["ingest"] = ingest_section(isa, torch, raw, args.steps, args.warmup, cpu=not args.no_cpu_baseline)
    if not args.no_extra and args.config == "full_fp32":
        # the same burst merged at 2x over the whole frame (12000 x 8000 output)
        p2 = isa.default_params(levels=cfg["levels"], super_resolution=2, prealign_max_shift=cfg["prealign"])
        m2 = isa.BurstMerger(w, h, n, p2)
        out2 = torch.empty((2 * h, 2 * w, 3), dtype=torch.float32, device="cuda")
        mean2, med2, _ = timed_runs(torch, lambda: m2.run(raw, out2), args.steps, args.warmup)
        gt2 = isa.ground_truth(scfg, m2.reference_index, 2 * w, 2 * h, 0.5, -0.25, -0.25)
        result["full_2x"] = {"workload": f"the same burst merged at 2x over the whole frame ({2 * w}x{2 * h})",
                             "value": round(n * (w * h / 1e6) / (med2 / 1000.0), 3), "median_ms": round(med2, 3),
                             "stages_ms": {k: round(v, 3) for k, v in m2.stage_ms().items()},
                             "psnr_vs_gt_db": psnr_t(torch, out2, gt2)}
        log(f"[bench] 2x: {med2:.2f} ms/merge, stages {result['full_2x']['stages_ms']}")
        del m2, out2, gt2
    del raw
    torch.cuda.empty_cache()
    if not args.no_extra and args.config == "full_fp32":
        log("[bench] c2 precision workload")
        result["c2"] = c2_section(isa, torch, args.steps, args.warmup)
        torch.cuda.empty_cache()
        log("[bench] c4 combined workload")
        result["c4"] = c4_section(isa, torch)
        torch.cuda.empty_cache()
    if not args.no_adjoint:
        log("[bench] differentiable merge")
        result["adjoint"] = adjoint_bench(isa, torch, args.steps, args.warmup)
        log(f"[bench] adjoint: {json.dumps(result['adjoint'])}")
    if not args.no_precision:
        log("[bench] errors against the FP64 oracle")
        result["precision_vs_fp64"] = precision_vs_fp64(isa, torch)
        log(f"[bench] precision: {json.dumps(result['precision_vs_fp64'])}")
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU baseline")
        result["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
