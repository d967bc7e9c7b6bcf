#!/usr/bin/env python3
"""Benchmark: Siddon forward + adjoint Grays/s at 512x512x256, 1024x768, 720 views.

One step = one exact Siddon forward projection of the seeded 30-ellipsoid
phantom over all 720 views, then the exact adjoint of that sinogram, with
every input already resident on the GPU.  Prints one JSON line on stdout
(progress goes to stderr).

Roofline numerator (algorithmic bytes per launch): sum over rays of
(4 * N_int + 4), N_int = voxels with nonzero intersection length on the ray,
counted by the GPU count kernel (pinned against the FP64 oracle in the tests).
Peak: 8.0 TB/s HBM3E.

cpu_baseline: the FP32 restatement of the reference algorithm
(oracle/oracle.c, reference-mode flags) on the host cores, on a strided
sample of the same workload (every 45th view = 16 views).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "ct-nef_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, img_cpu, stride):
    from oracle import oracle as O
    import numpy as np
    ov = O.vol_struct(wl.vol.grid, wl.vol.center, wl.vol.size)
    d = wl.det
    od = O.det_struct(d.cyl, d.SID, d.SAD, d.na, d.da, d.ai, d.nb, d.db, d.bi)
    ang = wl.angles()
    img = img_cpu.numpy()
    nsel = len(range(0, wl.nv, stride))
    rays = nsel * d.na * d.nb
    t0 = time.perf_counter()
    p = O.ref_fwd(ov, od, ang, img, flags=O.REF_FWD_FLAGS, v0=0, vstep=stride)
    t1 = time.perf_counter()
    O.ref_bp(ov, od, ang, p, flags=O.REF_BP_FLAGS, v0=0, vstep=stride)
    t2 = time.perf_counter()
    return {
        "value": rays / (t2 - t0) / 1e9,
        "unit": "Grays/s",
        "cores": O.threads(),
        "kind": "port",
        "sample": f"reference-mode forward+adjoint restated in C (oracle/oracle.c), views 0::{stride} "
                  f"({nsel} of {wl.nv}) of the same workload, {rays} rays; fwd {t1 - t0:.2f} s, adj {t2 - t1:.2f} s",
    }


def c1_full_leg(dev, reps=3):
    """C1 (128^3 / 256x256 / 180 views, 10 ellipsoids) in full: the C
    restatement of the reference algorithm (reference-mode forward + adjoint)
    on the host cores over every view, beside the GPU pair in reference mode
    and in exact mode on the same phantom."""
    import numpy as np
    import torch
    import nefct_hip as nh
    from nefct_hip import _lib
    from nefct_hip.configs import c1
    from nefct_hip.phantom import ellipsoid_phantom
    from oracle import oracle as O
    wl = c1()
    d = wl.det
    img = ellipsoid_phantom(wl.vol, wl.seed, wl.ellipsoids, device=dev)
    ang_np = wl.angles()
    ang = torch.tensor(ang_np, device=dev)
    ov = O.vol_struct(wl.vol.grid, wl.vol.center, wl.vol.size)
    od = O.det_struct(d.cyl, d.SID, d.SAD, d.na, d.da, d.ai, d.nb, d.db, d.bi)
    img_cpu = img.cpu().numpy()
    t0 = time.perf_counter()
    p = O.ref_fwd(ov, od, ang_np, img_cpu, flags=O.REF_FWD_FLAGS)
    t1 = time.perf_counter()
    O.ref_bp(ov, od, ang_np, p, flags=O.REF_BP_FLAGS)
    t2 = time.perf_counter()

    def pair(ff, fb):
        return nh.backproject(nh.project(img, wl.vol, d, ang, flags=ff), wl.vol, d, ang, flags=fb)

    t_ref, _ = _timed(lambda: pair(_lib.REF_FORWARD, _lib.REF_ADJOINT), reps)
    t_exact, _ = _timed(lambda: pair(0, 0), reps)
    rays = wl.rays
    log(f"[bench] C1 in full: CPU {t2 - t0:.2f} s, GPU reference-mode pair {t_ref * 1e3:.2f} ms")
    return {
        "workload": wl.name, "rays": rays, "cores": O.threads(),
        "cpu_grays_s": round(rays / (t2 - t0) / 1e9, 6),
        "cpu_s": {"forward": round(t1 - t0, 3), "adjoint": round(t2 - t1, 3)},
        "gpu_reference_mode_pair_grays_s": round(rays / t_ref / 1e9, 4),
        "gpu_exact_pair_grays_s": round(rays / t_exact / 1e9, 4),
        "gpu_over_cpu_reference_mode": round((t2 - t0) / t_ref, 1),
    }


def batched_leg(dev, iters, log_every=True):
    """64 uneven differentiable problems (sides 64-256, 60-360 views, detector
    2 side, SOD 500-1000 mm, own phantom and Gaussian-bump warp field each):
    one gradient pass = warp + forward + adjoint with gradients to the base
    volume and the displacement field, and the projector pair alone.  The
    plain path calls the per-problem autograd functions in a loop; the batched
    path runs one persistent launch per direction over all problems."""
    import torch
    import nefct_hip as nh
    from nefct_hip import _lib
    from nefct_hip.batch import SiddonBatch, batched_project
    from nefct_hip.configs import c5
    from nefct_hip.deform import warp
    from nefct_hip.fourd import warp_project
    from nefct_hip.phantom import ellipsoid_phantom, gaussian_dvf

    probs = c5()
    log(f"[bench] batched leg: {len(probs)} problems, building phantoms and warp fields")
    bases = [ellipsoid_phantom(p.vol, p.phantom_seed, 10, device=dev) for p in probs]
    dvfs = [gaussian_dvf(p.vol.shape_zyx, p.dvf_seed, voxel_mm=p.vol.voxel, device=dev) for p in probs]
    angs = [torch.tensor(p.angles, device=dev) for p in probs]
    batch = SiddonBatch([(p.vol, p.det, p.angles) for p in probs], device=dev)
    g = torch.Generator(device=dev).manual_seed(55)
    ys = [torch.randn(s.shape, generator=g, device=dev) for s in batch.sinograms]
    rays = batch.rays
    for dst, y in zip(batch.sinograms, ys):
        dst.copy_(y)
    tuned = batch.tune()             # batched adjoint lane spacing, untimed
    shape = batch.flags & _lib.F_LANE_STRIDE2   # the plain calls get the same lane spacing
    log(f"[bench] batched leg: tuned adjoint {tuned}")

    def plain(with_warp):
        for p, b, d, a, y in zip(probs, bases, dvfs, angs, ys):
            x = b.detach().requires_grad_(True)
            if with_warp:
                dv = d.detach().requires_grad_(True)
                v = warp(x, dv, 1)
            else:
                v = x
            s = nh.forward_project(v, p.vol, p.det, a, flags=shape)
            (s * y).sum().backward()

    def batched(with_warp):
        xs = [b.detach().requires_grad_(True) for b in bases]
        if with_warp:
            ds = [d.detach().requires_grad_(True) for d in dvfs]
            outs = warp_project(batch, xs, ds, boundary=1)
        else:
            outs = batched_project(batch, xs)
        sum((o * y).sum() for o, y in zip(outs, ys)).backward()

    res = {}
    for name, fn, ww in (("plain_pair", plain, False), ("batched_pair", batched, False),
                         ("plain_warp_pair", plain, True), ("batched_warp_pair", batched, True)):
        fn(ww)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters):
            fn(ww)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / iters
        res[name] = rays / dt / 1e9
        log(f"[bench] batched leg {name}: {dt * 1e3:.1f} ms per pass, {res[name]:.4f} Grays/s")
    return {
        "workload": "64 uneven problems: cubic sides {64..256} @ 1 mm, 60-360 views, detector (2 side)^2, "
                    "SOD 500-1000 mm; gradient pass to volume (and warp field)",
        "rays": rays,
        "unit": "Grays/s",
        "plain_pair": round(res["plain_pair"], 5),
        "batched_pair": round(res["batched_pair"], 5),
        "speedup_pair": round(res["batched_pair"] / res["plain_pair"], 3),
        "plain_warp_pair": round(res["plain_warp_pair"], 5),
        "batched_warp_pair": round(res["batched_warp_pair"], 5),
        "speedup_warp_pair": round(res["batched_warp_pair"] / res["plain_warp_pair"], 3),
        "adjoint_lane_spacing": tuned,
        "iters": iters,
    }


def c4_leg(dev, iters=3):
    """4-D CT with deformation: ten respiratory phases of one 256x256x160 base
    volume, each warped by its own Gaussian-bump displacement field and seen
    from 40-120 views of a 384x288 detector.  One gradient pass = warp every
    phase, project, backproject, transposed warps into the base gradient
    (and the displacement gradients).  Plain: per-phase autograd calls in a
    loop; batched: one warp per phase and one launch per direction for all
    phases.  Also one graph-captured descent iteration on the base volume
    (warps, batched pair, residual, transposed warps, update) replayed."""
    import torch
    import nefct_hip as nh
    from nefct_hip import _lib
    from nefct_hip.batch import SiddonBatch
    from nefct_hip.configs import c4
    from nefct_hip.deform import deform_adjoint_op, deform_op, warp
    from nefct_hip.fourd import warp_project
    from nefct_hip.phantom import ellipsoid_phantom, gaussian_dvf

    phases = c4()
    vol = phases[0].vol
    log(f"[bench] C4 leg: {len(phases)} phases, {sum(len(p.angles) for p in phases)} views")
    base = ellipsoid_phantom(vol, phases[0].phantom_seed, 10, device=dev)
    dvfs = [gaussian_dvf(vol.shape_zyx, p.dvf_seed, voxel_mm=vol.voxel, device=dev) for p in phases]
    angs = [torch.tensor(p.angles, device=dev) for p in phases]
    batch = SiddonBatch([(p.vol, p.det, p.angles) for p in phases], device=dev)
    g = torch.Generator(device=dev).manual_seed(56)
    ys = [torch.randn(s.shape, generator=g, device=dev) for s in batch.sinograms]
    rays = batch.rays
    for dst, y in zip(batch.sinograms, ys):
        dst.copy_(y)
    tuned = batch.tune()             # batched adjoint lane spacing, untimed
    shape = batch.flags & _lib.F_LANE_STRIDE2   # the plain calls get the same lane spacing
    log(f"[bench] C4 leg: tuned adjoint {tuned}")

    def plain():
        x = base.detach().requires_grad_(True)
        ds = [d.detach().requires_grad_(True) for d in dvfs]
        loss = 0.0
        for p, d, a, y in zip(phases, ds, angs, ys):
            loss = loss + (nh.forward_project(warp(x, d, 1), p.vol, p.det, a, flags=shape) * y).sum()
        loss.backward()

    def batched():
        x = base.detach().requires_grad_(True)
        ds = [d.detach().requires_grad_(True) for d in dvfs]
        outs = warp_project(batch, [x] * len(phases), ds, boundary=1)
        sum((o * y).sum() for o, y in zip(outs, ys)).backward()

    res = {}
    for name, fn in (("plain_warp_pair", plain), ("batched_warp_pair", batched)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / iters
        res[name] = (rays / dt / 1e9, dt)
        log(f"[bench] C4 leg {name}: {dt * 1e3:.1f} ms per pass, {res[name][0]:.4f} Grays/s")

    # one descent iteration on the base volume, captured in a graph
    x = (0.5 * base).contiguous()
    grad = torch.zeros_like(base)
    targets = [t.clone() for t in ys]

    def iteration():
        for p, d in enumerate(dvfs):
            batch.volumes[p].copy_(deform_op(x, d[0], d[1], d[2], 1))
        batch.project()
        for s_, t_ in zip(batch.sinograms, targets):
            s_.sub_(t_)
        batch.backproject()
        grad.zero_()
        for p, d in enumerate(dvfs):
            grad.add_(deform_adjoint_op(batch.volumes[p], d[0], d[1], d[2], 1))
        x.add_(grad, alpha=-1e-6)

    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        iteration()
    torch.cuda.current_stream().wait_stream(st)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        iteration()
    graph.replay()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    t_graph = e0.elapsed_time(e1) / 1e3 / iters
    log(f"[bench] C4 leg graph-captured iteration: {t_graph * 1e3:.1f} ms")
    return {
        "workload": "C4: 10 phases of a 256x256x160 @ 1 mm base, own Gaussian-bump warp each, 40-120 views, "
                    "384x288 detector; gradient pass to base volume and warp fields",
        "rays": rays,
        "unit": "Grays/s",
        "plain_warp_pair": round(res["plain_warp_pair"][0], 5),
        "batched_warp_pair": round(res["batched_warp_pair"][0], 5),
        "speedup_warp_pair": round(res["batched_warp_pair"][0] / res["plain_warp_pair"][0], 3),
        "ms_plain": round(res["plain_warp_pair"][1] * 1e3, 2),
        "ms_batched": round(res["batched_warp_pair"][1] * 1e3, 2),
        "graph_iteration_ms": round(t_graph * 1e3, 2),
        "graph_iteration_grays_s": round(rays / t_graph / 1e9, 5),
        "adjoint_lane_spacing": tuned,
        "iters": iters,
    }


def copy_bandwidth(dev, nbytes=2 << 30, reps=5):
    """Achievable streaming bandwidth: a device-to-device copy of nbytes
    (read nbytes + write nbytes per copy), HIP-event timed over reps copies."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    del src, dst
    torch.cuda.empty_cache()
    gbs = 2 * nbytes / t / 1e9
    return {"bytes_moved": 2 * nbytes, "ms": round(t * 1e3, 3), "GBs": round(gbs, 1),
            "pct_of_spec": round(100 * gbs / HBM_PEAK_GBS, 1)}


def _timed(fn, reps):
    """Mean HIP-event time (s) of fn() over reps calls on torch's stream, after one untimed call."""
    import torch
    fn()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        out = fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps, out


def fdk_leg(wl, img, sino, ang, reps=2):
    """FDK of the workload's projections: the short scan (180 degrees + fan
    angle) of the first views, cosine and Parker weights, ram-lak row filter,
    voxel-driven backprojection; each stage timed on its own."""
    import torch
    from nefct_hip import fdk as F
    vol, det = wl.vol, wl.det
    nv_full = ang.numel()
    dang = 2.0 * math.pi / nv_full
    gamma = math.atan(0.5 * det.na * det.da / det.SAD)
    nv = min(nv_full, int(math.ceil((math.pi + 2.0 * gamma) / dang)) + 1)
    p = sino[:nv].contiguous()
    a = ang[:nv].contiguous()
    scale = F.fdk_scale(det, nv, dang, True)
    log(f"[bench] FDK leg: short scan of {nv} views")
    t_w, q = _timed(lambda: F.preweight(p, det, parker=True, dang_rad=dang, direction=1, scale=scale), reps)
    t_f, f = _timed(lambda: F.ramp_filter(q, "ram-lak"), reps)
    t_b, rec = _timed(lambda: F.backproject(f, vol, det, a), reps)
    t_bg, _ = _timed(lambda: F.backproject(f, vol, det, a, mapping="gather"), reps)
    t_bs, _ = _timed(lambda: F.backproject(f, vol, det, a, mapping="staged"), reps)
    t_bw, _ = _timed(lambda: F.backproject(f, vol, det, a, mapping="wide"), reps)
    nvox = vol.grid[0] * vol.grid[1] * vol.grid[2]
    # inner cylinder (radius 0.4 of the width), central half of the slices: no truncation or
    # cone-beam edge effects
    nz, ny, nx = vol.shape_zyx
    yy, xx = torch.meshgrid(torch.arange(ny, device=img.device), torch.arange(nx, device=img.device), indexing="ij")
    inner = ((xx - (nx - 1) / 2) ** 2 + (yy - (ny - 1) / 2) ** 2) < (0.4 * nx) ** 2
    zs = slice(nz // 4, 3 * nz // 4)
    d = (rec[zs][:, inner] - img[zs][:, inner]).double()
    rel = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(img[zs][:, inner].double()))
    proj_bytes = 4 * p.numel()
    bp_bytes = proj_bytes + 4 * nvox + 8 * nv
    log(f"[bench] FDK leg: weight {t_w * 1e3:.2f} ms, filter {t_f * 1e3:.2f} ms, backprojection {t_b * 1e3:.2f} ms")
    return {
        "views": nv,
        "ms": round((t_w + t_f + t_b) * 1e3, 3),
        "stage_ms": {"preweight": round(t_w * 1e3, 3), "filter": round(t_f * 1e3, 3),
                     "backproject": round(t_b * 1e3, 3)},
        "backproject_voxel_views_per_s": nvox * nv / t_b,
        "backproject_mapping": "staged footprints, 16x16x16 blocks (auto)",
        "backproject_gather_ms": round(t_bg * 1e3, 3),
        "backproject_staged_ms": round(t_bs * 1e3, 3),
        "backproject_wide_ms": round(t_bw * 1e3, 3),
        "backproject_hbm": {"algorithmic_bytes": bp_bytes, "achieved_GBs": round(bp_bytes / t_b / 1e9, 1),
                            "frac": round(bp_bytes / t_b / 1e9 / HBM_PEAK_GBS, 4)},
        "filter_gflops": round(2.0 * p.numel() * det.na / t_f / 1e9, 1),
        "preweight_hbm": {"algorithmic_bytes": 2 * proj_bytes, "achieved_GBs": round(2 * proj_bytes / t_w / 1e9, 1),
                          "frac": round(2 * proj_bytes / t_w / 1e9 / HBM_PEAK_GBS, 4)},
        "rel_err_vs_phantom_inner": round(rel, 4),
    }


def backprojector_leg(wl, sino, ang, reps=2):
    """The reference's two other backprojectors at the workload's size: the
    voxel-driven one (nefct/cc/pixel_3d/pixel.cu.cc, four taps per voxel and
    view) over the FDK leg's short scan, and the pixel-driven FDK footprint one
    (nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc, float atomics into the
    volume) over the first 8 views, in that file's voxel units."""
    import torch
    from nefct_hip.fdk import footprint_backproject
    from nefct_hip.pixel import pixel_backproject
    vol, det = wl.vol, wl.det
    nv_full = ang.numel()
    gamma = math.atan(0.5 * det.na * det.da / det.SAD)
    nv = min(nv_full, int(math.ceil((math.pi + 2.0 * gamma) / (2.0 * math.pi / nv_full))) + 1)
    p = sino[:nv].contiguous()
    a = ang[:nv].contiguous()
    nvox = vol.grid[0] * vol.grid[1] * vol.grid[2]
    t_p, _ = _timed(lambda: pixel_backproject(p, vol, det, a), reps)
    b_p = 4 * p.numel() + 4 * nvox
    nf = 8
    vx = vol.voxel[0]
    pf = sino[:nf].contiguous()
    af = ang[:nf].contiguous()
    out = torch.zeros(vol.shape_zyx, dtype=torch.float32, device=sino.device)
    # the footprint op's naming: SAD = source to axis, SID = source to detector
    kw = dict(SID=det.SAD / vx, SAD=det.SID / vx, da=det.da / vx, ai=0.0, na=det.na, db=det.db / vx, bi=0.0,
              nb=det.nb)
    t_f, _ = _timed(lambda: footprint_backproject(pf, vol.grid, af, out=out, **kw), reps)
    log(f"[bench] backprojectors: voxel-driven {t_p * 1e3:.2f} ms ({nv} views), footprint {t_f * 1e3:.2f} ms "
        f"({nf} views)")
    return {
        "voxel_driven": {"views": nv, "ms": round(t_p * 1e3, 3), "voxel_views_per_s": nvox * nv / t_p,
                         "algorithmic_bytes": b_p, "achieved_GBs": round(b_p / t_p / 1e9, 1),
                         "frac": round(b_p / t_p / 1e9 / HBM_PEAK_GBS, 4)},
        "fdk_footprint": {"views": nf, "ms": round(t_f * 1e3, 3), "ms_per_view": round(t_f * 1e3 / nf, 3),
                          "rays_per_s": nf * det.na * det.nb / t_f},
    }


def warp_leg(wl, img, reps=3):
    """Trilinear warp of the workload's volume by a smooth displacement field
    (forward, transpose scatter, displacement gradient), HBM rate of each."""
    import torch
    from nefct_hip.deform import deform_adjoint_op, deform_dvf_grad_op, deform_op
    from nefct_hip.phantom import gaussian_dvf
    nvox = img.numel()
    dvf = gaussian_dvf(wl.vol.shape_zyx, 7, voxel_mm=wl.vol.voxel, device=img.device).float().contiguous()
    g = torch.ones_like(img)
    t_f, _ = _timed(lambda: deform_op(img, dvf[0], dvf[1], dvf[2], 1), reps)
    t_a, _ = _timed(lambda: deform_adjoint_op(g, dvf[0], dvf[1], dvf[2], 1), reps)
    t_g, _ = _timed(lambda: deform_dvf_grad_op(img, g, dvf[0], dvf[1], dvf[2], 1), reps)
    # algorithmic bytes per voxel: forward reads the image and 3 field components, writes 1;
    # the transpose reads g and the field, writes the image gradient; the field gradient reads
    # the image, g and the field, writes 3 components
    b_f, b_a, b_g = 20 * nvox, 20 * nvox, 32 * nvox
    log(f"[bench] warp leg: forward {t_f * 1e3:.2f} ms, transpose {t_a * 1e3:.2f} ms, field gradient {t_g * 1e3:.2f} ms")
    rate = lambda b, t: {"ms": round(t * 1e3, 3), "algorithmic_bytes": b, "achieved_GBs": round(b / t / 1e9, 1),
                         "frac": round(b / t / 1e9 / HBM_PEAK_GBS, 4)}
    return {"volume": list(wl.vol.grid), "forward": rate(b_f, t_f), "transpose": rate(b_a, t_a),
            "field_gradient": rate(b_g, t_g)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-stride", type=int, default=45)
    ap.add_argument("--views", type=int, default=0,
                    help="profiling only: project the first N views instead of all 720 (not the headline number)")
    ap.add_argument("--adj-flags", type=int, default=0,
                    help="profiling only: adjoint algorithm flags (16 deterministic, 32 slab windows)")
    ap.add_argument("--only", choices=["pair", "fwd", "adj", "fdk"], default="pair",
                    help="profiling only: run one direction, or only the FDK leg")
    ap.add_argument("--no-tune", action="store_true",
                    help="use the default launch shapes instead of the autotuner's (tuning is untimed)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain and C4 legs")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched-vs-plain legs (64 uneven problems and C4)")
    ap.add_argument("--batched-iters", type=int, default=2)
    ap.add_argument("--no-fdk", action="store_true", help="skip the FDK and warp legs")
    ap.add_argument("--calib", action="store_true",
                    help="profiling only: first copy 256 MiB device-to-device (a known-size dispatch that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(views would shard across GPUs with one all-reduce of the adjoint volume; not built)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import nefct_hip as nh
    from nefct_hip.configs import c2
    from nefct_hip.phantom import ellipsoid_phantom

    wl = c2()
    vol, det = wl.vol, wl.det
    dev = torch.device("cuda:0")
    if args.batched_only:
        print(json.dumps({"batched": batched_leg(dev, args.batched_iters), "c4": c4_leg(dev)}), flush=True)
        return
    if args.calib:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("[bench] calibration copy: 268435456 bytes read, 268435456 bytes written")
    log(f"[bench] workload {wl.name}: building phantom")
    img = ellipsoid_phantom(vol, wl.seed, wl.ellipsoids, device=dev)
    ang = torch.tensor(wl.angles(), device=dev)
    if args.views:
        ang = ang[: args.views].contiguous()
    rays = ang.numel() * det.na * det.nb
    if args.only == "fdk":
        log("[bench] FDK leg only")
        print(json.dumps({"fdk": fdk_leg(wl, img, nh.project(img, vol, det, ang), ang)}), flush=True)
        return

    log("[bench] counting intersections")
    counts = nh.count_intersections(vol, det, ang, device=dev)
    n_int = int(counts.sum(dtype=torch.int64).item())
    del counts
    bytes_fwd = 4 * n_int + 4 * rays
    bytes_adj = 4 * rays + 4 * n_int

    # launch shapes: the autotuner's choice for this geometry (timed on a view
    # sample before the timed region), unless fixed on the command line
    fwd_flags, adj_flags = 0, args.adj_flags
    tuned = None
    if not args.no_tune and not args.adj_flags:
        from nefct_hip import autotune
        log("[bench] autotuning the forward and adjoint launch shapes")
        fwd_flags, frec = autotune.tune_forward(vol, det, ang)
        adj_flags, arec = autotune.tune_adjoint(vol, det, ang)
        tuned = {"forward": frec["variant"], "adjoint": arec["variant"],
                 "forward_ms_on_sample": frec["ms_on_sample"], "adjoint_ms_on_sample": arec["ms_on_sample"],
                 "forward_sample_views": frec["sample_views"], "adjoint_sample_views": arec["sample_views"]}
        log(f"[bench] tuned: forward {frec['variant']}, adjoint {arec['variant']}")

    def step():
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e2 = torch.cuda.Event(enable_timing=True)
        e0.record()
        sino = nh.project(img, vol, det, ang, flags=fwd_flags) if args.only != "adj" else sino_fixed
        e1.record()
        back = nh.backproject(sino, vol, det, ang, flags=adj_flags) if args.only != "fwd" else None
        e2.record()
        return e0, e1, e2, sino, back

    sino_fixed = nh.project(img, vol, det, ang) if args.only == "adj" else None

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"[bench] warmup {w + 1}/{args.warmup} done")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs = []
    for k in range(args.steps):
        e0, e1, e2, sino, back = step()
        evs.append((e0, e1, e2))
        del sino, back
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    total_s = t1 - t0
    t_fwd = max(sum(e0.elapsed_time(e1) for e0, e1, _ in evs) / 1e3 / args.steps, 1e-9)
    t_adj = max(sum(e1.elapsed_time(e2) for _, e1, e2 in evs) / 1e3 / args.steps, 1e-9)
    log(f"[bench] timed {args.steps} steps: {total_s:.3f} s (fwd {t_fwd * 1e3:.1f} ms, adj {t_adj * 1e3:.1f} ms)")

    value = rays * args.steps / total_s / 1e9
    dominant = "siddon_adjoint" if t_adj > t_fwd else "siddon_forward"
    t_dom = max(t_adj, t_fwd)
    b_dom = bytes_adj if t_adj > t_fwd else bytes_fwd
    achieved = b_dom / t_dom / 1e9

    fdk = warp_rates = bps = None
    if not args.no_fdk and not args.views and args.only == "pair":
        sino_fdk = nh.project(img, vol, det, ang)
        fdk = fdk_leg(wl, img, sino_fdk, ang)
        bps = backprojector_leg(wl, sino_fdk, ang)
        del sino_fdk
        warp_rates = warp_leg(wl, img)
        torch.cuda.empty_cache()

    batched = c4 = None
    if not args.no_batched and not args.views and args.only == "pair":
        del sino_fixed
        torch.cuda.empty_cache()
        batched = batched_leg(dev, args.batched_iters)
        torch.cuda.empty_cache()
        c4 = c4_leg(dev)

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline on views 0::{args.cpu_stride}")
        cpu = cpu_baseline(wl, img.cpu(), args.cpu_stride)
        log(f"[bench] CPU baseline {cpu['value']:.4f} Grays/s on {cpu['cores']} threads")
        if not args.views and args.only == "pair":
            cpu["c1_full"] = c1_full_leg(dev)

    # HBM bytes per launch, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    # (scaled by a known-size copy) and committed; with this run's kernel
    # times they give each kernel's measured share of the HBM peak
    copy_bw = copy_bandwidth(dev) if not args.views else None
    traffic = None
    measured = {}
    tpath = os.path.join(ROOT, "profiles", "siddon_pair_c2_traffic.json")
    if not args.views and os.path.exists(tpath):
        with open(tpath) as f:
            tks = json.load(f)["kernels"]
        for name, t_k in (("siddon_forward", t_fwd), ("siddon_adjoint", t_adj)):
            tk = tks.get(name)
            if tk:
                gbs = tk["traffic_bytes"] / t_k / 1e9
                measured[name] = {"traffic_bytes": tk["traffic_bytes"], "GBs": round(gbs, 1),
                                  "frac_of_spec": round(gbs / HBM_PEAK_GBS, 4),
                                  "frac_of_copy": round(gbs / copy_bw["GBs"], 4) if copy_bw else None}
        if dominant in measured:
            traffic = measured[dominant]["traffic_bytes"]

    # VALU instructions per launch, measured with rocprofv3 SQ counters and
    # committed; with this run's kernel times they give the issue rate against
    # one wave64 instruction per SIMD per 4 cycles (plain FP32 and integer
    # ops; conversions, v_fract and compares partly co-issue, so above 1 is
    # possible)
    valu = None
    spath = os.path.join(ROOT, "profiles", "siddon_pair_c2_sq.json")
    if not args.views and os.path.exists(spath):
        with open(spath) as f:
            sq = json.load(f)
        peak = sq["issue_peak"]["wave_instructions_per_s"]
        valu = {"peak_wave_instr_per_s": peak}
        for name, t_k in (("siddon_forward", t_fwd), ("siddon_adjoint", t_adj)):
            k = sq["kernels"].get(name)
            if k:
                rate = k["SQ_INSTS_VALU"] / t_k
                valu[name] = {"valu_instr": k["SQ_INSTS_VALU"], "G_wave_instr_per_s": round(rate / 1e9, 1),
                              "issue_frac": round(rate / peak, 3),
                              "lane_valu_per_intersection": round(64 * k["SQ_INSTS_VALU"] / n_int, 2)}

    rec = {
        "metric": "Siddon fwd+adjoint Grays/s at 512x512x256, 720 views",
        "value": round(value, 5),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_s / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 30-ellipsoid phantom (numpy PCG64 seed 2), exact Siddon pair",
        "config": {
            "workload": wl.name,
            "volume": list(vol.grid), "voxel_mm": vol.voxel[0],
            "detector": [det.na, det.nb], "pitch_mm": det.da, "views": int(ang.numel()),
            "sod_mm": det.SID, "sdd_mm": det.SAD, "rays": rays, "sum_n_int": n_int,
            "parallelism": "single GPU",
            "launch_shapes": tuned or {"forward": "16x16/auto", "adjoint": "default" if not args.adj_flags
                                       else f"flags {args.adj_flags}"},
        },
        "fwd_grays_s": round(rays / t_fwd / 1e9, 5),
        "adj_grays_s": round(rays / t_adj / 1e9, 5),
        "voxel_intersections_per_s": (2 * n_int) / (t_fwd + t_adj),
        "pct_hbm_peak_pair": round(100 * (bytes_fwd + bytes_adj) / (t_fwd + t_adj) / 1e9 / HBM_PEAK_GBS, 2),
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": "profiles/siddon_pair_c2_traffic.json (rocprofv3 FETCH_SIZE/WRITE_SIZE x copy calibration)"
            if traffic else None,
            "algorithmic_bytes": b_dom,
            "avg_launch_ms": round(t_dom * 1e3, 3),
            "frac_of_copy": round(achieved / copy_bw["GBs"], 4) if copy_bw else None,
            "note": "achieved counts every voxel gather as HBM bytes; most are served by L2 "
                    "(the forward exceeds the HBM peak by this measure), so hbm_measured gives the real "
                    "HBM traffic; both Siddon kernels are bound by instruction issue",
        },
        "hbm_measured": measured or None,
        "valu_issue": valu,
        "copy_bandwidth": copy_bw,
        "batched": batched,
        "c4": c4,
        "fdk": fdk,
        "backprojectors": bps,
        "warp": warp_rates,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
