"""Benchmark: 4K deferred PBR+IBL+SSAO resolve on one MI355X.

Workload (the full-size plain method): 3840x2160 G-buffer of four RGBA32F
planes (fBm heightfield seen by a 60-degree pinhole camera), 64 point lights
with inverse-square falloff, 16-sample hemisphere SSAO (radius 0.5, 4x4
rotation noise), IBL from a 512^2 x 6 prefiltered cube with 6 mips (1024
Hammersley samples, from a 2048x1024 synthetic HDR map), 32^2 irradiance cube
and 128^2 BRDF LUT.  A step is one resolve of the frame; inputs are resident
in HBM before the timed region.

Also measured (reported under "secondary"): the IBL precompute kernels of the
small configuration (512x256 source -> 128^2 x 6 x 5 mips at 1024 samples,
32^2 irradiance, 128^2 LUT) and of the 512^2 cube.

Prints one JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

SEED0 = 0x414C0000
WIDTH, HEIGHT, LIGHTS = 3840, 2160, 64
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# Per-launch counters of the FP32-G-buffer resolve of this workload, from the
# rocprofv3 passes in profiles/resolve_kernel_4k_fp32_64lights_ssao16.md (the
# commands are there): HBM bytes from FETCH_SIZE and WRITE_SIZE scaled by the
# same counters on a 1 GiB copy (bench.py --only copy), and the VALU wave
# instructions (SQ_INSTS_VALU).  They describe the kernel source they were
# measured on and are refreshed with it.
RESOLVE_TRAFFIC_BYTES = {"fp32": 1.428e9}
RESOLVE_VALU_WAVE_INSTS = {"fp32": 360039216}
# The SSAO view-depth pre-pass of the same call (ssao_depth_kernel): the
# position plane read once, the 4-byte plane written once.
DEPTH_PASS_TRAFFIC_BYTES = {"fp32": 1.659e8}
# Nominal VALU issue: one wave64 instruction per SIMD per 4 cycles, 1024 SIMDs
# at 2.4 GHz.
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=100, help="untimed steps first: the GPU reaches its steady clock after about 50 frames")
    p.add_argument("--no-cpu-baseline", action="store_true", help="skip the oracle timing (profiling runs)")
    p.add_argument("--no-secondary", action="store_true", help="skip the precompute timings")
    p.add_argument("--fp16-gbuffer", action="store_true", help="shade from an FP16 G-buffer")
    p.add_argument("--only", choices=["resolve", "prefilter", "copy", "ssr", "shadow", "gbuffer", "irradiance", "mlp"],
                   help="profiling runs: time one kernel family and print its line only "
                        "(copy = a plain 1 GiB device copy, the calibration for HBM byte counters)")
    return p.parse_args()


def algorithmic_bytes(ibl, fp16: bool) -> int:
    """Compulsory HBM bytes of one resolve: every G-buffer plane read once,
    the output written once, the IBL state and lights read once."""
    px = WIDTH * HEIGHT
    gbuf = 4 * (8 if fp16 else 16)
    out = 16
    return px * (gbuf + out) + ibl.state_bytes() + LIGHTS * 32


def time_events(torch, fn, reps: int) -> float:
    """Average ms per call of fn() on the current stream, HIP events."""
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / reps


def secondary(torch, aa) -> dict:
    """Precompute timings.  nominal_samples_per_s counts texels x samples per
    texel; the roughness-0 mip evaluates one lookup per texel (its samples
    coincide), so the evaluated count is lower."""
    res = {}
    env0 = aa.synth.env(seed=SEED0, width=512, height=256)
    dev0 = torch.from_numpy(env0).cuda()
    pyr0 = aa.env_pyramid(dev0)
    texels0 = aa.cube_texels(128, 5)
    ms = time_events(torch, lambda: aa.prefilter_cube(pyr0, 512, 256, 128, 5, 1024), 5)
    res["prefilter_128x6x5_1024spp"] = {"ms": ms, "texels_per_s": texels0 / (ms * 1e-3),
                                        "nominal_samples_per_s": texels0 * 1024 / (ms * 1e-3)}
    ms = time_events(torch, lambda: aa.irradiance_cube(pyr0, 512, 256, 32), 5)
    res["irradiance_32x6_45000"] = {"ms": ms, "texels_per_s": 6 * 32 * 32 / (ms * 1e-3),
                                    "samples_per_s": 6 * 32 * 32 * 45000 / (ms * 1e-3)}
    ms = time_events(torch, lambda: aa.brdf_lut(128, 1024), 5)
    res["brdf_lut_128_1024"] = {"ms": ms}
    env1 = aa.synth.env(seed=SEED0 + 1, width=2048, height=1024)
    pyr1 = aa.env_pyramid(torch.from_numpy(env1).cuda())
    texels1 = aa.cube_texels(512, 6)
    ms = time_events(torch, lambda: aa.prefilter_cube(pyr1, 2048, 1024, 512, 6, 1024), 2)
    res["prefilter_512x6x6_1024spp"] = {"ms": ms, "texels_per_s": texels1 / (ms * 1e-3),
                                        "nominal_samples_per_s": texels1 * 1024 / (ms * 1e-3)}
    res["batched_prefilter_64maps"] = batched(torch, aa)
    return res


def batched(torch, aa) -> dict:
    """64 maps, widths uniform over {256..4096} (seeded by index), faces W/4,
    5 roughness mips, 1024 samples: one graph launch per run (pyramids, sample
    tables, persistent prefilter), launch shape from the autotuner."""
    from althea_amd.batch import PrefilterBatch, batch_widths

    widths = batch_widths(64)
    maps = [torch.from_numpy(aa.synth.env(seed=i, width=w, height=w // 2)).cuda() for i, w in enumerate(widths)]
    b = PrefilterBatch(maps, [w // 4 for w in widths], mips=5, samples=1024)
    t0 = time.perf_counter()
    tuning = b.autotune()
    tune_s = time.perf_counter() - t0
    b.run()  # builds and instantiates the graph
    ms = time_events(torch, b.run, 2)
    texels = b.texels()
    return {"maps": 64, "widths": {str(w): widths.count(w) for w in sorted(set(widths))}, "texels": texels,
            "ms_per_run": ms, "texels_per_s": texels / (ms * 1e-3),
            "nominal_samples_per_s": texels * 1024 / (ms * 1e-3),
            "tuning": tuning, "autotune_s": tune_s}


def mlp_errors_vs_oracle(torch, aa, ibl, scene, env, net, gbuf, lights, n_pix: int = 2000) -> dict:
    """Shading error of the cube-IBL frame and of the MLP-IBL frame with every
    cell accepted, against the FP64 oracle at sampled covered pixels.  The
    oracle shades each pixel with the exact prefiltered radiance at its own
    reflection query (R, roughness): the FP64 prefilter with 1024 Hammersley
    samples (the plain method's estimator) and with 16384 (converged)."""
    from tests import oracle_lib as ol

    rng = np.random.default_rng(SEED0 + 9)
    cov = np.nonzero(scene.position[..., 3].ravel() != 0)[0]
    lin = rng.choice(cov, n_pix, replace=False)
    pix = np.stack([lin % WIDTH, lin // WIDTH], 1).astype(np.int32)
    views, oenv = ibl.views_numpy(), ol.Env(env)
    _, _, margin, q = ol.resolve(scene, views, oenv, pix, query=True)
    dr = q.astype(np.float32)
    cube = aa.DeferredRendering.resolve(gbuf, ibl, lights, scene.globals, scene.ssao).cpu().numpy()[lin, :3]
    used = torch.zeros(WIDTH * HEIGHT, dtype=torch.uint8, device="cuda")
    mlp = aa.DeferredRendering.resolve_mlp(gbuf, ibl, net, lights, scene.globals, scene.ssao,
                                           mlp_used=used).cpu().numpy()[lin, :3]
    served = used.cpu().numpy()[lin] != 0
    clear = margin > 2e-3
    out = {"pixels": int(n_pix), "clear_of_ssao_thresholds": int(clear.sum()), "mlp_served": int(served.sum())}
    for spp in (1024, 16384):
        refl = oenv.prefilter_points(spp, dr)
        ref, _, _ = ol.resolve(scene, views, oenv, pix, refl=refl)
        for name, img in (("cube_ibl", cube), ("mlp_ibl", mlp)):
            e = (np.abs(img - ref) / np.maximum(np.abs(ref), 1e-2)).max(1)[clear]
            out[f"{name}_vs_fp64_prefilter_{spp}spp"] = {"mean": float(e.mean()), "max": float(e.max()),
                                                          "p99": float(np.percentile(e, 99))}
    return out


CONFIG4_EPOCHS = 32


def config4_recipe(torch, aa, maps_per_width: int = 1) -> dict:
    """The learned surrogate as configured (config 4): one MLP per environment
    of the 64-map batch (here the first map of each source width, 256 to
    4096), trained with Adam (step size 1e-3, batch 65536) on 2^22 queries,
    directions uniform on the sphere and roughness uniform in [0, 1],
    labelled by the plain 1024-sample prefilter of that map; 32 epochs (2048
    steps).  Per map: the training time, the loss, and the mean relative
    error over 2^14 held-out queries against 1024-sample labels and against
    converged 16384-sample labels, with the cube path (the map's own 5-mip
    cube, faces W/4) judged against the same labels."""
    from althea_amd import mlp as M
    from althea_amd.batch import batch_widths

    widths = batch_widths(64)
    picks = [widths.index(w) for w in sorted(set(widths))]
    res = {"recipe": {"queries": 1 << 22, "batch": 65536, "lr": 1e-3, "roughness": [0.0, 1.0], "label_spp": 1024,
                      "epochs": CONFIG4_EPOCHS, "steps": CONFIG4_EPOCHS * ((1 << 22) // 65536)}, "maps": {}}
    for i in picks:
        w = widths[i]
        env = torch.from_numpy(aa.synth.env(seed=i, width=w, height=w // 2)).cuda()
        pyr = aa.env_pyramid(env)
        t0 = time.perf_counter()
        dr, labels = M.training_set(pyr, w, w // 2, 1 << 22, seed=SEED0 + 40 + i, samples=1024)
        net = M.PrefilterMLP(seed=SEED0 + 50 + i, lr=1e-3,
                             init_output_bias=float(np.log(labels[:, :3].mean().item())))
        torch.cuda.synchronize()
        label_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        hist = M.fit(net, dr, labels, batch=65536, epochs=CONFIG4_EPOCHS)
        torch.cuda.synchronize()
        train_s = time.perf_counter() - t0
        q, t1024 = M.training_set(pyr, w, w // 2, 1 << 14, seed=SEED0 + 60 + i, samples=1024)
        t16k = aa.prefilter_points(pyr, w, w // 2, 16384, q)
        pred = net.infer(q)
        cube = aa.prefilter_cube(pyr, w, w // 2, w // 4, 5, 1024)
        qn = q.cpu().numpy()
        dir_lod = torch.from_numpy(np.concatenate([qn[:, :3], qn[:, 3:4] * 4.0], 1).astype(np.float32)).cuda()
        cube_pred = aa.cube_sample(cube, w // 4, 5, dir_lod)

        def err(a, b):
            return float(((a[:, :3] - b[:, :3]).abs() / b[:, :3].abs().clamp_min(1e-2)).amax(1).mean().item())

        res["maps"][f"map{i}_{w}x{w // 2}"] = {
            "label_s": label_s, "train_s": train_s, "ms_per_step": 1e3 * train_s / res["recipe"]["steps"],
            "loss_first": hist[0], "loss_last": hist[-1],
            "mlp_mean_rel_err_vs_1024spp": err(pred, t1024), "mlp_mean_rel_err_vs_16384spp": err(pred, t16k),
            "cube_mean_rel_err_vs_1024spp": err(cube_pred, t1024), "cube_mean_rel_err_vs_16384spp": err(cube_pred, t16k),
            "labels_1024_vs_16384spp": err(t1024, t16k),
            "ibl_state_bytes": {"mlp": net.state_bytes(), "cube": 16 * aa.cube_texels(w // 4, 5)}}
        del env, pyr, dr, labels, net, cube
    return res


def mlp_ibl(torch, aa, ibl, scene, gbuf, lights, env, steps: int) -> dict:
    """The learned stand-in on the same frame.  2^22 (direction, roughness)
    queries, roughness uniform in [1/(mips-1), 1] (the range whose cube
    lookups blend two mips the monitor can accept; mip 0 is a mirror of
    per-texel noise), labelled by the 1024-sample prefilter of the frame's
    environment; Adam, batch 16384, step size 3e-3 decaying to 1e-5 over 256
    epochs (65536 steps); then the error monitor against the 512^2 x6 cube
    and the MLP-IBL resolve (the resolve queues the pixels the MLP serves, a
    dense MFMA kernel completes them) at monitor thresholds 1e-3 (the target),
    3e-3, 5e-3, 1e-2 and with every cell accepted (the full-frame MLP
    inference).  Each MLP-IBL frame is compared with the cube-IBL frame over
    every pixel (relative difference, floor 1e-2), and the cube-IBL and
    full-frame MLP-IBL shading against the FP64 oracle."""
    from althea_amd import mlp as M

    t0 = time.perf_counter()
    pyr = ibl.environment_tensor()
    v = ibl.views()
    rmin = 1.0 / (v.prefiltered_mips - 1)
    dr, labels = M.training_set(pyr, env.shape[1], env.shape[0], 1 << 22, seed=SEED0 + 4, samples=1024, rmin=rmin)
    torch.cuda.synchronize()
    label_s = time.perf_counter() - t0
    net = M.PrefilterMLP(seed=SEED0 + 5, lr=3e-3, init_output_bias=float(np.log(labels[:, :3].mean().item())))
    epochs = 256
    t0 = time.perf_counter()
    batch = 16384
    hist = M.fit(net, dr, labels, batch=batch, epochs=epochs, lr=3e-3, final_lr=1e-5)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    train_steps = epochs * ((1 << 22) // batch)
    cube = ibl.prefiltered_tensor()
    out = torch.empty((WIDTH * HEIGHT, 4), dtype=torch.float32, device="cuda")
    used = torch.zeros(WIDTH * HEIGHT, dtype=torch.uint8, device="cuda")
    plain = aa.DeferredRendering.resolve(gbuf, ibl, lights, scene.globals, scene.ssao)
    plain_ms = time_events(torch, lambda: aa.DeferredRendering.resolve(gbuf, ibl, lights, scene.globals, scene.ssao,
                                                                        out=out), steps)
    per_threshold = {}
    for thr in (1e-3, 3e-3, 5e-3, 1e-2, float("inf")):
        t0 = time.perf_counter()
        cell_err, accepted = net.build_monitor(cube, v.prefiltered_size, v.prefiltered_mips, threshold=thr)
        torch.cuda.synchronize()
        monitor_s = time.perf_counter() - t0

        def step():
            aa.DeferredRendering.resolve_mlp(gbuf, ibl, net, lights, scene.globals, scene.ssao, out=out,
                                             mlp_used=used)

        step()
        ms = time_events(torch, step, steps)
        d = ((out[:, :3] - plain[:, :3]).abs() / plain[:, :3].abs().clamp_min(1e-2)).double()
        per_threshold["every_cell" if thr == float("inf") else f"{thr:g}"] = {
            "accepted_cells": accepted, "cells": cell_err.numel(), "monitor_s": monitor_s,
            "mlp_pixel_fraction": float(used.float().mean().item()),
            "ms": ms, "mpix_per_s": WIDTH * HEIGHT / (ms * 1e-3) / 1e6,
            "rel_diff_vs_cube_ibl_frame": {"mean": float(d.mean().item()), "max": float(d.max().item())}}
    e = cell_err.cpu().numpy()
    # The monitor now accepts every cell: the full-frame MLP inference.
    errors = mlp_errors_vs_oracle(torch, aa, ibl, scene, env, net, gbuf, lights)
    return {"cube_ibl": {"ms": plain_ms, "mpix_per_s": WIDTH * HEIGHT / (plain_ms * 1e-3) / 1e6},
            "mlp_ibl_by_monitor_threshold": per_threshold,
            "shading_error_vs_fp64_oracle": errors,
            "cell_error_quantiles_by_mip": {m: {q: float(np.quantile(e[m], float(q))) for q in ("0.1", "0.5", "0.9")}
                                            for m in range(e.shape[0])},
            "ibl_state_bytes": {"cube_ibl": ibl.state_bytes(), "mlp_weights": net.state_bytes(),
                                "monitor_flags": int(e.size)},
            "training": {"queries": 1 << 22, "label_spp": 1024, "label_s": label_s, "batch": batch,
                         "roughness_range": [rmin, 1.0], "steps": train_steps, "train_s": train_s,
                         "ms_per_step": 1e3 * train_s / train_steps, "loss_first": hist[0], "loss_last": hist[-1]}}


def ssr_frame(torch, aa, scene, ibl, gbuf, lights, steps: int) -> dict:
    """Screen-space reflections on the same frame: capture into a 5-level
    FP16 reflection buffer (Src/ReflectionBuffer.cpp:30-37), the glossy
    convolution of levels 1-4, and the resolve reading the buffer.  The
    capture's compulsory bytes: position and normal planes once, level 0
    written once, the IBL state and lights; the march's filtered position
    reads and the hit shading's albedo / MRO reads are gathers on top."""
    refl = aa.ReflectionBuffer(WIDTH, HEIGHT, mips=5, fp16=True)
    out = torch.empty((WIDTH * HEIGHT, 4), dtype=torch.float32, device="cuda")
    glob, ssao = scene.globals, scene.ssao

    def capture():
        aa.ScreenSpaceReflection.captureReflection(gbuf, ibl, lights, glob, refl)

    def convolve():
        aa.ScreenSpaceReflection.convolveReflectionBuffer(refl)

    def resolve():
        aa.DeferredRendering.resolve(gbuf, ibl, lights, glob, ssao, out=out, reflection=refl)

    def frame():
        capture()
        convolve()
        resolve()

    frame()
    cap_ms = time_events(torch, capture, steps)
    conv_ms = time_events(torch, convolve, steps)
    res_ms = time_events(torch, resolve, steps)
    frame_ms = time_events(torch, frame, steps)
    hits = float((refl.level(0)[..., 3] == 1.0).float().mean().item())
    px = WIDTH * HEIGHT
    gb = 16 if gbuf.getHandles().format == 0 else 8
    cap_bytes = px * (2 * gb + 8) + ibl.state_bytes() + LIGHTS * 32
    texels = [max(1, WIDTH >> m) * max(1, HEIGHT >> m) for m in range(5)]
    conv_bytes = sum(8 * (texels[m - 1] + texels[m]) for m in range(1, 5))  # read m-1, write m (FP16 RGBA)
    return {"capture_ms": cap_ms, "capture_mpix_per_s": px / (cap_ms * 1e-3) / 1e6,
            "capture_algorithmic_bytes": cap_bytes, "capture_achieved_gbs": cap_bytes / (cap_ms * 1e-3) / 1e9,
            "hit_fraction": hits, "convolve_ms": conv_ms, "convolve_levels": 4,
            "convolve_algorithmic_bytes": conv_bytes, "resolve_with_buffer_ms": res_ms,
            "frame_ms": frame_ms, "frame_mpix_per_s": px / (frame_ms * 1e-3) / 1e6}


def ssr_only(torch, aa, args) -> int:
    """Profiling run of the SSR pipeline on the 4K frame."""
    env = aa.synth.env(seed=SEED0 + 1, width=2048, height=1024)
    desc = aa.IBLDesc(irradiance_size=32, prefiltered_size=512, prefiltered_mips=6, prefilter_samples=1024,
                      brdf_lut_size=128, brdf_lut_samples=1024)
    ibl = aa.ImageBasedLighting.createResources(torch.from_numpy(env).cuda(), desc)
    scene = aa.synth.scene(seed=SEED0 + 1, width=WIDTH, height=HEIGHT, lights=LIGHTS)
    gbuf = aa.GBufferResources(*scene.gbuffer_tensors("cuda", fp16=args.fp16_gbuffer))
    r = ssr_frame(torch, aa, scene, ibl, gbuf, scene.lights_tensor("cuda"), args.steps)
    print(json.dumps({"kind": "ssr", **r}), flush=True)
    return 0


SHADOW_SIZE = 256  # the reference's cube extent (Src/PointLight.cpp:35)


def shadow_frame(torch, aa, scene, ibl, gbuf, steps: int) -> dict:
    """Point-light shadows on the same frame: the cubes of all 64 lights
    (256^2 x 6 each, the reference's extent) from the scene's heightfield as
    a 128 x 128-cell mesh (32 768 triangles), and the resolve reading them.
    The producer's compulsory bytes are the cubes written once (clear and
    depth) and the mesh read once; the resolve's are the plain resolve's plus
    the cubes read once."""
    pos, idx = aa.synth.heightfield_mesh(SEED0 + 1, n=128)
    pos_t = torch.from_numpy(pos).cuda()
    idx_t = torch.from_numpy(idx.view(np.int32)).cuda()
    coll = aa.PointLightCollection.from_lights(scene.lights_tensor("cuda"), shadow_size=SHADOW_SIZE)
    out = torch.empty((WIDTH * HEIGHT, 4), dtype=torch.float32, device="cuda")
    glob, ssao = scene.globals, scene.ssao

    def draw():
        coll.drawShadowMaps(pos_t, idx_t)

    def resolve():
        aa.DeferredRendering.resolve(gbuf, ibl, coll, glob, ssao, out=out)

    def frame():
        draw()
        resolve()

    frame()
    draw_ms = time_events(torch, draw, steps)
    res_ms = time_events(torch, resolve, steps)
    frame_ms = time_events(torch, frame, steps)
    cubes = coll.getShadowMapImage()
    covered = float((cubes < 1.0).float().mean().item())
    cube_bytes = coll.shadow_storage.numel() * 4  # faces with their one-texel borders
    px = WIDTH * HEIGHT
    # The same mesh as a glTF primitive with an alpha-masked base colour (a
    # quarter of a 512^2 sRGB texture below the 0.5 cut-off, 10 mips,
    # coordinates repeating 8 times): the shadow shaders' alpha test.
    rng = np.random.default_rng(SEED0 + 11)
    base = rng.integers(0, 256, (512, 512, 4), dtype=np.uint8)
    base[..., 3] = np.where(rng.uniform(size=(512, 512)) < 0.25, 60, 230)
    tex = aa.Texture(torch.from_numpy(base).cuda(), mips=10, srgb=True)
    uv = np.stack([(pos[:, 0] / 20.0 + 0.5) * 8.0, (pos[:, 2] / 20.0 + 0.5) * 8.0], 1).astype(np.float32)
    prim = aa.GltfPrimitive(pos_t, None, None, None, [torch.from_numpy(uv).cuda()], idx_t, 0)
    mat = aa.GltfMaterial(0, 0, 0)

    def draw_gltf():
        coll.drawShadowMapsGltf([prim], [mat], [tex])

    draw_gltf()
    gltf_ms = time_events(torch, draw_gltf, steps)
    gltf_covered = float((coll.getShadowMapImage() < 1.0).float().mean().item())
    return {"lights": LIGHTS, "cube_size": SHADOW_SIZE, "triangles": int(idx.shape[0]),
            "draw_ms": draw_ms, "draw_texels_per_s": cubes.numel() / (draw_ms * 1e-3),
            "draw_algorithmic_bytes": 2 * cube_bytes + pos.nbytes + idx.nbytes, "texels_covered": covered,
            "draw_gltf_alpha_tested_ms": gltf_ms, "texels_covered_gltf": gltf_covered,
            "resolve_shadowed_ms": res_ms, "resolve_shadowed_mpix_per_s": px / (res_ms * 1e-3) / 1e6,
            "frame_ms": frame_ms, "frame_mpix_per_s": px / (frame_ms * 1e-3) / 1e6}


def shadow_only(torch, aa, args) -> int:
    """Profiling run of the shadow pipeline on the 4K frame."""
    env = aa.synth.env(seed=SEED0 + 1, width=2048, height=1024)
    desc = aa.IBLDesc(irradiance_size=32, prefiltered_size=512, prefiltered_mips=6, prefilter_samples=1024,
                      brdf_lut_size=128, brdf_lut_samples=1024)
    ibl = aa.ImageBasedLighting.createResources(torch.from_numpy(env).cuda(), desc)
    scene = aa.synth.scene(seed=SEED0 + 1, width=WIDTH, height=HEIGHT, lights=LIGHTS)
    gbuf = aa.GBufferResources(*scene.gbuffer_tensors("cuda", fp16=args.fp16_gbuffer))
    r = shadow_frame(torch, aa, scene, ibl, gbuf, args.steps)
    print(json.dumps({"kind": "shadow", **r}), flush=True)
    return 0


def gbuffer_frame(torch, aa, scene, ibl, steps: int, fp16: bool = False) -> dict:
    """The G-buffer producer on the same camera: the scene's heightfield as a
    128 x 128-cell mesh (32 768 triangles) with per-vertex normals, colours
    and MRO, rasterized into the 4K G-buffer, then the resolve of that
    G-buffer (64 lights, SSAO 16).  Compulsory bytes of the producer: the four
    planes written once (64 B/px FP32), the mesh read once."""
    pos, idx = aa.synth.heightfield_mesh(SEED0 + 1, n=128)
    nrm, col, mro = aa.synth.mesh_attributes(SEED0 + 1, pos, idx)
    f = lambda a: torch.from_numpy(a).cuda()
    mesh = aa.GBufferMesh(f(pos), f(idx.view(np.int32)), f(col), f(mro), f(nrm))
    pas = aa.SceneToGBufferPass(WIDTH, HEIGHT, fp16=fp16)
    lights = scene.lights_tensor("cuda")
    out = torch.empty((WIDTH * HEIGHT, 4), dtype=torch.float32, device="cuda")
    glob, ssao = scene.globals, scene.ssao

    def draw():
        pas.begin(mesh, glob)

    def resolve():
        aa.DeferredRendering.resolve(pas.gbuffer, ibl, lights, glob, ssao, out=out)

    def frame():
        draw()
        resolve()

    frame()
    draw_ms = time_events(torch, draw, steps)
    res_ms = time_events(torch, resolve, steps)
    frame_ms = time_events(torch, frame, steps)
    covered = float((pas.gbuffer.planes[0][..., 3] == 1.0).float().mean().item())
    px = WIDTH * HEIGHT
    nbytes = px * (32 if fp16 else 64) + pos.nbytes * 4 + idx.nbytes
    return {"triangles": int(idx.shape[0]), "draw_ms": draw_ms, "draw_mpix_per_s": px / (draw_ms * 1e-3) / 1e6,
            "draw_algorithmic_bytes": nbytes, "draw_achieved_gbs": nbytes / (draw_ms * 1e-3) / 1e9,
            "covered": covered, "resolve_ms": res_ms, "frame_ms": frame_ms,
            "frame_mpix_per_s": px / (frame_ms * 1e-3) / 1e6}


def gltf_frame(torch, aa, scene, ibl, steps: int) -> dict:
    """The glTF subpass on the same camera: the heightfield mesh (32 768
    triangles) as a textured primitive (512^2 sRGB base colour with alpha,
    512^2 normal map, 256^2 metallic-roughness map, full mip chains, trilinear,
    texture coordinates repeating 8 times) over a floor primitive, into the
    4K G-buffer; then the resolve of it."""
    pos, idx = aa.synth.heightfield_mesh(SEED0 + 1, n=128)
    nrm, _, _ = aa.synth.mesh_attributes(SEED0 + 1, pos, idx)
    f = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    tan = np.tile(np.float32([1, 0, 0]), (len(pos), 1))
    bit = np.tile(np.float32([0, 0, 1]), (len(pos), 1))
    uv = np.stack([(pos[:, 0] / 20.0 + 0.5) * 8.0, (pos[:, 2] / 20.0 + 0.5) * 8.0], 1).astype(np.float32)
    rng = np.random.default_rng(SEED0 + 2)
    base = rng.integers(0, 256, (512, 512, 4), dtype=np.uint8)
    base[..., 3] = np.where(rng.uniform(size=(512, 512)) < 0.1, 40, 255)
    nmap = np.concatenate([rng.integers(96, 160, (512, 512, 2)), np.full((512, 512, 2), 255)], -1).astype(np.uint8)
    mrt = rng.integers(0, 256, (256, 256, 4), dtype=np.uint8)
    texs = [aa.Texture(f(base), mips=10, srgb=True), aa.Texture(f(nmap), mips=10), aa.Texture(f(mrt), mips=9),
            aa.Texture(f(np.full((1, 1, 4), 255, np.uint8))), aa.Texture(f(np.array([[[128, 128, 255, 255]]], np.uint8)))]
    mats = [aa.GltfMaterial(0, 1, 2, base_color_factor=(0.9, 0.85, 0.8, 1.0), metallic_factor=0.6, roughness_factor=0.8),
            aa.GltfMaterial(3, 4, 3)]
    terrain = aa.GltfPrimitive(f(pos), f(tan), f(bit), f(nrm), [f(uv)], f(idx.view(np.int32)), 0)
    fp = np.array([[-30, -2, -30], [30, -2, 30], [30, -2, -30], [-30, -2, 30]], np.float32)
    floor = aa.GltfPrimitive(f(fp), f(np.tile(np.float32([1, 0, 0]), (4, 1))), f(np.tile(np.float32([0, 0, 1]), (4, 1))),
                             f(np.tile(np.float32([0, 1, 0]), (4, 1))),
                             [f(np.array([[0, 0], [1, 1], [1, 0], [0, 1]], np.float32))],
                             f(np.array([[0, 1, 2], [0, 3, 1]], np.int32)), 1)
    pas = aa.SceneToGBufferPass(WIDTH, HEIGHT)
    lights = scene.lights_tensor("cuda")
    out = torch.empty((WIDTH * HEIGHT, 4), dtype=torch.float32, device="cuda")
    glob, ssao = scene.globals, scene.ssao

    def draw():
        pas.begin_gltf([terrain, floor], mats, texs, glob)

    def frame():
        draw()
        aa.DeferredRendering.resolve(pas.gbuffer, ibl, lights, glob, ssao, out=out)

    frame()
    draw_ms = time_events(torch, draw, steps)
    frame_ms = time_events(torch, frame, steps)
    px = WIDTH * HEIGHT
    return {"triangles": int(idx.shape[0]) + 2, "draw_ms": draw_ms, "draw_mpix_per_s": px / (draw_ms * 1e-3) / 1e6,
            "frame_ms": frame_ms, "frame_mpix_per_s": px / (frame_ms * 1e-3) / 1e6}


def full_frame(torch, aa, scene, ibl, steps: int) -> dict:
    """Every pass of the path on one 4K frame, as an engine runs them: the
    G-buffer from the 32 768-triangle heightfield mesh (pass-through), the
    shadow cubes of the 64 lights (256^2) from the same mesh, the SSR capture
    with shadowed hit shading into a 5-level FP16 buffer, its glossy
    convolution, and the resolve with shadows and the reflection buffer."""
    pos, idx = aa.synth.heightfield_mesh(SEED0 + 1, n=128)
    nrm, col, mro = aa.synth.mesh_attributes(SEED0 + 1, pos, idx)
    f = lambda a: torch.from_numpy(a).cuda()
    mesh = aa.GBufferMesh(f(pos), f(idx.view(np.int32)), f(col), f(mro), f(nrm))
    pas = aa.SceneToGBufferPass(WIDTH, HEIGHT)
    coll = aa.PointLightCollection.from_lights(scene.lights_tensor("cuda"), shadow_size=SHADOW_SIZE)
    refl = aa.ReflectionBuffer(WIDTH, HEIGHT, mips=5, fp16=True)
    out = torch.empty((WIDTH * HEIGHT, 4), dtype=torch.float32, device="cuda")
    glob, ssao = scene.globals, scene.ssao
    pos_t, idx_t = mesh.positions, mesh.indices

    def frame():
        g = pas.begin(mesh, glob)
        coll.drawShadowMaps(pos_t, idx_t)
        aa.ScreenSpaceReflection.captureReflection(g, ibl, coll, glob, refl)
        aa.ScreenSpaceReflection.convolveReflectionBuffer(refl)
        aa.DeferredRendering.resolve(g, ibl, coll, glob, ssao, out=out, reflection=refl)

    frame()
    ms = time_events(torch, frame, steps)
    px = WIDTH * HEIGHT
    return {"passes": "gbuffer + shadow cubes + ssr (shadowed) + convolve + resolve (shadowed, reflections)",
            "ms": ms, "mpix_per_s": px / (ms * 1e-3) / 1e6}


def gbuffer_only(torch, aa, args) -> int:
    """Profiling run of the G-buffer producer and its resolve on the 4K frame."""
    env = aa.synth.env(seed=SEED0 + 1, width=2048, height=1024)
    desc = aa.IBLDesc(irradiance_size=32, prefiltered_size=512, prefiltered_mips=6, prefilter_samples=1024,
                      brdf_lut_size=128, brdf_lut_samples=1024)
    ibl = aa.ImageBasedLighting.createResources(torch.from_numpy(env).cuda(), desc)
    scene = aa.synth.scene(seed=SEED0 + 1, width=WIDTH, height=HEIGHT, lights=LIGHTS)
    r = gbuffer_frame(torch, aa, scene, ibl, args.steps, fp16=args.fp16_gbuffer)
    r["gltf"] = gltf_frame(torch, aa, scene, ibl, args.steps)
    r["full_frame"] = full_frame(torch, aa, scene, ibl, args.steps)
    print(json.dumps({"kind": "gbuffer", **r}), flush=True)
    return 0


def fp16_resolve(torch, aa, scene, ibl, lights, steps: int) -> dict:
    """The same frame shaded from an FP16 G-buffer (32 B/px)."""
    g16 = aa.GBufferResources(*scene.gbuffer_tensors("cuda", fp16=True))
    out = torch.empty((WIDTH * HEIGHT, 4), dtype=torch.float32, device="cuda")
    fn = lambda: aa.DeferredRendering.resolve(g16, ibl, lights, scene.globals, scene.ssao, out=out)
    fn()
    ms = time_events(torch, fn, steps)
    nbytes = algorithmic_bytes(ibl, True)
    return {"ms": ms, "mpix_per_s": WIDTH * HEIGHT / (ms * 1e-3) / 1e6, "algorithmic_bytes": nbytes,
            "achieved_gbs": nbytes / (ms * 1e-3) / 1e9}


def cpu_baseline(aa, scene, ibl, env) -> dict:
    """The CPU baseline: the FP32 port of the resolve (oracle/port_resolve.cpp,
    the kernel's arithmetic restated for the host cores with OpenMP) on the
    whole 3840x2160 frame, 64 lights, SSAO 16 against the same IBL state; one
    warm-up, then the median of 5 frames.  The FP64 oracle resolving the same
    frame once is reported beside it as "oracle", never as the baseline."""
    from tests import oracle_lib as ol

    threads = min(16, os.cpu_count() or 1)
    views = ibl.views_numpy()
    ol.port_resolve(scene, views, env, threads=threads)
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        ol.port_resolve(scene, views, env, threads=threads)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    ys, xs = np.mgrid[0:HEIGHT, 0:WIDTH]
    pix = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
    t0 = time.perf_counter()
    ol.resolve(scene, views, ol.Env(env), pix)
    oracle_s = time.perf_counter() - t0
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), model)
    except OSError:
        pass
    px = WIDTH * HEIGHT
    return {"value": px / dt / 1e6, "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": f"whole {WIDTH}x{HEIGHT} frame, {LIGHTS} lights, SSAO 16: FP32 port with the kernel's df32 "
                      f"addressing, median of 5 frames ({dt:.2f} s each) after 1 warm-up, {threads} OpenMP threads "
                      f"of {os.cpu_count()} (nproc) on {model}",
            "oracle": {"value": px / oracle_s / 1e6, "unit": "Mpix/s", "seconds": oracle_s,
                       "threads": ol.max_threads(), "what": "FP64 oracle, same frame, once"}}


def valu_issue(ms: float, fp16: bool):
    """VALU wave instructions per launch (profiled) over this run's launch
    time: per SIMD per cycle at 2.4 GHz, and against one wave64 instruction
    per SIMD per 4 cycles (a ratio at or above 1 means the SIMDs issue VALU
    as fast as that model allows, or faster)."""
    n = RESOLVE_VALU_WAVE_INSTS.get("fp16" if fp16 else "fp32")
    if n is None:
        return None
    rate = n / (ms * 1e-3)
    return {"wave_insts_per_launch": n, "issue_per_s": rate, "wave_insts_per_simd_cycle": rate / (1024 * 2.4e9),
            "ratio_to_one_per_4_cycles": rate / VALU_ISSUE_PEAK}


def timed_steps(torch, fn, args) -> float:
    for _ in range(args.warmup):
        fn()
    return time_events(torch, fn, args.steps)


def copy_calibration(torch, args) -> int:
    """A plain elementwise copy of 1 GiB (read 1 GiB, write 1 GiB): what the HBM
    byte counters read for a kernel whose traffic is known."""
    n = 1 << 28
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    ms = timed_steps(torch, lambda: torch.neg(src, out=dst), args)
    nbytes = 2 * n * 4
    print(json.dumps({"kind": "copy", "bytes_per_launch": nbytes, "ms": ms, "gbs": nbytes / (ms * 1e-3) / 1e9}),
          flush=True)
    return 0


def prefilter_only(torch, aa, args) -> int:
    """The largest single prefilter: 4096x2048 source (8 sun lobes up to 1e4)
    -> 1024^2 x6 cube, 5 roughness mips, 4096 samples per texel."""
    env = aa.synth.env(seed=SEED0 + 2, width=4096, height=2048, suns=8, peak_max=1e4)
    pyr = aa.env_pyramid(torch.from_numpy(env).cuda())
    texels = aa.cube_texels(1024, 5)
    ms = timed_steps(torch, lambda: aa.prefilter_cube(pyr, 4096, 2048, 1024, 5, 4096), args)
    print(json.dumps({"kind": "prefilter", "workload": "4096x2048 -> 1024^2x6x5 mips, 4096 spp", "ms": ms,
                      "texels_per_s": texels / (ms * 1e-3), "nominal_samples_per_s": texels * 4096 / (ms * 1e-3)}),
          flush=True)
    return 0


def irradiance_only(torch, aa, args) -> int:
    """The irradiance convolution of the 512x256 source into a 32^2 x6 cube,
    45000 samples per texel (300 theta x 150 phi rows)."""
    env = aa.synth.env(seed=SEED0, width=512, height=256)
    pyr = aa.env_pyramid(torch.from_numpy(env).cuda())
    ms = timed_steps(torch, lambda: aa.irradiance_cube(pyr, 512, 256, 32), args)
    print(json.dumps({"kind": "irradiance", "workload": "512x256 -> 32^2x6, 45000 samples per texel", "ms": ms,
                      "samples_per_s": 6 * 32 * 32 * 45000 / (ms * 1e-3)}), flush=True)
    return 0


def mlp_only(torch, aa, args) -> int:
    """The MLP kernels at their configured sizes: training steps at batch
    65536 on 2^22 labelled queries of the headline 2048x1024 map, then the
    4K MLP-IBL resolve with every monitor cell accepted (the full-frame
    inference: every covered pixel with roughness above mip 0's band goes
    through the MFMA finishing kernel)."""
    from althea_amd import mlp as M

    env = aa.synth.env(seed=SEED0 + 1, width=2048, height=1024)
    desc = aa.IBLDesc(irradiance_size=32, prefiltered_size=512, prefiltered_mips=6, prefilter_samples=1024,
                      brdf_lut_size=128, brdf_lut_samples=1024)
    ibl = aa.ImageBasedLighting.createResources(torch.from_numpy(env).cuda(), desc)
    scene = aa.synth.scene(seed=SEED0 + 1, width=WIDTH, height=HEIGHT, lights=LIGHTS)
    gbuf = aa.GBufferResources(*scene.gbuffer_tensors("cuda"))
    lights = scene.lights_tensor("cuda")
    dr, labels = M.training_set(ibl.environment_tensor(), 2048, 1024, 1 << 22, seed=SEED0 + 4, samples=1024)
    net = M.PrefilterMLP(seed=SEED0 + 5, lr=1e-3, init_output_bias=float(np.log(labels[:, :3].mean().item())))
    batch = 65536
    x, y = dr[:batch].contiguous(), labels[:batch].contiguous()
    train_ms = timed_steps(torch, lambda: net.train_step(x, y), args)
    v = ibl.views()
    net.build_monitor(ibl.prefiltered_tensor(), v.prefiltered_size, v.prefiltered_mips, threshold=float("inf"))
    out = torch.empty((WIDTH * HEIGHT, 4), dtype=torch.float32, device="cuda")
    used = torch.zeros(WIDTH * HEIGHT, dtype=torch.uint8, device="cuda")
    ms = timed_steps(torch, lambda: aa.DeferredRendering.resolve_mlp(gbuf, ibl, net, lights, scene.globals, scene.ssao,
                                                                      out=out, mlp_used=used), args)
    print(json.dumps({"kind": "mlp", "train_batch": batch, "train_ms_per_step": train_ms,
                      "resolve_every_cell_ms": ms, "mlp_pixels": int(used.sum().item()),
                      "weights_bytes": net.state_bytes()}), flush=True)
    return 0


def main() -> int:
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the resolve and precompute are single-GPU kernels here "
              "(the resolve would shard by screen bands, the precompute by map, face and tile); use --gpus 1",
              file=sys.stderr)
        return 2
    import torch

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import althea_amd as aa

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.only == "copy":
        return copy_calibration(torch, args)
    if args.only == "prefilter":
        return prefilter_only(torch, aa, args)
    if args.only == "ssr":
        return ssr_only(torch, aa, args)
    if args.only == "shadow":
        return shadow_only(torch, aa, args)
    if args.only == "gbuffer":
        return gbuffer_only(torch, aa, args)
    if args.only == "irradiance":
        return irradiance_only(torch, aa, args)
    if args.only == "mlp":
        return mlp_only(torch, aa, args)
    env = aa.synth.env(seed=SEED0 + 1, width=2048, height=1024)
    desc = aa.IBLDesc(irradiance_size=32, prefiltered_size=512, prefiltered_mips=6, prefilter_samples=1024,
                      brdf_lut_size=128, brdf_lut_samples=1024)
    ibl = aa.ImageBasedLighting.createResources(torch.from_numpy(env).cuda(), desc)
    scene = aa.synth.scene(seed=SEED0 + 1, width=WIDTH, height=HEIGHT, lights=LIGHTS)
    gbuf = aa.GBufferResources(*scene.gbuffer_tensors("cuda", fp16=args.fp16_gbuffer))
    lights = scene.lights_tensor("cuda")
    glob, ssao = scene.globals, scene.ssao
    out = torch.empty((WIDTH * HEIGHT, 4), dtype=torch.float32, device="cuda")

    def step():
        aa.DeferredRendering.resolve(gbuf, ibl, lights, glob, ssao, out=out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        step()
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(end) / args.steps
    px = WIDTH * HEIGHT
    value = px / (ms * 1e-3) / 1e6
    nbytes = algorithmic_bytes(ibl, args.fp16_gbuffer)
    achieved = nbytes / (ms * 1e-3) / 1e9
    line = {
        "metric": "4K deferred PBR+IBL Mpix/s & prefilter texels/s; MLP vs cubemap IBL @<=1e-3 err",
        "value": value, "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "wall_s": wall, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded fBm G-buffer, log-normal + sun-lobe HDR map)",
        "config": {"workload": f"deferred resolve {WIDTH}x{HEIGHT}, {'FP16' if args.fp16_gbuffer else 'FP32'} "
                               f"G-buffer 4xvec4, {LIGHTS} point lights, SSAO 16, IBL 512^2x6 cube 6 mips",
                   "pixels": px, "lights": LIGHTS, "gbuffer": "fp16" if args.fp16_gbuffer else "fp32"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": RESOLVE_TRAFFIC_BYTES.get("fp16" if args.fp16_gbuffer else "fp32"),
                     "algorithmic_bytes": nbytes, "kernel": "resolve_kernel",
                     "traffic_depth_pass": DEPTH_PASS_TRAFFIC_BYTES.get("fp16" if args.fp16_gbuffer else "fp32"),
                     # What binds the kernel: VALU issue (64 lights per pixel), not HBM.
                     "limiter": "valu", "valu": valu_issue(ms, args.fp16_gbuffer),
                     "counters_from": "profiles/resolve_kernel_4k_fp32_64lights_ssao16.md"},
    }
    if not args.no_secondary and args.only is None:
        line["secondary"] = secondary(torch, aa)
        line["secondary"]["resolve_fp16_gbuffer"] = fp16_resolve(torch, aa, scene, ibl, lights, args.steps)
        line["secondary"]["ssr"] = ssr_frame(torch, aa, scene, ibl, gbuf, lights, args.steps)
        line["secondary"]["shadows"] = shadow_frame(torch, aa, scene, ibl, gbuf, args.steps)
        line["secondary"]["gbuffer"] = gbuffer_frame(torch, aa, scene, ibl, args.steps)
        line["secondary"]["gbuffer"]["gltf"] = gltf_frame(torch, aa, scene, ibl, args.steps)
        line["secondary"]["full_frame"] = full_frame(torch, aa, scene, ibl, args.steps)
        line["secondary"]["mlp_ibl_resolve"] = mlp_ibl(torch, aa, ibl, scene, gbuf, lights, env, args.steps)
        line["secondary"]["mlp_config4_recipe"] = config4_recipe(torch, aa)
    if not args.no_cpu_baseline and args.only is None:
        line["cpu_baseline"] = cpu_baseline(aa, scene, ibl, env)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
