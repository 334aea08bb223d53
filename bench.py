"""Wavefront path-tracing benchmark on one MI355X.

Workload: the 1024 x 1024-cell heightfield (2,097,152 triangles, seeded
6-octave value noise, 64 emissive triangles, per-tile GGX PreviewSurface
materials), 1920 x 1080, 1 sample per pixel per frame, max depth 6, next
event estimation with MIS. A step is one frame: camera rays, every bounce's
extension and shadow rays, and the resolve into the accumulation, on scene
data already resident in HBM.

Prints one JSON line: whole-frame Mrays/s (camera + extension + shadow rays
over the wall time of the timed frames), Mrays/s by ray kind from hipEvent
kernel times, the dominant kernel's roofline, the CPU oracle's rate on a
bounded sample of the same frame, and the equal-relMSE race: the GPU ms the
radiance-cache form (a fresh cache trained online, 65536 training paths per
frame, Adam lr 1e-2) needs to reach the relMSE that K plain 1-spp frames
reach, against a high-spp plain reference, on the same heightfield at 1080p
and on the Cornell box (1024^2, depth 8).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "hydra-lotus_amd"))

import numpy as np  # noqa: E402

from lotus_amd import hip, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0
# VALU issue capacity: 256 CUs x 2.4 GHz, one wave64 VALU instruction per CU
# per clock (4 SIMDs, 4 clocks each; packed FP32 is compiled out)
VALU_ISSUE_PER_S = 256 * 2.4e9
NODE_BYTES = 128
TRI_BYTES = 48
# Queue bytes per ray of each traversal kernel: camera writes a hit record
# (48 B) and the radiance word (16 B); intersect reads a ray (32 B) and writes
# a hit record (48 B); shadow reads a ray (32 B) and, unblocked, reads the
# delta and read-modify-writes the radiance (48 B).
QUEUE_BYTES = {"camera": 64, "intersect": 80, "shadow": 80}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(scene, width, height, depth, threads, target_s=25.0):
    """The CPU oracle on the same frame: enough samples per pixel for about
    `target_s` seconds of host work (first a 64-row band calibrates)."""
    import oracle

    o = oracle.OracleScene(scene)

    def run(rows, spp):
        y0 = max(0, height // 2 - rows // 2)
        f = oracle.make_frame(scene.world_to_clip, width, height, spp=spp, max_bounces=depth,
                              light_sampling=1, rect=(0, y0, width, rows))
        t0 = time.perf_counter()
        _, counts, shadow = o.render(f, threads=threads)
        dt = time.perf_counter() - t0
        return dt, int(counts.sum()) + int(shadow.sum())

    dt, _ = run(64, 1)
    spp = int(max(1, min(64, target_s / max(dt * height / 64, 1e-3))))
    dt, rays = run(height, spp)
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"the full {width}x{height} frame, {spp} spp, depth {depth}, NEE+MIS: "
                      f"{rays} rays in {dt:.1f} s"}


def relmse(img, ref):
    """Mean over pixels and channels of (I - R)^2 / (R^2 + 1e-2)."""
    return float(np.mean((img[..., :3].astype(np.float64) - ref) ** 2 / (ref * ref + 1e-2)))


def centred_crop(width, height, fraction):
    """A window of `fraction` of each side, centred, so its rows are the same
    counted from the top or from the bottom."""
    cw, ch = max(1, int(width * fraction)), max(1, int(height * fraction))
    x, y = (width - cw) // 2, (height - ch) // 2
    if (height - ch) % 2:
        ch += 1
    return x, y, cw, ch


def fp64_reference(scene, width, height, spp, crop, views=8, **kw):
    """The mean of `spp` samples per pixel over the crop (a data window of
    the full-size target). `views` renderers of the scene render it as the
    views of one replayed hipGraph batch per sample, each on its own sample
    stream; every 1024 samples each renderer's float32 sums are added in
    float64 on the host and its accumulation restarts (as the reference's
    --write-reference sums its batches, main.cpp:3406-3452)."""
    x, y, cw, ch = crop
    acc = np.zeros((ch, cw, 3), np.float64)
    rs = [hip.Renderer(0) for _ in range(views)]
    try:
        for v in rs:
            v.apply(scene, builder=BUILDER)
        done = 0
        while done < spp:
            k = min(1024, (spp - done) // views)
            if k == 0:
                raise ValueError("spp must be a multiple of the view count")
            params = [dict(world_to_clip=scene.world_to_clip, width=width, height=height, data_window=crop,
                           sample_index=(1 << 24) + done + j * k, **kw) for j in range(views)]
            for _ in range(k):
                hip.render_batch(rs, params, graph=True, want_color=False)
            for v in rs:
                sums, n = v.accumulation(width, height)
                assert n == k, (n, k)
                acc += sums[y:y + ch, x:x + cw, :3].astype(np.float64)
            done += views * k
    finally:
        for v in rs:
            v.close()
    return acc / spp


def variance_share(r, scene, width, height, depth, crop, ref, frames=256, **kw):
    """The share of a 1-spp frame's relMSE that survives when everything
    after the second vertex's emission is replaced by its mean: per-pixel
    sample variances of the full estimator and of the estimator stopped
    there (max_bounces 1: emission, next event estimation at the camera
    hit, the BSDF-sampled emission at the second vertex), weighted as
    relMSE weighs them. A radiance cache queried at the second vertex can
    at best remove the rest."""
    x, y, cw, ch = crop
    wgt = 1.0 / (ref * ref + 1e-2)
    out = {}
    for name, bounces in (("full", depth), ("first_two_vertices", 1)):
        s1 = np.zeros((ch, cw, 3))
        s2 = np.zeros((ch, cw, 3))
        for k in range(frames):
            c = r.render(scene.world_to_clip, width, height, sample_index=(1 << 25) + k, max_bounces=bounces,
                         data_window=crop, **kw)["color"][y:y + ch, x:x + cw, :3].astype(np.float64)
            s1 += c
            s2 += c * c
        var = (s2 - s1 * s1 / frames) / (frames - 1)
        out[name] = float(np.mean(var * wgt))
        if name == "full":
            per_pixel = np.sort((var * wgt).sum(-1).ravel())[::-1]
            top = per_pixel[: max(1, per_pixel.size // 1000)].sum() / max(per_pixel.sum(), 1e-30)
    return out["first_two_vertices"] / out["full"], dict(out, frames=frames,
                                                         share_of_the_top_tenth_percent_of_pixels=float(top))


def relmse_race(r, scene, width, height, depth, frames, ref_spp, cache_cfg, checkpoints=(4, 16, 64),
                crop_fraction=0.25):
    """Plain frames and radiance-cache frames of the full target, each
    accumulated from the start (the cache from a fresh initialisation, its
    training inside every frame's GPU time), scored over a centred crop
    against a ref_spp-sample float64 reference of that crop."""
    w2c = scene.world_to_clip
    kw = dict(max_bounces=depth, light_sampling=hip.LIGHT_MIS)
    crop = centred_crop(width, height, crop_fraction)
    x, y, cw, ch = crop
    t0 = time.perf_counter()
    ref = fp64_reference(scene, width, height, ref_spp, crop, **kw)
    ref_s = time.perf_counter() - t0
    share, variances = variance_share(r, scene, width, height, depth, crop, ref,
                                      light_sampling=hip.LIGHT_MIS)

    def curve(cache):
        t, out_c = 0.0, []
        last = None
        for f in range(frames):
            out = r.render(w2c, width, height, sample_index=0, cache=cache, **kw)
            t += out["frame_ms"]
            out_c.append((t, relmse(out["color"][y:y + ch, x:x + cw], ref)))
            last = out
        return out_c, last

    plain, _ = curve(False)
    r.cache_configure(**cache_cfg)
    cached, last = curve(True)
    race = {}
    for k in checkpoints:
        if k > frames:
            continue
        t_plain, level = plain[k - 1]
        hit = next((i for i, (_, e) in enumerate(cached) if e <= level), None)
        race[f"plain_{k}_frames"] = {
            "relmse": round(level, 7), "plain_gpu_ms": round(t_plain, 2),
            "cache_gpu_ms": round(cached[hit][0], 2) if hit is not None else None,
            "cache_frames": hit + 1 if hit is not None else None,
            "ratio": round(cached[hit][0] / t_plain, 3) if hit is not None else None}
    cost = (cached[-1][0] / frames) / (plain[-1][0] / frames)
    bound = {"relmse_share_of_first_two_vertices": round(share, 4),
             "variance_weighted": {k: float(f"{v:.4g}") for k, v in variances.items()},
             "cache_over_plain_frame_cost": round(cost, 3),
             "best_possible_ratio": round(share * cost, 3)}
    q, smp = last["cache_queries"], last["cache_samples"]
    steps = cache_cfg.get("train_steps", hip.CACHE_DEFAULTS["train_steps"])
    # MFMA work: 2 x 16384 multiply-adds per query forwards; a training
    # sample costs a forward, the backward and the weight gradient (3x)
    infer_tflops = 2 * 16384 * q / (last["cache_infer_ms"] * 1e-3) / 1e12 if last["cache_infer_ms"] else None
    train_tflops = 3 * 2 * 16384 * min(smp, 4 * cache_cfg.get("batch_capacity", 65536)) / \
        (last["cache_train_ms"] * 1e-3) / 1e12 if last["cache_train_ms"] else None
    return {"reference": {"spp": ref_spp, "accumulation": "float64 sums of float32 batches of 1024",
                          "crop": {"x": x, "y": y, "width": cw, "height": ch}, "seconds": round(ref_s, 1)},
            "frames": frames, "race": race, "cache_bound": bound,
            "plain_relmse": [float(f"{e:.4g}") for _, e in plain[:: max(1, frames // 16)]],
            "cache_relmse": [float(f"{e:.4g}") for _, e in cached[:: max(1, frames // 16)]],
            "plain_ms_per_frame": round(plain[-1][0] / frames, 3),
            "cache_ms_per_frame": round(cached[-1][0] / frames, 3),
            "cache_last_frame": {"queries": q, "samples": smp, "train_steps": steps,
                                 "infer_ms": round(last["cache_infer_ms"], 3),
                                 "train_ms": round(last["cache_train_ms"], 3), "loss": round(last["cache_loss"], 5),
                                 "infer_mfma_tflops": round(infer_tflops, 2) if infer_tflops else None,
                                 "train_mfma_tflops": round(train_tflops, 2) if train_tflops else None,
                                 "mfma_peak_tflops": 2500.0}}


def restir_svgf_run(r, width, height, frames, depth=8, candidates=32, history=20, radius=30):
    """ReSTIR DI (32 initial candidates, temporal reuse capped at 20, spatial
    reuse of the four axis neighbours at 30 px) and SVGF (32-sample history,
    5 a-trous iterations) over a seeded camera path through the Cornell box
    with 100k emissive ceiling triangles; GPU ms per temporal frame by pass
    and, at the last camera, relMSE of the output and of a plain NEE+MIS
    frame against a 512-spp plain reference."""
    sc = scenes.cornell_many_lights(camera_aspect=width / height)
    r.apply(sc, builder=BUILDER)
    tot = {"frame_ms": 0.0, "restir_ms": 0.0, "trace_ms": 0.0, "shade_ms": 0.0, "camera_ms": 0.0,
           "temporal_surface_ms": 0.0, "svgf_ms": 0.0}
    rays = 0
    out = None
    for f in range(frames):
        view = scenes.camera_spline(f, frames)
        w2c = scenes.multiply(sc.projection, view)
        out = r.render_temporal(w2c, view, width, height, frame_index=f + 1, accumulate=True,
                                max_history=32, filter_iterations=5, restir_temporal=True, restir_spatial=True,
                                radius=radius, max_history_candidates=history, max_bounces=depth,
                                light_sampling=hip.LIGHT_RESTIR, restir_candidates=candidates)
        for k in tot:
            tot[k] += out[k]
        rays += out["camera_rays"] + out["extension_rays"] + out["shadow_rays"]
    final = out["color"]
    ref = None
    for _ in range(512):
        ref = r.render(w2c, width, height, sample_index=1 << 22, max_bounces=depth, want_color=False)
    ref = r.render(w2c, width, height, sample_index=1 << 22, frame_count=0, max_bounces=depth)["color"][..., :3]
    ref = ref.astype(np.float64)
    plain = r.render(w2c, width, height, sample_index=77, max_bounces=depth)["color"]
    return {"workload": f"Cornell box + {sc.triangle_count - 34} emissive ceiling tris, {width}x{height}, "
                        f"{frames}-frame camera path, depth {depth}",
            "frames": frames,
            "gpu_ms_per_frame": {k.replace("_ms", ""): round(v / frames, 3) for k, v in tot.items()},
            "mrays_per_s": round(rays / (tot["frame_ms"] * 1e3), 1),
            "relmse_last_frame": {"restir_svgf": round(relmse(final, ref), 5), "plain_mis_1spp": round(relmse(plain, ref), 5)},
            "reference_spp": 512}


BUILDER = hip.BVH_SAH  # the BVH topology every workload's scene is built with (--builder)


def scene_update_run(r, scene, width, height, depth, steps, frames=4):
    """The heightfield animated (scenes.heightfield_wave): every step's plan
    keeps each mesh's vertex count and indices, so the renderer re-places the
    triangles and refits the BVH instead of building one (GpuScene::Apply's
    point refit, gpu_scene.cpp:1638-1671). Reports the refit's GPU time beside
    a build of the same plan, and frame times on the refitted tree beside a
    tree built for the last step's heights (the refit's quality cost)."""
    kw = dict(frame_count=1, max_bounces=depth, light_sampling=hip.LIGHT_MIS, want_color=False)
    refit_ms, upload_ms, uploaded = [], [], []
    apply_wall = {"compared": [], "vouched": []}
    moved = scene
    every = hip.UNCHANGED_TOPOLOGY | hip.UNCHANGED_POSITIONS | hip.UNCHANGED_TRANSFORMS | \
        hip.UNCHANGED_MATERIALS | hip.UNCHANGED_LIGHTS
    for k in range(steps):
        moved = scenes.heightfield_wave(scene, 0.35 * (k + 1))
        moved.packed()  # the host's plan, built before the call is timed
        # odd steps vouch for everything but the moved positions, as a host
        # that knows its SceneUpdate's flags does; even steps let the
        # library compare the whole plan
        vouch = k % 2 == 1
        t0 = time.perf_counter()
        before = r.apply(moved, builder=BUILDER, unchanged=every if vouch else 0,
                         moved_geometries=np.arange(len(moved.geometries)) if vouch else None)
        apply_wall["vouched" if vouch else "compared"].append((time.perf_counter() - t0) * 1e3)
        if before.update != hip.SCENE_REFIT:
            raise RuntimeError(f"animated step {k} was applied as update {before.update}, not as a refit")
        refit_ms.append(before.build_gpu_ms)
        upload_ms.append(before.upload_ms)
        uploaded.append(before.uploaded_bytes)
        r.render(moved.world_to_clip, width, height, **kw)

    def frame_ms(renderer):
        renderer.render(moved.world_to_clip, width, height, **kw)  # rounds from a previous sample
        t = 0.0
        for _ in range(frames):
            t += renderer.render(moved.world_to_clip, width, height, **kw)["frame_ms"]
        return t / frames

    refit_frame = frame_ms(r)
    fresh = hip.Renderer(0)
    try:
        ev = fresh.apply(moved, builder=BUILDER)
        built_frame = frame_ms(fresh)
        # the GPU build (Morton order), the fast one beside the host SAH topology
        gpu_build = fresh.apply(moved, builder=hip.BVH_LBVH).build_gpu_ms
    finally:
        fresh.close()
    r.apply(scene, builder=BUILDER)  # back to the still heightfield (a refit)
    return {"workload": f"heightfield {ev.triangle_count} tris, {steps} animated steps "
                        f"(heights + 0.05 sin(4x + 3z + t)), {width}x{height}, depth {depth}",
            "build_ms": round(ev.build_gpu_ms, 3),
            "apply_wall_ms": {k: round(float(np.mean(v)), 3) for k, v in apply_wall.items() if v},
            "build_gpu_ms_lbvh": round(gpu_build, 3),
            "refit_gpu_ms": round(float(np.mean(refit_ms)), 3),
            "refit_upload_wall_ms": round(float(np.mean(upload_ms)), 2),
            "refit_uploaded_mb": round((uploaded[-1] - uploaded[0]) / max(len(uploaded) - 1, 1) / 1e6, 1),
            "frame_gpu_ms_refitted_tree": round(refit_frame, 3),
            "frame_gpu_ms_built_tree": round(built_frame, 3)}


def lights_run(r, scene, width, height, depth, frames):
    """The heightfield lit by analytic lights as well as its emissive
    triangles: 16 seeded sphere lights above it and a sun (distant light of
    0.53 degrees), NEE+MIS. Every light is in the light table and every
    continuation ray tests the crossing hierarchy and, escaping, the sun."""
    lit = scene.copy()
    rng = np.random.default_rng(11)
    for k in range(16):
        x, z = rng.uniform(-0.9, 0.9, 2)
        lit.add_light(scenes.analytic_light(scenes.LIGHT_SPHERE, scenes.translation(x, rng.uniform(0.3, 0.6), z),
                                            radiance=tuple(rng.uniform(2.0, 8.0, 3)), radius=rng.uniform(0.01, 0.04)))
    sun = scenes.identity()
    sun[8:11] = (0.3, 1.0, 0.4)  # the frame's Z: towards the sun
    lit.add_light(scenes.analytic_light(scenes.LIGHT_DISTANT, sun, radiance=(3.0, 2.8, 2.5), angle=0.53,
                                        normalize=True))
    r.apply(lit, builder=BUILDER)
    kw = dict(frame_count=1, max_bounces=depth, light_sampling=hip.LIGHT_MIS, want_color=False)
    r.render(lit.world_to_clip, width, height, **kw)
    tot = {"rays": 0, "frame_ms": 0.0, "shade_ms": 0.0, "trace_ms": 0.0, "tail_ms": 0.0, "camera_ms": 0.0}
    for _ in range(frames):
        out = r.render(lit.world_to_clip, width, height, **kw)
        tot["rays"] += out["camera_rays"] + out["extension_rays"] + out["shadow_rays"]
        for k in ("frame_ms", "shade_ms", "trace_ms", "tail_ms", "camera_ms"):
            tot[k] += out[k]
    result = {"workload": f"heightfield {scene.triangle_count} tris + 16 sphere lights + a 0.53-degree sun, "
                          f"{width}x{height}, depth {depth}, NEE+MIS",
              "frames": frames,
              "mrays_per_s_gpu": round(tot["rays"] / tot["frame_ms"] / 1e3, 1),
              "gpu_ms_per_frame": {k.replace("_ms", ""): round(v / frames, 3) for k, v in tot.items() if k != "rays"}}
    # the same heightfield under a 1024 x 512 sky image instead of the sun
    sky = scene.copy()
    sky.environment_map = scenes.sky_image(1024, 512, sun=(0.3, 0.3), sun_radiance=200.0)
    r.apply(sky, builder=BUILDER)
    r.render(sky.world_to_clip, width, height, **kw)
    rays, ms = 0, 0.0
    for _ in range(frames):
        out = r.render(sky.world_to_clip, width, height, **kw)
        rays += out["camera_rays"] + out["extension_rays"] + out["shadow_rays"]
        ms += out["frame_ms"]
    result["sky_image"] = {"workload": "the heightfield under a 1024x512 sky image with a sun, NEE+MIS",
                           "mrays_per_s_gpu": round(rays / ms / 1e3, 1), "gpu_ms_per_frame": round(ms / frames, 3)}
    # the same heightfield with every tile's base colour and roughness read
    # from one 1024x1024 sRGB texture through per-corner coordinates
    tex = scene.copy()
    texels = np.random.default_rng(12).integers(0, 256, (1024, 1024, 4), dtype=np.uint8)
    t = tex.add_texture(texels, scenes.TEXTURE_RGBA8_SRGB)
    first = tex.add_binding(scenes.texture_binding(t, st_s=(4.0, 0.0, 0.0), st_t=(0.0, 4.0, 0.0)))
    tex.add_binding(scenes.texture_binding(t, channel=1, scale=(0.8,) * 4, bias=(0.1,) * 4))
    tiled = set()
    for g, _, m in tex.instances:
        if not tex.materials[m][2] and tex.geometries[g][1].shape[0] > 1:  # the tiles, not the emitters
            tex.materials[m] = scenes.with_lookups(tex.materials[m], first,
                                                   [scenes.SLOT_BASE_COLOR, scenes.SLOT_ROUGHNESS])
            tiled.add(g)
    for g in tiled:
        pos, tri = tex.geometries[g]
        tex.set_texcoords(g, pos[tri.reshape(-1)][:, [0, 2]] * 0.5 + 0.5)
    tex._packed = None
    r.apply(tex, builder=BUILDER)
    r.render(tex.world_to_clip, width, height, **kw)
    rays, ms = 0, 0.0
    for _ in range(frames):
        out = r.render(tex.world_to_clip, width, height, **kw)
        rays += out["camera_rays"] + out["extension_rays"] + out["shadow_rays"]
        ms += out["frame_ms"]
    result["textured"] = {"workload": "the heightfield with base colour and roughness looked up in a 1024x1024 "
                                      "sRGB texture (per-corner coordinates), NEE+MIS",
                          "mrays_per_s_gpu": round(rays / ms / 1e3, 1), "gpu_ms_per_frame": round(ms / frames, 3)}
    r.apply(scene, builder=BUILDER)  # back to the untextured scene (a rebuild)
    return result


def render_buffer_run(r, width, height, calls=200):
    """The colour product written into a host's UNorm8 render buffer (rows
    flipped, RGBA32F to UNorm8) and the depth product into a Float32 one,
    each call synchronised; algorithmic bytes: 16 + 4 and 4 + 4 per pixel."""
    import ctypes
    rt = ctypes.CDLL("libamdhip64.so")
    px = width * height
    ptrs = []

    def alloc(nbytes, fill):
        p = ctypes.c_void_p()
        if rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0 or \
                rt.hipMemset(p, fill, ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc / hipMemset failed")
        ptrs.append(p)
        return p.value

    # 0x3e bytes: every float 0.186 (inside [0, 1], so the conversion rounds)
    colour, unorm = alloc(px * 16, 0x3E), alloc(px * 4, 0)
    depth, dbuf = alloc(px * 4, 0x3E), alloc(px * 4, 0)
    res = {}
    try:
        for name, args, nbytes in (("colour_rgba32f_to_unorm8", (colour, hip.BUFFER_FLOAT32X4, unorm,
                                                                 hip.BUFFER_UNORM8X4), 20),
                                   ("depth_float32", (depth, hip.BUFFER_FLOAT32, dbuf, hip.BUFFER_FLOAT32), 8)):
            src, sf, dst, df = args
            r.write_render_buffer(src, sf, width, height, dst, df)
            t0 = time.perf_counter()
            for _ in range(calls):
                r.write_render_buffer(src, sf, width, height, dst, df)
            us = (time.perf_counter() - t0) / calls * 1e6
            res[name] = {"us_per_call": round(us, 2), "gb_per_s": round(nbytes * px / us / 1e3, 1)}
    finally:
        for p in ptrs:
            rt.hipFree(p)
    return {"workload": f"{width}x{height} products into bottom-up render buffers, host-synchronised calls",
            **res}


def batch_run(views, frames):
    """64 seeded views of 64^2..512^2 over their own icosphere rigs (1k..200k
    triangles, depth 2..12): each view autotuned, then frames rendered view
    by view (one call each), as one batch, as one hipGraph per frame, and as
    one fused batch (every view's paths in one path-megakernel launch fed by
    one device work queue)."""
    rs = []
    t_build = time.perf_counter()
    for v in views:
        r = hip.Renderer(0)
        r.apply(scenes.icosphere_rig(v["scene_triangles"], v["scene_seed"]), builder=BUILDER)
        rs.append(r)
    build_s = time.perf_counter() - t_build
    kws = [dict(world_to_clip=v["world_to_clip"], width=v["size"], height=v["size"], max_bounces=v["depth"],
                light_sampling=hip.LIGHT_MIS) for v in views]
    tunings = [r.autotune(trials=1, **kw) for r, kw in zip(rs, kws)]
    res = {}
    batch = hip.Batch(rs, kws)  # the views' records built once, as a serving loop keeps them
    graph_kinds = []
    for mode in ("per_view", "batch", "graph", "fused"):
        if mode == "fused":
            for r in rs:
                r.set_tuning(megakernel=1)
        for r, kw in zip(rs, kws):  # warm-up frame: rounds and targets settle
            r.render(want_color=False, sample_index=0, **kw)
        if mode == "graph":
            batch.run(graph=True, sample_index=0)  # captured and instantiated once; the timed frames replay it
        rays, gpu_ms = 0, 0.0
        t0 = time.perf_counter()
        for f in range(frames):
            if mode == "per_view":
                for r, kw in zip(rs, kws):
                    out = r.render(want_color=False, sample_index=f + 1, **kw)
                    rays += out["camera_rays"] + out["extension_rays"] + out["shadow_rays"]
                    gpu_ms += out["frame_ms"]
            else:
                n, ms, kind = batch.run(graph=mode == "graph", sample_index=f + 1, fused=mode == "fused")
                rays += n
                gpu_ms += ms
                if mode == "graph":
                    graph_kinds.append(kind)
        wall = time.perf_counter() - t0
        res[mode] = {"gpu_ms_per_frame": round(gpu_ms / frames, 3), "wall_ms_per_frame": round(wall / frames * 1e3, 3),
                     "mrays_per_s_gpu": round(rays / (gpu_ms * 1e3), 1), "views_per_s_wall": round(len(rs) * frames / wall, 1)}
    for r in rs:
        r.close()
    return {"workload": f"{len(views)} views {min(v['size'] for v in views)}^2..{max(v['size'] for v in views)}^2 over "
                        f"their own icosphere rigs ({min(v['scene_triangles'] for v in views)}.."
                        f"{max(v['scene_triangles'] for v in views)} tris), depth 2..12, NEE+MIS",
            "frames": frames, "scene_build_s": round(build_s, 2),
            "autotuned_dynamic_trace": sum(t["trace_dynamic"] for t in tunings),
            "autotuned_half_resident": sum(t["resident_percent"] == 50 for t in tunings),
            "autotuned_megakernel": sum(t["megakernel"] for t in tunings),
            "autotuned_camera_tiles": {n: sum(t["camera_tile"] == k and not t["megakernel"] for t in tunings)
                                       for k, n in enumerate(("8x8", "16x4", "4x16"))},
            "graph_frames_replayed": sum(k == 3 for k in graph_kinds), **res}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=6)
    ap.add_argument("--builder", choices=("lbvh", "sah"), default="sah",
                    help="BVH topology: Morton order on the GPU, or binned SAH on the host")
    ap.add_argument("--leaf-size", type=int, default=0, help="BVH leaf size (0: the default, 2)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kind-frames", type=int, default=16,
                    help="frames rendered after the timed ones with each ray kind in launches of its own")
    ap.add_argument("--race-frames", type=int, default=64,
                    help="frames of each form in the equal-relMSE race (0 skips it)")
    ap.add_argument("--race-reference-spp", type=int, default=65536)
    ap.add_argument("--batch-frames", type=int, default=16,
                    help="frames of the 64-view batched workload (0 skips it)")
    ap.add_argument("--restir-frames", type=int, default=120,
                    help="temporal frames of the ReSTIR + SVGF run (0 skips it)")
    ap.add_argument("--update-steps", type=int, default=8,
                    help="animated heightfield steps applied as BVH refits (0 skips them)")
    ap.add_argument("--lights-frames", type=int, default=16,
                    help="frames of the heightfield with analytic lights (0 skips them)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times: the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        a = torch.ones(256 * 1024 * 1024, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by image "
              "tiles, see the design notes)", file=sys.stderr)
        return 2

    log(f"building heightfield {args.cells}^2")
    scene = scenes.heightfield(cells=args.cells)
    r = hip.Renderer(0)
    global BUILDER
    BUILDER = hip.BVH_SAH if args.builder == "sah" else hip.BVH_LBVH
    ev = r.apply(scene, builder=BUILDER, leaf_size=args.leaf_size)
    log(f"scene: {ev.triangle_count} tris, {ev.node_count} nodes, depth {ev.bvh_depth}, "
        f"build {ev.build_gpu_ms:.2f} ms ({args.builder})")
    w2c = scene.world_to_clip
    kw = dict(frame_count=1, max_bounces=args.depth, light_sampling=hip.LIGHT_MIS, want_color=False)
    for _ in range(args.warmup):
        r.render(w2c, args.width, args.height, **kw)
    log(f"warmup done ({args.warmup} frames)")
    keys = ("camera_rays", "extension_rays", "shadow_rays", "camera_ms", "intersect_ms", "shade_ms",
            "shadow_ms", "trace_ms", "tail_ms", "resolve_ms", "frame_ms")
    tot = {k: 0.0 for k in keys}
    nodes = np.zeros(4)
    tris = np.zeros(4)
    walked = np.zeros(4)
    trace_launches = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        out = r.render(w2c, args.width, args.height, **kw)
        for key in tot:
            tot[key] += out[key]
        nodes += out["nodes"]
        tris += out["tris"]
        walked += out["walked"]
        trace_launches += out["trace_launches"]
    wall = time.perf_counter() - t0
    rays = tot["camera_rays"] + tot["extension_rays"] + tot["shadow_rays"]
    value = rays / wall / 1e6
    # rays by kind: the same frames with extension and shadow rays in launches
    # of their own, so each kind's kernel time is its own (not in the timed loop)
    split = {k: 0.0 for k in keys}
    split_walked = np.zeros(4)
    for k in range(args.kind_frames):
        out = r.render(w2c, args.width, args.height, split_kinds=True, **kw)
        for key in split:
            split[key] += out[key]
        split_walked += out["walked"]
    # each kind's rays walked by its own launches (the tail's are its own)
    by_kind = {
        "camera": tot["camera_rays"] / tot["camera_ms"] / 1e3,
        "extension": split_walked[1] / split["intersect_ms"] / 1e3 if split["intersect_ms"] else None,
        "shadow": split_walked[2] / split["shadow_ms"] / 1e3 if split["shadow_ms"] else None,
        "tail": (walked[3] / tot["tail_ms"] / 1e3) if tot["tail_ms"] else None,
        "frame_gpu": rays / tot["frame_ms"] / 1e3,
    }
    # roofline of the dominant kernel, k_trace (the rounds' extension + shadow
    # rays; the tail kernel walks the later bounces): algorithmic bytes =
    # queue records + the wide nodes and triangles the traversal tests
    # (counted by the traversal itself in the timed frames), over the
    # kernel's hipEvent time
    ext, sh = walked[1], walked[2]
    node_visits = nodes[1] + nodes[2]
    tri_tests = tris[1] + tris[2]
    bytes_total = node_visits * NODE_BYTES + tri_tests * TRI_BYTES + ext * QUEUE_BYTES["intersect"] + \
        sh * QUEUE_BYTES["shadow"]
    ms = tot["trace_ms"]
    achieved = bytes_total / (ms / 1e3) / 1e9
    roofline = {"bound": None, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_trace",
                "algorithmic_bytes_per_launch": int(bytes_total / max(trace_launches, 1)),
                "avg_launch_ms": round(ms / max(trace_launches, 1), 4),
                "nodes_per_ray": {"camera": round(nodes[0] / max(walked[0], 1), 2),
                                  "extension": round(nodes[1] / max(ext, 1), 2),
                                  "shadow": round(nodes[2] / max(sh, 1), 2),
                                  "tail": round(nodes[3] / max(walked[3], 1), 2)},
                "tris_per_ray": {"camera": round(tris[0] / max(walked[0], 1), 2),
                                 "extension": round(tris[1] / max(ext, 1), 2),
                                 "shadow": round(tris[2] / max(sh, 1), 2),
                                 "tail": round(tris[3] / max(walked[3], 1), 2)}}
    # What bounds it, from the committed counter summary of the same kernel
    # on the same workload (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE
    # bytes, SQ_INSTS_VALU and active lanes per launch): the measured HBM
    # fraction against the VALU issue fraction over this run's launch time.
    # Neither near its peak means the kernel is bound by memory latency (a
    # walk's dependent node fetches) and lane divergence.
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            t = json.load(f)
        if t.get("kernel") == roofline["kernel"]:
            launch_s = roofline["avg_launch_ms"] / 1e3
            roofline["traffic"] = t.get("hbm_bytes_per_launch")
            hbm_frac = t["hbm_bytes_per_launch"] / launch_s / 1e9 / HBM_PEAK_GBS
            valu_frac = t["valu_insts_per_launch"] / (VALU_ISSUE_PER_S * launch_s)
            lanes = t["valu_active_lanes"]
            bound = "latency" if max(hbm_frac, valu_frac) < 0.5 else ("hbm" if hbm_frac >= valu_frac else "valu")
            roofline["bound"] = bound
            roofline["measured"] = {"hbm_frac": round(hbm_frac, 4), "valu_issue_frac": round(valu_frac, 4),
                                    "valu_active_lanes_of_64": round(lanes, 1), "l2_hit": t.get("l2_hit"),
                                    "achieved_is": "demand bytes of node, triangle and queue touches, most "
                                                   "served by L2 and the Infinity Cache"}
    line = {
        "metric": "Mrays/s by ray kind + GPU ms to reach equal relMSE (plain vs MLP cache), 1080p",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural heightfield)",
        "config": {"workload": f"wavefront PT, {args.cells}x{args.cells} heightfield "
                               f"({ev.triangle_count} tris), {args.width}x{args.height}, 1 spp/frame, "
                               f"depth {args.depth}, NEE+MIS",
                   "frames": args.steps, "bvh": f"4-wide, {ev.node_count} nodes, depth {ev.bvh_depth}, "
                                                f"{args.builder} topology, build {ev.build_gpu_ms:.1f} ms"},
        "mrays_by_kind": {k: (round(v, 1) if v is not None else None) for k, v in by_kind.items()},
        "rays_per_frame": {"camera": tot["camera_rays"] / args.steps,
                           "extension": tot["extension_rays"] / args.steps,
                           "shadow": tot["shadow_rays"] / args.steps},
        "gpu_ms_per_frame": {k.replace("_ms", ""): round(tot[k] / args.steps, 3)
                             for k in ("camera_ms", "shade_ms", "trace_ms", "tail_ms", "resolve_ms", "frame_ms")},
        "roofline": roofline,
    }
    if args.update_steps > 0:
        log("scene updates (refits) ...")
        line["scene_update"] = scene_update_run(r, scene, args.width, args.height, args.depth, args.update_steps)
    if args.lights_frames > 0:
        log("analytic lights ...")
        line["analytic_lights"] = lights_run(r, scene, args.width, args.height, args.depth, args.lights_frames)
    if args.lights_frames > 0:
        line["render_buffer_writes"] = render_buffer_run(r, args.width, args.height)
    if args.race_frames > 0:
        cache_cfg = dict(log2_table=19, train_paths=65536, batch_capacity=65536, train_steps=4, learning_rate=1e-2)
        log("equal-relMSE race: heightfield ...")
        race = {"heightfield_1080p": relmse_race(r, scene, args.width, args.height, args.depth, args.race_frames,
                                                 args.race_reference_spp, cache_cfg)}
        log("equal-relMSE race: Cornell box ...")
        box = scenes.cornell_lambert()
        r.apply(box, builder=BUILDER)
        race["cornell_1024"] = relmse_race(r, box, 1024, 1024, 8, args.race_frames, args.race_reference_spp,
                                           cache_cfg)
        line["equal_relmse"] = race
    if args.batch_frames > 0:
        log("batched views ...")
        line["batched_views"] = batch_run(scenes.batch_views(), args.batch_frames)
    if args.restir_frames > 0:
        log("ReSTIR + SVGF run ...")
        line["restir_svgf"] = restir_svgf_run(r, args.width, args.height, args.restir_frames)
    r.close()
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline(scene, args.width, args.height, args.depth, args.cpu_threads)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
