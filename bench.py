#!/usr/bin/env python3
"""Throughput of the gfx950 wavefront path tracer on the 737K-triangle workload.

One step renders the whole frame: 6x6 displaced icospheres (737,290 triangles),
4 quad lights, a 1024x512 sky + sun environment, 1024x1024 pixels, 64 samples
per pixel, max depth 8 (adaptive sampling off, no clamps).  Inputs are resident
in HBM before the timed region.  Prints one JSON line.

  value     samples/s = W * H * spp * steps / wall time, the steps run as a
            caller runs them (no per-kernel events, graph capture on)
  roofline  the closest-hit (extend) kernel: algorithmic bytes = 48 B per
            active closest ray (32 B ray read + 16 B hit written), divided by
            the kernel's HIP-event time, taken in further steps of the same
            frame that record an event pair around every launch
  frame_roofline  the whole frame's algorithmic bytes over the step time
  cpu_baseline  the oracle (CPU restatement of the reference integrator) on a
            band of rows of the same frame, on this host's cores
  gpu_baseline  the same frame rendered with the reference's algorithmic
            structure on this GPU (phx_render_reference_structure), and the
            speedup of the wavefront backend over it
  batched   16 views of uneven size in one phx_render_batch against the same
            16 views as separate renders
  adjoint   path-replay adjoint time over primal time on the differentiable
            workload
  volume    the 256^3 heterogeneous-medium workload: frame time, and 8 orbit
            views batched against 8 separate renders
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def batched_form(pt_c1, c1, args):
    """16 views in one phx_render_batch against the same 16 views as plain
    launches: resolutions from {64..1024}^2 and spp from {4..64} (seeded); views
    0-7 are the 737K-triangle scene from orbit cameras, views 8-15 Cornell
    boxes with random wall colours.  Returns the figures for the JSON line."""
    import numpy as np
    import torch

    from photon_amd import abi, scenes
    from photon_amd.backend import BatchRenderer, HipBackend, PathTracer, make_camera

    draw = scenes.batch_draw(2024, 16)
    res = [d["res"] for d in draw]
    spp = [d["spp"] for d in draw]
    views = []
    for d in draw:
        w = h = d["res"]
        if d["scene"] == "spheres":
            pt, s = pt_c1, c1
        else:
            s = scenes.cornell_box(wall_colors=d["walls"])
            pt = PathTracer(HipBackend(device=0))
            pt.set_scene(s)
        cam = make_camera(scenes.orbit_eye(s, d["angle"]), s.camera["look_at"], s.camera["vup"], s.camera["vfov"], 1.0)
        views.append((pt, cam, abi.default_render_params(w, h, d["spp"], args.depth)))
    total = sum(p.width * p.height * p.samples_per_pixel for _, _, p in views)

    def plain_pass():
        per = []
        for pt, cam, p in views:
            t = time.perf_counter()
            pt.camera, pt.params = cam, p
            pt.render(read=False)
            torch.cuda.synchronize()
            per.append(time.perf_counter() - t)
        return sum(per), per

    plain_pass()  # warm-up: allocations
    seq, per_view = min((plain_pass() for _ in range(args.batch_reps)), key=lambda x: x[0])
    # views of 2^22 samples or more fill the GPU alone (the batch streams them
    # in runs of their own): their plain time bounds the batch from below
    solo = [v for v, (_, _, p) in enumerate(views) if p.width * p.height * p.samples_per_pixel >= 1 << 22]
    solo_ms = sum(per_view[v] for v in solo) * 1e3
    br = BatchRenderer(device=0)
    tuning = None
    if args.autotune:
        t, ms = br.autotune(views, repeats=1)
        tuning = {f: getattr(t, f) for f, _ in t._fields_}
        print(f"[bench] autotuned batch shape {tuning}: {ms:.1f} ms", file=sys.stderr, flush=True)
    br.render(views)  # warm-up: allocations, graph capture
    torch.cuda.synchronize()
    times = []
    for i in range(args.batch_reps):
        t = time.perf_counter()
        br.render(views)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        print(f"[bench] batched step {i} {times[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    tb = min(times)
    # each view of the batch equals its plain render (checked on the smallest)
    small = min(range(16), key=lambda v: views[v][2].width * views[v][2].samples_per_pixel)
    pt, cam, p = views[small]
    pt.camera, pt.params = cam, p
    same = bool(np.array_equal(pt.render().color, br.read(small).color))
    return {"views": 16, "samples": total, "ms": tb * 1e3, "samples_per_s": total / tb,
            "sequential_ms": seq * 1e3, "sequential_samples_per_s": total / seq, "speedup": seq / tb,
            "plain_ms_per_view": [round(x * 1e3, 3) for x in per_view],
            "gpu_filling_views": solo, "gpu_filling_plain_ms": solo_ms,
            "speedup_bound": seq * 1e3 / solo_ms if solo_ms > 0 else None,
            "resolutions": [int(x) for x in res], "spp": [int(x) for x in spp], "tuning": tuning,
            "equal_to_plain": same}


def volume_form(args):
    """The volume workload: a 256^3 fBm density grid (threshold 0.4, scale 20)
    with a 32^3 majorant grid and a 256-entry transfer function in an open box
    lit by a quad light and the sky environment; delta tracking with NEE ratio
    tracking, albedo 0.9, 1024x1024, 32 spp, max depth 16.  Plain frame time,
    then 8 orbit views in one batched launch against 8 plain renders."""
    import numpy as np
    import torch

    from photon_amd import abi, scenes
    from photon_amd.backend import BatchRenderer, HipBackend, PathTracer, make_camera

    t = time.time()
    s = scenes.fog_box(n=256, macro=32)
    gen = time.time() - t
    pt = PathTracer(HipBackend(device=0))
    pt.set_scene(s)

    def cam_at(ang):
        return make_camera(scenes.orbit_eye(s, ang), s.camera["look_at"], s.camera["vup"], s.camera["vfov"], 1.0)

    params = abi.default_render_params(1024, 1024, 32, 16)
    pt.camera, pt.params = cam_at(0.0), params
    pt.render(read=False)
    times = []
    for _ in range(args.volume_reps):
        pt.render(read=False)
        times.append(pt.ctx.stats().last_render_ms)
    ms = min(times)
    st = pt.ctx.stats()
    samples = 1024 * 1024 * 32
    views = [(pt, cam_at(a), params) for a in np.linspace(-0.6, 0.6, 8)]
    seq = 0.0
    for p_, cam, prm in views:
        p_.camera, p_.params = cam, prm
        p_.render(read=False)
        seq += p_.ctx.stats().last_render_ms
    br = BatchRenderer(device=0)
    br.render(views)
    torch.cuda.synchronize()
    br.render(views)
    bms = br.ctx.stats().last_render_ms
    print(f"[bench] volume: frame {ms:.1f} ms, 8 views batched {bms:.1f} ms vs {seq:.1f} ms", file=sys.stderr,
          flush=True)
    return {"workload": "fog box 256^3 fBm (32^3 majorants), quad light + sky, 1024x1024, 32 spp, max depth 16",
            "ms": ms, "samples_per_s": samples / (ms * 1e-3),
            "mrays_per_s": (st.closest_rays + st.shadow_rays) / (ms * 1e-3) / 1e6,
            "batched_8_views_ms": bms, "sequential_8_views_ms": seq, "batched_speedup": seq / bms,
            "scene_generation_s": gen}


def adjoint_form(args):
    """Path-replay adjoint against its primal on the differentiable workload:
    Cornell box with a textured floor (512^2 RGB albedo, initialised to 0.5)
    and 32 trainable Disney materials; target from hidden seeded parameters and
    a 3-octave coloured-noise texture; L2 loss at 256^2, 32 primal spp, 32
    decorrelated adjoint spp, max depth 6.  Times are the library's own (the
    call, synchronised)."""
    import torch

    from photon_amd import scenes
    from photon_amd.backend import make_camera
    from photon_amd.diff import DiffRenderer

    tex_target = scenes.colored_noise_texture(512, 99)
    target_scene = scenes.textured_cornell(n_materials=32, tex_size=512, texture=tex_target, seed=77)
    scene = scenes.textured_cornell(n_materials=32, tex_size=512, seed=5)
    cam = make_camera(scene.camera["look_from"], scene.camera["look_at"], scene.camera["vup"], scene.camera["vfov"],
                      1.0)
    tgt = DiffRenderer(target_scene, cam, 256, 256, spp=32, max_depth=6, n_train=32)
    target = tgt.render(*tgt.initial_params(target_scene)).detach()
    dr = DiffRenderer(scene, cam, 256, 256, spp=32, max_depth=6, adjoint_spp=32, n_train=32)
    params, tex = dr.initial_params(scene)
    params.requires_grad_(True)
    tex.requires_grad_(True)
    primal, adjoint = [], []
    for i in range(1 + args.adjoint_reps):
        params.grad = None
        tex.grad = None
        loss = ((dr.render(params, tex) - target) ** 2).mean()
        loss.backward()
        torch.cuda.synchronize()
        if i > 0:  # the first pass warms up allocations
            primal.append(dr.primal_ms)
            adjoint.append(dr.adjoint_ms)
    p_ms, a_ms = min(primal), min(adjoint)
    print(f"[bench] adjoint: primal {p_ms:.2f} ms, adjoint {a_ms:.2f} ms", file=sys.stderr, flush=True)
    return {"workload": "cornell + 512^2 textured floor, 32 Disney materials, 256x256, 32+32 spp, max depth 6",
            "primal_ms": p_ms, "adjoint_ms": a_ms, "ratio": a_ms / p_ms, "loss": float(loss.item()),
            "params": int(params.numel() + tex.numel())}


def adaptive_form(pt, args, plain_ms):
    """The plain workload's view with adaptive sampling on (pathtracer.cpp:
    676-723; the workload itself is quoted with it off): the frame time, the
    samples the pixels took and the rate of taken samples."""
    import numpy as np
    import torch

    from photon_amd import abi
    keep = pt.params
    p = abi.default_render_params(args.width, args.height, args.spp, args.depth)
    p.adaptive_sampling = 1
    pt.params = p
    pt.render(read=False)  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.adaptive_reps):
        pt.render(read=False)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / args.adaptive_reps
    counts = pt.sample_counts().astype(np.int64)
    pt.params = keep
    taken = int(counts.sum())
    total = args.width * args.height * args.spp
    print(f"[bench] adaptive: {ms:.1f} ms, {taken / total:.3f} of the samples taken", file=sys.stderr, flush=True)
    return {"ms_per_frame": ms, "reps": args.adaptive_reps, "plain_ms_per_frame": plain_ms,
            "samples_taken": taken, "fraction_taken": taken / total,
            "pixels_stopped_at_16": float(np.mean(counts == 16)),
            "taken_samples_per_s": taken / (ms / 1e3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=256)
    ap.add_argument("--baseline-steps", type=int, default=2,
                    help="timed frames of the reference-structure GPU baseline (0: skip)")
    ap.add_argument("--batch-reps", type=int, default=3,
                    help="timed repetitions of the batched 16-view launch and of its 16 plain launches (0: skip)")
    ap.add_argument("--autotune", action="store_true",
                    help="batched form with the launch shape phx_autotune picks (default: the library's default "
                         "shape, which sets the refill threshold per run from its tree's depth)")
    ap.add_argument("--adjoint-reps", type=int, default=3, help="timed primal+adjoint passes (0: skip)")
    ap.add_argument("--volume-reps", type=int, default=2, help="timed frames of the volume workload (0: skip)")
    ap.add_argument("--adaptive-reps", type=int, default=2,
                    help="timed frames of the plain view with adaptive sampling on (0: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 1 << 28  # 1 GiB of float32
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not offered: this build runs on one GPU "
              "(the path shards by sample_offset ranges)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from photon_amd import abi, scenes
    from photon_amd.backend import HipBackend, PathTracer, make_camera

    t = time.time()
    scene = scenes.sphere_grid(seed=1234)
    gen_s = time.time() - t
    pt = PathTracer(HipBackend(device=0))
    t = time.time()
    pt.set_scene(scene)
    build_s = time.time() - t
    cam = make_camera(scene.camera["look_from"], scene.camera["look_at"], scene.camera["vup"], scene.camera["vfov"],
                      args.width / args.height)
    pt.camera = cam
    pt.params = abi.default_render_params(args.width, args.height, args.spp, args.depth)
    print(f"[bench] scene {scene.triangle_count} tris generated in {gen_s:.1f}s, set_scene {build_s:.1f}s",
          file=sys.stderr, flush=True)

    for i in range(args.warmup):
        pt.render(read=False)
        print(f"[bench] warmup {i} {pt.ctx.stats().last_render_ms:.1f} ms", file=sys.stderr, flush=True)
    # the timed steps: the default path (no per-launch events, the graph on)
    closest = shadow = shade = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        pt.render(read=False)
        st = pt.ctx.stats()
        closest += st.closest_rays
        shadow += st.shadow_rays
        shade += st.shade_events
        print(f"[bench] step {i} {st.last_render_ms:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the kernel split, in separate steps that record an event pair around
    # every kernel launch on the launch's stream: the roofline's kernel time
    pt.ctx.set_tuning(time_kernels=1)
    k_steps = max(2, args.steps // 4)
    k_closest = 0
    ext_ms = shd_ms = sdw_ms = 0.0
    ext_launches = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(k_steps):
        pt.render(read=False)
        st = pt.ctx.stats()
        k_closest += st.closest_rays
        ext_ms += st.extend_ms
        shd_ms += st.shade_ms
        sdw_ms += st.shadow_ms
        ext_launches += st.extend_launches
    torch.cuda.synchronize()
    ms_instrumented = (time.perf_counter() - t0) * 1e3 / k_steps
    pt.ctx.set_tuning(time_kernels=0)

    npix = args.width * args.height
    samples = npix * args.spp * args.steps
    ms_per_step = wall * 1e3 / args.steps
    res = pt.read_result()
    assert np.all(np.isfinite(res.color)), "non-finite pixels in the benchmark frame"

    alg_bytes = 48.0 * k_closest  # over all extend launches of the kernel-timing steps
    achieved = alg_bytes / (ext_ms * 1e-3) / 1e9 if ext_ms > 0 else None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                "kernel": "k_extend", "launches": ext_launches,
                "avg_launch_ms": ext_ms / max(ext_launches, 1),
                "alg_bytes_per_launch": alg_bytes / max(ext_launches, 1)}
    traffic_file = os.path.join(ROOT, "profiles", "extend_traffic.json")
    if os.path.exists(traffic_file):
        tr = json.load(open(traffic_file))
        roofline["traffic"] = tr.get("hbm_bytes_per_launch")
        roofline["traffic_source"] = ("stored: calibrated rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE of k_extend on this "
                                      "workload (" + tr.get("profile", "profiles/extend_traffic.json") + "), not "
                                      "measured in this run")
    # the whole frame against the same roofline: algorithmic bytes of every
    # kernel (48 B per closest ray, 36 B per shadow ray, 128 B per shading
    # event, 64 B per sample, 16 B per pixel) plus the scene read once
    st = pt.ctx.stats()
    frame_bytes = (48.0 * closest + 36.0 * shadow + 128.0 * shade) / args.steps + 64.0 * npix * args.spp + \
        16.0 * npix + float(st.scene_bytes)
    frame = {"alg_bytes": frame_bytes, "achieved": frame_bytes / (ms_per_step * 1e-3) / 1e9, "unit": "GB/s"}
    frame["frac"] = frame["achieved"] / HBM_PEAK_GBS

    # Same frame with the reference's algorithmic structure on this GPU (the
    # stand-in for the reference's Kokkos-HIP build): binary median-split BVH,
    # full-width bounce loop with a host readback per iteration.
    gpu_base = None
    if args.baseline_steps > 0:
        pt.render(read=False, structure="reference")  # builds its BVH, warms up
        torch.cuda.synchronize()
        tb = time.perf_counter()
        for i in range(args.baseline_steps):
            pt.render(read=False, structure="reference")
            print(f"[bench] reference-structure step {i} {pt.ctx.stats().last_render_ms:.1f} ms",
                  file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        bwall = time.perf_counter() - tb
        bval = npix * args.spp * args.baseline_steps / bwall
        gpu_base = {"kind": "reference-structure proxy (binary median BVH, 64-entry stack, full-width "
                            "max_depth+9 bounce loop, per-iteration host readback)",
                    "value": bval, "unit": "samples/s", "ms_per_step": bwall * 1e3 / args.baseline_steps,
                    "steps": args.baseline_steps, "speedup": (samples / wall) / bval}

    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        cores = min(16, os.cpu_count() or 1)
        o = oracle.OracleScene(scene)
        rows = min(args.cpu_rows, args.height)
        y0 = (args.height - rows) // 2
        ref, _, secs = o.render(cam, pt.params, threads=cores,
                                pixel_range=(y0 * args.width, (y0 + rows) * args.width))
        cpu_samples = rows * args.width * args.spp
        # the oracle's rows against the same rows of the GPU frame
        g, r = res.color[y0:y0 + rows].astype(np.float64), ref[y0:y0 + rows].astype(np.float64)
        rel_mse = float(np.mean((g - r) ** 2 / (r ** 2 + 1e-2)))
        equal = float(np.mean(np.abs(g - r) <= 1e-3 * (1 + np.abs(r))))
        cpu = {"value": cpu_samples / secs, "unit": "samples/s", "cores": cores, "kind": "port",
               "sample": f"rows {y0}-{y0 + rows - 1} of the same 1024x1024x64spp frame ({cpu_samples} samples, "
                         f"{secs:.1f} s)",
               "parity": {"rel_mse_vs_gpu": rel_mse, "pixels_equal_1e-3": equal}}

    adaptive = adaptive_form(pt, args, ms_per_step) if args.adaptive_reps > 0 else None
    batched = batched_form(pt, scene, args) if args.batch_reps > 0 else None
    adjoint = adjoint_form(args) if args.adjoint_reps > 0 else None
    volume = volume_form(args) if args.volume_reps > 0 else None

    line = {
        "metric": "samples/s & Mrays/s at 1024^2 64spp (plain)",
        "value": samples / wall,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"spheres6x6_l5 ({scene.triangle_count} tris, 4 quad lights, 1024x512 sky+sun env), "
                               f"{args.width}x{args.height}, {args.spp} spp, max depth {args.depth}",
                   "width": args.width, "height": args.height, "spp": args.spp, "max_depth": args.depth,
                   "parallelism": "single GPU"},
        "mrays_per_s": (closest + shadow) / wall / 1e6,
        "rays": {"closest": closest, "shadow": shadow, "shade_events": shade},
        "kernel_ms_per_step": {"extend": ext_ms / k_steps, "shade": shd_ms / k_steps,
                               "shadow": sdw_ms / k_steps, "steps": k_steps,
                               "note": "HIP-event spans on each launch's stream; the shadow pass runs on a second "
                                       "stream beside the next extend and shade launches, so the shade and shadow spans include waiting for CUs"},
        "roofline": roofline,
        "frame_roofline": frame,
        "ms_per_step_with_kernel_events": ms_instrumented,
        "cpu_baseline": cpu,
        "gpu_baseline": gpu_base,
        "batched": batched,
        "adjoint": adjoint,
        "volume": volume,
        "adaptive": adaptive,
        "bvh_build_s": build_s,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
