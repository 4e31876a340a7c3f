#!/usr/bin/env python3
"""Frame benchmark of RendererHIP on one MI355X.

Workload (default, `--config c2`): 1920x1080 MSAA-4x, reversed Z, a
1,003,520-triangle displaced icosphere with PBR metallic-roughness (2048^2
albedo/normal/metal-rough maps, full mips, Morton layout) and IBL cubes, a
20,480-triangle translucent sphere blended over it, after a 2048^2 shadow-map
depth pass. Inputs are synthetic and seeded; all resources are resident on the
GPU before timing. One step = one whole frame (every pass from the shadow
pass's beginRenderPass to the main pass's endRenderPass, MSAA resolve
included), like SoftGLRender's Viewer::drawFrame minus resource creation.

Prints one JSON line. `value` is ms/frame (lower is better).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _host_cpus():
    """CPUs this process may run on, and the count std::thread::hardware_concurrency() reports (the
    reference's thread pool size, ThreadPool.h:22)."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return allowed, os.cpu_count() or 1


def _time_reference(oracle, data, out_bytes, frames, threaded):
    """Median ms/frame of SoftGLRender's RendererSoft over `frames` whole frames after one warm-up."""
    _, times = oracle.ref_replay(data, out_bytes, repeats=1 + frames, threaded=threaded)
    return statistics.median(times[1:]), times[1:]


def _reference_baseline(oracle, data, out_bytes, frames, scale, what, ms_per_frame):
    """cpu_baseline: the shipped multithreaded RendererSoft (RASTER_MULTI_THREAD, RendererSoft.cpp:20; a pool of
    hardware_concurrency() workers, ThreadPool.h:22) on this host's cores, with the serial build beside it."""
    allowed, online = _host_cpus()
    res = None
    if oracle.have_reference(threaded=True):
        med, times = _time_reference(oracle, data, out_bytes, frames, True)
        res = {"value": round(med * scale, 2), "unit": "ms/frame", "cores": allowed, "kind": "reference",
               "threads": online,
               "sample": f"{what}: median of {frames} frames after 1 warm-up ({', '.join(f'{t:.0f}' for t in times)} "
                         f"ms) of SoftGLRender RendererSoft as shipped (multithreaded, {online}-worker pool) built "
                         f"from its sources" + (f", times {scale} (the area ratio)" if scale != 1 else "") +
                         f"; host CPU '{_cpu_model()}', {allowed} CPUs available, nproc {online}",
               "gpu_speedup": round(med * scale / ms_per_frame, 1)}
    if oracle.have_reference(threaded=False):
        sframes = max(1, min(frames, 3))
        med, times = _time_reference(oracle, data, out_bytes, sframes, False)
        serial = {"value": round(med * scale, 2), "unit": "ms/frame", "cores": 1, "kind": "reference",
                  "sample": f"{what}: median of {sframes} frames after 1 warm-up of the serial build (the parity "
                            f"oracle)" + (f", times {scale}" if scale != 1 else ""),
                  "gpu_speedup": round(med * scale / ms_per_frame, 1)}
        if res is None:
            res = serial
        else:
            res["serial"] = serial
    return res


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


# varying floats the fragment shader reads per vertex, by shading model and normal-map define
def _vary_floats(shading, defines):
    from softglrender_amd import stream as S
    nm = bool(defines & S.DEF_NORMAL)
    if shading == S.SHADING_PBR:
        return 14 if nm else 11
    if shading in (S.SHADING_BLINN_PHONG, S.SHADING_BLINN_PHONG_DIR):
        return 14 if nm else 11
    if shading == S.SHADING_VERTEX_COLOR:
        return 4
    return 0


def frame_draws(data):
    """Per draw of the frame part of a stream: triangles, vertices, framebuffer size and samples, the varying
    floats its shader reads, whether it has a colour target, and the raster kernel it runs."""
    import struct
    from softglrender_amd import stream as S
    from softglrender_amd.replay import parse
    tex, vao, fbo, draws = {}, {}, {}, []
    cur = None
    in_frame = False
    for op, p in parse(data):
        if op == S.OP_TEXTURE:
            tid, typ, fmt, w, h, usage, mips, msaa = struct.unpack_from("<8i", p, 0)
            tex[tid] = (w, h, 4 if msaa else 1)
        elif op == S.OP_VAO:
            vid, stride, nv, ni = struct.unpack_from("<4i", p, 0)
            vao[vid] = (ni // 3, nv)
        elif op == S.OP_FBO:
            fid, color, layer, level, depth = struct.unpack_from("<5i", p, 0)
            fbo[fid] = (color, depth)
        elif op == S.OP_FRAME:
            in_frame = True
        elif op == S.OP_BEGIN:
            cur = struct.unpack_from("<i", p, 0)[0]
        elif op == S.OP_DRAW and in_frame:
            vid, shading, defines = struct.unpack_from("<iiI", p, 0)
            st = struct.unpack_from("<13i", p, 12)
            color, depth = fbo[cur]
            w, h, smp = tex[color if color >= 0 else depth]
            two_phase = (not st[0]) and depth >= 0 and st[7] and st[8] and st[9] in (1, 3, 4, 6)
            writes_depth = depth >= 0 and st[7] and st[8]
            sfx = "_msaa4" if smp == 4 else ""
            # the fine raster of the draw, and the kernel that runs its fragment shader
            if two_phase:
                kern, shade = "k_raster_vis" + sfx, ("k_shade" + sfx if color >= 0 else None)
            elif not writes_depth:
                kern, shade = "k_ord_list" + sfx, ("k_ord_shade" + sfx if color >= 0 else None)
            else:
                kern, shade = "k_raster_ord" + sfx, None
            t, v = vao[vid]
            draws.append(dict(kernel=kern, shade_kernel=shade, tris=t, verts=v, w=w, h=h, samples=smp,
                              color=color >= 0, vary=_vary_floats(shading, defines) if color >= 0 else 0))
    return draws


def raster_algorithmic_bytes(d):
    """Compulsory HBM bytes of one fine-raster launch: every colour and depth sample read and written once, the
    resolved image written once, 12 bytes of indices per triangle, and per vertex its screen position and the
    varyings the shader reads, once (texel reads are not counted)."""
    px = d["w"] * d["h"]
    fb = px * d["samples"] * 4 * 2 + (px * d["samples"] * 4 * 2 if d["color"] else 0)
    if d["color"] and d["samples"] == 4:
        fb += px * 4
    return fb + 12 * d["tris"] + d["verts"] * (16 + 4 * d["vary"])


def shade_algorithmic_bytes(d):
    """Compulsory HBM bytes of one shading launch: each queued fragment's 8-byte record read once, each sample
    it writes (4 bytes) written once, per vertex its screen position and the varyings the shader reads, once,
    and each distinct texel the shader reads, once (d["texel_bytes"], counted by the checker; absent without
    the CPU leg). d["fs"] and d["samples_written"] are the draw's own counts."""
    return (8 * d["fs"] + 4 * d["samples_written"] + d["verts"] * (16 + 4 * d["vary"]) +
            d.get("texel_bytes", 0))


class _HipEvents:
    """HIP events on the renderer's stream (ctypes on the HIP runtime librhip links)."""

    def __init__(self, r):
        import ctypes
        self.ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.stream = ctypes.c_void_p(r.lib.rhip_device_stream(r.dev))
        self.events = []

    def record(self):
        e = self.ct.c_void_p()
        if self.hip.hipEventCreate(self.ct.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")
        self.hip.hipEventRecord(e, self.stream)
        self.events.append(e)
        return e

    def ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        out = self.ct.c_float()
        self.hip.hipEventElapsedTime(self.ct.byref(out), a, b)
        return float(out.value)

    def close(self):
        for e in self.events:
            self.hip.hipEventDestroy(e)
        self.events = []


def _measured_traffic(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 counters on the C2 frame, or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "c2_hbm_traffic.json")) as f:
            return json.load(f)["kernels"][kernel]["hbm_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def _sq_counters(kernel):
    """Per-dispatch SQ counters of `kernel` on the C2 frame from rocprofv3 --pmc runs, or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "c2_sq_counters.json")) as f:
            return json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None


# MI355X: 256 CUs at 2.4 GHz; a CU issues one wave64 VALU instruction per cycle (4 SIMD16 units, 4 cycles each)
VALU_PEAK_GINSTR = 256 * 2.4


def dominant_roofline(prof, draws, config):
    """Roofline of the dominant kernel (largest summed time over the event-timed frames). Its algorithmic bytes
    are those of the draws it ran for, averaged over its launches. A shading kernel is bound by its VALU issue:
    then `bound` is "valu" with the wave64 VALU instructions per launch (SQ_INSTS_VALU) over the CU issue peak, and
    the HBM line (texel bytes included where the checker counted them) sits beside it."""
    if not prof:
        return None
    name, (tot_ms, launches) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_ms = tot_ms / launches
    mine = [d for d in draws if d["kernel"] == name]
    shading = [d for d in draws if d.get("shade_kernel") == name]
    algo = None
    if mine:
        algo = sum(raster_algorithmic_bytes(d) for d in mine) / len(mine)
    elif shading:
        algo = sum(shade_algorithmic_bytes(d) for d in shading) / len(shading)
    if not algo:
        return None
    achieved = algo / (avg_ms * 1e-3) / 1e9
    hbm = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 5),
           "traffic": _measured_traffic(name) if config == "c2" else None,
           "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE runs of this command, scaled "
                           "on k_resolve (profiles/c2_hbm_traffic.md)",
           "algorithmic_bytes_per_launch": int(algo),
           "texel_bytes_counted": bool(shading) and all("texel_bytes" in d for d in shading)}
    sq = _sq_counters(name) if (shading and config == "c2") else None
    if sq and sq.get("valu_insts"):
        ginstr = sq["valu_insts"] / (avg_ms * 1e-3) / 1e9
        return {"kernel": name, "bound": "valu", "achieved": round(ginstr, 2), "peak": VALU_PEAK_GINSTR,
                "unit": "G wave64 VALU instructions/s", "frac": round(ginstr / VALU_PEAK_GINSTR, 5),
                "traffic": hbm["traffic"], "avg_launch_ms": round(avg_ms, 4),
                "valu_insts_per_launch": sq["valu_insts"],
                "counter_note": "SQ_INSTS_VALU per dispatch from rocprofv3 --pmc on this command "
                                "(profiles/c2_sq_counters.md)",
                "hbm": hbm}
    return {"kernel": name, **hbm, "avg_launch_ms": round(avg_ms, 4)}


def _pct(x, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(x), q)), 4)


def run_batched(r, quick=False, reps=5, log=lambda m: None):
    """Batched views: 256 views over 64 seeded scenes (displaced icospheres of 1k-500k triangles, 1024^2
    Tiled textures, views of 256^2/512^2/1024^2, Blinn-Phong, no MSAA). The same render passes run once as
    256 separate draws and once as one captured batch, after the autotuner picked the batch's tile size
    and wavefronts per tile. Reports views/s and per-view latency (p50/p99) of both."""
    import hashlib
    from softglrender_amd import scenes
    from softglrender_amd.replay import Replayer, split

    t0 = time.time()
    if quick:
        sc = scenes.scene_c4(n_scenes=16, n_views=64, tri_hi=100_000)
    else:
        sc = scenes.scene_c4()
    data = sc.stream.bytes()
    setup, frame, _ = split(data)
    log(f"[bench] batched views: {len(sc.meta['view_sizes'])} views, {sc.meta['view_triangles']} view-triangles, "
        f"{sc.meta['pixels']} pixels; generated in {time.time() - t0:.1f}s")
    rp = Replayer(r, layouts=sc.meta["layouts"])
    sink = []
    for op, p in setup:
        rp.exec(op, p, sink)
    r.waitIdle()
    # the frame is one (begin, viewport, draw, end) group per view
    views = [frame[i:i + 4] for i in range(0, len(frame), 4)]
    nv = len(views)
    ev = _HipEvents(r)

    def separate(per_view):
        marks = []
        for cmds in views:
            a = ev.record() if per_view else None
            for op, p in cmds:
                rp.exec(op, p, sink)
            marks.append((a, ev.record() if per_view else None))
        return marks

    separate(False)  # warm-up
    r.waitIdle()
    a = ev.record()
    for _ in range(reps):
        separate(False)
    b = ev.record()
    sep_ms = ev.ms(a, b) / reps
    marks = separate(True)
    r.waitIdle()
    own = [ev.ms(x, y) for x, y in marks]
    since = [ev.ms(marks[0][0], y) for _, y in marks]
    ev.close()
    # every view's colour and depth target (the textures after the scenes' own)
    out_tex = sorted(t for t in rp.tex if t > len(sc.meta["scene_triangles"]))
    digest = lambda: [hashlib.sha256(rp.tex[t].readImage().tobytes()).hexdigest() for t in out_tex]
    ref_digest = digest()
    log(f"[bench] separate draws: {sep_ms:.3f} ms per {nv} views")

    r.beginBatch()
    for op, p in frame:
        rp.exec(op, p, sink)
    batch = r.endBatch()
    t_tune = time.time()
    best, table = batch.tune(reps=3)
    t_tune = time.time() - t_tune
    log(f"[bench] autotuner: best tile {best[0]} waves {best[1]} in {t_tune:.1f}s: {table}")
    batch.run()  # warm-up
    r.waitIdle()
    a = ev.record()
    for _ in range(reps):
        batch.run()
    b = ev.record()
    bat_ms = ev.ms(a, b) / reps
    ev.close()
    batch.run()
    lat = batch.view_times()
    got_digest = digest()
    same_targets = sum(a == b for a, b in zip(got_digest, ref_digest))
    same = same_targets == len(ref_digest)
    log(f"[bench] batch: {bat_ms:.3f} ms per {nv} views; identical to separate draws: {same} "
        f"({same_targets} of {len(ref_digest)} colour and depth targets)")

    # the same views as k batches run back to back (the tuned configuration): a view completes with its own
    # batch, so smaller batches trade throughput for per-view completion
    sub = {}
    for k in (4, 16):
        if nv % k:
            continue
        per = nv // k
        parts = []
        for i in range(k):
            r.beginBatch()
            for cmds in views[i * per:(i + 1) * per]:
                for op, p in cmds:
                    rp.exec(op, p, sink)
            bt = r.endBatch()
            bt.set_config(*best)
            parts.append(bt)
        for bt in parts:  # warm-up
            bt.run()
        r.waitIdle()
        ev = _HipEvents(r)
        a = ev.record()
        for _ in range(reps):
            for bt in parts:
                bt.run()
        b = ev.record()
        k_ms = ev.ms(a, b) / reps
        starts = []
        for bt in parts:
            starts.append(ev.record())
            bt.run()
        r.waitIdle()
        comp = []
        for i, bt in enumerate(parts):
            comp += list(ev.ms(starts[0], starts[i]) + bt.view_times())
        ev.close()
        k_same = sum(x == y for x, y in zip(digest(), ref_digest))
        for bt in parts:
            bt.destroy()
        sub[f"{k}x{per}"] = {"ms_per_256_views" if nv == 256 else "ms_per_views": round(k_ms, 4),
                             "views_per_s": round(nv / (k_ms * 1e-3), 1),
                             "completion_ms_p50": _pct(comp, 50), "completion_ms_p99": _pct(comp, 99),
                             "targets_identical": k_same}
        log(f"[bench] {k} batches of {per} views: {k_ms:.3f} ms, completion p50 {_pct(comp, 50)} ms "
            f"p99 {_pct(comp, 99)} ms; {k_same} of {len(ref_digest)} targets identical")
    draws = frame_draws(data)
    algo = sum(raster_algorithmic_bytes(d) for d in draws)
    res = {
        "workload": f"{nv} views over {len(sc.meta['scene_triangles'])} scenes, "
                    f"{sc.meta['view_triangles']} triangles and {sc.meta['pixels']} pixels per batch",
        "views": nv,
        "batched": {"ms_per_batch": round(bat_ms, 4), "views_per_s": round(nv / (bat_ms * 1e-3), 1),
                    "latency_ms_p50": _pct(lat, 50), "latency_ms_p99": _pct(lat, 99),
                    "latency_note": "completion of each view, from the batch's first kernel (device clock)",
                    "achieved_gbs": round(algo / (bat_ms * 1e-3) / 1e9, 2)},
        "separate_draws": {"ms_per_batch": round(sep_ms, 4), "views_per_s": round(nv / (sep_ms * 1e-3), 1),
                           "latency_ms_p50": _pct(own, 50), "latency_ms_p99": _pct(own, 99),
                           "completion_ms_p50": _pct(since, 50), "completion_ms_p99": _pct(since, 99),
                           "latency_note": "each draw's own duration (HIP events around its commands); "
                                           "completion = from the first draw's start"},
        "speedup": round(sep_ms / bat_ms, 3),
        "sub_batches": sub,
        "autotuner": {"best": {"tile": best[0], "waves": best[1]},
                      "ms_per_batch": {f"tile{t}_waves{w}": round(ms, 4) for (t, w), ms in table},
                      "seconds": round(t_tune, 1)},
        "algorithmic_bytes_per_batch": int(algo),
        "identical_to_separate_draws": same,
        "targets_compared": len(ref_digest),
        "targets_identical": same_targets,
    }
    batch.destroy()
    return res


def run_ibl(r, reps=5, cpu=True, log=lambda m: None):
    """The C2 environment's IBL precompute on the GPU (IBLIrradianceSoft.h:74-101, IBLPrefilterSoft.h:88-168
    drawn into cube faces as Environment.cpp:48-103,152-195 does): a 512^2 RGBA8 radiance cube with mips ->
    32^2 irradiance cube and 128^2 x 5-level prefiltered cube, 36 render passes. Timed from resident
    resources to the maps written; with `cpu`, SoftGLRender's RendererSoft as shipped (multithreaded) runs
    the same 36 passes once on the host beside it (10-20 s)."""
    from softglrender_amd import scenes
    from softglrender_amd.replay import Replayer, parse
    from softglrender_amd import stream as S
    sc = scenes.scene_ibl_precompute()
    rp = Replayer(r)
    setup, frame = [], []
    in_frame = False
    for op, p in parse(sc.stream.bytes()):
        if op == S.OP_FRAME:
            in_frame = True
        elif op == S.OP_READ:
            continue
        elif in_frame:
            frame.append((op, p))
        else:
            setup.append((op, p))
    sink = []
    for op, p in setup:
        rp.exec(op, p, sink)
    for op, p in frame:
        rp.exec(op, p, sink)
    r.waitIdle()
    t = time.perf_counter()
    for _ in range(reps):
        for op, p in frame:
            rp.exec(op, p, sink)
    r.waitIdle()
    ms = (time.perf_counter() - t) * 1e3 / reps
    r.profile(True)
    for op, p in frame:
        rp.exec(op, p, sink)
    r.waitIdle()
    prof = r.profileResults()
    r.profile(False)
    m = sc.meta
    # fragments: every texel of the 6 irradiance faces and of the prefiltered levels' faces
    frags = 6 * (m["irr_size"] ** 2 + sum(max(1, m["pref_size"] >> lev) ** 2 for lev in range(m["pref_levels"])))
    log(f"[ibl] precompute {ms:.3f} ms ({frags} fragments)")
    res = {"workload": f"{m['src_size']}^2 RGBA8 mipmapped radiance cube -> {m['irr_size']}^2 irradiance, "
                       f"{m['pref_size']}^2 x {m['pref_levels']} levels prefiltered (36 render passes)",
           "ms": round(ms, 3), "fragments": frags,
           "kernels_ms": {k: round(v[0], 4) for k, v in prof.items()}}
    if cpu:
        try:
            import oracle
            if oracle.have_reference(threaded=True):
                allowed, online = _host_cpus()
                _, times = oracle.ref_replay(sc.stream.bytes(), sc.stream.output_bytes(), repeats=1,
                                             threaded=True)
                res["cpu_baseline"] = {
                    "value": round(times[0], 1), "unit": "ms", "cores": allowed, "kind": "reference",
                    "threads": online,
                    "sample": f"the same 36 passes once, SoftGLRender RendererSoft as shipped ({online}-worker "
                              f"pool) built from its sources",
                    "gpu_speedup": round(times[0] / ms, 1)}
                log(f"[ibl] reference {times[0]:.0f} ms")
        except Exception as e:  # report, do not hide
            res["cpu_baseline"] = {"error": repr(e)}
    return res


def run_c3(r, steps=5, warmup=1, cpu=True, log=lambda m: None):
    """C3 beside the headline frame: 3840x2160 MSAA-4x, 2,000,000 random small triangles and 2,000 screen-covering
    ones, vertex colour, LESS (the hierarchical-Z stress scene). ms/frame between synchronisations after `warmup`
    frames, the kernel breakdown from a second pass with events, and, with `cpu`, the shipped multithreaded
    RendererSoft on a 240x135 crop of the same triangle density with all 2,000 screen-covering triangles, times
    256 (a whole frame takes the reference about half an hour)."""
    from softglrender_amd import scenes
    from softglrender_amd.replay import Replayer, parse
    from softglrender_amd import stream as S
    sc = scenes.scene_c3()
    rp = Replayer(r, layouts=sc.meta.get("layouts"))
    setup, frame = [], []
    in_frame = False
    for op, p in parse(sc.stream.bytes()):
        if op == S.OP_FRAME:
            in_frame = True
        elif op == S.OP_READ:
            continue
        elif in_frame:
            frame.append((op, p))
        else:
            setup.append((op, p))
    sink = []
    for op, p in setup:
        rp.exec(op, p, sink)

    def run_frame():
        for op, p in frame:
            rp.exec(op, p, sink)

    for _ in range(warmup):
        run_frame()
    r.waitIdle()
    t = time.perf_counter()
    for _ in range(steps):
        run_frame()
    r.waitIdle()
    ms = (time.perf_counter() - t) * 1e3 / steps
    r.profile(True)
    run_frame()
    r.waitIdle()
    prof = r.profileResults()
    r.profile(False)
    log(f"[c3] {ms:.3f} ms/frame")
    res = {"workload": "c3_2160p_msaa4x_2M_random_tris_plus_2k_screen_covering_vertex_colour_less",
           "ms_per_frame": round(ms, 4), "steps": steps, "warmup": warmup,
           "triangles": sc.meta.get("triangles"),
           "kernels_ms_per_frame": {k: round(v[0], 4) for k, v in prof.items()}}
    if cpu:
        try:
            import oracle
            if oracle.have_reference(threaded=True):
                allowed, online = _host_cpus()
                crop = scenes.scene_c3(240, 135, 2_000_000 // 256, 2000, seed=5)
                med, times = _time_reference(oracle, crop.stream.bytes(), crop.stream.output_bytes(), 3, True)
                res["cpu_baseline"] = {
                    "value": round(med * 256, 1), "unit": "ms/frame", "cores": allowed, "kind": "reference",
                    "threads": online,
                    "sample": f"one 240x135 crop with {2_000_000 // 256} random triangles and the 2000 "
                              f"screen-covering ones: median of 3 frames after 1 warm-up "
                              f"({', '.join(f'{t:.0f}' for t in times)} ms) of SoftGLRender RendererSoft as shipped "
                              f"({online}-worker pool), times 256 (the area ratio)",
                    "gpu_speedup": round(med * 256 / ms, 1)}
                log(f"[c3] reference crop {med:.0f} ms x 256")
        except Exception as e:  # report, do not hide
            res["cpu_baseline"] = {"error": repr(e)}
    return res


def run_neural(r, quick=False, log=lambda m: None):
    """C5: the learned PBR shader on the C2 scene (point light off for both shaders). Trains on fragments of
    64 random cameras x 8 random materials, evaluates per material on 32 held-out cameras against the FP64
    plain shader, and times learned against plain shading of the C2 view."""
    import numpy as np
    from softglrender_amd import neural, scenes
    from softglrender_amd import stream as S
    t0 = time.time()
    albedo, normal, mr = scenes.pbr_textures(2048, 3)
    mats = neural.neural_materials(8, 1024, 11, base=(albedo, mr))
    verts = scenes.icosphere(224, 1.0, displace=0.1, seed=3)
    n_train, n_held, shift = (8, 4, 5) if quick else (64, 32, 5)
    log(f"[neural] materials and geometry in {time.time() - t0:.1f}s")
    r.neural_records(24_000_000)
    t1 = time.time()
    tr = neural.capture(r, neural.scene_views(neural.random_cameras(n_train, 1), mats, normal, verts=verts), shift)
    ho = neural.capture(r, neural.scene_views(neural.random_cameras(n_held, 2), mats, normal, verts=verts), shift)
    t_cap = time.time() - t1
    log(f"[neural] captured {tr[1]} training and {ho[1]} held-out fragments in {t_cap:.1f}s")
    r.neural_init(1)
    steps = 6000 if quick else 60000
    t2 = time.time()
    losses = r.neural_train(tr[0], tr[1], steps=steps, batch=16384, lr_start=1e-2, lr_end=2e-5, seed=1)
    t_train = time.time() - t2
    log(f"[neural] {steps} steps in {t_train:.1f}s, loss {losses[0]:.5f} -> {losses[-1]:.6f}")
    table = neural.fallback_table(r, ho[0], ho[1], len(mats))
    fallback = neural.apply_fallback(r, table)
    e_nn, e_pl, _ = r.neural_evaluate(ho[0], ho[1])
    r.neural_material(0)

    # shading throughput on the C2 view (camera at (0, 0, 2.6), material 0)
    cam = np.array([[0.0, 0.0, 2.6]])
    times = {}
    for name, sh in (("plain", S.SHADING_PBR), ("neural", S.SHADING_NEURAL_PBR)):
        ns = neural.scene_views(cam, mats[:1], normal, verts=verts, shading=sh)
        from softglrender_amd.replay import Replayer, parse
        rp = Replayer(r, layouts=ns.scene.meta.get("layouts"))
        ops = list(parse(ns.scene.stream.bytes()))
        k = [i for i, (op, _) in enumerate(ops) if op == S.OP_FRAME][0]
        sink = []
        for op, p in ops[:k]:
            rp.exec(op, p, sink)
        frame = [(op, p) for op, p in ops[k + 1:] if op != S.OP_READ]
        for _ in range(2):
            for op, p in frame:
                rp.exec(op, p, sink)
        r.waitIdle()
        r.profile(True)
        reps = 20
        t = time.perf_counter()
        for _ in range(reps):
            for op, p in frame:
                rp.exec(op, p, sink)
        r.waitIdle()
        frame_ms = (time.perf_counter() - t) * 1e3 / reps
        prof = r.profileResults()
        r.profile(False)
        # a material that falls back draws its learned-shader draws with the plain shader
        kerns = ["k_shade_msaa4"] if name == "plain" or 0 in fallback else ["k_nn_shade_msaa4"]
        times[name] = {"shade_kernel_ms": round(sum(prof[k][0] / prof[k][1] for k in kerns), 4),
                       "frame_ms": round(frame_ms, 4)}
        if name == "neural":
            times[name].update({k: round(prof[k][0] / prof[k][1], 4) for k in kerns})
    speed = times["plain"]["shade_kernel_ms"] / times["neural"]["shade_kernel_ms"]
    rnd = lambda e: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in e.items()}
    return {
        "network": "24-64-64-3 ReLU MLP, BF16 MFMA (v_mfma_f32_16x16x32_bf16), FP32 accumulation",
        "scene": "c2 geometry and IBL, image-based PBR (point light off for both shaders)",
        "train_fragments": int(tr[1]), "heldout_fragments": int(ho[1]),
        "train_cameras": n_train, "heldout_cameras": n_held, "materials": len(mats),
        "steps": steps, "batch": 16384, "train_s": round(t_train, 2),
        "train_us_per_step": round(t_train * 1e6 / max(steps, 1), 2), "loss_first_last": [losses[0], losses[-1]],
        "error_vs_fp64_255": {"network": rnd(e_nn), "plain_gpu_shader": rnd(e_pl)},
        "per_material": [{"material": t["material"], "network_mean": round(t["network"]["mean"], 3),
                          "network_p99": round(t["network"]["p99"], 3), "fallback": t["fallback"]} for t in table],
        "fallback_materials": fallback,
        "shading": times, "shading_speedup": round(speed, 3),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4"],
                    help="c4: only the batched-views measurement (views/s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-neural", action="store_true", help="skip the learned-shader (c5) measurement")
    ap.add_argument("--neural-quick", action="store_true", help="reduced c5 run (fewer cameras and steps)")
    ap.add_argument("--cpu-frames", type=int, default=5, help="timed reference frames (after one warm-up)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-views (c4) measurement")
    ap.add_argument("--no-ibl", action="store_true", help="skip the IBL precompute measurement")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3 frame measured beside the C2 frame")
    ap.add_argument("--batched-quick", action="store_true", help="reduced c4 run (64 views over 16 scenes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build renders on one GPU; --gpus must be 1 (multi-GPU sharding is a design "
              "note only, see the sharding section of the design document)", file=sys.stderr)
        sys.exit(2)

    from softglrender_amd import scenes
    from softglrender_amd.renderer import RendererHIP
    from softglrender_amd.replay import Replayer, parse
    from softglrender_amd import stream as S

    if args.config == "c4":
        r = RendererHIP(0)
        r.create()
        res = run_batched(r, quick=args.batched_quick, reps=args.steps,
                          log=lambda m: print(m, file=sys.stderr, flush=True))
        b = res["batched"]
        print(json.dumps({"metric": "views/s, 256 views of 64 scenes in one batch", "value": b["views_per_s"],
                          "unit": "views/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
                          "ms_per_step": b["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scenes)",
                          "config": {"workload": "c4_256_views_64_scenes_blinnphong_tiled_textures"},
                          "batched_views": res}), flush=True)
        r.destroy()
        return

    t0 = time.time()
    if args.config == "c2":
        sc = scenes.scene_c2()
        workload = "c2_1080p_msaa4x_reverseZ_pbr_ibl_1M_tris_plus_translucent_plus_2048_shadow"
    elif args.config == "c3":
        sc = scenes.scene_c3()
        workload = "c3_2160p_msaa4x_2M_random_tris_plus_2k_screen_covering_vertex_colour_less"
    else:
        sc = scenes.scene_c1()
        workload = "c1_640x480_uvsphere_4k_tris_blinnphong_trilinear"
    data = sc.stream.bytes()
    print(f"[bench] scene {sc.name} generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    r = RendererHIP(0)
    r.create()
    rp = Replayer(r, layouts=sc.meta.get("layouts"))
    setup, frame, reads = [], [], []
    in_frame = False
    for op, p in parse(data):
        if op == S.OP_FRAME:
            in_frame = True
        elif op == S.OP_READ:
            reads.append((op, p))
        elif in_frame:
            frame.append((op, p))
        else:
            setup.append((op, p))
    sink = []
    for op, p in setup:
        rp.exec(op, p, sink)
    r.waitIdle()
    print(f"[bench] resources resident; warmup {args.warmup}", file=sys.stderr, flush=True)

    def run_frame():
        for op, p in frame:
            rp.exec(op, p, sink)

    for _ in range(args.warmup):
        run_frame()
    r.waitIdle()
    # the timed frames: nothing but the frame's own work between the synchronisations
    r.resetStats()
    r.waitIdle()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        run_frame()
    r.waitIdle()
    t_end = time.perf_counter()
    ms_per_frame = (t_end - t_start) * 1e3 / args.steps
    stats = r.stats()
    # the same frames again with HIP events around every launch, for the per-kernel breakdown
    r.profile(True)
    r.waitIdle()
    t_a = time.perf_counter()
    for _ in range(args.steps):
        run_frame()
    r.waitIdle()
    ms_with_events = (time.perf_counter() - t_a) * 1e3 / args.steps
    prof = r.profileResults()
    r.profile(False)
    fs_per_frame = stats["fs_invocations"] / args.steps
    print(f"[bench] {args.steps} frames: {ms_per_frame:.3f} ms/frame; kernels {prof}", file=sys.stderr, flush=True)

    # one more frame, draw by draw, for each draw's own fragment-shader and sample-write counts
    draws = frame_draws(data)
    k = 0
    r.resetStats()
    prev = r.stats()
    for op, p in frame:
        rp.exec(op, p, sink)
        if op == S.OP_DRAW:
            r.waitIdle()
            cur = r.stats()
            draws[k]["fs"] = cur["fs_invocations"] - prev["fs_invocations"]
            draws[k]["samples_written"] = cur["samples_written"] - prev["samples_written"]
            prev, k = cur, k + 1
    r.waitIdle()

    cpu = None
    ref_frags = None
    if not args.no_cpu_baseline and args.config == "c3":
        # the reference needs about an hour for a whole c3 frame: time it on a 240x135 crop (1/256 of the
        # area) holding the same triangle density and all 2000 screen-covering triangles, and scale
        try:
            import oracle
            crop = scenes.scene_c3(240, 135, 2_000_000 // 256, 2000, seed=5)
            st = {}
            oracle.cpu_ref_replay(crop.stream.bytes(), crop.stream.output_bytes(), stats=st,
                                  threads=min(16, os.cpu_count() or 1))
            ref_frags = st["fs_invocations"] * 256
            cpu = _reference_baseline(oracle, crop.stream.bytes(), crop.stream.output_bytes(), 1, 256,
                                      f"one 240x135 crop with {2_000_000 // 256} random triangles and the 2000 "
                                      f"screen-covering ones", ms_per_frame)
        except Exception as e:  # report, do not hide
            cpu = {"error": repr(e)}
    if not args.no_cpu_baseline and args.config != "c3":
        try:
            import oracle
            # fragments the reference shades per frame (its early-Z survivors), counted by the
            # checker on the same frame: the work Gfrag/s is quoted against
            st, tb = {}, []
            oracle.cpu_ref_replay(data, sc.stream.output_bytes(), stats=st, texel_bytes=tb)
            ref_frags = st["fs_invocations"]
            # the distinct texels each draw's shader reads: the texture part of its algorithmic bytes
            for d, t in zip(draws, tb):
                d["texel_bytes"] = t
            cpu = _reference_baseline(oracle, data, sc.stream.output_bytes(), args.cpu_frames, 1,
                                      f"whole {sc.name} frames", ms_per_frame)
        except Exception as e:  # report, do not hide
            cpu = {"error": repr(e)}

    roofline = dominant_roofline(prof, draws, args.config)

    neural_res = None
    if args.config == "c2" and not args.no_neural:
        try:
            neural_res = run_neural(r, quick=args.neural_quick,
                                    log=lambda m: print(m, file=sys.stderr, flush=True))
        except Exception as e:  # report, do not hide
            neural_res = {"error": repr(e)}

    ibl_res = None
    if args.config == "c2" and not args.no_ibl:
        try:
            ibl_res = run_ibl(r, cpu=not args.no_cpu_baseline, log=lambda m: print(m, file=sys.stderr, flush=True))
        except Exception as e:  # report, do not hide
            ibl_res = {"error": repr(e)}

    c3_res = None
    if args.config == "c2" and not args.no_c3:
        try:
            c3_res = run_c3(r, cpu=not args.no_cpu_baseline, log=lambda m: print(m, file=sys.stderr, flush=True))
        except Exception as e:  # report, do not hide
            c3_res = {"error": repr(e)}

    batched_res = None
    if args.config == "c2" and not args.no_batched:
        try:
            batched_res = run_batched(r, quick=args.batched_quick, log=lambda m: print(m, file=sys.stderr, flush=True))
        except Exception as e:  # report, do not hide
            batched_res = {"error": repr(e)}

    line = {
        "metric": "ms/frame 1080p MSAA4x 1M tris; Gfrag/s; neural vs plain PBR at <=1/255 err",
        "value": round(ms_per_frame, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_frame, 4),
        "ms_per_frame_with_kernel_events": round(ms_with_events, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": workload, "width": sc.width, "height": sc.height, "msaa": 4 if sc.msaa else 1,
                   "triangles": sc.meta.get("triangles"), "parallelism": "single_gpu"},
        "gfrag_per_s": round(ref_frags / (ms_per_frame * 1e-3) / 1e9, 4) if ref_frags else None,
        "reference_fragments_note": ("counted by the checker on the 1/256 crop, times 256"
                                     if args.config == "c3" and ref_frags else None),
        "reference_fragments_per_frame": ref_frags,
        "gpu_fs_invocations_per_frame": int(fs_per_frame),
        "kernels_ms_per_frame": {k: round(v[0] / args.steps, 4) for k, v in prof.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "neural": neural_res,
        "ibl_precompute": ibl_res,
        "c3": c3_res,
        "batched_views": batched_res,
    }
    print(json.dumps(line), flush=True)
    r.destroy()


if __name__ == "__main__":
    main()
