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


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def raster_algorithmic_bytes(w, h, samples, nprims, npairs, nvary_floats):
    """Compulsory HBM bytes of one fine-raster launch: colour and depth samples read
    and written once, the resolved image written once, each tile pair and
    primitive setup record read once, and the screen position and the varyings
    the shader reads for each vertex read once (3 vertices per primitive is the
    upper bound; shared vertices are counted once per primitive here)."""
    fb = w * h * samples * (4 + 4) * 2 + (w * h * 4 if samples == 4 else 0)
    geom = npairs * 4 + nprims * 64 + nprims * 3 * (16 + nvary_floats * 4) // 6
    return fb + geom


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=3, help="timed reference frames (after one warm-up)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build renders on one GPU; --gpus must be 1 (multi-GPU sharding is a design "
              "note only, see the sharding section of the design document)", file=sys.stderr)
        sys.exit(2)

    from softglrender_amd import scenes
    from softglrender_amd.renderer import RendererHIP
    from softglrender_amd.replay import Replayer, parse
    from softglrender_amd import stream as S

    t0 = time.time()
    if args.config == "c2":
        sc = scenes.scene_c2()
        workload = "c2_1080p_msaa4x_reverseZ_pbr_ibl_1M_tris_plus_translucent_plus_2048_shadow"
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
    r.resetStats()
    r.profile(True)
    r.waitIdle()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        run_frame()
    r.waitIdle()
    t_end = time.perf_counter()
    ms_per_frame = (t_end - t_start) * 1e3 / args.steps
    stats = r.stats()
    prof = r.profileResults()
    r.profile(False)
    fs_per_frame = stats["fs_invocations"] / args.steps
    print(f"[bench] {args.steps} frames: {ms_per_frame:.3f} ms/frame; kernels {prof}", file=sys.stderr, flush=True)

    # roofline of the dominant kernel (largest summed time over the timed frames)
    dom = max(prof.items(), key=lambda kv: kv[1][0]) if prof else None
    roofline = None
    if dom:
        name, (tot_ms, launches) = dom
        avg_ms = tot_ms / launches
        w, h = sc.width, sc.height
        samples = 4 if sc.msaa else 1
        if name.startswith("k_raster") and name.endswith("msaa4"):
            # launches alternate between passes; count the bytes of the average launch
            pairs = stats["tile_pairs"] / max(stats["draws"], 1)
            prims = stats["triangles_binned"] / max(stats["draws"], 1)
            algo = raster_algorithmic_bytes(w, h, samples, prims, pairs, 11)
        else:
            algo = None
        if algo:
            achieved = algo / (avg_ms * 1e-3) / 1e9
            roofline = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                        "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": int(algo)}

    cpu = None
    ref_frags = None
    if not args.no_cpu_baseline:
        try:
            import oracle
            # fragments the reference shades per frame (its early-Z survivors), counted by the
            # checker on the same frame: the work Gfrag/s is quoted against
            st = {}
            oracle.cpu_ref_replay(data, sc.stream.output_bytes(), stats=st)
            ref_frags = st["fs_invocations"]
            if oracle.have_reference(threaded=False):
                _, times = oracle.ref_replay(data, sc.stream.output_bytes(), repeats=1 + args.cpu_frames)
                med = statistics.median(times[1:])
                cpu = {"value": round(med, 2), "unit": "ms/frame", "cores": 1, "kind": "reference",
                       "sample": f"{args.cpu_frames} whole {sc.name} frames (median, after 1 warm-up) of "
                                 f"SoftGLRender RendererSoft built serially from its sources; host CPU "
                                 f"'{_cpu_model()}', nproc {os.cpu_count()}",
                       "gpu_speedup": round(med / ms_per_frame, 1)}
        except Exception as e:  # report, do not hide
            cpu = {"error": repr(e)}

    line = {
        "metric": "ms/frame 1080p MSAA4x 1M tris; Gfrag/s; neural vs plain PBR at <=1/255 err",
        "value": round(ms_per_frame, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_frame, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": workload, "width": sc.width, "height": sc.height, "msaa": 4 if sc.msaa else 1,
                   "triangles": sc.meta.get("triangles"), "parallelism": "single_gpu"},
        "gfrag_per_s": round(ref_frags / (ms_per_frame * 1e-3) / 1e9, 4) if ref_frags else None,
        "reference_fragments_per_frame": ref_frags,
        "gpu_fs_invocations_per_frame": int(fs_per_frame),
        "kernels_ms_per_frame": {k: round(v[0] / args.steps, 4) for k, v in prof.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    r.destroy()


if __name__ == "__main__":
    main()
