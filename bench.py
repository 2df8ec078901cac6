#!/usr/bin/env python3
"""Benchmark of the MI355X drawable-processing pass.

Default run (one GPU): the full10m workload — a 3-level transform tree of
10,101,000 nodes (1,000 roots -> 100 children -> 100 leaves), 10M leaf
drawables in 64 StateSets, 3 LODs — propagated level by level to world
matrices and world spheres, then frustum-culled, LOD-selected and compacted
per StateSet for one view.  A step is one such frame; `value` is drawables
culled per second over the timed frames.

Further legs, reported beside it:
  batched  batch256: the full10m scene culled for 256 views in one persistent
           launch replayed from a HIP graph, against the plain pass launched
           once per view;
  uneven   uneven64: 64 scenes (10k..5M drawables, depth 1..4, 1..32 views),
           every tree propagated and every view culled in one graph per frame;
  served   served128: 128 resident scenes under Poisson cull requests through
           the serving loop, p50/p99/p999 latency and throughput per rate;
  resident facade frames at 10M drawables with 0.1% of the drawables edited
           per frame: the reference's full per-frame copy of the drawable
           list against the resident list with delta uploads.
Kernel shapes come from the autotuner (cadrCullAutotune) unless --no-tune.

Prints one JSON line (metric / value / roofline / cpu_baseline ...).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def measured_per_launch(what: str, kernel: str):
    """A per-launch figure of `kernel` measured with rocprofv3 --pmc and
    recorded in a profiles/*.md summary as a line `<what>_per_launch
    <kernel> <value>`; None if not recorded."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(PROFILES, "*.md"))):
        with open(path) as f:
            for line in f:
                m = re.match(r"\s*" + re.escape(what) + r"_per_launch\s+(\S+)\s+(\d+)", line)
                if m and m.group(1) == kernel:
                    return int(m.group(2))
    return None


def measured_traffic(kernel: str):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel
    calibration), from a `traffic_bytes_per_launch` line."""
    return measured_per_launch("traffic_bytes", kernel)


def valu_issue(kernel: str, ms: float, cus: int, clock_ghz: float = 2.4):
    """VALU issue rate of a launch: SQ_INSTS_VALU wave-instructions per
    launch (a `valu_wave_instructions_per_launch` line) against one wave64
    VALU instruction per CU per clock (4 SIMD16 units, 4 cycles each), at
    the peak engine clock; None if not recorded."""
    n = measured_per_launch("valu_wave_instructions", kernel)
    if n is None:
        return None
    cap = cus * clock_ghz * 1e9 * ms * 1e-3
    return {"wave_instructions": n, "issue_frac": n / cap, "clock_ghz": clock_ghz, "cus": cus,
            "lane_ops_per_s": n * 64 / (ms * 1e-3),
            "note": "no packed FP32 (disabled at compile time), so the non-packed lane rate is the ceiling"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--block", type=int, default=512, help="batched cull workgroup size (with --no-tune)")
    p.add_argument("--items", type=int, default=4, help="batched cull drawables per thread (with --no-tune)")
    p.add_argument("--view-block", type=int, default=256, help="single-view cull workgroup size (with --no-tune)")
    p.add_argument("--view-items", type=int, default=16, help="single-view cull drawables per thread (with --no-tune)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                   help="also copy BYTES device-to-device 3 times (reference kernel for HBM counter scaling)")
    p.add_argument("--batch-views", type=int, default=256)
    p.add_argument("--batch-steps", type=int, default=10)
    p.add_argument("--no-batch", action="store_true")
    p.add_argument("--no-uneven", action="store_true")
    p.add_argument("--no-served", action="store_true")
    p.add_argument("--no-gather", action="store_true")
    p.add_argument("--no-resident", action="store_true")
    p.add_argument("--no-plain100k", action="store_true")
    p.add_argument("--no-facade", action="store_true", help="skip the facade cull frames")
    p.add_argument("--no-hbm-probe", action="store_true", help="skip the measured copy / read / write bandwidth")
    p.add_argument("--resident-drawables", type=int, default=10_000_000)
    p.add_argument("--served-rates", default="5000,10000,20000,50000,65000,80000,90000,100000,200000",
                   help="Poisson arrival rates (req/s)")
    p.add_argument("--served-overload-gib", type=float, default=1.0,
                   help="output budget of the overload run (0: skip it)")
    p.add_argument("--served-seconds", type=float, default=1.0, help="arrival window per rate")
    p.add_argument("--served-budget-gib", type=float, default=64.0, help="output memory budget of the server")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-tune", action="store_true", help="use --block/--items instead of the autotuner's choice")
    p.add_argument("--no-plain", action="store_true", help="skip the per-view plain pass of the batch legs")
    p.add_argument("--scale", type=float, default=1.0, help="root-count scale of full10m (1.0 = 10M drawables)")
    return p.parse_args()


class KernelEnds:
    """One HIP event per timed step that a kernel launch records at the end of
    its own dispatch (cadrTimeNextLaunch with a stop event only).  A kernel's
    time in a chain of launches on one stream is its end minus the end of the
    launch before it: its execution plus the dispatch gap in front of it
    (under a microsecond), an upper bound on its duration.  Stop events bound
    to a dispatch add nothing to the chain; an event recorded as a command of
    its own (or a start event) adds about 4 us per kernel."""

    def __init__(self, n):
        from cadr_amd.device import Event
        self.ends = [Event() for _ in range(n)]

    @staticmethod
    def mean_ms(before: "KernelEnds", after: "KernelEnds") -> float:
        n = len(after.ends)
        return sum(b.elapsed_ms(a) for b, a in zip(before.ends, after.ends)) / max(n, 1)


def main():
    args = parse()
    # keep large numpy temporaries in the heap instead of fresh mmaps (first-touch page faults)
    try:
        libc = ctypes.CDLL("libc.so.6")
        libc.mallopt(-1, 1 << 40)   # M_TRIM_THRESHOLD
        libc.mallopt(-3, 32 << 20)  # M_MMAP_THRESHOLD
    except OSError:
        pass
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU only (the pass would shard by scene range "
                                   "or by view); no multi-GPU figure is produced", "n_gpus": args.gpus}))
        sys.exit(2)

    from cadr_amd import _native as N
    from cadr_amd import cull, device, scenes

    from cadr_amd.tune import autotune, describe
    device.set_device(0)
    stream = device.Stream()
    s = stream.handle

    t0 = time.time()
    spec = scenes.full10m_spec(args.scale)
    ds = cull.DeviceScene(spec)  # streamed to HBM level by level
    log(f"scene generated and uploaded in {time.time() - t0:.1f}s: {ds.n} drawables, {sum(spec.sizes)} nodes")
    n = ds.n
    config = cull.make_config(args.view_block, args.view_items)
    views1 = cull.make_views([scenes.DEFAULT_CAMERA])
    out1 = cull.CullOutputs(1, n, ds.K)
    tuned1 = None
    if not args.no_tune:
        ds.propagate(s)
        config, table = autotune([(ds, views1, out1)], single_view=True, reps=5, stream=s)
        tuned1 = {"choice": describe(config), "candidates_ms": {tune_key(t): t["ms"] for t in table}}
        log(f"single-view autotune: {describe(config)}  {tuned1['candidates_ms']}")
    batch1 = cull.CullBatch(config, [(ds, views1, out1)])

    lib = N.load()
    level_descs = ds.level_descs(spheres=True)
    leaf_fused = level_descs[-1].copy()
    leaf_fused["spheres"] = 0  # the one-view frame needs no world spheres in HBM
    from cadr_amd.device import Event, time_next_launch

    # the levels above the leaf level; a root level's world matrices are its
    # local ones, so the level below it reads the root's local matrices as its
    # parents' world matrices and both go in one launch
    upper = [lv.copy() for lv in level_descs[:-1]]
    first_pair, first_pair_dev = None, None
    if len(upper) >= 2 and int(upper[0]["parentWorld"][0]) == 0:
        upper[1]["parentWorld"] = upper[0]["localMatrices"]
        first_pair = np.concatenate(upper[:2])
        first_pair_dev = device.DeviceBuffer.from_array(first_pair)
        upper = upper[2:]

    def propagate_inner(end=None):
        """end: event that the last of these launches records at its end"""
        if end is not None and first_pair is None and not upper:
            end.record(s)
        if first_pair is not None:
            if end is not None and not upper:
                time_next_launch(None, end)
            N.check(lib.cadrPropagateTransformsBatch(first_pair_dev.address, N.ptr(first_pair), 2, s))
        for i, lv in enumerate(upper):
            if end is not None and i == len(upper) - 1:
                time_next_launch(None, end)
            N.check(lib.cadrPropagateTransforms(N.ptr(lv), s))

    # kernel shape of the fused leaf + cull launch: the fastest of the sweep
    fused_config, fused_table = cull.make_config(256, args.view_items), None
    if not args.no_tune:
        propagate_inner()
        fused_table = {}
        for block, items in ((256, 2), (256, 4), (256, 8), (256, 16), (512, 2), (512, 4), (1024, 1), (1024, 2)):
            cfg = cull.make_config(block, items)
            b = cull.CullBatch(cfg, [(ds, views1, out1)])
            b.run_fused(leaf_fused, s)
            e0, e1 = Event(), Event()
            e0.record(s)
            for _ in range(5):
                b.run_fused(leaf_fused, s)
            e1.record(s)
            stream.synchronize()
            fused_table[f"{block}x{items}"] = e0.elapsed_ms(e1) / 5
            b.close()
        best = min(fused_table, key=fused_table.get)
        fused_config = cull.make_config(*(int(x) for x in best.split("x")))
        log(f"fused autotune: {best}  {fused_table}")
    batch_fused = cull.CullBatch(fused_config, [(ds, views1, out1)])

    def run_frames(fused):
        """warmup + steps frames; returns (ms per frame, mean ms of the leaf
        launch, mean ms of the separate cull launch or None, counts)"""
        e_inner, e_leaf, e_cull = KernelEnds(args.steps), KernelEnds(args.steps), KernelEnds(args.steps)

        def frame(k=None):
            """k: index of a timed frame, whose launches record their ends"""
            propagate_inner(None if k is None else e_inner.ends[k])
            if k is not None:
                time_next_launch(None, e_leaf.ends[k])
            if fused:
                batch_fused.run_fused(leaf_fused, s)
            else:
                N.check(lib.cadrPropagateTransforms(N.ptr(level_descs[-1]), s))
                if k is not None:
                    time_next_launch(None, e_cull.ends[k])
                batch1.run_view(s)

        for _ in range(args.warmup):
            frame()
        stream.synchronize()
        # the timed region: exactly `steps` frames, synchronised on both sides
        device.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            frame(k)
        stream.synchronize()
        device.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / args.steps
        return (ms, KernelEnds.mean_ms(e_inner, e_leaf), (None if fused else KernelEnds.mean_ms(e_leaf, e_cull)),
                out1.download_counts())

    un_ms, leaf_ms, cull_ms, counts_un = run_frames(False)
    ms_per_step, fused_ms, _, counts1 = run_frames(True)
    survivors1 = int(counts1.sum())
    same = bool(np.array_equal(counts1, counts_un))

    leaves = ds.levels[-1][0]
    parents = ds.levels[-2][0] if len(ds.levels) > 1 else 0
    inner_nodes = sum(l[0] for l in ds.levels[:-1])
    # algorithmic bytes of the fused leaf launch (the frame's dominant kernel):
    # per leaf 64 B local + 4 B parent index + 4 B mesh index read and 64 B
    # world written, each parent's world matrix read once, 20 B per visible
    # drawable and 4 B per StateSet written
    fused_bytes = leaves * (64 + 4 + 4 + 64) + parents * 64 + survivors1 * 20 + ds.K * 4
    fused_gbs = fused_bytes / (fused_ms * 1e-3) / 1e9
    # the separate launches: leaf propagation also writes 16-B spheres; the cull
    # reads 16-B spheres + 4-B mesh indexes
    leaf_bytes = leaves * (64 + 4 + 4 + 64 + 16) + parents * 64
    leaf_gbs = leaf_bytes / (leaf_ms * 1e-3) / 1e9
    cull_bytes = n * 20 + survivors1 * 20 + ds.K * 4
    frame_bytes = fused_bytes + inner_nodes * (64 + 4 + 64)
    result = {
        "metric": "culled drawables/s + % of HBM roofline; p50/p99 view latency (served)",
        "value": n / (ms_per_step * 1e-3),
        "unit": "drawables/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scene generator)",
        "config": {"workload": "full10m: 1000x100x100 transform tree, 10M drawables, 64 StateSets, 3 LODs, "
                               "1 view (propagation + cull + LOD + per-StateSet compaction)",
                   "drawables": n, "nodes": int(sum(l[0] for l in ds.levels)), "statesets": ds.K, "views": 1,
                   "frame": "1 propagation launch (root and inner level together) + 1 fused leaf propagation + "
                            "cull launch",
                   "block": fused_config.blockSize, "items_per_thread": fused_config.itemsPerThread,
                   "autotune": {"fused_ms": fused_table, "single_view": tuned1}},
        "roofline": {"bound": "hbm", "kernel": f"cullFusedKernel<{fused_config.blockSize},{fused_config.itemsPerThread}> "
                                               "(leaf propagation + cull + LOD + compaction)",
                     "achieved": fused_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": fused_gbs / HBM_PEAK_GBS,
                     "duration": "mean over the timed frames of the kernel's end minus the end of the launch "
                                 "before it, both from events its dispatches record (cadrTimeNextLaunch)",
                     "traffic": measured_traffic(f"cullFusedKernel<{fused_config.blockSize},{fused_config.itemsPerThread}>"),
                     "algorithmic_bytes": fused_bytes, "avg_ms": fused_ms},
        "frame": {"algorithmic_bytes": frame_bytes, "hbm_frac": frame_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                  "visible": survivors1, "counts_identical_to_unfused": same},
        "frame_unfused": {"ms_per_frame": un_ms, "culled_drawables_per_s": n / (un_ms * 1e-3),
                          "leaf_ms": leaf_ms, "leaf_gbs": leaf_gbs, "leaf_frac": leaf_gbs / HBM_PEAK_GBS,
                          "leaf_algorithmic_bytes": leaf_bytes, "cull_ms": cull_ms,
                          "cull_gbs": cull_bytes / (cull_ms * 1e-3) / 1e9,
                          "cull_config": {"block": config.blockSize, "items_per_thread": config.itemsPerThread}},
    }
    # reference layout: the gather's DrawablePointers of the visible drawables
    # beside their commands (cadrCompactDrawablePointers), a synthetic
    # 32-B record per drawable as its input
    ptr_in = device.DeviceBuffer(32 * n)
    N.check(lib.cadrFillU32(ptr_in.address, 0x5A5A5A5A, 32 * n, s))
    ptr_out = device.DeviceBuffer(32 * n)
    ss_base = device.DeviceBuffer.from_array(
        np.concatenate([[0], np.cumsum(ds.stateset_counts)[:-1]]).astype(np.uint32))

    def compact_pointers():
        N.check(lib.cadrCompactDrawablePointers(out1.commands.address, out1.counts.address, ss_base.address, 1, ds.K,
                                                out1.stride, ptr_in.address, n, ptr_out.address, s))

    ptr_ms = time_graph(compact_pointers, stream, max(args.steps, 3))
    ptr_bytes = survivors1 * (4 + 32 + 32) + ds.K * 4
    result["reference_layout"] = {
        "what": "cadrCompactDrawablePointers after the frame: each visible drawable's 32-B DrawablePointers "
                "beside its command (the gl_DrawID contract of the reference's shaders)",
        "pointers_ms": ptr_ms, "ms_per_frame": ms_per_step + ptr_ms,
        "pointers_gbs": ptr_bytes / (ptr_ms * 1e-3) / 1e9, "algorithmic_bytes": ptr_bytes}
    for x in (ptr_in, ptr_out, ss_base):
        x.close()
    log(f"full10m fused: {ms_per_step:.3f} ms/frame, fused kernel {fused_ms:.3f} ms ({fused_gbs:.0f} GB/s); "
        f"unfused {un_ms:.3f} ms/frame (leaf {leaf_ms:.3f} ms {leaf_gbs:.0f} GB/s, cull {cull_ms:.3f} ms); "
        f"visible {survivors1}, identical={same}")

    if not args.no_plain100k:
        result["plain100k"] = bench_plain100k(args, stream, lib, N, cull, scenes, device)
    served_config = cull.make_config(args.block, args.items)
    if not args.no_batch:
        result["batched"], served_config = bench_batched(args, ds, config, stream, lib, N, cull, scenes, device,
                                                         propagate_inner, level_descs[-1])
    if not args.no_uneven or not args.no_served or not args.no_gather:
        batch1.close(); batch_fused.close(); out1.close(); ds.close()
    if not args.no_gather:
        result["gather"] = bench_gather(args, stream, lib, N, scenes, device)
    if not args.no_uneven:
        result["uneven"] = bench_uneven(args, config, stream, lib, N, cull, scenes, device)
    if not args.no_served:
        result["served"] = bench_served(args, served_config, N, cull, scenes, device)

    if not args.no_resident:
        result["resident"] = bench_resident(args)
    if not args.no_facade:
        result["facade_cull"] = bench_facade_cull(args)
    if not args.no_hbm_probe:
        result["hbm_probe"] = hbm_probe()
        result["roofline"]["frac_of_measured_copy"] = fused_gbs / result["hbm_probe"]["copy_gbs"]
        if "batched" in result:  # its bytes are almost all command writes
            rb = result["batched"]["roofline"]
            rb["frac_of_measured_write"] = rb["achieved"] / result["hbm_probe"]["write_gbs"]

    if args.copy_calibration:
        copy_calibration(args.copy_calibration, stream, lib, N, device)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)


def cu_count(lib):
    n = ctypes.c_int(0)
    return n.value if lib.cadrGetComputeUnitCount(ctypes.byref(n)) == 0 and n.value > 0 else 256


def hbm_probe():
    """Measured HBM ceiling of this GPU (cadr_amd/bin/hbm_probe, HIP): a
    streaming copy, read-only and write-only pass over 1 GiB, GB/s."""
    import subprocess
    exe = os.path.join(ROOT, "cadr_amd", "bin", "hbm_probe")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built (make -C cadr_amd/csrc)")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    if p.returncode != 0:
        raise RuntimeError(f"hbm_probe failed ({p.returncode}): {p.stderr[-2000:]}")
    r = json.loads(p.stdout.strip().splitlines()[-1])
    r["what"] = "one float4 per thread over 1 GiB, 20 launches after 2 warm-up; copy counts read + write bytes"
    log(f"hbm probe: copy {r['copy_gbs']:.0f} GB/s, read {r['read_gbs']:.0f} GB/s, write {r['write_gbs']:.0f} GB/s")
    return r


def bench_facade_cull(args, frames=10):
    """Cull frames through the CadR facade (cadr_amd/bin/facade_cull, C++): a
    1000 x 100 x 100 tree whose leaf level is propagated on the GPU straight
    into the MatrixLists of 4096 instanced Drawables (one parent-ordered
    cadrTransformPlan launch; the 1024-level batch launches beside it), then
    Renderer::recordDrawableProcessing + recordDrawableCulling for one view;
    medians of HIP-event and host wall-clock frame times."""
    import subprocess
    exe = os.path.join(ROOT, "cadr_amd", "bin", "facade_cull")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built (make -C cadr_amd/csrc)")
    roots = max(1, int(round(1000 * args.scale)))

    def run(mode):
        p = subprocess.run([exe, str(roots), str(frames), mode], capture_output=True, text=True, timeout=900)
        if p.returncode != 0:
            raise RuntimeError(f"facade_cull {mode} failed ({p.returncode}): {p.stderr[-2000:]}")
        return json.loads(p.stdout.strip().splitlines()[-1])

    r = run("plan")
    b = run("batch")
    r["batch_leaf_propagation"] = {"gpu_ms": b["gpu_ms"], "host_ms": b["host_ms"],
                                   "visible_identical": b["visible"] == r["visible"]}
    r["workload"] = (f"facade cull frames: {roots} x 100 x 100 transform tree; the leaves of each (StateSet, "
                     f"Geometry) pair are the instances of one Drawable ({r['drawables']} Drawables, "
                     f"{r['cull_items']} cull items, 64 StateSets, 3 LODs); frame = inner level + leaf level into "
                     f"the MatrixLists + recordDrawableProcessing + recordDrawableCulling(1 view), median of "
                     f"{frames}")
    r["culled_drawables_per_s"] = r["cull_items"] / (r["gpu_ms"] * 1e-3)
    log(f"facade cull: {r['gpu_ms']:.3f} ms/frame on the GPU ({r['host_ms']:.3f} ms host), "
        f"{r['cull_items']} items, {r['visible']} visible; leaf levels in 1024-level batches instead of the "
        f"parent-ordered plan: {b['gpu_ms']:.3f} ms, same visible count: {b['visible'] == r['visible']}")
    return r


def bench_resident(args, frames=10):
    """Facade frames (cadr_amd/bin/facade_frames, C++): the same scene and
    edits with the full per-frame copy of the drawable list and with the
    resident list; medians of host wall clock per frame."""
    import subprocess
    exe = os.path.join(ROOT, "cadr_amd", "bin", "facade_frames")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built (make -C cadr_amd/csrc)")
    n = args.resident_drawables
    edits = max(1, n // 1000)
    p = subprocess.run([exe, str(n), str(frames), str(edits)], capture_output=True, text=True, timeout=900)
    if p.returncode != 0:
        raise RuntimeError(f"facade_frames failed ({p.returncode}): {p.stderr[-2000:]}")
    r = json.loads(p.stdout.strip().splitlines()[-1])
    r["workload"] = (f"facade frames: {n} drawables in 64 StateSets, {edits} drawables re-created in place per "
                     f"frame (Drawable::create), gather at handle level {r['handle_level']}; frame = beginFrame .. "
                     f"stream done, host wall clock, median of {frames}")
    r["speedup"] = r["full_copy_ms"] / r["resident_ms"]
    log(f"resident: full copy {r['full_copy_ms']:.3f} ms/frame ({r['full_copy_bytes']} B), resident "
        f"{r['resident_ms']:.3f} ms/frame ({r['resident_bytes']} B), {r['speedup']:.1f}x")
    return r


def bench_batched(args, ds, view_config, stream, lib, N, cull, scenes, device, propagate_inner, leaf_desc):
    """batch256: all views in one persistent launch (graph-replayed) vs the plain
    pass (cadrCullView with the single-view config) launched once per view;
    then whole frames (propagation + all views): the leaf level as its own
    launch before the batched cull, and fused into it (cadrPropagateCullBatch).
    Returns (result, batched config)."""
    from cadr_amd.tune import autotune, describe
    s = stream.handle
    V = args.batch_views
    cams = scenes.batch_cameras(V)
    views = cull.make_views(cams)
    n = ds.n
    out = cull.CullOutputs(V, n, ds.K)
    config, tuned = cull.make_config(args.block, args.items), None
    if not args.no_tune:
        config, table = autotune([(ds, views, out)], reps=2, stream=s)
        tuned = {"choice": describe(config), "candidates_ms": {tune_key(t): t["ms"] for t in table}}
        log(f"batch{V} autotune: {describe(config)}  {tuned['candidates_ms']}")
    batch = cull.CullBatch(config, [(ds, views, out)])
    # plain: one single-view launch per view, with tiles built for its own
    # tile size, writing straight into the batched output arrays (same layout)
    vtiles, vnt = ds.tiles(cull.tile_size(view_config))
    ws = cull.Workspace(max(vnt, 1), 4)
    plain_jobs = []
    for v in range(V):
        j = batch.host_jobs.copy()
        j["tiles"], j["numTiles"] = vtiles.address, vnt
        j["views"] = batch.host_jobs["views"][0] + v * N.VIEW_DTYPE.itemsize
        j["commands"] = out.commands.address + v * n * 20
        j["counts"] = out.counts.address + v * ds.K * 4
        j["numViews"] = 1
        plain_jobs.append((j, device.DeviceBuffer.from_array(j)))

    def run_plain():
        for j, dj in plain_jobs:
            N.check(lib.cadrCullView(ctypes.byref(view_config), dj.address, N.ptr(j), ws.handle, s))

    from cadr_amd.device import Event
    plain_ms, plain_mean_ms, counts_plain = None, None, None
    if not args.no_plain:
        plain_ms, plain_mean_ms = time_plain(run_plain, stream, max(2, args.batch_steps // 2))
        counts_plain = out.download_counts()

    # batched, captured once into a HIP graph and replayed
    graph = Graph()
    graph.begin(s)
    batch.run(s)
    graph.end(s)
    for _ in range(max(2, args.warmup)):  # the clocks settle after the latency-bound plain pass
        graph.launch(s)
    stream.synchronize()
    e0, e1 = Event(), Event()
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(args.batch_steps):
        graph.launch(s)
    e1.record(s)
    stream.synchronize()
    batch_ms = 1e3 * (time.perf_counter() - t0) / args.batch_steps
    batch_kernel_ms = e0.elapsed_ms(e1) / args.batch_steps
    counts = out.download_counts()
    same = None if counts_plain is None else bool(np.array_equal(counts, counts_plain))
    survivors = int(counts.sum())
    tests = V * n
    alg_bytes = n * 20 + survivors * 20 + V * ds.K * 4
    gbs = alg_bytes / (batch_kernel_ms * 1e-3) / 1e9
    res = {
        "workload": f"batch256: full10m scene x {V} views (eye U[-600,600]^3, random direction, vfov U[40,90] deg), "
                    "one persistent launch per frame, HIP-graph replay",
        "views": V, "tests_per_frame": tests, "visible": survivors,
        "ms_per_frame": batch_ms, "kernel_ms": batch_kernel_ms,
        "culled_drawables_per_s": tests / (batch_ms * 1e-3),
        "roofline": {"bound": "hbm", "kernel": f"cullKernel<{config.blockSize},{config.itemsPerThread}>",
                     "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"cullKernel<{config.blockSize},{config.itemsPerThread}>"),
                     "algorithmic_bytes": alg_bytes,
                     "valu": valu_issue(f"cullKernel<{config.blockSize},{config.itemsPerThread}>", batch_kernel_ms,
                                        cu_count(lib))},
        "config": {"block": config.blockSize, "items_per_thread": config.itemsPerThread, "flags": config.flags,
                   "autotune": tuned},
        "plain_config": {"block": view_config.blockSize, "items_per_thread": view_config.itemsPerThread},
        "plain_ms_per_frame": plain_ms,
        "plain_ms_per_frame_mean": plain_mean_ms,
        "plain_timing": "HIP events around the 256 launches of a frame, fastest of the repetitions",
        "plain_culled_drawables_per_s": None if plain_ms is None else tests / (plain_ms * 1e-3),
        "speedup_vs_plain": None if plain_ms is None else plain_ms / batch_ms,
        "plain_and_batched_counts_identical": same,
    }
    log(f"batch{V}: {batch_ms:.3f} ms/frame ({tests / (batch_ms * 1e-3) / 1e9:.1f} G tests/s, {gbs:.0f} GB/s alg), "
        f"plain {plain_ms} ms, identical={same}")
    res["frame"] = batched_frames(args, [(ds, views, out)], [leaf_desc], propagate_inner, batch, stream, lib, N, cull,
                                  f"batch{V}", inner_nodes=int(sum(lv[0] for lv in ds.levels[:-1])))
    if not args.no_cpu:
        ds.propagate(s)
        stream.synchronize()
        sph, mi = ds.download_spheres(), ds.mesh_index.download(np.uint32, n)
        res["cpu_baseline"] = cpu_cull_rate([(sph, mi, ds.scene.meshes, ds.stateset_counts, views[:16])], 8.0,
                                            f"batch{V} scene (10M drawables) x its first 16 views")
        del sph, mi
    graph.close()
    batch.close(); out.close(); ws.close()
    for _, dj in plain_jobs:
        dj.close()
    return res, config


def batched_frames(args, entries, leaf_descs, propagate_inner, batch, stream, lib, N, cull, name, unfused=None,
                   inner_nodes=0):
    """Whole frames of a batched workload, each one HIP graph: the levels
    above the leaves (propagate_inner), then either every job's leaf level as
    its own launch (world spheres written) followed by the batched cull (or
    the given unfused frame), or the leaf levels fused into the batched launch
    (cadrPropagateCullBatch, no spheres in HBM), for each fused shape; counts
    of both compared.  inner_nodes: nodes of the levels above the leaves."""
    s = stream.handle
    steps = max(3, args.batch_steps)
    leaves = np.concatenate(leaf_descs)

    def separate():
        propagate_inner()
        for lv in leaf_descs:
            N.check(lib.cadrPropagateTransforms(N.ptr(lv), s))
        batch.run(s)

    unfused = unfused or separate

    un_ms = time_graph(unfused, stream, steps)
    counts_un = [e[2].download_counts() for e in entries]
    fused_leaves = leaves.copy()
    fused_leaves["spheres"] = 0
    sweep, same = {}, True
    for block, items in FUSED_BATCH_SHAPES:
        b = cull.CullBatch(cull.make_config(block, items), entries)
        b.set_leaves(fused_leaves)
        for e in entries:
            e[2].counts.upload(np.zeros(e[0].K * e[1].shape[0], np.uint32))
        sweep[f"{block}x{items}"] = time_graph(lambda: (propagate_inner(), b.run_fused_batch(s)), stream, steps)
        same = same and all(np.array_equal(e[2].download_counts(), c) for e, c in zip(entries, counts_un))
        b.close()
    best = min(sweep, key=sweep.get)
    # the fused launch alone (the levels above are already propagated), for its roofline
    b = cull.CullBatch(cull.make_config(*(int(x) for x in best.split("x"))), entries)
    b.set_leaves(fused_leaves)
    fused_ms = time_graph(lambda: b.run_fused_batch(s), stream, steps)
    b.close()
    # per leaf: 64-B local + 4-B parent index + 4-B mesh index read, 64-B world
    # written; each parent's world matrix read once; 20 B per visible command
    # and 4 B per (view, StateSet) count written
    leaves_n = int(leaves["count"].sum())
    parents_n = int(leaves["parentCount"][leaves["parentWorld"] != 0].sum())
    survivors = int(sum(c.sum() for c in counts_un))
    fused_bytes = (leaves_n * (64 + 4 + 4 + 64) + parents_n * 64 + survivors * 20 +
                   sum(e[0].K * e[1].shape[0] * 4 for e in entries))
    fused_gbs = fused_bytes / (fused_ms * 1e-3) / 1e9
    # the unfused frame: 64 + 4 B read and 64 B written per node above the
    # leaves, 152 B per leaf (its world sphere written too) + 64 B per parent,
    # then the batched cull: 20 B per drawable read once, 20 B per visible
    # command and 4 B per (view, StateSet) count written
    unfused_bytes = (inner_nodes * (64 + 4 + 64) + leaves_n * (64 + 4 + 4 + 64 + 16) + parents_n * 64 +
                     leaves_n * 20 + survivors * 20 + sum(e[0].K * e[1].shape[0] * 4 for e in entries))
    r = {"unfused_ms_per_frame": un_ms, "unfused_algorithmic_bytes": unfused_bytes,
         "unfused_hbm_frac": unfused_bytes / (un_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "fused_ms_per_frame": sweep[best], "fused_shape": best,
         "fused_candidates_ms": sweep, "speedup_fused": un_ms / sweep[best], "fused_counts_identical": bool(same),
         "fused_roofline": {"bound": "hbm", "kernel": f"cullKernel<{best.replace('x', ',')},true>",
                            "kernel_ms": fused_ms, "achieved": fused_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": fused_gbs / HBM_PEAK_GBS, "algorithmic_bytes": fused_bytes},
         "frame": "levels above the leaves + (leaf propagation launch(es), batched cull) or the leaf levels fused "
                  "into the batched launch; one HIP graph per frame"}
    log(f"{name} frame: unfused {un_ms:.3f} ms, fused {sweep[best]:.3f} ms ({best}) {sweep}, identical={same}; "
        f"fused kernel {fused_ms:.3f} ms ({fused_gbs:.0f} GB/s alg)")
    return r


def bench_uneven(args, view_config, stream, lib, N, cull, scenes, device):
    """uneven64: 64 resident scenes (generated in HBM), each with 1..32 views;
    a frame propagates every scene's tree and culls all views of all scenes in
    one persistent launch, the whole frame captured in one HIP graph.  Against
    it: the plain pass, one cadrCullView launch per (scene, view)."""
    from cadr_amd.device import Event
    from cadr_amd.tune import autotune, describe
    s = stream.handle
    t0 = time.time()
    specs = scenes.uneven_specs()
    dss, entries, views_l = [], [], []
    for i, (n, depth, lods, nv) in enumerate(specs):
        ds = cull.DeviceScene(scenes.uneven_synth(i, n, depth, lods, nv))
        views = cull.make_views(ds.scene.cameras)
        out = cull.CullOutputs(nv, n, ds.K)
        dss.append(ds); views_l.append(views); entries.append((ds, views, out))
    device.synchronize()
    log(f"uneven64: {len(specs)} scenes, {sum(d.n for d in dss)} drawables, "
        f"{sum(v.shape[0] for v in views_l)} views generated in {time.time() - t0:.1f}s")
    levels = [ds.level_descs(spheres=True) for ds in dss]
    for lvs in levels:
        for lv in lvs:
            N.check(lib.cadrPropagateTransforms(N.ptr(lv), s))
    config, tuned = cull.make_config(args.block, args.items), None
    if not args.no_tune:
        config, table = autotune(entries, reps=2, stream=s)
        tuned = {"choice": describe(config), "candidates_ms": {tune_key(t): t["ms"] for t in table}}
        log(f"uneven64 autotune: {describe(config)}  {tuned['candidates_ms']}")
    batch = cull.CullBatch(config, entries)

    # one propagation launch per tree depth, over all scenes that have it
    by_depth = {}
    for lvs in levels:
        for d, lv in enumerate(lvs):
            by_depth.setdefault(d, []).append(lv)
    depth_batches = []
    for d in sorted(by_depth):
        host = np.concatenate(by_depth[d])
        depth_batches.append((device.DeviceBuffer.from_array(host), host))

    def propagate_all():
        for dev, host in depth_batches:
            N.check(lib.cadrPropagateTransformsBatch(dev.address, N.ptr(host), host.shape[0], s))

    # the levels above each scene's leaf level, one launch per depth
    inner_by_depth = {}
    for lvs in levels:
        for d, lv in enumerate(lvs[:-1]):
            inner_by_depth.setdefault(d, []).append(lv)
    inner_batches = []
    for d in sorted(inner_by_depth):
        host = np.concatenate(inner_by_depth[d])
        inner_batches.append((device.DeviceBuffer.from_array(host), host))

    def propagate_inner():
        for dev, host in inner_batches:
            N.check(lib.cadrPropagateTransformsBatch(dev.address, N.ptr(host), host.shape[0], s))

    # plain: one single-view launch per (scene, view), same outputs, tiles
    # built for the single-view tile size
    vtile = cull.tile_size(view_config)
    ws = cull.Workspace(max(max(ds.tiles(vtile)[1] for ds in dss), 1), 4)
    plain = []
    for j, (ds, views, out) in enumerate(entries):
        vtiles, vnt = ds.tiles(vtile)
        for v in range(views.shape[0]):
            hj = batch.host_jobs[j:j + 1].copy()
            hj["tiles"], hj["numTiles"] = vtiles.address, vnt
            hj["views"] = batch.host_jobs["views"][j] + v * N.VIEW_DTYPE.itemsize
            hj["commands"] = out.commands.address + v * ds.n * 20
            hj["counts"] = out.counts.address + v * ds.K * 4
            hj["numViews"] = 1
            plain.append((hj, device.DeviceBuffer.from_array(hj)))

    def run_plain():
        for hj, dj in plain:
            N.check(lib.cadrCullView(ctypes.byref(view_config), dj.address, N.ptr(hj), ws.handle, s))

    propagate_all()
    e0, e1 = Event(), Event()
    plain_ms, plain_mean_ms, counts_plain = None, None, None
    if not args.no_plain:
        plain_ms, plain_mean_ms = time_plain(run_plain, stream, 3)
        counts_plain = [e[2].download_counts() for e in entries]

    frame = Graph()
    frame.begin(s)
    propagate_all()
    batch.run(s)
    frame.end(s)
    cull_only = Graph()
    cull_only.begin(s)
    batch.run(s)
    cull_only.end(s)
    for g in (frame, cull_only):
        g.launch(s)
        g.launch(s)
    stream.synchronize()
    steps = max(3, args.batch_steps)
    e0.record(s)
    for _ in range(steps):
        frame.launch(s)
    e1.record(s)
    stream.synchronize()
    frame_ms = e0.elapsed_ms(e1) / steps
    e0.record(s)
    for _ in range(steps):
        cull_only.launch(s)
    e1.record(s)
    stream.synchronize()
    cull_ms = e0.elapsed_ms(e1) / steps
    counts = [e[2].download_counts() for e in entries]
    same = None if counts_plain is None else all(np.array_equal(a, b) for a, b in zip(counts, counts_plain))
    survivors = int(sum(c.sum() for c in counts))
    tests = sum(ds.n * v.shape[0] for ds, v in zip(dss, views_l))
    nodes = sum(sum(l[0] for l in ds.levels) for ds in dss)
    leaves = sum(ds.n for ds in dss)
    cull_bytes = sum(ds.n * 20 for ds in dss) + survivors * 20 + sum(ds.K * 4 * v.shape[0] for ds, v in zip(dss, views_l))
    prop_bytes = nodes * (64 + 64 + 4) + leaves * (4 + 16)
    cull_gbs = cull_bytes / (cull_ms * 1e-3) / 1e9
    res = {
        "workload": "uneven64: 64 scenes of round(exp(U[ln 1e4, ln 5e6])) drawables, tree depth U{1..4}, LOD count "
                    "U{1..4}, 64 StateSets, U{1..32} views each; propagation of every scene (one launch per tree "
                    "depth) + all views of all scenes in one persistent cull launch, one HIP graph per frame",
        "scenes": len(dss), "drawables": leaves, "nodes": nodes, "views": int(sum(v.shape[0] for v in views_l)),
        "tests_per_frame": tests, "visible": survivors,
        "ms_per_frame": frame_ms, "cull_ms": cull_ms,
        "culled_drawables_per_s": tests / (frame_ms * 1e-3),
        "cull_roofline": {"bound": "hbm", "kernel": f"cullKernel<{config.blockSize},{config.itemsPerThread}>",
                          "achieved": cull_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": cull_gbs / HBM_PEAK_GBS,
                          "algorithmic_bytes": cull_bytes},
        "frame_hbm_frac": (cull_bytes + prop_bytes) / (frame_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "config": {"block": config.blockSize, "items_per_thread": config.itemsPerThread, "flags": config.flags,
                   "autotune": tuned},
        "plain_cull_ms": plain_ms,
        "plain_cull_ms_mean": plain_mean_ms,
        "speedup_cull_vs_plain": None if plain_ms is None else plain_ms / cull_ms,
        "plain_and_batched_counts_identical": same,
    }
    log(f"uneven64: frame {frame_ms:.3f} ms, cull {cull_ms:.3f} ms ({cull_gbs:.0f} GB/s alg), "
        f"plain cull {plain_ms} ms, identical={same}")
    res["fused_frame"] = batched_frames(args, entries, [lvs[-1] for lvs in levels], propagate_inner, batch, stream,
                                        lib, N, cull, "uneven64", unfused=lambda: (propagate_all(), batch.run(s)),
                                        inner_nodes=nodes - leaves)
    res["fused_frame"]["culled_drawables_per_s"] = tests / (res["fused_frame"]["fused_ms_per_frame"] * 1e-3)
    if not args.no_cpu:
        # the first 8 scenes' inputs (generated on the device) read back
        samples = []
        for ds, views in zip(dss[:8], views_l[:8]):
            lv = [(lb.download(np.float32, 16 * c).reshape(c, 16), None if pb is None else pb.download(np.uint32, c))
                  for c, lb, pb, _ in ds.levels]
            samples.append((lv, ds.mesh_index.download(np.uint32, ds.n), ds.scene.meshes, ds.stateset_counts, views))
        res["cpu_baseline"] = cpu_frame_rate(samples, 10.0, f"uneven64 scenes 0-7 ({sum(x[1].shape[0] for x in samples)} "
                                                            f"drawables, {sum(x[4].shape[0] for x in samples)} views)")
        del samples
    frame.close(); cull_only.close()
    for _, dj in plain:
        dj.close()
    batch.close(); ws.close()
    for ds, _, out in entries:
        out.close(); ds.close()
    for dev, _ in depth_batches + inner_batches:
        dev.close()
    return res


# Launches in flight in the serving loop.  Measured on served128 (256 x 8, no
# tile spheres, p99 at 50k / 65k / 80k / 100k req/s): 1 in flight 0.92 /
# 1.36 / 1.75 / 2.59 ms, 2 in flight 0.98 / 1.26 / 1.66 / 2.49 ms, 3 0.85 /
# 1.32 / 1.82 / 2.81 ms, 4 0.97 / 1.71 / 2.39 / 3.89 ms, 8 1.11 / 2.08 /
# 3.00 / 5.13 ms: persistent launches running at once share the CUs, so
# beyond two the extra overlap costs more than it hides.
SERVED_INFLIGHT = 2


def bench_served(args, config, N, cull, scenes, device):
    """served128: 128 resident flat scenes (generated in HBM); Poisson cull
    requests, 80% to the 16 hottest scenes, admitted by output memory and
    batched across scenes by the serving loop; latency from scheduled arrival
    to the result on the host."""
    from cadr_amd.serve import Server
    from cadr_amd.tune import autotune, describe
    t0 = time.time()
    sizes = scenes.served_specs()
    dss = [cull.DeviceScene(scenes.served_synth(i, n)) for i, n in enumerate(sizes)]
    device.synchronize()
    views = cull.make_views(scenes.batch_cameras(256))
    tuned = None
    if not args.no_tune:
        # a launch as the server forms it at moderate load: the 16 hot scenes,
        # two views each
        outs = [cull.CullOutputs(2, ds.n, ds.K) for ds in dss[:16]]
        entries = [(ds, views[2 * i:2 * i + 2], out) for i, (ds, out) in enumerate(zip(dss[:16], outs))]
        config, table = autotune(entries, reps=3)
        tuned = {"choice": describe(config), "candidates_ms": {tune_key(t): t["ms"] for t in table}}
        log(f"served128 autotune (16 scenes x 2 views): {describe(config)}  {tuned['candidates_ms']}")
        for out in outs:
            out.close()
    srv = Server(block=config.blockSize, items=config.itemsPerThread, flags=config.flags, max_batch_views=256,
                 inflight=SERVED_INFLIGHT, budget_bytes=int(args.served_budget_gib * (1 << 30)))
    for ds in dss:
        srv.add_scene(ds)
    srv.start()
    log(f"served128: {len(dss)} scenes, {sum(sizes)} drawables resident, set up in {time.time() - t0:.1f}s")
    srv.run_load(views, rate=2000.0, duration=0.2, seed=99, hot=16, hot_fraction=0.8)  # warm-up
    runs = []
    for rate in [float(r) for r in args.served_rates.split(",") if r]:
        st = srv.run_load(views, rate=rate, duration=args.served_seconds, seed=int(rate), hot=16, hot_fraction=0.8)
        st.pop("samples")
        st["offered_rate"] = rate
        # algorithmic bytes of the run's launches: each job's scene arrays
        # once, plus the commands written, over the run's span
        alg = st["sceneBytesRead"] + 20 * st["visible"]
        st["algorithmic_gbs"] = alg / st["elapsedSeconds"] / 1e9 if st["elapsedSeconds"] > 0 else 0.0
        st["hbm_frac"] = st["algorithmic_gbs"] / HBM_PEAK_GBS
        runs.append(st)
        log(f"served128 @ {rate:.0f} req/s: {st['requestsPerSecond']:.0f} req/s, "
            f"{st['drawablesPerSecond'] / 1e9:.1f} G tests/s, p50 {st['p50Ms']:.3f} ms, p99 {st['p99Ms']:.3f} ms, "
            f"p999 {st['p999Ms']:.3f} ms, {st['meanViewsPerLaunch']:.1f} views/launch, "
            f"{st['algorithmic_gbs']:.0f} GB/s algorithmic ({100 * st['hbm_frac']:.0f}% of HBM)")
    # saturation: the highest offered rate whose p99 stays under 2 ms
    under = [r["offered_rate"] for r in runs if r["p99Ms"] < 2.0 and r["refused"] == 0]
    srv.close()
    # overload: a small output budget binds, so requests queue for memory
    # and the bounded queue refuses the excess
    overload = None
    if args.served_overload_gib > 0:
        small = Server(block=config.blockSize, items=config.itemsPerThread, flags=config.flags, max_batch_views=256,
                       inflight=SERVED_INFLIGHT, budget_bytes=int(args.served_overload_gib * (1 << 30)), max_queued=4096)
        for ds in dss:
            small.add_scene(ds)
        small.start()
        overload = small.run_load(views, rate=100_000.0, duration=args.served_seconds, seed=777, hot=16,
                                  hot_fraction=0.8)
        overload.pop("samples")
        overload.update({"offered_rate": 100_000.0, "output_budget_gib": args.served_overload_gib,
                         "max_queued_requests": 4096})
        log(f"served128 overload (budget {args.served_overload_gib} GiB, queue 4096) @ 100k req/s: "
            f"{overload['requestsPerSecond']:.0f} req/s done, {overload['refused']} refused, "
            f"p50 {overload['p50Ms']:.3f} ms, p99 {overload['p99Ms']:.3f} ms")
        small.close()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_cull_rate([(ds.download_spheres(), ds.mesh_index.download(np.uint32, ds.n), ds.scene.meshes,
                              ds.stateset_counts, views[i:i + 1]) for i, ds in enumerate(dss[:16])], 5.0,
                            "served128: the 16 hottest scenes, one request (view) each")
    for ds in dss:
        ds.close()
    return {"max_rate_p99_under_2ms": max(under) if under else None, "overload": overload, "cpu_baseline": cpu,
            "workload": "served128: 128 flat scenes of round(exp(U[ln 1e4, ln 2e7])) drawables (64 StateSets, 3 LODs), "
                        "Poisson arrivals, 80% of requests to the 16 hottest scenes, views as batch256",
            "drawables_resident": int(sum(sizes)), "output_budget_gib": args.served_budget_gib,
            "max_batch_views": 256, "inflight_batches": SERVED_INFLIGHT,
            "config": {"block": config.blockSize, "items_per_thread": config.itemsPerThread, "flags": config.flags,
                       "autotune": tuned},
            "runs": runs}


def copy_calibration(nbytes, stream, lib, N, device):
    """A plain device-to-device copy of nbytes, 3 times: under rocprofv3 --pmc
    its FETCH_SIZE / WRITE_SIZE per dispatch against nbytes gives the scale
    factor of those counters."""
    src, dst = device.DeviceBuffer(nbytes), device.DeviceBuffer(nbytes)
    N.check(lib.cadrFillU32(src.address, 0x3F800000, nbytes, stream.handle))
    for _ in range(3):
        N.check(lib.cadrMemcpyDtoD(dst.address, src.address, nbytes, stream.handle))
    stream.synchronize()
    log(f"copy calibration: 3 x {nbytes} bytes device to device")
    src.close(); dst.close()


def bench_gather(args, stream, lib, N, scenes, device, n=10_000_000):
    """The reference's own pass (processDrawables.comp, handle level 3) at
    10M drawables: 48-B DrawableGpuData records, 3-level handle lookups,
    PrimitiveSet and MatrixList-header reads, 16-B indirect + 32-B pointer
    writes.  Outputs checked against the closed form; the oracle (one
    thread) timed on a 1M-drawable level-3 sample."""
    from cadr_amd.device import DeviceBuffer, Event
    s = stream.handle
    t0 = time.time()
    g = scenes.GatherScene(n)
    size = g.layout(0)["size"]
    img_buf = DeviceBuffer(size)
    img, push, level = g.image(img_buf.address)
    img_buf.upload(img)
    ind, ptr = DeviceBuffer(16 * n), DeviceBuffer(32 * n)
    pc = N.PushConstants(push[0], push[1], ind.address, ptr.address)
    leaf_bytes = (g.layout(0)["leaf_range"][1] - g.layout(0)["leaf_range"][0] + 1) * 2048 * 8
    del img
    log(f"gather: {n} drawables, handle level {level}, image {size / 1e9:.2f} GB built in {time.time() - t0:.1f}s")

    def run():
        N.check(lib.cadrProcessDrawables(ctypes.byref(pc), level, n, s))

    for _ in range(2):
        run()
    stream.synchronize()
    e0, e1 = Event(), Event()
    steps = 10
    e0.record(s)
    for _ in range(steps):
        run()
    e1.record(s)
    stream.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    ei, ep = g.expected(img_buf.address)
    same = bool(np.array_equal(ind.download(np.uint32, 4 * n).reshape(n, 4), ei) and
                np.array_equal(ptr.download(np.uint64, 4 * n).reshape(n, 4), ep))
    del ei, ep
    # per drawable: 48-B record + the MatrixList header line (64 B) read, 16 + 32 B
    # written; the last-level handle tables read once
    alg = n * (48 + 64 + 16 + 32) + leaf_bytes
    gbs = alg / (ms * 1e-3) / 1e9
    img_buf.close(); ind.close(); ptr.close()
    # CPU: the oracle's gather, one thread, on a 1M-drawable level-3 image
    import oracle
    m = 1_000_000
    gs = scenes.GatherScene(m, first_handle=1 << 22)
    base, ib, pb = 1 << 44, 1 << 45, 3 << 44
    simg, spush, slevel = gs.image(base)
    sind, sptr = np.zeros(4 * m, np.uint32), np.zeros(4 * m, np.uint64)
    regions = [(base, simg), (ib, sind.view(np.uint8)), (pb, sptr.view(np.uint8))]
    times = []
    for _ in range(3):
        t1 = time.perf_counter()
        oracle.process_drawables(regions, (spush[0], spush[1], ib, pb), slevel, m)
        times.append(time.perf_counter() - t1)
    cpu_rate = m / min(times)
    res = {"workload": f"gather: processDrawables at {n} drawables, handle level {level} (64 geometries, one "
                       "MatrixList per drawable, 4 PrimitiveSets per geometry)",
           "ms": ms, "drawables_per_s": n / (ms * 1e-3),
           "roofline": {"bound": "hbm", "kernel": f"processDrawablesKernel<{level}>", "achieved": gbs,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": alg,
                        "traffic": measured_traffic(f"processDrawablesKernel<{level}>")},
           "outputs_match_closed_form": same,
           "cpu_oracle": {"value": cpu_rate, "unit": "drawables/s", "cores": 1,
                          "sample": f"{m} drawables, handle level 3, oracle/cadr_oracle.c, best of 3"}}
    log(f"gather: {ms:.3f} ms ({n / (ms * 1e-3) / 1e9:.1f} G drawables/s, {gbs:.0f} GB/s alg), match={same}, "
        f"cpu oracle {cpu_rate / 1e6:.1f} M drawables/s")
    return res


def time_plain(run_plain, stream, reps):
    """The plain pass (one launch per view) timed with HIP events on its stream:
    one warm-up frame, then `reps` frames; returns (fastest, mean) ms."""
    from cadr_amd.device import Event
    s = stream.handle
    run_plain()
    stream.synchronize()
    times = []
    for _ in range(reps):
        e0, e1 = Event(), Event()
        e0.record(s)
        run_plain()
        e1.record(s)
        stream.synchronize()
        times.append(e0.elapsed_ms(e1))
    return min(times), sum(times) / len(times)


def time_graph(fn, stream, steps):
    """fn's launches captured in one HIP graph; 2 warm-up replays, then
    `steps` replays timed with HIP events on the stream; ms per replay."""
    from cadr_amd.device import Event
    s = stream.handle
    g = Graph()
    g.begin(s)
    fn()
    g.end(s)
    try:
        for _ in range(2):
            g.launch(s)
        stream.synchronize()
        e0, e1 = Event(), Event()
        e0.record(s)
        for _ in range(steps):
            g.launch(s)
        e1.record(s)
        stream.synchronize()
        return e0.elapsed_ms(e1) / steps
    finally:
        g.close()


FUSED_BATCH_SHAPES = ((256, 4), (256, 8), (512, 4))


def tune_key(t):
    return f"{t['block']}x{t['items']}{'' if t['flags'] else '+ts'}" + ("" if t["agrees"] else "!mismatch")


class Graph:
    """hipStreamBeginCapture / hipGraphInstantiate / hipGraphLaunch through libamdhip64."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.graph = ctypes.c_void_p(0)
        self.exec = ctypes.c_void_p(0)

    def _ck(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: {rc}")

    def begin(self, stream):
        self._ck(self.hip.hipStreamBeginCapture(ctypes.c_void_p(stream), 0), "hipStreamBeginCapture")

    def end(self, stream):
        self._ck(self.hip.hipStreamEndCapture(ctypes.c_void_p(stream), ctypes.byref(self.graph)), "hipStreamEndCapture")
        self._ck(self.hip.hipGraphInstantiate(ctypes.byref(self.exec), self.graph, None, None, ctypes.c_size_t(0)),
                 "hipGraphInstantiate")

    def launch(self, stream):
        self._ck(self.hip.hipGraphLaunch(self.exec, ctypes.c_void_p(stream)), "hipGraphLaunch")

    def close(self):
        if self.exec:
            self.hip.hipGraphExecDestroy(self.exec)
        if self.graph:
            self.hip.hipGraphDestroy(self.graph)
        self.exec = self.graph = None


def cpu_threads() -> int:
    """Threads of the CPU baselines: the host's cores, at most 16 (one GPU's
    share of a shared host)."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            return next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "unknown CPU")
    except OSError:
        return "unknown CPU"


def cpu_cull_rate(scenes_views, budget_s: float, what: str):
    """The oracle's cull (OpenMP over (view, StateSet) pairs) on the host cores
    over a bounded sample: scenes_views = [(spheres, mesh_index, meshes,
    stateset_counts, views)], all culled once per repetition into reused output
    arrays; median repetition; (drawable, view) tests per second."""
    import oracle
    from cadr_amd import _native as N
    threads = cpu_threads()
    outs = [(np.zeros((v.shape[0], sp.shape[0]), N.CMD_DTYPE), np.zeros((v.shape[0], len(c)), np.uint32))
            for sp, _, _, c, v in scenes_views]
    tests = sum(sp.shape[0] * v.shape[0] for sp, _, _, _, v in scenes_views)
    times = []
    start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - start < budget_s and len(times) < 200):
        t0 = time.perf_counter()
        for (sp, mi, me, c, v), o in zip(scenes_views, outs):
            oracle.cull(sp, mi, me, c, v, cmd_dtype=N.CMD_DTYPE, nthreads=threads, out=o)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    log(f"cpu baseline ({what}): {tests / t / 1e6:.0f} M tests/s ({len(times)} repetitions, {threads} threads)")
    return {"value": tests / t, "unit": "culled drawables/s", "cores": threads, "kind": "port",
            "sample": f"{what}: {tests} (drawable, view) tests per repetition, median of {len(times)}; "
                      f"oracle/cadr_oracle.c cull, gcc -O2, {threads} OpenMP threads of {os.cpu_count()} "
                      f"({cpu_model()})"}


def cpu_frame_rate(samples, budget_s: float, what: str):
    """Whole frames on the host cores: samples = [(levels [(local, parent)],
    mesh_index, meshes, stateset_counts, views)], each propagated level by
    level (the last level writing world spheres) and culled for its views once
    per repetition; median repetition; (drawable, view) tests per second."""
    import oracle
    from cadr_amd import _native as N
    threads = cpu_threads()
    oracle.set_threads(threads)
    outs = [(np.zeros((v.shape[0], mi.shape[0]), N.CMD_DTYPE), np.zeros((v.shape[0], len(c)), np.uint32))
            for _, mi, _, c, v in samples]
    tests = sum(mi.shape[0] * v.shape[0] for _, mi, _, _, v in samples)
    times = []
    start = time.perf_counter()
    try:
        while len(times) < 3 or (time.perf_counter() - start < budget_s and len(times) < 200):
            t0 = time.perf_counter()
            for (levels, mi, me, c, v), o in zip(samples, outs):
                prev = None
                for d, (local, parent) in enumerate(levels):
                    last = d == len(levels) - 1
                    prev, sph = oracle.propagate_level(local, parent, prev, mi if last else None, me if last else None)
                oracle.cull(sph, mi, me, c, v, cmd_dtype=N.CMD_DTYPE, nthreads=threads, out=o)
            times.append(time.perf_counter() - t0)
    finally:
        oracle.set_threads(1)
    t = float(np.median(times))
    log(f"cpu baseline ({what}): {tests / t / 1e6:.0f} M tests/s ({len(times)} repetitions, {threads} threads)")
    return {"value": tests / t, "unit": "culled drawables/s", "cores": threads, "kind": "port",
            "sample": f"{what}: whole frames (propagation + cull), {tests} (drawable, view) tests per repetition, "
                      f"median of {len(times)}; oracle/cadr_oracle.c, gcc -O2, {threads} OpenMP threads of "
                      f"{os.cpu_count()} ({cpu_model()})"}


def bench_plain100k(args, stream, lib, N, cull, scenes, device):
    """plain100k: the workload's 100k-drawable flat scene (one StateSet, the
    default camera) culled for one view by the single-view kernel; launch-
    bound at this size.  Per-launch time from HIP events around single
    launches; back-to-back throughput from one event pair around 200 launches
    and from a HIP graph of 100 launches; the oracle on the same scene and
    view beside it."""
    from cadr_amd.device import Event
    from cadr_amd.tune import autotune, describe
    import oracle
    s = stream.handle
    scene = scenes.plain_scene(100_000)
    ds = cull.DeviceScene(scene)
    ds.propagate(s)
    views = cull.make_views([scenes.DEFAULT_CAMERA])
    out = cull.CullOutputs(1, ds.n, ds.K)
    config, table = cull.make_config(256, 4), None
    if not args.no_tune:
        config, table = autotune([(ds, views, out)], single_view=True, reps=20, stream=s)
    b = cull.CullBatch(config, [(ds, views, out)])
    for _ in range(20):
        b.run_view(s)
    stream.synchronize()
    pairs = [(Event(), Event()) for _ in range(200)]
    for e0, e1 in pairs:
        e0.record(s)
        b.run_view(s)
        e1.record(s)
    stream.synchronize()
    single = sorted(e0.elapsed_ms(e1) for e0, e1 in pairs)
    e0, e1 = Event(), Event()
    e0.record(s)
    for _ in range(200):
        b.run_view(s)
    e1.record(s)
    stream.synchronize()
    back_to_back = e0.elapsed_ms(e1) / 200
    graph_ms = time_graph(lambda: [b.run_view(s) for _ in range(100)], stream, 5) / 100
    counts = out.download_counts()
    survivors = int(counts.sum())
    _, spheres = oracle.propagate_scene(scene)
    exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                                       cmd_dtype=N.CMD_DTYPE)
    cmds = out.download()[0]
    identical = bool(np.array_equal(counts, exp_counts) and
                     cmds[0, :survivors].tobytes() == exp_cmds[0, :survivors].tobytes())
    alg = ds.n * 20 + survivors * 20 + ds.K * 4
    med = single[len(single) // 2]
    res = {"workload": "plain100k: 100k drawables (64 meshes, translation U[-500,500]^3, scale log-uniform [0.1,2]), "
                       "1 StateSet, camera at the origin looking -z, vfov 60 deg, 16:9; one single-view launch",
           "drawables": ds.n, "visible": survivors, "identical_to_oracle": identical,
           "launch_us_median": 1e3 * med, "launch_us_min": 1e3 * single[0],
           "back_to_back_us": 1e3 * back_to_back, "graph_us": 1e3 * graph_ms,
           "culled_drawables_per_s": ds.n / (med * 1e-3),
           "culled_drawables_per_s_graph": ds.n / (graph_ms * 1e-3),
           "algorithmic_bytes": alg, "hbm_frac_graph": alg / (graph_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "bound": "launch / latency: 2.2 MB per launch stays in the caches",
           "config": {"block": config.blockSize, "items_per_thread": config.itemsPerThread,
                      "autotune": None if table is None else {tune_key(t): t["ms"] for t in table}},
           "cpu_baseline": cpu_cull_rate([(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views)],
                                         3.0, "plain100k, 1 view")}
    log(f"plain100k: {1e3 * med:.1f} us per launch (median), {1e3 * graph_ms:.1f} us per launch in a graph, "
        f"{survivors} visible, identical to the oracle: {identical}")
    b.close(); out.close(); ds.close()
    return res


def cpu_baseline(scale: float = 1.0, budget_s: float = 15.0):
    """The oracle (C restatement, gcc -O2, OpenMP over nodes and over
    (view, StateSet) pairs) on the host cores on the headline workload itself
    (full10m, 10M drawables): whole frames (propagation + cull of one view)
    repeated until about budget_s seconds; median frame time.  Threads: the
    host's cores, at most 16 (one GPU's share of a shared host)."""
    import oracle
    from cadr_amd import _native as N
    from cadr_amd import cull, scenes
    threads = cpu_threads()
    oracle.set_threads(threads)
    scene = scenes.full10m_scene(scale)
    views = cull.make_views([scenes.DEFAULT_CAMERA])
    times = []
    start = time.perf_counter()
    try:
        while len(times) < 3 or (time.perf_counter() - start < budget_s and len(times) < 500):
            t0 = time.perf_counter()
            _, spheres = oracle.propagate_scene(scene)
            oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                        cmd_dtype=N.CMD_DTYPE, nthreads=threads)
            times.append(time.perf_counter() - t0)
    finally:
        oracle.set_threads(1)
    t = float(np.median(times))
    model = cpu_model()
    log(f"cpu baseline: {t:.3f} s per {scene.num_drawables}-drawable frame ({len(times)} frames, {threads} threads)")
    return {"value": scene.num_drawables / t, "unit": "drawables/s", "cores": threads, "kind": "port",
            "sample": f"full10m at {scale:g} of the roots ({scene.num_drawables} drawables), whole frames "
                      f"(propagation + cull of 1 view), median of {len(times)}; oracle/cadr_oracle.c, gcc -O2, "
                      f"{threads} OpenMP threads of {os.cpu_count()} ({model})"}


if __name__ == "__main__":
    main()
