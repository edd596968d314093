"""Benchmark: ray-batch closest-hit throughput on one MI355X.

Workload (one "step" = one pass of the hot path over one batch): the level-8
displaced icosphere (1,310,720 triangles, unit radius, 3-octave value noise of
amplitude 0.05) inside a 4-unit cube of 12 triangles, kd-tree built with the
reference's SAH builder, 2^24 rays with origins uniform in the cube and
directions uniform on the sphere; every step is one closest-hit launch
(tn_intersect) over all rays, inputs and outputs resident in HBM.  Hit ids
are bit-exact against the reference semantics (checked on a sample here and
in full by tests/test_gpu_intersect.py).

Prints one JSON line.  `roofline` is computed from the algorithmic bytes of
the traversal kernel (counted by an extra, untimed counted launch over the
same batch) divided by the average launch duration measured with HIP events
on the launch stream.  `cpu_baseline` times the oracle's restatement of the
reference traversal (KDTreeIntersection::intersect semantics, no early exit)
on a bounded sample of the same rays on all host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def host_cpu():
    """Host CPU model, the logical CPUs this process may run on, and the thread
    count the CPU baselines use: all of them, capped at OMP_NUM_THREADS when
    it is set (the host's CPU share per GPU)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else avail
    return {"model": model, "logical_cpus": avail, "nproc": os.cpu_count(), "threads": threads}
METRIC = "Mrays/s + time-to-relMSE 0.01, guided vs plain tracer, 1 MI355X"


def note(msg):
    """a progress line on stderr (stdout carries only the JSON line)"""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays-log2", type=int, default=24)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-log2", type=int, default=23)
    ap.add_argument("--no-guided", action="store_true",
                    help="skip the guided-vs-plain time-to-relMSE measurement")
    ap.add_argument("--no-render", action="store_true",
                    help="skip the wavefront render throughput measurement")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched heterogeneous rendering measurement")
    ap.add_argument("--guided-ref-spp", type=int, default=8192)
    ap.add_argument("--guided-width", type=int, default=512)
    ap.add_argument("--guided-min-pass", type=int, default=8,
                    help="image passes after the first are sized from the error so far (relMSE ~ 1/spp), "
                         "in multiples of this many spp, at most 32 (0: fixed 32-spp passes)")
    ap.add_argument("--render-spp", type=int, default=16)
    ap.add_argument("--render-repeats", type=int, default=3)
    ap.add_argument("--mlp-probe", action="store_true",
                    help="only the guide MLP kernels: 3 forward launches over 2^20 points and 3 "
                         "training steps of 2^16 samples (for rocprofv3 MFMA counters), then exit")
    ap.add_argument("--tree-probe", action="store_true",
                    help="reference vs rendering kd-tree on render-like ray batches of the "
                         "occluded-light scene (camera, shadow, bounce), with work counters, then exit")
    ap.add_argument("--launch-probe", action="store_true",
                    help="the 32-scene batch and a small 4-scene batch rendered with each launch form "
                         "(direct launches on two pipelines, graph replay, one persistent cooperative "
                         "launch) at fixed chunk sizes, then exit")
    ap.add_argument("--guide-pass-probe", action="store_true",
                    help="1-spp guided training passes of the occluded-light scene (512x512, depth 6) timed "
                         "with the records trained in order and deferred beside the next render, then exit")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="one traversal launch plus a 1 GiB device copy (the calibration "
                         "kernel for rocprofv3 FETCH_SIZE / WRITE_SIZE), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import turner_amd as T
    from turner_amd import scenes
    from turner_amd._lib import TN_ISECT_COUNT, TN_ISECT_REFINED

    if args.mlp_probe:
        print(json.dumps(mlp_probe()), flush=True)
        return
    if args.tree_probe:
        print(json.dumps(tree_probe()), flush=True)
        return
    if args.launch_probe:
        print(json.dumps(launch_probe()), flush=True)
        return
    if args.guide_pass_probe:
        print(json.dumps(guide_pass_probe()), flush=True)
        return
    dev = torch.device("cuda:0")
    name, cus = T.device_info()
    sc = scenes.intersect_scene(level=args.level)
    t0 = time.perf_counter()
    tree = T.KDTree(sc["verts"], refined=True)
    build_s = time.perf_counter() - t0
    n = 1 << args.rays_log2
    o, d = scenes.random_rays(n, seed=2024)
    to = torch.from_numpy(o).to(dev)
    td = torch.from_numpy(d).to(dev)
    ids = torch.empty(n, dtype=torch.int32, device=dev)
    r = torch.empty(n, dtype=torch.float32, device=dev)
    s = torch.empty_like(r)
    t = torch.empty_like(r)
    stream = torch.cuda.current_stream(dev)
    L = T.lib()

    def launch(flags=TN_ISECT_REFINED):
        T.check(L.tn_intersect(tree.handle, T.ptr(to), T.ptr(td), n, T.ptr(ids), T.ptr(r),
                               T.ptr(s), T.ptr(t), flags, stream.cuda_stream))

    def counted(flags):
        """algorithmic bytes of one launch and the per-ray work, from an
        untimed counted launch: ray in (24) + hit out (16) + 8 per node
        visited (inner + leaf) + 32 per test + 32 more for full-record tests
        (the records are stored per triangle reference, so a test reads no
        separate reference word)"""
        cnt = np.zeros(8, np.uint64)
        T.check(L.tn_get_counters(tree.handle, cnt.ctypes.data, 1))
        launch(flags | TN_ISECT_COUNT)
        T.check(L.tn_get_counters(tree.handle, cnt.ctypes.data, 1))
        inner, leaves, refs, tests, rays, restarts, full = (int(x) for x in cnt[:7])
        per_ray = {"inner_nodes": round(inner / rays, 2), "leaves": round(leaves / rays, 2),
                   "triangle_tests": round(tests / rays, 2), "full_record_tests": round(full / rays, 2),
                   "restarts": restarts}
        return 40 * n + 8 * (inner + leaves) + 32 * tests + 32 * full, per_ray

    def timed(flags, steps):
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            launch(flags)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps, time.perf_counter() - w0

    def traffic_of(fname):
        """HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
        runs (calibrated against a copy kernel), recorded with the profile
        summary"""
        tp = os.path.join(ROOT, "profiles", fname)
        if os.path.exists(tp) and args.level == 8 and args.rays_log2 == 24:
            with open(tp) as f:
                for line in f:
                    if line.startswith("hbm_bytes_per_launch:"):
                        return int(line.split(":")[1])
        return None

    if args.pmc_probe:
        launch()
        torch.cuda.synchronize()
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": "done", "copy_bytes": src.numel() * 4}), flush=True)
        return

    # the reference tree's own traversal: the closest hits every refined
    # launch must reproduce bit for bit
    launch(0)
    ref_hits = [x.clone() for x in (ids, r, s, t)]
    for _ in range(args.warmup):
        launch(0)
    ref_bytes, ref_per_ray = counted(0)
    ref_ms, _ = timed(0, args.steps)
    for _ in range(args.warmup):
        launch()
    alg_bytes, per_ray = counted(TN_ISECT_REFINED)
    ms_per_step, wall = timed(TN_ISECT_REFINED, args.steps)
    value = n / (ms_per_step * 1e-3) / 1e6
    # ids, r, s, t of the last timed launch == the reference traversal's, bit for bit
    exact = all(torch.equal(a.view(torch.int32), b.view(torch.int32)) for a, b in zip((ids, r, s, t), ref_hits))
    assert exact, "refined traversal differs from the reference tree's"
    # the cube is closed: every ray must hit
    hid = ids.cpu().numpy().view(np.uint32)
    assert np.all(hid != T.TN_INVALID_ID), "a ray escaped the closed scene"
    achieved = alg_bytes / (ms_per_step * 1e-3) / 1e9
    ref_achieved = ref_bytes / (ref_ms * 1e-3) / 1e9
    note(f"intersect: refined {value:.1f} Mrays/s, reference tree {n / (ref_ms * 1e-3) / 1e6:.1f} Mrays/s, "
         f"bit-exact {exact}")
    traffic = traffic_of("traffic_intersect_refined_level8_2p24.txt")
    ref_traffic = traffic_of("traffic_intersect_level8_2p24.txt")
    info = tree.info

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("ray-batch closest hit, level-%d displaced icosphere + 12-tri cube "
                         "(%d triangles), reference SAH kd-tree with its leaves refined into sub-trees "
                         "(TN_SCENE_REFINED), 2^%d rays; ids, r, s, t of every timed launch equal to "
                         "the reference tree's traversal bit for bit (checked on the full batch)"
                         % (args.level, tree.num_triangles(), args.rays_log2)),
            "triangles": tree.num_triangles(),
            "rays_per_step": n,
            "kdtree_nodes": tree.num_nodes(),
            "kdtree_height": tree.height(),
            "refined_nodes": info.refined_nodes,
            "refined_refs": info.refined_refs,
            "host_build_s": round(build_s, 2),
            "device": f"{name}, {cus} CUs",
            "per_ray": per_ray,
            "bit_exact_vs_reference_traversal": exact,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": ("HBM bytes per launch of this workload from rocprofv3 FETCH_SIZE and WRITE_SIZE "
                               "runs (calibrated on a 1 GiB copy), read from "
                               "profiles/traffic_intersect_refined_level8_2p24.txt; not measured in this run"
                               if traffic is not None else None),
            "algorithmic_bytes_per_launch": alg_bytes,
            "kernel": "k_rays_coop<closest hit, refined tree> (tn_intersect, TN_ISECT_REFINED)",
            "note": ("the refined tree's traversal reads %.0f algorithmic bytes per ray against the reference "
                     "tree's %.0f (reference_tree_traversal), so it reaches a lower fraction of the HBM roofline "
                     "at %.2fx the rays/s: its launch is bound by the latency of dependent node fetches"
                     % (alg_bytes / n, ref_bytes / n, value / (n / (ref_ms * 1e-3) / 1e6))),
        },
        "reference_tree_traversal": {
            "value": round(n / (ref_ms * 1e-3) / 1e6, 2), "unit": "Mrays/s", "ms_per_step": round(ref_ms, 4),
            "per_ray": ref_per_ray,
            "roofline": {"bound": "hbm", "achieved": round(ref_achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ref_achieved / HBM_PEAK_GBS, 4), "traffic": ref_traffic,
                         "algorithmic_bytes_per_launch": ref_bytes,
                         "kernel": "k_rays_coop<closest hit> (tn_intersect, default flags)"},
        },
        "wall_s_timed_region": round(wall, 4),
    }

    if not args.no_cpu_baseline:
        import oracle as O
        m = 1 << args.cpu_sample_log2
        cpu = host_cpu()
        ncores = cpu["threads"]
        pre = tree.triangle_records()
        nodes = tree.flat_nodes()
        box = tree.box()
        c0 = time.perf_counter()
        oid, _, _, _ = O.intersect(pre, nodes, box, o[:m], d[:m], nthreads=ncores)
        csec = time.perf_counter() - c0
        assert np.array_equal(oid, hid[:m]), "GPU ids differ from the oracle on the CPU sample"
        out["cpu_baseline"] = {
            "value": round(m / csec / 1e6, 4),
            "unit": "Mrays/s",
            "cores": ncores,
            "kind": "port",
            "sample": (f"first 2^{args.cpu_sample_log2} of the 2^{args.rays_log2} rays, "
                       f"oracle restatement of KDTreeIntersection::intersect (exhaustive), "
                       f"{ncores} threads, {csec:.1f} s; GPU ids identical on the sample"),
            "host": cpu,
        }
        note(f"intersect CPU baseline: {out['cpu_baseline']['value']} Mrays/s on {ncores} threads")
    if not args.no_render:
        out["render"] = render_throughput(args)
    if not args.no_batch:
        out["batch"] = batch_throughput(args)
        note("batch done")
    if not args.no_guided:
        out["guided"] = guided_vs_plain(args)
        note("guided done")
    print(json.dumps(out), flush=True)


def batch_throughput(args):
    """Batched heterogeneous rendering: 32 scenes in one packed forest and one
    launch sequence per chunk.  Scene 0 is the white furnace (enclosure
    emitting 1, albedo-0.8 sphere); scenes 1..31 are seeded displaced
    icospheres of subdivision level uniform in 3..8 (1,280 to 1.31M
    triangles, log-uniform) in the Cornell enclosure with a ceiling emitter;
    square images of width uniform in 128..1024, 16..64 pixel samples,
    max-depth 3, one path per sample, cosine + area-light NEE; chunk size and
    grid from the autotuner (probe excluded from the timed render)."""
    import numpy as np
    import torch

    import turner_amd as T
    from turner_amd import scenes
    from turner_amd._lib import TN_RENDER_AREA_LIGHTS, TN_RENDER_AUTOTUNE, TN_RENDER_COSINE, TN_RENDER_GRAPH
    from turner_amd.batch import Batch, heterogeneous_workload
    scs, levels, jobs = heterogeneous_workload(seed=2024)
    t0 = time.perf_counter()
    trees = [T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], empty_leaves=True) for sc in scs]
    build_s = time.perf_counter() - t0
    b = Batch(trees)
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS | TN_RENDER_GRAPH | TN_RENDER_AUTOTUNE
    # tuning call (probes + one full render), then the timed render with the cached shape
    b.render(jobs, max_depth=3, mc_samples=1, flags=flags, seed=1)
    tune = b.tuning()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    imgs, st, _ = b.render(jobs, max_depth=3, mc_samples=1, flags=flags, seed=2, sample_offset=1 << 20)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    # furnace check: pixels whose four corner rays hit the sphere see rho * L = 0.8,
    # pixels whose corners all miss it see the enclosure, exactly 1
    f = imgs[0].cpu().numpy()
    w0 = jobs[0]["width"]
    cam = scs[0]["camera"]
    ys, xs = np.mgrid[0:w0 + 1, 0:w0 + 1]
    cx = -cam["delta_x"] * (1 - 2 * xs / w0)
    cy = cam["delta_y"] * (1 - 2 * ys / w0)
    d = (np.stack([cx, cy, -np.ones_like(cx)], -1).reshape(-1, 3) @ np.asarray(cam["trafo"]).T).astype(np.float32)
    o = np.tile(np.asarray(cam["pos"], np.float32), (len(d), 1))
    ids = T.KDTreeIntersection(trees[0]).intersect_batch(o, d)[0].reshape(w0 + 1, w0 + 1)
    nsph = 20 * 4 ** 4
    on = (ids[:-1, :-1] < nsph) & (ids[1:, :-1] < nsph) & (ids[:-1, 1:] < nsph) & (ids[1:, 1:] < nsph)
    off = (ids[:-1, :-1] >= nsph) & (ids[1:, :-1] >= nsph) & (ids[:-1, 1:] >= nsph) & (ids[1:, 1:] >= nsph)
    sph = f[on][:, 0]
    samples = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    b.close()
    return {
        "workload": ("32 scenes in one batch: white furnace + 31 displaced icospheres (levels 3..8) in "
                     "Cornell enclosures, widths 128..1024, 16..64 spp, max-depth 3, m 1, cosine + "
                     "area-light NEE, rendering kd-trees, autotuned shape: "
                     + ("graph-replayed chunks, one pipeline" if tune["graph"] else
                        "direct launches, two pipelines (faster than graph replay on this batch)")),
        "scenes": len(scs), "levels": levels, "triangles": int(sum(len(sc["verts"]) for sc in scs)),
        "samples": samples, "host_build_s": round(build_s, 2), "seconds": round(sec, 3),
        "closest_hit_mrays_s": round(st["rays"] / sec / 1e6, 1),
        "all_query_mrays_s": round((st["rays"] + st["shadow_rays"]) / sec / 1e6, 1),
        "tuning": {k: (round(v, 2) if isinstance(v, float) else v) for k, v in tune.items()},
        "furnace": {"sphere_mean": round(float(sph.mean()), 5),
                    "sphere_sem": round(float(sph.std() / np.sqrt(max(sph.size, 1))), 5),
                    "expected": 0.8, "sphere_pixels": int(on.sum()), "spp": jobs[0]["spp"],
                    "enclosure_max_abs_err": round(float(np.abs(f[off][:, :3] - 1.0).max()), 7)},
    }


def guide_pass_probe():
    """Six 1-spp training passes (2^18 records each, 5 epochs of 16,384-sample
    Adam steps) and one image pass of the guided tracer on the occluded-light
    scene, 512x512, depth 6, rendering kd-tree: wall time per call (torch-
    synchronised) with the records trained at the end of each call and with
    TN_RENDER_DEFER_TRAINING (trained on the guide's stream beside the next
    call's render).  Two rounds each; the first call of a round allocates."""
    import torch

    import turner_amd as T
    from turner_amd import scenes
    from turner_amd._lib import TN_RENDER_AREA_LIGHTS, TN_RENDER_COSINE
    from turner_amd.guide import Guide, render_guided
    sc = scenes.occluded_light_cornell(level=8)
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], empty_leaves=True)
    w, depth = 512, 6
    conf = T.TracerConfig(width=w, max_recursion_depth=depth, num_pixel_samples=1, num_monte_carlo_samples=1,
                          flags=TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS, seed=12)
    out = {}
    for rnd in range(2):
        for defer in (False, True):
            g = Guide.for_tree(tree, seed=2, epochs=5, mis_guided=0.7)
            acc = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
            ms = []
            for k in range(7):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                render_guided(tree, sc["camera"], sc["lights"], conf, g, acc,
                              train_samples=(1 << 18) if k < 6 else 0, sample_offset=k, defer=defer)
                torch.cuda.synchronize()
                ms.append(round((time.perf_counter() - t0) * 1e3, 2))
            g.close()
            out[f"{'deferred' if defer else 'in_order'}_round{rnd}"] = {"ms_per_call": ms, "sum_ms": round(sum(ms), 2)}
    note("guide pass probe done")
    return {"workload": "occluded-light Cornell box + level-8 icosphere, 512x512, depth 6, cosine + area-light "
                        "NEE, guide share 0.7; calls 0-5 record 2^18 samples each and train them (5 epochs, "
                        "batch 16384), call 6 is an image pass (it trains the last set when deferred)",
            "calls": out}


def launch_probe():
    """Launch forms of a batch render at fixed chunk sizes: direct launches
    (two pipelines, shadow tests on side streams), graph replay of the
    per-chunk sequence (one pipeline) and TN_RENDER_PERSISTENT (one
    cooperative launch walking every chunk and level behind grid barriers).
    Best wall time of a few renders after one untimed render per form; the
    segment and shadow-ray counts of the forms must agree."""
    import torch

    import turner_amd as T
    from turner_amd import scenes
    from turner_amd._lib import (TN_RENDER_AREA_LIGHTS, TN_RENDER_COSINE, TN_RENDER_GRAPH,
                                 TN_RENDER_PERSISTENT)
    from turner_amd.batch import Batch, heterogeneous_workload
    base = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    forms = (("direct", 0), ("graph", TN_RENDER_GRAPH), ("persistent", TN_RENDER_PERSISTENT))

    def timed(b, jobs, chunk, reps):
        out, counts = {}, set()
        for name, f in forms:
            b.render(jobs, max_depth=3, mc_samples=1, flags=base | f, seed=1, chunk_paths=chunk)
            best = float("inf")
            for r in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                _, st, _ = b.render(jobs, max_depth=3, mc_samples=1, flags=base | f, seed=2 + r, chunk_paths=chunk)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
                if r == 0:
                    counts.add((st["rays"], st["shadow_rays"]))
            out[name] = {"ms": round(best * 1e3, 2), "closest_hit_mrays_s": round(st["rays"] / best / 1e6, 1)}
        assert len(counts) == 1, counts  # same paths whatever the launch form
        return out

    res = {}
    scs, _, jobs = heterogeneous_workload(seed=2024)
    trees = [T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], empty_leaves=True) for sc in scs]
    b = Batch(trees)
    note("launch probe: 32-scene batch")
    res["batch_32_scenes_chunk_2p22"] = timed(b, jobs, 1 << 22, 2)
    b.close()
    small = [scenes.cornell_box(), scenes.sphere_in_cornell(level=5), scenes.furnace(level=4),
             scenes.occluded_light_cornell(level=5)]
    trees = [T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], empty_leaves=True) for sc in small]
    sjobs = [{"camera": sc["camera"], "lights": sc["lights"], "width": 128, "height": 128, "spp": 4} for sc in small]
    b = Batch(trees)
    for c in (14, 16, 18):
        note(f"launch probe: 4 scenes of 128x128x4, chunk 2^{c}")
        res[f"batch_4_small_scenes_chunk_2p{c}"] = timed(b, sjobs, 1 << c, 5)
    b.close()
    return {"workload": "max-depth 3, m 1, cosine + area-light NEE, rendering kd-trees; 32 scenes = the batch "
                        "section's heterogeneous workload (seed 2024); 4 small scenes = Cornell box, sphere in "
                        "Cornell (level 5), furnace (level 4), occluded light (level 5) at 128x128, 4 spp",
            "forms": res}


def tree_probe():
    """tn_intersect / tn_occluded on render-like batches of the occluded-light
    scene (512x512 at 16 spp): camera rays, shadow rays from the camera hits to
    uniform emitter points, cosine bounces from the camera hits; reference kd-
    tree vs rendering kd-tree; HIP-event times and per-ray work counters."""
    import numpy as np
    import torch

    import turner_amd as T
    from turner_amd import scenes
    from turner_amd._lib import TN_ISECT_COUNT
    sc = scenes.occluded_light_cornell(level=8)
    rng = np.random.default_rng(1)
    w, spp = 512, 16
    cam = sc["camera"]
    ys, xs = np.mgrid[0:w, 0:w]
    xs = np.repeat(xs.ravel(), spp) + rng.uniform(size=w * w * spp)
    ys = np.repeat(ys.ravel(), spp) + rng.uniform(size=w * w * spp)
    cx = -cam["delta_x"] * (1 - 2 * xs / w)
    cy = cam["delta_y"] * (1 - 2 * ys / w)
    cd = (np.stack([cx, cy, -np.ones_like(cx)], -1) @ np.asarray(cam["trafo"]).T).astype(np.float32)
    co = np.tile(np.asarray(cam["pos"], np.float32), (len(cd), 1))
    verts = sc["verts"].reshape(-1, 3, 3)
    L = T.lib()
    out = {}
    batches = None
    for label, el in (("reference", False), ("rendering", True)):
        tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], empty_leaves=el)

        def run(o, d, tmax=None):
            n = len(o)
            to, td = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
            ids = torch.empty(n, dtype=torch.int32, device="cuda")
            r, s_, t_ = (torch.empty(n, device="cuda") for _ in range(3))
            tm = torch.from_numpy(tmax).cuda() if tmax is not None else None
            occ = torch.empty(n, dtype=torch.uint8, device="cuda")
            st = torch.cuda.current_stream().cuda_stream

            def go(f):
                if tm is None:
                    T.check(L.tn_intersect(tree.handle, T.ptr(to), T.ptr(td), n, T.ptr(ids), T.ptr(r), T.ptr(s_),
                                           T.ptr(t_), f, st))
                else:
                    T.check(L.tn_occluded(tree.handle, T.ptr(to), T.ptr(td), T.ptr(tm), n, T.ptr(occ), f, st))
            go(0)
            cnt = np.zeros(8, np.uint64)
            T.check(L.tn_get_counters(tree.handle, cnt.ctypes.data, 1))
            go(TN_ISECT_COUNT)
            torch.cuda.synchronize()
            T.check(L.tn_get_counters(tree.handle, cnt.ctypes.data, 1))
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            go(0)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            # the first 2^18 rays alone: a launch that small is bound by its slowest rays
            ns = min(n, 1 << 18)
            ts, ds_ = to[:ns].contiguous(), td[:ns].contiguous()
            tms = tm[:ns].contiguous() if tm is not None else None

            def go_small():
                if tm is None:
                    T.check(L.tn_intersect(tree.handle, T.ptr(ts), T.ptr(ds_), ns, T.ptr(ids), T.ptr(r),
                                           T.ptr(s_), T.ptr(t_), 0, st))
                else:
                    T.check(L.tn_occluded(tree.handle, T.ptr(ts), T.ptr(ds_), T.ptr(tms), ns, T.ptr(occ), 0, st))
            go_small()
            e0.record()
            go_small()
            e1.record()
            torch.cuda.synchronize()
            ms_small = e0.elapsed_time(e1)
            go(0)  # restore the full batch's outputs
            torch.cuda.synchronize()
            c = cnt.astype(np.float64) / n
            res = {"n": n, "ms": round(ms, 3), "mrays_s": round(n / ms / 1e3, 1),
                   "first_2p18_ms": round(ms_small, 3)}
            if tm is None:  # tn_occluded keeps no counters
                res.update({"inner_per_ray": round(c[0], 2), "leaves_per_ray": round(c[1], 2),
                            "tests_per_ray": round(c[3], 2), "max_leaves_one_ray": int(cnt[7])})
            return res, ids.cpu().numpy().view(np.uint32), r.cpu().numpy()

        res_cam, ids, r = run(co, cd)
        if batches is None:  # secondary batches from the reference tree's camera hits
            hit = ids != T.TN_INVALID_ID
            p = co[hit] + r[hit, None] * cd[hit]
            v = verts[ids[hit]]
            nrm = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
            nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
            nrm *= -np.sign(np.sum(nrm * cd[hit], axis=1))[:, None]
            p2 = (p + 1e-4 * nrm).astype(np.float32)
            q = np.stack([0.35 + 0.3 * rng.uniform(size=len(p)), np.full(len(p), 0.999),
                          0.35 + 0.3 * rng.uniform(size=len(p))], 1)
            ld = q - p2
            dist = np.linalg.norm(ld, axis=1)
            ld /= dist[:, None]
            u1, u2 = rng.uniform(size=len(p)), rng.uniform(size=len(p))
            a = np.where(np.abs(nrm[:, 0:1]) > 0.5, np.array([[0.0, 1, 0]]), np.array([[1.0, 0, 0]]))
            b1 = np.cross(nrm, a)
            b1 /= np.linalg.norm(b1, axis=1, keepdims=True)
            b2 = np.cross(nrm, b1)
            rr, ph = np.sqrt(u1), 2 * np.pi * u2
            bd = (rr * np.cos(ph))[:, None] * b1 + (rr * np.sin(ph))[:, None] * b2 + np.sqrt(1 - u1)[:, None] * nrm
            batches = (p2, ld.astype(np.float32), (dist * (1 - 1e-4)).astype(np.float32), bd.astype(np.float32))
        p2, ld, tmx, bd = batches
        res_sh, _, _ = run(p2, ld, tmx)
        res_b, _, _ = run(p2, bd)
        out[label] = {"camera": res_cam, "shadow_any_hit": res_sh, "bounce": res_b,
                      "device_nodes": tree.num_nodes(), "height": tree.height()}
        tree.close()
    return out


def mlp_probe():
    """Guide MLP kernels alone: forward over 2^20 points, training steps of
    2^16 samples, timed with HIP events on the launch stream."""
    import torch

    from turner_amd.guide import Guide
    g = Guide([0.0, 0.0, 0.0], [1.0, 1.0, 1.0], seed=1)
    n, nb = 1 << 20, 1 << 16
    gen = torch.Generator(device="cuda").manual_seed(5)
    p = torch.rand((n, 3), device="cuda", generator=gen)
    nrm = torch.nn.functional.normalize(torch.randn((n, 3), device="cuda", generator=gen), dim=1)
    w = torch.nn.functional.normalize(torch.randn((n, 3), device="cuda", generator=gen), dim=1)
    c = torch.rand(n, device="cuda", generator=gen)
    stream = torch.cuda.current_stream()
    out = {}
    for name, fn, flop in (("forward", lambda: g.forward(p, nrm), 31488.0 * n),
                           ("train_step", lambda: g.train_step(p[:nb], nrm[:nb], w[:nb], c[:nb], want_loss=False),
                            3 * 31488.0 * nb)):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(3):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 3
        out[name] = {"ms": round(ms, 4), "tflops": round(flop / (ms * 1e-3) / 1e12, 2)}
    g.close()
    return out


def render_throughput(args):
    """Plain wavefront path tracing at scale: the displaced level-8 icosphere
    (albedo 0.8) in the Cornell enclosure with a 0.2 x 0.2 emitter of
    radiance 15; 1920 x 1080, max-depth 3, monte-carlo-samples 4, cosine
    sampling + area-light next-event estimation.  Times --render-spp pixel
    samples per render, --render-repeats renders, with HIP events on the
    render's stream, on the rendering kd-tree, on the refined reference tree
    and on the reference kd-tree (the last two give the reference's closest
    hits bit for bit).  The CPU leg runs
    the oracle's trace() over the reference tree on every 68th image row (16
    rows, 2 pixel samples) on the host cores and extrapolates to the full
    render."""
    import types

    import numpy as np
    import torch

    import turner_amd as T
    from turner_amd import scenes
    from turner_amd._lib import TN_RENDER_AREA_LIGHTS, TN_RENDER_COSINE
    sc = scenes.sphere_in_cornell(level=8)
    w, h = 1920, 1080
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    stream = torch.cuda.current_stream()
    out = {"workload": ("wavefront path tracing, level-8 icosphere (albedo 0.8) in the Cornell enclosure, "
                        "0.2x0.2 emitter L=15, 1920x1080, max-depth 3, m 4, cosine + area-light NEE, "
                        f"{args.render_spp} pixel samples per timed render, {args.render_repeats} renders")}
    keep = None
    ref_tree = None
    hits = {"rendering_tree": ("same distances except where a clipped triangle sliver is missing from a "
                               "leaf (of order 1e-7 of the rays); ids may differ on exact ties"),
            "refined_reference_tree": ("bit-exact ids vs the reference: the reference tree with its leaves refined "
                                       "into sub-trees, exact ties and rays with a zero direction component "
                                       "re-traced on the reference tree"),
            "reference_tree": "bit-exact ids vs the reference"}
    for label, kw in (("rendering_tree", {"empty_leaves": True}), ("refined_reference_tree", {"refined": True}),
                      ("reference_tree", {})):
        el = label == "rendering_tree"
        t0 = time.perf_counter()
        tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], **kw)
        build_s = time.perf_counter() - t0
        conf = T.TracerConfig(width=w, aspect=w / h, max_recursion_depth=3, num_pixel_samples=1,
                              num_monte_carlo_samples=4, flags=flags, seed=1)
        acc = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
        T.render_accumulate(tree, sc["camera"], sc["lights"], conf, acc)  # warm-up (queue allocation)
        conf.num_pixel_samples = args.render_spp
        rates, walls, rays, shadow = [], [], 0, 0
        for rep in range(args.render_repeats):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            w0 = time.perf_counter()
            e0.record(stream)
            st = T.render_accumulate(tree, sc["camera"], sc["lights"], conf, acc,
                                     sample_offset=1 + rep * args.render_spp)
            e1.record(stream)
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - w0)
            ms = e0.elapsed_time(e1)
            rates.append((st["rays"] / (ms * 1e-3) / 1e6, (st["rays"] + st["shadow_rays"]) / (ms * 1e-3) / 1e6, ms))
            rays, shadow = st["rays"], st["shadow_rays"]
        med = sorted(rates)[len(rates) // 2]
        samples = w * h * args.render_spp
        out[label] = {
            "closest_hit_mrays_s": round(med[0], 1), "all_query_mrays_s": round(med[1], 1),
            "closest_hit_mrays_s_each": [round(r[0], 1) for r in rates],
            "ms_per_render": round(med[2], 2), "wall_s_each": [round(x, 3) for x in walls],
            "rays_per_render": rays, "shadow_rays_per_render": shadow,
            "full_128spp_render_s_extrapolated": round(med[2] * 1e-3 * 128 * w * h / samples, 2),
            "host_build_s": round(build_s, 2), "device_nodes": tree.num_nodes(),
            "hits": hits[label],
        }
        if label == "refined_reference_tree":
            out[label]["refined_nodes"] = tree.info.refined_nodes
            out[label]["refined_refs"] = tree.info.refined_refs
        note(f"render {label}: {out[label]['closest_hit_mrays_s']} M closest-hit rays/s")
        if el:
            keep = tree
        elif label == "reference_tree":
            ref_tree = tree
        else:
            tree.close()
    # compatibility keys: the bit-exact (refined reference tree) numbers
    out["closest_hit_mrays_s"] = out["refined_reference_tree"]["closest_hit_mrays_s"]
    out["all_query_mrays_s"] = out["refined_reference_tree"]["all_query_mrays_s"]
    if not args.no_cpu_baseline:
        import oracle as O
        cpu = host_cpu()
        rows, spp_c = 16, 2
        step = h // rows + (1 if h % rows else 0)  # 68: rows 0, 68, ..., 1020
        ot = types.SimpleNamespace(pre=ref_tree.triangle_records(), nodes=ref_tree.flat_nodes(), box=ref_tree.box())
        c0 = time.perf_counter()
        _, _, nr, ns = O.render(sc, ot, w, h, 3, 4, spp_c, flags=O.FLAG_COSINE | O.FLAG_AREA_LIGHTS, seed=3,
                                row_step=step, row_offset=0, nthreads=min(cpu["threads"], rows))
        csec = time.perf_counter() - c0
        nrows = len(range(0, h, step))
        scale = (h / nrows) * (128 / spp_c)
        out["cpu_baseline"] = {
            "rays_per_s": round(nr / csec, 1), "all_query_per_s": round((nr + ns) / csec, 1),
            "cores": min(cpu["threads"], rows), "kind": "port", "host": cpu,
            "sample": (f"rows 0, {step}, ... ({nrows} of {h}) at {spp_c} spp of the same render, oracle "
                       f"restatement of trace() over the reference kd-tree (exhaustive traversal), "
                       f"{csec:.1f} s; the full 1920x1080x128 render extrapolates linearly to "
                       f"{csec * scale / 3600:.1f} h"),
            "full_128spp_render_s_extrapolated": round(csec * scale, 1),
        }
        out["gpu_over_cpu_reference_tree"] = round(out["reference_tree"]["closest_hit_mrays_s"] * 1e6 /
                                                   out["cpu_baseline"]["rays_per_s"], 1)
    ref_tree.close()
    out["cornell_320"] = cornell_gpu_vs_cpu(args)
    out["plugins"] = plugin_throughput(args, sc, keep)
    return out


def plugin_throughput(args, sc, tree):
    """The reference's other trace() plugins on the same mesh (rendering
    kd-tree), 1920 x 1080, 16 pixel samples: the raycaster (rays = hits, as
    raycaster.cpp:17 counts) and the raytracer at max-depth 3 with a point
    light and the icosphere a half mirror (rays = trace() calls, as
    raytracer.cpp:9 counts; shadow rays on top)."""
    import numpy as np
    import torch

    import turner_amd as T
    n = len(sc["verts"])
    tree.set_reflection(np.tile(np.array([0.9, 0.9, 0.9, 1.0], np.float32), (n, 1)),
                        np.where(np.arange(n) >= 10, 0.5, 0.0).astype(np.float32))
    lights = [((0.5, 0.9, 0.3), (1.0, 1.0, 1.0, 1.0))]
    w, h = 1920, 1080
    acc = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
    out = {}
    for name, renderer in (("raycaster", T.TN_RENDERER_RAYCASTER), ("raytracer", T.TN_RENDERER_RAYTRACER)):
        conf = T.TracerConfig(width=w, aspect=w / h, max_recursion_depth=3, num_pixel_samples=1,
                              num_monte_carlo_samples=1, renderer=renderer, max_visibility=3.0)
        T.render_accumulate(tree, sc["camera"], lights, conf, acc)  # warm-up
        conf.num_pixel_samples = 16
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = T.render_accumulate(tree, sc["camera"], lights, conf, acc, sample_offset=1)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        out[name] = {"mrays_s": round(st["rays"] / sec / 1e6, 1),
                     "all_query_mrays_s": round((st["rays"] + st["shadow_rays"]) / sec / 1e6, 1),
                     "rays": st["rays"], "shadow_rays": st["shadow_rays"], "seconds": round(sec, 4)}
    out["workload"] = ("raycaster and raytracer (max-depth 3, point light, icosphere reflectivity 0.5) on the "
                       "level-8 icosphere in the Cornell enclosure, 1920x1080, 16 pixel samples, rendering kd-tree")
    tree.set_reflection(np.zeros((n, 4), np.float32), np.zeros(n, np.float32))
    return out


def cornell_gpu_vs_cpu(args):
    """The reference's own CPU-runnable case: the 36-triangle Cornell box,
    320x320, --max-depth 3, --monte-carlo-samples 1, --pixel-samples 8, the
    reference estimator (uniform hemisphere, point light) with the reference's
    pixel jitter; GPU render vs the oracle's restatement of trace() on the
    host cores (the full render; reference rays/s = closest-hit segments per
    second as lib/output.h:101-112 prints)."""
    import torch

    import oracle as O
    import turner_amd as T
    from turner_amd import scenes
    from turner_amd._lib import TN_RENDER_REF_JITTER
    sc = scenes.cornell_box()
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])
    conf = T.TracerConfig(width=320, max_recursion_depth=3, num_pixel_samples=8, num_monte_carlo_samples=1,
                          flags=TN_RENDER_REF_JITTER)
    acc = torch.zeros((320, 320, 4), dtype=torch.float32, device="cuda")
    T.render_accumulate(tree, sc["camera"], sc["lights"], conf, acc)  # warm-up
    best = None
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = T.render_accumulate(tree, sc["camera"], sc["lights"], conf, acc)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        best = sec if best is None else min(best, sec)
    out = {"gpu_seconds": round(best, 4), "gpu_rays_per_s": round(st["rays"] / best, 1),
           "gpu_all_query_per_s": round((st["rays"] + st["shadow_rays"]) / best, 1)}
    if not args.no_cpu_baseline:
        cpu = host_cpu()
        ncores = cpu["threads"]
        ot = O.Tree(sc["verts"])
        t0 = time.perf_counter()
        _, _, nr, ns = O.render(sc, ot, 320, 320, 3, 1, 8, flags=O.FLAG_REF_RNG, nthreads=ncores)
        csec = time.perf_counter() - t0
        out["cpu_baseline"] = {"rays_per_s": round(nr / csec, 1), "all_query_per_s": round((nr + ns) / csec, 1),
                               "cores": ncores, "kind": "port", "host": cpu,
                               "sample": f"the full 320x320x8 render ({csec:.2f} s), oracle restatement of trace()"}
        out["gpu_over_cpu"] = round(out["gpu_rays_per_s"] / out["cpu_baseline"]["rays_per_s"], 1)
    return out


def guided_vs_plain(args):
    """Occluded-light Cornell box (0.3 x 0.3 emitter of radiance 50 behind a
    ceiling baffle, the level-8 displaced icosphere inside), 512 x 512,
    max-depth 6, cosine sampling + area-light NEE: wall time until the
    running image reaches relMSE 0.01 against a --guided-ref-spp reference,
    for the plain tracer and for the learned-guided tracer (online training
    included in its time), plus the MFMA rate of the guide kernels."""
    import ctypes as C

    import numpy as np
    import torch

    import turner_amd as T
    from turner_amd import scenes
    from turner_amd._lib import TN_RENDER_AREA_LIGHTS, TN_RENDER_COSINE
    from turner_amd.guide import Guide, render_guided
    from turner_amd.progressive import time_to_relmse
    from turner_amd.relmse import reference_image
    sc = scenes.occluded_light_cornell(level=8)
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], empty_leaves=True)
    w, depth = args.guided_width, 6
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    t0 = time.perf_counter()
    ref, ref_var = reference_image(tree, sc, w, depth, flags, spp=args.guided_ref_spp, seed=777, with_var=True)
    torch.cuda.synchronize()
    ref_s = time.perf_counter() - t0
    ref_noise = float((ref_var[..., :3].double() / (ref[..., :3].double() ** 2 + 0.01)).mean())
    # warm-up of both paths (allocations), untimed
    g = Guide.for_tree(tree, seed=1)
    tiny = T.TracerConfig(width=w, max_recursion_depth=depth, num_pixel_samples=1, num_monte_carlo_samples=1,
                          flags=flags)
    scratch = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
    T.render_accumulate(tree, sc["camera"], sc["lights"], tiny, scratch)
    render_guided(tree, sc["camera"], sc["lights"], tiny, g, scratch, train_samples=1 << 14)
    g.close()
    plain = time_to_relmse(tree, sc, ref, depth, flags, seed=11, ref_var=ref_var, spp_per_pass=32,
                           min_pass=args.guided_min_pass)
    # guide options and training schedule (swept on this scene over five seed pairs: guide share 0.5..0.9,
    # 4..10 training passes, 4..6 epochs, training in order or deferred beside the next pass's render)
    gopt = {"epochs": 5, "mis_guided": 0.7}
    sched = {"train_passes": 6, "train_spp": 1, "train_samples": 1 << 18, "keep_training_passes": False,
             "defer_training": True}
    g = Guide.for_tree(tree, seed=2, **gopt)
    guided = time_to_relmse(tree, sc, ref, depth, flags, guide=g, seed=12, ref_var=ref_var, spp_per_pass=32,
                            min_pass=args.guided_min_pass, **sched)
    ginfo = g.info()
    # MFMA rate of the guide kernels (HIP events on the launch stream)
    dev = torch.device("cuda:0")
    npts = 1 << 20
    p = torch.rand((npts, 3), device=dev)
    nrm = torch.nn.functional.normalize(torch.randn((npts, 3), device=dev), dim=1)
    wd = torch.nn.functional.normalize(torch.randn((npts, 3), device=dev), dim=1)
    cw = torch.rand(npts, device=dev)
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    g.forward(p, nrm)
    e0.record(stream)
    for _ in range(5):
        g.forward(p, nrm)
    e1.record(stream)
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / 5
    nb = 1 << 16
    g.train_step(p[:nb], nrm[:nb], wd[:nb], cw[:nb], want_loss=False)
    e0.record(stream)
    for _ in range(5):
        g.train_step(p[:nb], nrm[:nb], wd[:nb], cw[:nb], want_loss=False)
    e1.record(stream)
    torch.cuda.synchronize()
    train_ms = e0.elapsed_time(e1) / 5
    g.close()
    fwd_flop = 31488.0  # 2 (78*64 + 64*64 + 64*64 + 64*40) per point
    fwd_tf = fwd_flop * npts / (fwd_ms * 1e-3) / 1e12
    train_tf = 3 * fwd_flop * nb / (train_ms * 1e-3) / 1e12
    ratio = guided["seconds"] / plain["seconds"] if plain["reached"] and guided["reached"] else None
    # four more seed pairs: the ratio moves with the stopping granularity and the training noise
    seeds = [{"plain_s": round(plain["seconds"], 5), "plain_spp": plain["spp"], "guided_s": round(guided["seconds"], 5),
              "guided_spp": guided["spp"], "ratio": None if ratio is None else round(ratio, 4)}]
    for k in (1, 2, 3, 4):
        pl = time_to_relmse(tree, sc, ref, depth, flags, seed=11 + 100 * k, ref_var=ref_var, spp_per_pass=32,
                            min_pass=args.guided_min_pass)
        g2 = Guide.for_tree(tree, seed=2 + 100 * k, **gopt)
        gu = time_to_relmse(tree, sc, ref, depth, flags, guide=g2, seed=12 + 100 * k, ref_var=ref_var,
                            spp_per_pass=32, min_pass=args.guided_min_pass, **sched)
        g2.close()
        rk = gu["seconds"] / pl["seconds"] if pl["reached"] and gu["reached"] else None
        seeds.append({"plain_s": round(pl["seconds"], 5), "plain_spp": pl["spp"], "guided_s": round(gu["seconds"], 5),
                      "guided_spp": gu["spp"], "ratio": None if rk is None else round(rk, 4)})
    rs = [x["ratio"] for x in seeds if x["ratio"] is not None]
    return {
        "workload": (f"occluded-light Cornell box + level-8 icosphere, {w}x{w}, max-depth 6, cosine + "
                     f"area-light NEE, rendering kd-tree, relMSE target 0.01 vs a {args.guided_ref_spp}-spp "
                     f"plain reference ({ref_s:.1f} s; its own per-pixel variance, relMSE "
                     f"{ref_noise:.4f}, is subtracted)"),
        "guided_schedule": ("6 training passes of 1 spp, each training on a uniform 2^18-sample subset of "
                            "its radiance records for 5 epochs (Adam lr 1e-3, batch 16384) beside the next "
                            "training pass's render (deferred training; the last set trains before the image "
                            "passes), then image passes; training passes are timed but left out of the "
                            "image; one-sample MIS "
                            "0.7 guide / 0.3 cosine; both tracers add image passes of 32 spp, each after "
                            "the first sized from the error so far (relMSE ~ 1/spp: the missing samples, "
                            f"rounded up to a multiple of {args.guided_min_pass} spp, at most 32), and sample "
                            "emitters by solid angle"),
        "plain": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in plain.items() if k != "history"},
        "guided": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in guided.items() if k != "history"},
        "guided_time_over_plain": None if ratio is None else round(ratio, 4),
        "guided_time_over_plain_5_seeds": {"each": seeds, "mean": round(float(np.mean(rs)), 4) if rs else None},
        "guide_steps": ginfo["steps"], "guide_last_loss": round(ginfo["loss"], 4),
        "mfma": {"forward_tflops": round(fwd_tf, 2), "train_tflops": round(train_tf, 2),
                 "peak_bf16_tflops": 2500.0, "forward_frac": round(fwd_tf / 2500.0, 5),
                 "train_frac": round(train_tf / 2500.0, 5),
                 "forward_ms_per_2p20_points": round(fwd_ms, 4), "train_ms_per_2p16_batch": round(train_ms, 4)},
    }


if __name__ == "__main__":
    main()
