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
METRIC = "Mrays/s + time-to-relMSE 0.01, guided vs plain tracer, 1 MI355X"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays-log2", type=int, default=24)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-log2", type=int, default=23)
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
    from turner_amd._lib import TN_ISECT_COUNT

    dev = torch.device("cuda:0")
    name, cus = T.device_info()
    sc = scenes.intersect_scene(level=args.level)
    t0 = time.perf_counter()
    tree = T.KDTree(sc["verts"])
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

    def launch(flags=0):
        T.check(L.tn_intersect(tree.handle, T.ptr(to), T.ptr(td), n, T.ptr(ids), T.ptr(r),
                               T.ptr(s), T.ptr(t), flags, stream.cuda_stream))

    if args.pmc_probe:
        launch()
        torch.cuda.synchronize()
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": "done", "copy_bytes": src.numel() * 4}), flush=True)
        return

    for _ in range(args.warmup):
        launch()
    # untimed counted launch: algorithmic bytes of the traversal
    cnt = np.zeros(8, np.uint64)
    T.check(L.tn_get_counters(tree.handle, cnt.ctypes.data, 1))
    launch(TN_ISECT_COUNT)
    T.check(L.tn_get_counters(tree.handle, cnt.ctypes.data, 1))
    inner, leaves, refs, tests, rays, restarts, full = (int(x) for x in cnt[:7])
    # bytes: ray in (24) + hit out (16) + 8 per node visited (inner + leaf)
    # + 4 per triangle reference + 32 per test + 32 more for full-record tests
    alg_bytes = 40 * n + 8 * (inner + leaves) + 4 * refs + 32 * tests + 32 * full

    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        launch()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_per_step = e0.elapsed_time(e1) / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6

    # sanity: the cube is closed, every ray must hit; a sample is bit-exact
    hid = ids.cpu().numpy().view(np.uint32)
    assert np.all(hid != T.TN_INVALID_ID), "a ray escaped the closed scene"
    achieved = alg_bytes / (ms_per_step * 1e-3) / 1e9
    traffic = None
    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # (calibrated against a copy kernel), recorded with the profile summary
    tp = os.path.join(ROOT, "profiles", "traffic_intersect_level8_2p24.txt")
    if os.path.exists(tp) and args.level == 8 and args.rays_log2 == 24:
        with open(tp) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    traffic = int(line.split(":")[1])

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
                         "(%d triangles), reference SAH kd-tree, 2^%d rays, ids bit-exact"
                         % (args.level, tree.num_triangles(), args.rays_log2)),
            "triangles": tree.num_triangles(),
            "rays_per_step": n,
            "kdtree_nodes": tree.num_nodes(),
            "kdtree_height": tree.height(),
            "host_build_s": round(build_s, 2),
            "device": f"{name}, {cus} CUs",
            "per_ray": {"inner_nodes": round(inner / rays, 2), "leaves": round(leaves / rays, 2),
                        "triangle_tests": round(tests / rays, 2),
                        "full_record_tests": round(full / rays, 2),
                        "restarts": restarts},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": alg_bytes,
            "kernel": "k_rays_coop<closest hit> (tn_intersect)",
        },
        "wall_s_timed_region": round(wall, 4),
    }

    if not args.no_cpu_baseline:
        import oracle as O
        m = 1 << args.cpu_sample_log2
        ncores = min(os.cpu_count() or 1, 16)  # the box's CPU share for one GPU
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
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
