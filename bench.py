"""Benchmark: batched feasibility checks on one MI355X.

Workload: the 7-DOF arm (512-triangle capsule links) against a ~200k-triangle
clutter soup terrain, 65,536 configurations uniform in the joint limits; one
step = one kg_feasible_batch_dev call (IsFeasible semantics: joint limits,
environment, self collision) over the whole batch with inputs resident in
HBM.  Prints one JSON line.

Also reported: the roofline of the traversal kernel (algorithmic bytes from
its own box/triangle-test counters over its HIP-event duration), and the FP64
CPU path (IsFeasible with Klampt's early returns) timed on the host cores.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# Bytes each test reads (record sizes of the HBM layout): a box test reads two
# 64 B node records; a triangle test reads two triangles of three 16 B
# vertices; a pair item reads its 12 B record plus two 96 B FP64 transforms.
BYTES_PER_BOX_TEST = 128
BYTES_PER_TRI_TEST = 96
BYTES_PER_ITEM = 12 + 192
HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--configs", type=int, default=65536)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: multi-GPU runs are not part of this build "
              "(configurations shard independently; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from klampt_amd import engine, synthetic as S

    t0 = time.time()
    sc = S.soup_scene(0)
    Q = S.uniform_configs(sc.robot, args.configs, seed=0)
    ctx = engine.Context(0)
    world, robot = engine.world_from_scene(ctx, sc)
    print(f"[bench] scene {sc.terrains[0].ntris} tris, {args.configs} configs, setup {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    dQ = torch.from_numpy(Q).to(dev)
    dfeas = torch.empty(args.configs, dtype=torch.uint8, device=dev)
    for _ in range(args.warmup):
        world.feasible_batch_dev(dQ, feasible=dfeas)
    torch.cuda.synchronize()

    trav_ns, redo, boxes, tris, items = [], [], [], [], []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        world.feasible_batch_dev(dQ, feasible=dfeas)
        st = world.last_stats()
        trav_ns.append(st["t_trav32_ns"] + st["t_trav64_ns"])
        redo.append(st["fp64_rechecks"])
        boxes.append(st["box_tests"])
        tris.append(st["tri_tests"])
        items.append(st["pair_items"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed / args.steps * 1e3
    value = args.configs * args.steps / elapsed
    feas_frac = float(dfeas.float().mean().item())
    print(f"[bench] {value:.4g} checks/s, {ms_per_step:.3f} ms/step, feasible {feas_frac:.3f}, "
          f"fp64 rechecks/step {np.mean(redo):.1f}", file=sys.stderr, flush=True)

    kernel_s = float(np.mean(trav_ns)) * 1e-9
    alg_bytes = (float(np.mean(boxes)) * BYTES_PER_BOX_TEST + float(np.mean(tris)) * BYTES_PER_TRI_TEST
                 + float(np.mean(items)) * BYTES_PER_ITEM)
    achieved = alg_bytes / kernel_s / 1e9 if kernel_s > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "k_collide<float>", "kernel_ms": round(kernel_s * 1e3, 4),
                "box_tests": int(np.mean(boxes)), "tri_tests": int(np.mean(tris)),
                "pair_items": int(np.mean(items))}

    cpu = None
    if not args.no_cpu:
        import oracle
        ow, orb = oracle.world_from_scene(sc)
        nthr = oracle.nthreads_default()
        Q64 = Q.astype(np.float64)
        passes, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds or passes == 0:
            t1 = time.perf_counter()
            f_cpu = ow.isfeasible_batch(orb, Q64, nthreads=nthr)
            t_cpu += time.perf_counter() - t1
            passes += 1
        cpu_rate = passes * len(Q) / t_cpu
        agree = bool(np.array_equal(f_cpu, dfeas.cpu().numpy()))
        cpu = {"value": round(cpu_rate, 1), "unit": "checks/s", "cores": nthr, "kind": "port",
               "sample": f"{passes} pass(es) over the same {len(Q)} configurations, FP64 IsFeasible with "
                         f"Klampt's early returns, {nthr} threads, {t_cpu:.1f}s",
               "gpu_over_cpu": round(value / cpu_rate, 2), "results_equal": agree}
        print(f"[bench] cpu {cpu_rate:.4g} checks/s on {nthr} threads; equal={agree}", file=sys.stderr, flush=True)

    line = {
        "metric": "feasibility checks/sec (7-DOF, 200k-tri env)",
        "value": round(value, 1),
        "unit": "checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "7-DOF arm vs 200k-triangle soup terrain, feasibility (IsFeasible)",
                   "configs": args.configs, "scene_tris": int(sc.terrains[0].ntris), "link_tris": 512,
                   "feasible_fraction": round(feas_frac, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
