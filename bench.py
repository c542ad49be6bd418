"""Benchmark: batched feasibility checks on one MI355X.

Workload: the 7-DOF arm (512-triangle capsule links) against a ~200k-triangle
clutter soup terrain, 65,536 configurations uniform in the joint limits; one
step = one kg_feasible_batch_dev call (IsFeasible semantics: joint limits,
environment, self collision) over the whole batch with inputs resident in
HBM.  Prints one JSON line.

Also reported: the roofline of the traversal kernel -- algorithmic bytes from
the oracle's sequential IsFeasible (early returns, first hit) counted on the
engine's own OBB trees over the same configurations (64 B per environment
node visited, 36 B per environment triangle tested, 4 B per joint value in,
2 B out) over the kernel's average HIP-event duration in the timed steps;
beside it the fraction from the kernel's own box/triangle-test counters, the
HBM traffic fraction (committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary)
and the FP32 flop fraction -- and the FP64 CPU path (IsFeasible with
Klampt's early returns) timed on the host cores, all threads and one thread.

``python bench.py --copy-probe`` only runs a plain 256 MiB device copy five
times: the calibration workload for the FETCH_SIZE / WRITE_SIZE counters.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# Algorithmic bytes of one feasibility check: 4 B per joint value in, 2 B of
# result out, and per candidate pair 64 B per environment OBB node visited
# (one per box test) and 36 B per environment triangle tested (three FP32
# vertices).  Robot-side nodes and triangles are a few hundred KB shared by
# every configuration and are not counted.
BYTES_PER_ENV_NODE = 64
BYTES_PER_ENV_TRI = 36
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
TRAFFIC_FILE = os.path.join(HERE, "profiles", "k_collide_f32_soup200k_65536.traffic.json")


def copy_probe():
    import torch
    x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")   # 256 MiB
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print("copied 256 MiB x5", flush=True)


def host_info():
    """Host cores and CPU model of the machine the CPU baseline ran on."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return info


def fk_pose_error(robot, ow, orb, Q):
    """Largest FP32 forward-kinematics error against the FP64 oracle over Q:
    link origin position (m) and rotation entries (max |R32 - R64|)."""
    import numpy as np
    T32 = robot.fk_batch(Q).astype(np.float64)
    dp = dr = 0.0
    for c in range(Q.shape[0]):
        T64 = ow.fk(orb, Q[c].astype(np.float64))
        dr = max(dr, float(np.abs(T32[c, :, :9] - T64[:, :9]).max()))
        dp = max(dp, float(np.abs(T32[c, :, 9:] - T64[:, 9:]).max()))
    return {"configs": int(Q.shape[0]), "max_position_error_m": dp, "max_rotation_entry_error": dr}


def engine_trees(world, robot):
    """The engine's OBB trees of a world's meshes, in the order the oracle
    numbers them (terrains, rigid objects, then the robot's link meshes)."""
    geoms = list(world._keep) + [g for g in robot.link_geoms if g is not None]
    return [g.bvh() for g in geoms]


def measured_traffic():
    """HBM bytes per launch of the traversal kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return float(t["bytes_per_dispatch"]), os.path.relpath(TRAFFIC_FILE, HERE)
    except (OSError, KeyError, ValueError):
        return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--configs", type=int, default=65536)
    ap.add_argument("--streams", type=int, default=2,
                    help="streams the timed batches alternate over (1: one batch at a time)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--dist-configs", type=int, default=65536)
    ap.add_argument("--edges", type=int, default=8192)
    ap.add_argument("--traj", type=int, default=16384)
    ap.add_argument("--copy-probe", action="store_true")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
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

    # (1) one stream: batch latency, and the traversal kernel's duration per
    # launch (HIP events on the launch stream, launches not overlapping)
    ctx.kernel_times(reset=True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        world.feasible_batch_dev(dQ, feasible=dfeas)
    torch.cuda.synchronize()
    elapsed1 = time.perf_counter() - t_start
    ksum, kcnt = ctx.kernel_times(reset=True)
    kernel_s = ksum / max(1, kcnt)
    single = {"checks_per_s": round(args.configs * args.steps / elapsed1, 1),
              "ms_per_batch": round(elapsed1 / args.steps * 1e3, 4)}
    # (2) the timed region: the same batches issued alternately on S streams
    # (the context runs batches of different streams concurrently, each on
    # its own workspace slot): whole-job throughput
    streams = [torch.cuda.Stream() for _ in range(args.streams)]
    outs = [dfeas] + [torch.empty_like(dfeas) for _ in range(args.streams - 1)]
    for i in range(2 * args.streams):
        with torch.cuda.stream(streams[i % args.streams]):
            world.feasible_batch_dev(dQ, feasible=outs[i % args.streams])
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        with torch.cuda.stream(streams[i % args.streams]):
            world.feasible_batch_dev(dQ, feasible=outs[i % args.streams])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed / args.steps * 1e3
    value = args.configs * args.steps / elapsed
    for o in outs[1:]:
        assert torch.equal(o, dfeas), "batches on different streams disagree"
    # counters are deterministic per batch: read those of the last step
    st = world.last_stats()
    # broad-phase duration from one more (untimed) batch with per-phase events
    ctx.set_timing(2)
    world.feasible_batch_dev(dQ, feasible=dfeas)
    t_prep = world.last_stats()["t_prep_ns"]
    ctx.set_timing(1)
    feas_frac = float(dfeas.float().mean().item())
    print(f"[bench] {value:.4g} checks/s on {args.streams} streams ({ms_per_step:.3f} ms/step); one stream "
          f"{single['checks_per_s']:.4g} checks/s ({single['ms_per_batch']:.3f} ms/batch), traversal kernel "
          f"{kernel_s * 1e3:.4f} ms "
          f"(avg of {kcnt}), broad phase {t_prep / 1e6:.4f} ms, feasible {feas_frac:.3f}, "
          f"fp64 rechecks/step {st['fp64_rechecks']}",
          file=sys.stderr, flush=True)

    nj = sc.robot.n
    # the kernel's own work (its box / triangle-test counters, every box test
    # charged as one environment node)
    gpu_bytes = (args.configs * (4 * nj + 2) + st["box_tests"] * BYTES_PER_ENV_NODE
                 + st["tri_tests"] * BYTES_PER_ENV_TRI)
    traffic, traffic_src = measured_traffic()
    roofline = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "k_collide<float, false, 256, false, true>", "kernel_ms": round(kernel_s * 1e3, 4),
                "kernel_launches_timed": kcnt,
                "gpu_counters": {"box_tests": int(st["box_tests"]), "tri_tests": int(st["tri_tests"]),
                                 "pair_items": int(st["pair_items"]), "fp64_rechecks": int(st["fp64_rechecks"]),
                                 "bytes": int(gpu_bytes),
                                 "frac": round(gpu_bytes / kernel_s / 1e9 / HBM_PEAK_GBS, 5) if kernel_s else None}}
    if traffic and kernel_s:
        roofline["traffic_frac"] = round(traffic / kernel_s / 1e9 / HBM_PEAK_GBS, 5)

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
        # one thread, on the first 8192 configurations
        Q1 = Q64[:8192]
        p1, t1s = 0, 0.0
        while t1s < args.cpu_seconds / 3 or p1 == 0:
            t1 = time.perf_counter()
            ow.isfeasible_batch(orb, Q1, nthreads=1)
            t1s += time.perf_counter() - t1
            p1 += 1
        rate1 = p1 * len(Q1) / t1s
        cpu = {"value": round(cpu_rate, 1), "unit": "checks/s", "cores": nthr, "kind": "port",
               "sample": f"{passes} pass(es) over the same {len(Q)} configurations, FP64 IsFeasible with "
                         f"Klampt's early returns, {nthr} threads, {t_cpu:.1f}s; single thread: {p1} pass(es) "
                         f"over the first {len(Q1)}, {t1s:.1f}s",
               "single_thread_value": round(rate1, 1),
               "gpu_over_cpu": round(value / cpu_rate, 2), "gpu_over_one_thread": round(value / rate1, 1),
               "gpu_over_cpu_one_stream": round(single["checks_per_s"] / cpu_rate, 2),
               "results_equal": agree}
        print(f"[bench] cpu {cpu_rate:.4g} checks/s on {nthr} threads, {rate1:.4g} on one; equal={agree}",
              file=sys.stderr, flush=True)
        # algorithmic bytes and flops: the oracle's sequential IsFeasible (early
        # returns, first hit) on the engine's own OBB trees, counted per check
        for g, tree in zip(ow.meshes, engine_trees(world, robot)):
            ow.set_obb_tree(g, tree)
        f_cnt, cnt = ow.feasible_counted(orb, Q64, nthreads=nthr)
        tot = cnt.sum(axis=0)
        alg_bytes = args.configs * (4 * nj + 2) + int(tot[1]) * BYTES_PER_ENV_NODE + int(tot[2]) * BYTES_PER_ENV_TRI
        checked = int((cnt.sum(axis=1) > 0).sum())
        flops = 600 * args.configs + 200 * int(tot[0] + tot[3]) + 150 * int(tot[2] + tot[4])
        achieved = alg_bytes / kernel_s / 1e9 if kernel_s > 0 else 0.0
        roofline.update({"achieved": round(achieved, 2), "frac": round(achieved / HBM_PEAK_GBS, 5),
                         "algorithmic_bytes": int(alg_bytes),
                         "oracle_counters": {"env_box_tests": int(tot[0]), "env_node_visits": int(tot[1]),
                                             "env_tri_tests": int(tot[2]), "self_box_tests": int(tot[3]),
                                             "self_tri_tests": int(tot[4]), "checks_traversed": checked,
                                             "answers_equal": bool(np.array_equal(f_cnt, f_cpu))},
                         "fp32_flops": int(flops),
                         "fp32_flop_frac": round(flops / kernel_s / 1e12 / FP32_PEAK_TFLOPS, 5) if kernel_s else None})
        print(f"[bench] oracle counters {roofline['oracle_counters']}: {alg_bytes / 1e6:.1f} MB algorithmic, "
              f"{achieved:.1f} GB/s", file=sys.stderr, flush=True)
        # contact band (1e-6 m): booleans that differ from the oracle, those of
        # them in the band, and the band's population in a sample (FP64 signed
        # separations of every pair of the first configurations)
        gpu_f = dfeas.cpu().numpy()
        diff = np.nonzero(gpu_f != f_cpu)[0]
        in_band_diff = 0
        if len(diff):
            sep = ow.distance_batch(orb, Q64[diff], signed=True)
            in_band_diff = int((np.abs(sep).min(axis=1) < 1e-6).sum())
        nb = min(512, len(Q64))
        t1 = time.perf_counter()
        sep = ow.distance_batch(orb, Q64[:nb], signed=True)
        band_sample = int((np.abs(sep).min(axis=1) < 1e-6).sum())
        cpu["contact_band"] = {"band_m": 1e-6, "mismatches": int(len(diff)), "mismatches_in_band": in_band_diff,
                               "fp32_undecided_rechecked": int(st["fp64_rechecks"]),
                               "fp32_undecided_pairs": int(st["exact_pairs"]),
                               "sample_configs": nb, "sample_in_band": band_sample,
                               "sample_seconds": round(time.perf_counter() - t1, 2)}
        cpu["fk_fp32"] = fk_pose_error(robot, ow, orb, Q[:4096])
        cpu["host"] = host_info()
        print(f"[bench] contact band {cpu['contact_band']}, FK {cpu['fk_fp32']}, host {cpu['host']}",
              file=sys.stderr, flush=True)

    secondary = None if args.no_secondary else run_secondary(args, ctx, sc, world, robot)

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
                   "feasible_fraction": round(feas_frac, 4), "streams": args.streams,
                   "one_stream": single},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(line), flush=True)


def _time_dev(fn, reps, warm=1):
    """Seconds per call of fn (device calls on torch's current stream): HIP
    events recorded on that stream around reps calls, after warm warm-up
    calls (enough to touch every workspace slot the calls rotate over)."""
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / reps


def run_secondary(args, ctx, sc, world, robot):
    """Distances (C2), edge checks (C4), SDF penalty gradients (C5),
    heterogeneous batches (C3), planners over the batched queries."""
    import numpy as np
    import torch
    from klampt_amd import engine, synthetic as S
    out = {}
    dev = torch.device("cuda:0")

    # C2 distances: FP32 min distance per link-env and self pair
    Qd = torch.from_numpy(S.uniform_configs(sc.robot, args.dist_configs, seed=3)).to(dev)
    npairs = len(world.pair_list(0))
    dd = torch.empty((args.dist_configs, npairs), dtype=torch.float32, device=dev)
    t = _time_dev(lambda: world.distance_batch_dev(Qd, out=dd), 3)
    out["c2_distance"] = {"configs": args.dist_configs, "pairs_per_config": npairs,
                          "pair_distances_per_s": round(args.dist_configs * npairs / t, 1),
                          "configs_per_s": round(args.dist_configs / t, 1), "ms": round(t * 1e3, 3)}
    if not args.no_cpu:
        import oracle
        ow, orb = oracle.world_from_scene(sc)
        Qc = Qd[:512].cpu().numpy().astype(np.float64)
        t1 = time.perf_counter()
        ow.distance_batch(orb, Qc)
        tc = time.perf_counter() - t1
        out["c2_distance"]["cpu_pair_distances_per_s"] = round(len(Qc) * npairs / tc, 1)
        out["c2_distance"]["cpu_sample"] = f"{len(Qc)} configurations, FP64 oracle, {oracle.nthreads_default()} threads"
    print(f"[bench] C2 distances: {out['c2_distance']}", file=sys.stderr, flush=True)

    # C8 distance Jacobians (KineTrajOpt's point-collision step): the same
    # configurations and pairs, bound dcheck = 0.1 m, closest points and
    # joint-space gradients of every pair within it
    dj = (torch.empty((args.dist_configs, npairs), dtype=torch.float32, device=dev),
          torch.empty((args.dist_configs, npairs, 6), dtype=torch.float32, device=dev),
          torch.empty((args.dist_configs, npairs, sc.robot.n), dtype=torch.float32, device=dev))
    t = _time_dev(lambda: world.distance_jacobian_batch_dev(Qd, 0.1, out=dj), 3)
    within = int(torch.isfinite(dj[0]).sum().item())
    out["c8_distance_jacobian"] = {"configs": args.dist_configs, "pairs_per_config": npairs, "dcheck": 0.1,
                                   "pairs_within": within, "ms": round(t * 1e3, 3),
                                   "pair_queries_per_s": round(args.dist_configs * npairs / t, 1)}
    if not args.no_cpu:
        # the FP64 oracle's pair distances on the same configurations (no
        # bound, no closest points or Jacobians): a lower bound on the CPU cost
        out["c8_distance_jacobian"]["cpu_pair_queries_per_s"] = out["c2_distance"]["cpu_pair_distances_per_s"]
        out["c8_distance_jacobian"]["cpu_sample"] = ("the C2 line's oracle distances (distances only; a lower bound "
                                                     "on the CPU cost of this query)")
    print(f"[bench] C8 distance Jacobians: {out['c8_distance_jacobian']}", file=sys.stderr, flush=True)

    # C4 edges: 8192 segments, L-infinity step 0.01 rad, 10..500 samples;
    # device rounds on endpoints resident in HBM (kg_edge_batch_dev), and the
    # host-pointer call (copies included) beside it
    A, B = S.edge_segments(sc.robot, args.edges, eps=0.01, smin=10, smax=500, seed=8)
    dA, dB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    dfu = torch.empty(args.edges, dtype=torch.float32, device=dev)
    te = _time_dev(lambda: world.edge_batch_dev(dA, dB, eps=0.01, norm_inf=True, out=dfu), 5, warm=2)
    est = ctx.status_dev().cpu().numpy()
    fu = dfu.cpu().numpy()
    world.edge_batch(A, B, eps=0.01, norm_inf=True)   # sizes the context's own workspaces
    t1 = time.perf_counter()
    fh = world.edge_batch(A, B, eps=0.01, norm_inf=True)
    th = time.perf_counter() - t1
    out["c4_edges"] = {"segments": args.edges, "edges_per_s": round(args.edges / te, 1), "ms": round(te * 1e3, 3),
                       "free_fraction": round(float((fu < 0).mean()), 4), "undecided": int(est[1]),
                       "host_call_ms": round(th * 1e3, 3), "host_equal": bool(np.array_equal(fh, fu)),
                       "note": "kg_edge_batch_dev: device rounds, no host synchronization; host_call_ms is the "
                               "host-pointer call with its copies"}
    # whole-job rate with two batches in flight: batches alternating over two
    # streams of the same context run on its two edge workspace slots; the
    # rounds of one batch leave most of the GPU idle between them
    e_streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    e_outs = [torch.empty_like(dfu), torch.empty_like(dfu)]

    def two_in_flight(reps):
        for k in range(reps):
            with torch.cuda.stream(e_streams[k % 2]):
                world.edge_batch_dev(dA, dB, eps=0.01, norm_inf=True, out=e_outs[k % 2])
    two_in_flight(4)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    two_in_flight(20)
    torch.cuda.synchronize()
    t2 = (time.perf_counter() - t1) / 20
    out["c4_edges"]["two_in_flight"] = {
        "ms_per_batch": round(t2 * 1e3, 3), "edges_per_s": round(args.edges / t2, 1),
        "results_equal": bool(all(np.array_equal(o.cpu().numpy(), fu) for o in e_outs)),
        "note": "one context, batches alternating over two streams on its edge workspace slots "
                "(whole-job rate, host clock)"}
    if not args.no_cpu:
        t1 = time.perf_counter()
        fo, _, nchk = ow.edge_batch(orb, A, B, eps=0.01, norm_inf=True)
        tc = time.perf_counter() - t1
        out["c4_edges"]["cpu_edges_per_s"] = round(len(A) / tc, 1)
        out["c4_edges"]["cpu_samples_checked"] = int(nchk.sum())
        out["c4_edges"]["cpu_sample"] = f"all {len(A)} segments, FP64 oracle, {oracle.nthreads_default()} threads"
        out["c4_edges"]["results_equal"] = bool(np.array_equal(fu, fo.astype(np.float32)))
    print(f"[bench] C4 edges: {out['c4_edges']}", file=sys.stderr, flush=True)

    # C5 SDF hinge penalty, forward + backward
    t0 = time.perf_counter()
    P, inside = S.noisy_cloud(0, npoints=1_000_000)
    grid = S.sdf_from_cloud(P, inside, dims=256, voxel=0.01)
    sdf = engine.Geom.sdf(ctx, grid.bmin, grid.bmax, grid.values)
    pts = np.concatenate([S.surface_samples(m, 64, seed=i) for i, m in enumerate(sc.robot.link_meshes)])
    ppl = np.full(sc.robot.n, 64, np.int32)
    Qw = torch.from_numpy(S.trajectories(sc.robot, args.traj, 32, seed=4)).to(dev)
    dpts = torch.from_numpy(pts.astype(np.float32)).to(dev)
    Nw = Qw.shape[0]
    loss = torch.empty(Nw, dtype=torch.float32, device=dev)
    dQ = torch.empty_like(Qw)
    ones = torch.ones(Nw, dtype=torch.float32, device=dev)
    print(f"[bench] SDF grid built in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    def fwd_bwd():
        robot.sdf_penalty_dev(sdf, Qw, dpts, ppl, 0.05, out=loss)
        robot.sdf_penalty_grad_dev(sdf, Qw, dpts, ppl, 0.05, dloss=ones, out=dQ)
    tb = _time_dev(fwd_bwd, 5)
    # plain per-configuration path: one drop-in call per waypoint
    nplain = 512
    q1 = [Qw[i:i + 1] for i in range(nplain)]
    l1 = torch.empty(1, dtype=torch.float32, device=dev)
    g1 = torch.empty((1, sc.robot.n), dtype=torch.float32, device=dev)
    o1 = torch.ones(1, dtype=torch.float32, device=dev)

    def plain():
        for q in q1:
            robot.sdf_penalty_dev(sdf, q, dpts, ppl, 0.05, out=l1)
            robot.sdf_penalty_grad_dev(sdf, q, dpts, ppl, 0.05, dloss=o1, out=g1)
    tp = _time_dev(plain, 1)
    bytes_per_wp = 448 * 8 * 4 * 2 + 28 + 28
    out["c5_sdf_grad"] = {"waypoints": Nw, "evals_per_s": round(Nw / tb, 1), "ms": round(tb * 1e3, 3),
                          "plain_evals_per_s": round(nplain / tp, 1),
                          "batched_over_plain": round((Nw / tb) / (nplain / tp), 1),
                          "active_fraction": round(float((loss > 0).float().mean().item()), 4),
                          "roofline": {"bound": "hbm", "achieved": round(Nw * bytes_per_wp / tb / 1e9, 1),
                                       "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": round(Nw * bytes_per_wp / tb / 1e9 / HBM_PEAK_GBS, 4),
                                       "bytes_per_waypoint": bytes_per_wp}}
    if not args.no_cpu:
        orw = oracle.OracleWorld()
        orr = orw.add_robot(sc.robot)
        g64 = S.SDFGrid(grid.bmin, grid.bmax, grid.values.astype(np.float64))
        Qc = Qw[:4096].cpu().numpy().astype(np.float64)
        t1 = time.perf_counter()
        orw.sdf_penalty(orr, g64, Qc, pts, ppl, dsafe=0.05, grad=True)
        tc = time.perf_counter() - t1
        out["c5_sdf_grad"]["cpu_evals_per_s"] = round(len(Qc) / tc, 1)
        out["c5_sdf_grad"]["cpu_sample"] = (f"first {len(Qc)} waypoints, FP64 oracle loss and analytic gradient, "
                                            f"{oracle.nthreads_default()} threads")
        del g64
    print(f"[bench] C5 SDF: {out['c5_sdf_grad']}", file=sys.stderr, flush=True)

    # PointCloud vs ImplicitSurface distance: the 1M-point cloud, moved by a
    # transform, against the same 256^3 grid (Geometry3D.distance)
    pcg = engine.Geom.pointcloud(ctx, P)
    Rz = np.array([[0.8, -0.6, 0.0], [0.6, 0.8, 0.0], [0.0, 0.0, 1.0]])
    pcg.set_transform(Rz.T.reshape(9), np.array([0.01, -0.02, 0.03]))
    pcg.distance(sdf)
    t0 = time.perf_counter()
    for _ in range(10):
        rpc = pcg.distance(sdf)
    tpc = (time.perf_counter() - t0) / 10
    out["c7_pointcloud_sdf"] = {"points": int(P.shape[0]), "grid": list(grid.values.shape), "ms": round(tpc * 1e3, 3),
                                "points_per_s": round(P.shape[0] / tpc, 1), "min_value": round(float(rpc.d), 6),
                                "note": "host call: one kernel, a 64-bit result copy and a gradient evaluation"}
    if not args.no_cpu:
        g64 = S.SDFGrid(grid.bmin, grid.bmax, grid.values.astype(np.float64))
        sub = P[:200_000] @ Rz.T + np.array([0.01, -0.02, 0.03])
        t1 = time.perf_counter()
        dref, _ = oracle.sdf_eval(g64, sub)
        tc = time.perf_counter() - t1
        out["c7_pointcloud_sdf"]["cpu_points_per_s"] = round(len(sub) / tc, 1)
        out["c7_pointcloud_sdf"]["cpu_sample"] = f"first {len(sub)} points, FP64 oracle trilinear values, 1 thread"
        del g64
    pcg.close()
    print(f"[bench] C7 pointcloud-SDF: {out['c7_pointcloud_sdf']}", file=sys.stderr, flush=True)

    # C3 heterogeneous robots in one launch
    soup_geom = world._keep[0]
    robots, counts = S.hetero_robots(seed=0, nrobots=64)
    worlds = []
    for spec in robots:
        w = engine.World(ctx)
        w.add_terrain(soup_geom)
        w.add_robot(engine.Robot.from_spec(ctx, spec))
        worlds.append(w)
    Qh = torch.from_numpy(np.concatenate([S.uniform_configs(sp, c, seed=20 + i).ravel()
                                          for i, (sp, c) in enumerate(zip(robots, counts))])).to(dev)
    choice = engine.autotune(worlds, [0] * 64, counts, Qh)
    fh = torch.empty(int(sum(counts)), dtype=torch.uint8, device=dev)
    th = _time_dev(lambda: engine.feasible_multi_dev(worlds, [0] * 64, counts, Qh, out=fh), 3)
    out["c3_hetero"] = {"robots": 64, "configs": int(sum(counts)), "checks_per_s": round(sum(counts) / th, 1),
                        "ms": round(th * 1e3, 3), "dof_range": [int(min(r.n for r in robots)), int(max(r.n for r in robots))],
                        "autotune": {"cpw": choice[0], "leaf_min": choice[1], "scap": choice[2],
                                     "waves_per_cu": choice[3]},
                        "feasible_fraction": round(float(fh.float().mean().item()), 4)}
    engine.set_tuning(ctx)
    print(f"[bench] C3 hetero: {out['c3_hetero']}", file=sys.stderr, flush=True)

    # serving form: the C3 batch and the C4 edges as one library-captured HIP
    # graph (kg_serve_*), replayed on resident buffers; equal to the eager calls
    sv = engine.Serve(ctx, feas=[(w, 0, c) for w, c in zip(worlds, counts)], edges=[(world, 0, args.edges)],
                      eps=0.01, norm_inf=True)
    sv.Q.copy_(Qh)
    sv.A.copy_(dA.reshape(-1))
    sv.B.copy_(dB.reshape(-1))
    ts = _time_dev(sv.run, 5)
    out["serve"] = {"feasibility_configs": int(sum(counts)), "robots": 64, "edge_segments": args.edges,
                    "ms": round(ts * 1e3, 3), "graph_captures": sv.captures(),
                    "eager_ms": round((th + te) * 1e3, 3),
                    "results_equal": bool(torch.equal(sv.feasible, fh) and torch.equal(sv.first_u, dfu)),
                    "note": "one kg_serve_run (graph launch) per planner iteration, the feasibility batch and the "
                            "edge rounds as two branches of the graph; eager_ms: the C3 and C4 calls issued one "
                            "after the other"}
    sv.close()
    print(f"[bench] serve: {out['serve']}", file=sys.stderr, flush=True)

    # KineTrajOpt sweep terms: trajectory steps of the arm against the 16 boxes
    scb = S.boxes_scene(0)
    wb, rb = engine.world_from_scene(ctx, scb)
    nseg_sw = 4096
    rng = np.random.default_rng(9)
    Q0 = S.uniform_configs(scb.robot, nseg_sw, seed=9).astype(np.float64)
    stp = rng.normal(size=Q0.shape)
    stp *= rng.uniform(0.02, 0.3, (nseg_sw, 1)) / np.abs(stp).max(axis=1, keepdims=True)
    Q1 = Q0 + stp
    d0, d1 = torch.from_numpy(Q0).to(dev), torch.from_numpy(Q1).to(dev)
    tsw = _time_dev(lambda: wb.sweep_batch_dev(d0, d1, 0.25), 3)
    dsw, _, _ = wb.sweep_batch_dev(d0, d1, 0.25)
    nq = nseg_sw * scb.robot.n * len(scb.objects)
    out["c9_sweep"] = {"segments": nseg_sw, "queries": nq, "ms": round(tsw * 1e3, 3),
                       "queries_per_s": round(nq / tsw, 1),
                       "below_dcheck": int(torch.isfinite(dsw).sum().item()), "dcheck": 0.25}
    if not args.no_cpu:
        import oracle
        ow9, or9 = oracle.world_from_scene(scb)
        t1 = time.perf_counter()
        nsample = 0
        for k in range(64):
            X0, X1 = ow9.fk(or9, Q0[k]), ow9.fk(or9, Q1[k])
            for l in range(scb.robot.n):
                for mesh, xf in scb.objects:
                    oracle.hull_distance(scb.robot.link_meshes[l].verts, X0[l], mesh.verts, xf, xa2=X1[l])
                    nsample += 1
        tc = time.perf_counter() - t1
        out["c9_sweep"]["cpu_queries_per_s"] = round(nsample / tc, 1)
        out["c9_sweep"]["cpu_sample"] = (f"first 64 segments ({nsample} queries), FP64 oracle GJK distance only, "
                                         "one thread, called per query from Python")
    print(f"[bench] C9 sweep terms: {out['c9_sweep']}", file=sys.stderr, flush=True)

    # planners on the batched queries: PRM roadmap over the 16-box scene
    # (every candidate edge validated in batched calls) and a lazy PRM query
    from klampt_amd import plan, robotsim
    bs = S.boxes_scene(0)
    W = robotsim.WorldModel()
    for mesh, xf in bs.objects:
        W.addRigidObject(robotsim.Geometry3D(robotsim.TriangleMesh(mesh.verts, mesh.tris)), xf[:9], xf[9:])
    R = robotsim.RobotModel.from_spec(bs.robot)
    W.addRobot(R)
    space = plan.RobotCSpace(R, W, eps=0.02)
    space.rng = np.random.default_rng(5)
    Qs = S.uniform_configs(bs.robot, 4096, seed=63)
    Qf = Qs[space.feasible_batch(Qs)]
    dq = np.linalg.norm(Qf[:256, None] - Qf[None, :256], axis=-1)
    cand = np.argwhere((dq > 1.5) & (dq < 3.0))
    cand = cand[cand[:, 0] < cand[:, 1]]
    blocked = ~space.visible_batch(Qf[cand[:, 0]], Qf[cand[:, 1]])
    qa, qb = Qf[cand[int(np.argmax(blocked))]].astype(float)
    # warm-up (first use of the nearest-neighbour and edge-check kernels)
    warm = plan.MotionPlan(space, type="prm", knn=10)
    warm.setEndpoints(qa, qb)
    warm.planMore(2048)
    space.rng = np.random.default_rng(5)   # the timed roadmap draws the same samples as without it
    mp = plan.MotionPlan(space, type="prm", knn=10)
    mp.setEndpoints(qa, qb)
    tp0 = time.perf_counter()
    for _ in range(4):
        mp.planMore(8192)
    prm_path = mp.getPath()   # the shortest path over the grown roadmap, searched once
    tprm = time.perf_counter() - tp0
    sprm = mp.getStats()
    lz = plan.MotionPlan(space, type="lazyprm", knn=10)
    lz.setEndpoints(qa, qb)
    tl0 = time.perf_counter()
    while lz.getPath() is None and lz.getStats()["samples"] < 65536:
        lz.planMore(4096)
    tlazy = time.perf_counter() - tl0
    slz = lz.getStats()
    out["planner"] = {"scene": "7-DOF arm among 16 boxes, edge resolution 0.02 rad (L2)",
                      "prm_samples": sprm["samples"], "prm_milestones": sprm["milestones"],
                      "prm_edges_validated": sprm["edges_checked"], "prm_ms": round(tprm * 1e3, 1),
                      "prm_edges_per_s": round(sprm["edges_checked"] / tprm, 1),
                      "prm_path_found": prm_path is not None,
                      "lazyprm_ms_to_path": round(tlazy * 1e3, 1), "lazyprm_edges_checked": slz["edges_checked"],
                      "lazyprm_samples": slz["samples"], "lazyprm_path_found": lz.getPath() is not None}
    if not args.no_cpu:
        V, E = mp.getRoadmap()
        V, E = np.asarray(V), np.asarray([e for e in mp.edges][:512])
        obw, obr = oracle.world_from_scene(bs)
        t1 = time.perf_counter()
        obw.edge_batch(obr, V[E[:, 0]], V[E[:, 1]], eps=0.02, norm_inf=False)
        tc = time.perf_counter() - t1
        out["planner"]["cpu_edges_per_s"] = round(len(E) / tc, 1)
        out["planner"]["cpu_sample"] = f"first {len(E)} roadmap edges, FP64 oracle, {oracle.nthreads_default()} threads"
    print(f"[bench] planner: {out['planner']}", file=sys.stderr, flush=True)

    # TriangleMesh -> ImplicitSurface of the 200k-triangle soup terrain at
    # 1 cm cells (Geometry3D.convert; the reference's fast-marching
    # conversion of its complex mesh takes 14.06 s, Manual-Geometry.md:134)
    soup = sc.terrains[0]
    gm = engine.Geom.trimesh(ctx, soup.verts, soup.tris)
    hs = 0.01
    gm.to_sdf(hs).close()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        gs = gm.to_sdf(hs)
        if _ < reps - 1:
            gs.close()
    tconv = (time.perf_counter() - t0) / reps
    dims, lo, hi, vals = gs.sdf_grid(values=True)
    gs.close()
    ncell = int(np.prod(dims))
    out["c6_mesh_to_sdf"] = {"mesh_tris": int(soup.tris.shape[0]), "resolution": hs, "dims": list(dims),
                             "cells": ncell, "ms": round(tconv * 1e3, 2),
                             "cells_per_s": round(ncell / tconv, 1)}
    gm.to_sdf(0.0).close()
    t0 = time.perf_counter()
    gdd = gm.to_sdf(0.0)
    tdef = time.perf_counter() - t0
    out["c6_mesh_to_sdf"]["default_resolution"] = {"dims": list(gdd.sdf_grid()[0]), "ms": round(tdef * 1e3, 2)}
    gdd.close()
    if not args.no_cpu:
        # the oracle on 8 x-slabs through the middle of the same grid
        i0 = dims[0] // 2 - 4
        sub = (8, dims[1], dims[2])
        blo = lo + np.array([i0 * hs, 0.0, 0.0])
        t1 = time.perf_counter()
        ov, ow = oracle.mesh_to_sdf(soup, blo, hs, sub, with_winding=True)
        tc = time.perf_counter() - t1
        gv = vals[i0:i0 + 8]
        out["c6_mesh_to_sdf"].update({
            "cpu_cells_per_s": round(ov.size / tc, 1),
            "cpu_sample": f"8 x-slabs ({ov.size} cells) through the grid's middle, FP64 oracle, "
                          f"{oracle.nthreads_default()} threads",
            "max_abs_err_vs_oracle": float(np.abs(gv.astype(np.float64) - ov).max()),
            "classification_equal": bool(np.array_equal(np.signbit(gv), ow != 0))})
    print(f"[bench] C6 mesh->SDF: {out['c6_mesh_to_sdf']}", file=sys.stderr, flush=True)

    # on-disk formats: the headline world as files (arm .rob + link .off,
    # 200k-triangle .off terrain, world XML) loaded to a first answer, and 64
    # rigid objects from one mesh file (one BVH build, 63 shared handles)
    # against 64 separately built meshes
    import shutil
    import tempfile
    from klampt_amd import io as kio
    tmp = tempfile.mkdtemp(prefix="kg_bench_io_")
    try:
        files = []
        for l, m in enumerate(sc.robot.link_meshes):
            files.append(os.path.join(tmp, "link%d.off" % l))
            kio.save_mesh(files[-1], m)
        kio.save_rob(os.path.join(tmp, "arm.rob"), kio.RobotFile(sc.robot, geom_files=files))
        kio.save_mesh(os.path.join(tmp, "soup.off"), sc.terrains[0])
        with open(os.path.join(tmp, "world.xml"), "w") as f:
            f.write('<world>\n  <robot name="arm" file="arm.rob"/>\n  <terrain file="soup.off"/>\n</world>\n')
        robotsim.clear_geometry_cache()
        kio._MESH_CACHE.clear()
        t0 = time.perf_counter()
        wf = kio.load_world(os.path.join(tmp, "world.xml"))
        tparse = time.perf_counter() - t0
        kio._MESH_CACHE.clear()
        t0 = time.perf_counter()
        Wl = robotsim.WorldModel()
        assert Wl.loadFile(os.path.join(tmp, "world.xml"))
        tload = time.perf_counter() - t0
        q1 = S.uniform_configs(sc.robot, 1, seed=9)
        robotsim.RobotCSpace(Wl.robot(0), Wl).feasible_batch(q1)
        torch.cuda.synchronize()
        tfirst = time.perf_counter() - t0
        box = S.ellipsoid_mesh([0.1, 0.15, 0.2], edge=0.01)
        kio.save_mesh(os.path.join(tmp, "part.off"), box)
        with open(os.path.join(tmp, "part.obj"), "w") as f:
            f.write('mesh "part.off"\n')
        robotsim.clear_geometry_cache()
        t0 = time.perf_counter()
        Ws = robotsim.WorldModel()
        for _ in range(64):
            Ws.loadRigidObject(os.path.join(tmp, "part.obj"))
        tshared = time.perf_counter() - t0
        t0 = time.perf_counter()
        Wsep = robotsim.WorldModel()
        for _ in range(64):
            Wsep.addRigidObject(robotsim.Geometry3D(robotsim.TriangleMesh(box.verts, box.tris)))
        tsep = time.perf_counter() - t0
        out["io"] = {"world": "arm .rob + 7 link .off + %d-triangle .off terrain + world XML"
                              % wf.terrains[0].mesh.tris.shape[0],
                     "parse_ms": round(tparse * 1e3, 1), "load_ms": round(tload * 1e3, 1),
                     "load_to_first_answer_ms": round(tfirst * 1e3, 1),
                     "objects_from_one_file": 64, "triangles_per_object": int(box.tris.shape[0]),
                     "shared_bvh_load_ms": round(tshared * 1e3, 1), "separate_build_ms": round(tsep * 1e3, 1)}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    print(f"[bench] io: {out['io']}", file=sys.stderr, flush=True)
    return out


if __name__ == "__main__":
    main()
