#!/usr/bin/env python
"""Throughput of the shallow-water evolve() step on one MI355X.

Workload: seeded Delaunay mesh of 500k uniform points in a 10 km x 10 km
square (~1.0M triangles, Morton order), 5-octave fBm terrain (amplitude 20 m),
stage max(z, 5 m) + 3 m mound of radius 500 m (wet/dry fronts), Manning
n = 0.025, transmissive boundaries, FP64, second order, Euler steps.

A step is one full Euler step of the evolve loop (compute_fluxes + gravity +
friction + CFL timestep + update + protect + extrapolation + balance +
vertex->edge interpolation) over the whole mesh, with every input resident in
HBM.  The timed region runs --steps steps of the fused device loop (hipGraph
replay, no host synchronisation between steps).  A second, untimed-for-value
pass records HIP events around each kernel on the library's stream to get the
per-kernel durations behind the roofline numbers.  The CPU baseline runs the
FP64 C oracle (serial) on a bounded sample of steps of the same workload.

Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per triangle of each fused kernel launch: every array
# the kernel reads or writes, counted once; neighbour gathers (edge records,
# centroid records, centroid coordinates) counted as cache hits.
KERNEL_BYTES_PER_TRIANGLE = {
    # reads: edge geometry records 3x32 96, own edge state records 3x32 96,
    # {radius, area} 16, gravity stage-gradient term 16 | writes: explicit
    # updates 24, max_speed 8
    "flux": 96 + 96 + 16 + 16 + 24 + 8,
    # reads: centroid record 32, explicit 24, semi-implicit 24, Manning n 8,
    # mean bed 8 | writes: centroid record 32, next step's semi-implicit
    # momentum updates (friction) 16
    "update": 32 + 24 + 24 + 8 + 8 + 32 + 16,
    # reads: centroid record 32, centroid xy 16, surrogates 12, flags 4, vertex
    # xy 48, bed vertex 24 | writes: vertex w/uh/vh 72, edge state records 96,
    # next step's gravity term 16
    "distribute": 32 + 16 + 12 + 4 + 48 + 24 + 72 + 96 + 16,
}
# The mixed-precision loop (FP32 edge state and geometry, FP64 centroids).
KERNEL_BYTES_PER_TRIANGLE_MIXED = {
    # reads: FP32 edge geometry records 3x16 48, FP32 edge state records 3x16
    # 48, {radius, area} 16, gravity term 16 | writes: explicit updates 24,
    # max_speed 8
    "flux": 48 + 48 + 16 + 16 + 24 + 8,
    "update": KERNEL_BYTES_PER_TRIANGLE["update"],
    # as FP64, but the edge state written is the FP32 records 3x16 48
    "distribute": 32 + 16 + 12 + 4 + 48 + 24 + 72 + 48 + 16,
}
# One Euler step = the three kernels above (the timestep reduction moves 8
# bytes per 64 triangles).
STEP_BYTES_PER_TRIANGLE = sum(KERNEL_BYTES_PER_TRIANGLE.values())

# The vertex-record Euler step (the default FP64 path; sw_step_vertex.h):
# k_flux_v ("flux"), k_update_v ("update"), k_extrap_v ("distribute").  Node
# arrays (coordinates 16 B and bed 8 B per node, about half a node per
# triangle) are gathered through the node indices and counted once.
KERNEL_BYTES_PER_TRIANGLE_VERTEX = {
    # reads: node indices 12, node coordinates and bed 12, neighbour words 12,
    # flags 4, {radius, area} 16, the triangle's vertex records 3x24 72,
    # centroid depth 8 | writes: update record 24, max_speed 8
    "flux": 12 + 12 + 12 + 4 + 16 + 72 + 8 + 24 + 8,
    # reads: centroid record 32, update record 24, Manning n 8, mean bed 8 |
    # writes: centroid record 32
    "update": 32 + 24 + 8 + 8 + 32,
    # reads: centroid record 32, flags 4, centroid 16, node indices 12,
    # surrogates 12, node coordinates and bed 12 | writes: vertex records 3x24
    # 72, centroid depth 8
    "distribute": 32 + 4 + 16 + 12 + 12 + 12 + 72 + 8,
}
# The unique footprint of one FP64 Euler step (each array read once, each
# mutated array written once, the edge-value exchange counted as one write
# plus one read, neighbour gathers as cache hits): 604 bytes per triangle,
# split by the kernel of the vertex-record step that touches each array.
FOOTPRINT_BYTES_PER_TRIANGLE_STEP = 604
FOOTPRINT_SHARE_VERTEX = {
    # neighbours 12, neighbour edges 12, full flag 4, normals 48, edge lengths
    # 24, areas 8, radii 8, bed edge 24, edge values read 72, max_speed 8
    "flux": 12 + 12 + 4 + 48 + 24 + 8 + 8 + 24 + 72 + 8,
    # centroid state read+write 48, friction 8, bed centroid 8
    "update": 48 + 8 + 8,
    # surrogates 12, boundary counts 4, centroid xy 16, vertex xy 48, bed
    # vertex 24, vertex state read+write 144, edge values written 72
    "distribute": 12 + 4 + 16 + 48 + 24 + 144 + 72,
}
assert sum(FOOTPRINT_SHARE_VERTEX.values()) == FOOTPRINT_BYTES_PER_TRIANGLE_STEP
VERTEX_STEP_KERNELS = "vertex-record (k_flux_v, k_update_v, k_extrap_v)"


def kernel_bytes(precision):
    return KERNEL_BYTES_PER_TRIANGLE if precision == "fp64" else KERNEL_BYTES_PER_TRIANGLE_MIXED
KERNELS = ("flux", "timestep", "update", "distribute")


def workload_tag(args):
    tag = "fv_%s_%s_%dk_points" % (args.precision, args.order, args.points // 1000)
    return tag if args.locality == "hilbert" else tag + "_keep_order"


def kernel_symbol(kernel, precision, launch, vstep=False, limited=False):
    """Name of a fused kernel as rocprofv3 reports it; ``launch`` = workgroup
    sizes of (flux, update, distribute); ``limited``: the per-quantity
    limiters' variant of the vertex-record extrapolation."""
    if vstep:
        extrap = "k_extrap_v<%d, %s>" % (launch[2], "true" if limited else "false")
        return {"flux": "k_flux_v<%d>" % launch[0], "update": "k_update_v<%d>" % launch[1],
                "distribute": extrap}[kernel]
    r = "double" if precision == "fp64" else "float"
    fb, ub, db = launch
    flux = "k_flux<%s, true, %d>" % (r, fb) if precision == "fp64" else "k_flux_mixed<%d>" % fb
    return {"flux": flux, "timestep": "k_timestep<true>", "update": "k_update<true, %d>" % ub,
            "distribute": "k_distribute<%s, true, %d>" % (r, db)}[kernel]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def copy_peak_gbs(mib=1024, reps=20):
    """Bandwidth of a plain device-to-device copy of ``mib`` MiB (bytes read
    plus bytes written per second), timed with HIP events: the measured
    ceiling beside the vendor's 8 TB/s."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * n * 8 / sec / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="fv", choices=["fv", "ensemble", "learned", "merimbula", "cpu-plan"],
                    help="fv: the 1M-triangle plain FV step (default); ensemble: the batched ensemble; "
                         "learned: train the edge-GNN surrogate on FV trajectories and roll it out at 1M")
    ap.add_argument("--delta", type=float, default=0.1,
                    help="learned: seconds per learned step when training (a loaded network brings its own)")
    ap.add_argument("--merge-radius", type=float, default=2.5,
                    help="learned: inscribed radius (m) below which triangles are merged for the learned update")
    ap.add_argument("--train-iters", type=int, default=3000, help="learned: training iterations")
    ap.add_argument("--train-meshes", type=int, default=40,
                    help="learned: training windows taken at a time (50 trajectories each)")
    ap.add_argument("--train-trajectories", type=int, default=2000,
                    help="learned: take training windows until this many trajectories are kept")
    ap.add_argument("--save-net", default=None, help="learned: write the trained network (.npz)")
    ap.add_argument("--load-net", default=None,
                    help="learned: use this trained network (.npz); without --save-net the committed --network is used")
    ap.add_argument("--judge-precision", default="fp64", choices=["fp64", "mixed"],
                    help="learned: precision of the FV judge's step (the reference and fallbacks are FP64; the "
                         "other choice is reported beside it)")
    ap.add_argument("--members", type=int, default=64, help="ensemble members")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 200; ensemble 300)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--points", type=int, default=500000)
    ap.add_argument("--order", default="morton", choices=["morton", "random", "none"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--locality", default="hilbert", choices=["hilbert", "keep"],
                    help="device numbering of the triangles (library renumbering or the input order)")
    ap.add_argument("--timestepping", default="euler", choices=["euler", "rk2", "rk3"],
                    help="fv: step of the timed loop (a step of rk2 / rk3 runs 2 / 3 stages)")
    ap.add_argument("--limiter", default="optimised", choices=["optimised", "vertex", "edge"],
                    help="fv: distribute path (vertex: optimised_gradient_limiter off; edge: use_edge_limiter)")
    ap.add_argument("--mesh-build", default="host", choices=["host", "device"],
                    help="fv: build the mesh geometry on the GPU (sw_build_mesh) or the host; same arrays")
    ap.add_argument("--cpu-steps", type=int, default=20, help="oracle steps timed for the CPU baseline")
    ap.add_argument("--step-path", default="auto", choices=["auto", "edge"],
                    help="FV: vertex-record step where it applies (auto) or the edge-record step")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="FV: skip the learned rollout that the default line carries")
    ap.add_argument("--network", default=os.path.join(ROOT, "anuga_cuda_amd", "networks", "edge_gnn_cfg5.npz"),
                    help="FV / learned: the trained network of the learned rollout")
    ap.add_argument("--timing-steps", type=int, default=64)
    ap.add_argument("--traffic-json", default=None,
                    help="rocprof traffic summary (default: profiles/<workload>.traffic.json)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (scale for rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one MI355X only (the multi-GPU sharding is a design, not built); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from anuga_cuda_amd import scenarios

    if args.steps is None:
        args.steps = 300 if args.workload == "ensemble" else 200
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "merimbula":
        return bench_merimbula(args)
    if args.workload == "cpu-plan":
        return bench_cpu_plan(args)

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: %d MiB read + %d MiB written per copy kernel" % (args.copy_calibration,
                                                                             args.copy_calibration))
        del a, b

    t0 = time.time()
    limiter_kw = {"vertex": {"optimised_gradient_limiter": 0}, "edge": {"use_edge_limiter": 1}}.get(args.limiter, {})
    d = scenarios.delaunay_terrain(npoints=args.points, order=args.order, precision=args.precision,
                                   locality_order=args.locality, build=args.mesh_build, **limiter_kw)
    d.set_timestepping_method(args.timestepping)
    stages = {"euler": 1, "rk2": 2, "rk3": 3}[args.timestepping]
    N = d.number_of_elements
    log("mesh: %d triangles, %d boundary edges, built in %.1f s" % (N, d.number_of_boundary_elements,
                                                                  time.time() - t0))
    cpu_snapshot = None
    if not args.no_cpu:
        from oracle.oracle import OracleDomain
        cpu_snapshot = OracleDomain.from_domain(d)

    h = d._dev()
    from anuga_cuda_amd import _lib
    if args.step_path == "edge":
        h.set_step_path(_lib.SW_STEP_EDGE)
    h.call("sw_evolve_prepare", 1e9, 0.0, 0)
    vstep = h.step_path() == _lib.SW_STEP_VERTEX
    log("device prepared (%s-record step)" % ("vertex" if vstep else "edge"))
    # the autotuner first (10 timed steps per candidate workgroup size): it
    # times launches one by one and drops the captured step graph, which the
    # warmup then captures again
    launch = h.autotune(10)
    done, ev = h.evolve_steps(args.warmup)
    h.call("sw_synchronize")
    log("warmup %d steps done; autotuned workgroup sizes (flux, update, distribute): %s" % (done, launch))

    t_start = time.perf_counter()
    done, ev = h.evolve_steps(args.steps)
    h.call("sw_synchronize")
    elapsed = time.perf_counter() - t_start
    if done != args.steps:
        raise SystemExit("expected %d steps, ran %d (event %d)" % (args.steps, done, ev))
    s = h.get_state()
    value = N * args.steps / elapsed
    log("timed: %d steps in %.4f s -> %.3e triangle-steps/s (t=%.3f s, dt=%.4e)" % (
        args.steps, elapsed, value, s.time, s.timestep))

    # per-kernel durations from HIP events on the library stream
    h.call("sw_enable_timing", 1)
    h.evolve_steps(args.timing_steps)
    h.call("sw_enable_timing", 0)
    ms, cnt = h.kernel_times(4)
    avg_us = {k: (1000.0 * ms[i] / cnt[i] if cnt[i] else None) for i, k in enumerate(KERNELS)}
    if vstep:
        avg_us["timestep"] = None  # k_update_v reduces the timestep; the interval holds no kernel
    log("kernel avg us: %s" % {k: round(v, 2) if v else v for k, v in avg_us.items()})
    kb = KERNEL_BYTES_PER_TRIANGLE_VERTEX if vstep else kernel_bytes(args.precision)
    dominant = max(kb, key=lambda k: avg_us[k] or 0.0)
    # a Runge-Kutta step runs every kernel once per stage, plus the copy of the
    # starting centroid records (32 B written, then read by each later stage)
    design_bytes = sum(kb.values()) * stages + (32 * stages if stages > 1 else 0)
    # roofline bytes: the step's unique footprint (604 B per triangle and
    # stage) and, for the dominant kernel of the vertex-record step, its share
    # of it; the edge-record step's kernels are counted by their design bytes
    step_bytes = FOOTPRINT_BYTES_PER_TRIANGLE_STEP * stages
    kernel_roof_bytes = (FOOTPRINT_SHARE_VERTEX if vstep else kb)[dominant]
    achieved = kernel_roof_bytes * N / (avg_us[dominant] * 1e-6) / 1e9
    step_gbs = step_bytes * N * args.steps / elapsed / 1e9
    copy_gbs = copy_peak_gbs()
    log("device copy: %.0f GB/s" % copy_gbs)
    # HBM bytes per launch of the dominant kernel, measured by rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs of this workload (copy-calibrated; written by
    # anuga_cuda_amd.rocprof_summary), if such a profile is present.
    traffic = None
    tfile = args.traffic_json or os.path.join(ROOT, "profiles", "%s.traffic.json" % workload_tag(args))
    step_traffic = None
    if os.path.exists(tfile):
        kernels = json.load(open(tfile)).get("kernels", {})
        if vstep:
            recs = {k: kernels.get(kernel_symbol(k, args.precision, launch, True, args.limiter != "optimised"))
                    for k in kb}
            rec = recs[dominant] or {}
            if all(recs.values()):
                step_traffic = sum(r["traffic_bytes"] for r in recs.values())
        else:
            kname = kernel_symbol(dominant, args.precision, launch)
            if dominant == "distribute":
                # k_distribute<R, FUSED, BLOCK, LIMITED>: profiles name the limiter variant
                limited = kname[:-1] + (", true>" if args.limiter != "optimised" else ", false>")
                rec = kernels.get(limited) or (kernels.get(kname) if args.limiter == "optimised" else None)
            else:
                rec = kernels.get(kname)
            rec = rec or {}
        traffic = rec.get("traffic_bytes")

    cpu = None
    if cpu_snapshot is not None:
        od = cpu_snapshot
        od.evolve_prepare(1e9, None)
        od.evolve_step()  # first step outside the sample
        c0 = time.perf_counter()
        for _ in range(args.cpu_steps):
            od.evolve_step()
        c_el = time.perf_counter() - c0
        cpu_model = cpu_model_name()
        cpu = {"value": N * args.cpu_steps / c_el, "unit": "triangle-steps/s", "cores": 1, "kind": "port",
               "sample": "%d %s steps of the same %d-triangle workload, serial FP64 C oracle "
                         "(restated reference C), %.2f s on %s (nproc %d)" % (
                             args.cpu_steps, "Euler" if stages == 1 else args.timestepping, N, c_el, cpu_model,
                             os.cpu_count() or 0)}
        log("cpu baseline: %.3e triangle-steps/s" % cpu["value"])

    out = {
        "metric": "triangle-steps/s at 1M tris, FV vs learned at stated error; % HBM peak",
        "value": value,
        "unit": "triangle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "plain FV, seeded Delaunay %d points in 10 km x 10 km, %s order, fBm terrain "
                               "20 m, wet/dry mound, Manning 0.025, transmissive" % (args.points, args.order),
                   "triangles": N, "precision": args.precision, "order": 2, "timestepping": args.timestepping,
                   "limiter": args.limiter, "mesh_build": args.mesh_build,
                   "device_numbering": args.locality, "workgroup_sizes": list(launch)},
        "roofline": {"bound": "hbm", "kernel": kernel_symbol(dominant, args.precision, launch, vstep,
                                                               args.limiter != "optimised"),
                     "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_triangle": kernel_roof_bytes,
                     "design_bytes_per_triangle": kb[dominant],
                     "avg_launch_us": avg_us[dominant],
                     "measured_copy_peak": copy_gbs, "frac_of_copy": achieved / copy_gbs},
        "step_roofline": {"bytes_per_triangle_step": step_bytes, "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS, "frac_of_copy": step_gbs / copy_gbs,
                          "design_bytes_per_triangle_step": design_bytes,
                          "traffic_per_step": step_traffic,
                          "step_kernels": VERTEX_STEP_KERNELS if vstep else "edge-record (four kernels)",
                          "kernel_avg_us": avg_us},
        "cpu_baseline": cpu,
    }
    d.close()
    if not args.no_learned and args.precision == "fp64" and args.points == 500000 and args.order == "morton":
        out["learned"] = learned_rollout(args)
    print(json.dumps(out), flush=True)


def learned_rollout(args):
    """The learned form of the same workload: the committed network (trained
    by this package's kernels; provenance in anuga_cuda_amd/networks/)
    rolled out for 100 learned steps of the length it was trained for (0.1 s)
    from the workload's initial state, the FV judge (--judge-precision: the
    FP64 or the mixed-precision step) every 8 steps with the FP64 FV fallback
    above 1% stage L2 error, against the FP64 FV path over the same simulated
    time.  speedup = FV wall time / judged rollout wall time."""
    from anuga_cuda_amd import learned, scenarios

    net = learned.EdgeGNN.load(args.network)
    dom = scenarios.delaunay_terrain(npoints=args.points, order=args.order)
    res = learned.evaluate_rollout(net, dom, 100, net.scales["delta"] or 0.1, judge_every=8, threshold=0.01,
                                   judge_precision=args.judge_precision, log=log)
    dom.close()
    keys = ("speedup", "learned_rate", "fv_rate", "steps", "delta", "fv_steps", "final_stage_rel_l2",
            "max_stage_rel_l2", "mass_balance_rel", "mass_change_vs_fv", "fallback_windows", "judge_precision",
            "learned_step_ms", "fv_seconds", "learned_seconds")
    r = {k: res[k] for k in keys}
    r.update(froude_max=learned.FROUDE_MAX,
             unit="FV-equivalent triangle-steps/s (triangles x FV steps for the same simulated time / s)",
             windows_judged=len(res["judged"]), network=os.path.relpath(args.network, ROOT),
             stage_error="area-weighted L2 of the stage error over the depth norm, against FP64 FV",
             mass_balance="(volume change + open-boundary outflow) / volume over one window of 8 learned steps")
    return r


# The reference's only published timing of this path: its Merimbula scenario
# (43,200 triangles, reflective walls) evolved to 500 s in 1,301 Euler steps in
# 2.124 s on a GTX 480 (docs/profilling/CUDA/profile_on_GTX480_prepared_500fs_2.123959s,
# the time in the file name), i.e. 43,200 x 1,301 / 2.124 s triangle-steps/s.
REFERENCE_MERIMBULA_GTX480 = 43200 * 1301 / 2.123959


def bench_merimbula(args):
    """The reference's Merimbula scenario (tests/CUDA/merimbula/merimbula.py:50,
    86-173: yields every 50 s) evolved to 500 s through Domain.evolve, wall
    clock of the whole evolve (graph capture, yields and host syncs
    included).  The first run warms the process; the second is reported."""
    from anuga_cuda_amd import scenarios

    cpu = None
    if not args.no_cpu:
        # the serial FP64 C oracle over the same run, on one host core
        from oracle.oracle import OracleDomain

        od = OracleDomain.from_domain(scenarios.merimbula())
        c0 = time.perf_counter()
        csteps = 0
        for _ in od.evolve(50.0, 500.0):
            csteps += od.d.number_of_steps
        c_el = time.perf_counter() - c0
        cpu = {"value": 43200 * csteps / c_el, "unit": "triangle-steps/s", "cores": 1, "kind": "port",
               "sample": "the whole run (%d steps to 500 s) with the serial FP64 C oracle, %.2f s (nproc %d)" % (
                   csteps, c_el, os.cpu_count() or 0)}
        log("cpu baseline: %d steps in %.2f s (%.3e triangle-steps/s)" % (csteps, c_el, cpu["value"]))
    runs = []
    for rep in range(2):
        d = scenarios.merimbula()
        d._dev()
        t0 = time.perf_counter()
        steps = 0
        for _ in d.evolve(yieldstep=50.0, finaltime=500.0):
            steps += d.number_of_steps
        elapsed = time.perf_counter() - t0
        runs.append((elapsed, steps, d.get_time()))
        log("merimbula run %d: %d steps to t = %.1f s in %.4f s (%.3e triangle-steps/s)" % (
            rep, steps, d.get_time(), elapsed, d.number_of_elements * steps / elapsed))
        N = d.number_of_elements
        d.close()
    elapsed, steps, t_end = runs[-1]
    value = N * steps / elapsed
    out = {
        "metric": "triangle-steps/s on the reference's Merimbula scenario (43,200 triangles, to 500 s)",
        "value": value,
        "unit": "triangle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": 1000.0 * elapsed / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / REFERENCE_MERIMBULA_GTX480,
        "dtype": "f64",
        "data": "the reference's own mesh fixture (merimbula_43200.tsh) and initial condition",
        "config": {"workload": "Merimbula, %d triangles, reflective walls, yields every 50 s to t = %.0f s, "
                               "whole Domain.evolve wall time; vs_baseline against the reference's 2.124 s "
                               "(1,301 steps) on a GTX 480" % (N, t_end), "triangles": N,
                   "evolve_seconds": elapsed, "reference_gtx480_seconds": 2.123959},
        "roofline": None,  # 43,200 triangles sit in the 256 MB last-level cache: % of HBM is not meaningful
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def cpu_model_name():
    name = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return name


def bench_cpu_plan(args):
    """The serial FP64 C oracle timed beside the GPU on each plain-FV
    configuration, in one run on the same host: the 40,000-triangle dam break
    in full (t = 10 s, yields every 1 s, whole Domain.evolve on the GPU, the
    second of two runs), the 1M-triangle Delaunay workload in Morton and in
    random order (20 oracle steps after one untimed step; 200 GPU steps after
    autotuning and 20 warmup steps), and the 4M-triangle dam break (5 oracle
    steps; 20 GPU steps in FP64 and in mixed precision).  One JSON line with a
    record per case."""
    from anuga_cuda_amd import scenarios
    from oracle.oracle import OracleDomain

    def oracle_steps(od, n):
        od.evolve_prepare(1e9, None)
        od.evolve_step()  # outside the sample
        c0 = time.perf_counter()
        for _ in range(n):
            od.evolve_step()
        return time.perf_counter() - c0

    def gpu_steps(d, n, warmup):
        h = d._dev()
        h.call("sw_evolve_prepare", 1e9, 0.0, 0)
        h.autotune(10)
        h.evolve_steps(warmup)
        h.call("sw_synchronize")
        t0 = time.perf_counter()
        done, _ = h.evolve_steps(n)
        h.call("sw_synchronize")
        if done != n:
            raise SystemExit("expected %d steps, ran %d" % (n, done))
        return time.perf_counter() - t0

    cases = []

    def record(name, N, cpu_n, cpu_s, gpu_n, gpu_s, **extra):
        rec = dict(config=name, triangles=N, cpu_steps=cpu_n, cpu_seconds=cpu_s, cpu_rate=N * cpu_n / cpu_s,
                   gpu_steps=gpu_n, gpu_seconds=gpu_s, gpu_rate=N * gpu_n / gpu_s, **extra)
        rec["gpu_over_cpu"] = rec["gpu_rate"] / rec["cpu_rate"]
        cases.append(rec)
        log("%s: CPU %.3e, GPU %.3e triangle-steps/s (%.0fx)" % (name, rec["cpu_rate"], rec["gpu_rate"],
                                                                 rec["gpu_over_cpu"]))

    # 40,000-triangle dam break, the whole run
    od = OracleDomain.from_domain(scenarios.dam_break_channel(100, 100))
    c0 = time.perf_counter()
    cpu_n = 0
    for _ in od.evolve(1.0, 10.0):
        cpu_n += od.d.number_of_steps
    cpu_s = time.perf_counter() - c0
    for rep in range(2):
        d = scenarios.dam_break_channel(100, 100)
        d._dev()
        t0 = time.perf_counter()
        gpu_n = 0
        for _ in d.evolve(yieldstep=1.0, finaltime=10.0):
            gpu_n += d.number_of_steps
        gpu_s = time.perf_counter() - t0
        d.close()
    record("40,000-triangle dam break to t = 10 s (whole Domain.evolve, 11 yields)", 40000, cpu_n, cpu_s, gpu_n,
           gpu_s)
    del od

    # the 1M-triangle Delaunay workload in both orders
    for order in ("morton", "random"):
        d = scenarios.delaunay_terrain(npoints=500000, order=order)
        od = OracleDomain.from_domain(d)
        cpu_s = oracle_steps(od, 20)
        del od
        gpu_s = gpu_steps(d, 200, 20)
        record("1M-triangle Delaunay, %s order, FP64" % order, d.number_of_elements, 20, cpu_s, 200, gpu_s)
        d.close()

    # the 4M-triangle dam break: FP64 and mixed precision on the GPU
    d = scenarios.dam_break_channel(1000, 1000)
    od = OracleDomain.from_domain(d)
    cpu_s = oracle_steps(od, 5)
    del od
    gpu_s = gpu_steps(d, 20, 5)
    N = d.number_of_elements
    d.close()
    dm = scenarios.dam_break_channel(1000, 1000, precision="mixed")
    gpu_m = gpu_steps(dm, 20, 5)
    dm.close()
    record("4M-triangle dam break", N, 5, cpu_s, 20, gpu_s, gpu_mixed_seconds=gpu_m,
           gpu_mixed_rate=N * 20 / gpu_m)

    print(json.dumps({"metric": "serial FP64 C oracle and GPU, triangle-steps/s per configuration",
                      "unit": "triangle-steps/s", "cores": 1, "kind": "port", "cpu_model": cpu_model_name(),
                      "nproc": os.cpu_count() or 0, "cases": cases}))


def bench_ensemble(args):
    """Batched ensemble: ``--members`` seeded Delaunay scenarios with triangle
    counts log-uniform in [20k, 500k], Manning n ~ U[0.01, 0.06], dam height
    ~ U[1, 4] m, inflow ~ U[0, 1] m^2/s, in one handle (every fused launch and
    one hipGraph cover all members; per-member adaptive timesteps).  Every
    member runs exactly --steps timed steps."""
    from anuga_cuda_amd import Ensemble, scenarios

    t0 = time.time()
    domains, specs = scenarios.ensemble_scenarios(count=args.members, precision=args.precision)
    ens = Ensemble(domains)
    N = ens.number_of_elements
    log("ensemble: %d members, %d triangles (%d .. %d), built in %.1f s" % (
        len(domains), N, min(d.number_of_elements for d in domains), max(d.number_of_elements for d in domains),
        time.time() - t0))
    ens.prepare(1e9, None)
    launch = ens.autotune(10)
    ens.evolve_steps(args.warmup)
    log("autotuned workgroup sizes (flux, update, distribute): %s" % (launch,))
    h = ens.handle
    h.call("sw_synchronize")
    t_start = time.perf_counter()
    done, ev = h.evolve_steps(args.steps)
    h.call("sw_synchronize")
    elapsed = time.perf_counter() - t_start
    steps = [d._handle.get_state().number_of_steps for d in domains]
    if done != args.steps or min(steps) != max(steps):
        raise SystemExit("expected %d steps per member, ran %d (%d .. %d)" % (args.steps, done, min(steps),
                                                                             max(steps)))
    value = N * args.steps / elapsed
    dts = [d._handle.get_state().timestep for d in domains]
    log("timed: %d steps of %d members in %.4f s -> %.3e triangle-steps/s (member dt %.3g .. %.3g s)" % (
        args.steps, len(domains), elapsed, value, min(dts), max(dts)))
    h.call("sw_enable_timing", 1)
    h.evolve_steps(min(args.timing_steps, 32))
    h.call("sw_enable_timing", 0)
    ms, cnt = h.kernel_times(4)
    avg_us = {k: (1000.0 * ms[i] / cnt[i] if cnt[i] else None) for i, k in enumerate(KERNELS)}
    from anuga_cuda_amd import _lib
    vstep = h.step_path() == _lib.SW_STEP_VERTEX
    if vstep:
        avg_us["timestep"] = None  # k_update_v reduces the timestep; the interval holds no kernel
    log("kernel avg us: %s" % {k: round(v, 2) if v else v for k, v in avg_us.items()})
    # the step's unique footprint (604 B per triangle; the same figure as the
    # single-domain line) when the vertex-record step runs, and the design bytes
    # of the kernels that run beside it
    design = sum((KERNEL_BYTES_PER_TRIANGLE_VERTEX if vstep else kernel_bytes(args.precision)).values())
    step_bytes = FOOTPRINT_BYTES_PER_TRIANGLE_STEP if args.precision == "fp64" else design
    out = {
        "metric": "triangle-steps/s at 1M tris, FV vs learned at stated error; % HBM peak",
        "value": value,
        "unit": "triangle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "batched ensemble: %d seeded Delaunay dam-break scenarios, triangles log-uniform "
                               "in [20k, 500k], Manning n ~ U[0.01, 0.06], dam height ~ U[1, 4] m, inflow "
                               "~ U[0, 1] m^2/s, one handle, per-member adaptive timesteps" % len(domains),
                   "triangles": N, "members": len(domains), "precision": args.precision,
                   "workgroup_sizes": list(launch)},
        "step_roofline": {"bytes_per_triangle_step": step_bytes,
                          "achieved": step_bytes * N * args.steps / elapsed / 1e9,
                          "frac": step_bytes * N * args.steps / elapsed / 1e9 / HBM_PEAK_GBS,
                          "design_bytes_per_triangle_step": design,
                          "step_kernels": VERTEX_STEP_KERNELS if vstep else "edge-record (four kernels)",
                          "kernel_avg_us": avg_us},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    ens.close()


def bench_learned(args):
    """Learned edge-flux surrogate: trained on FV trajectories of seeded
    configuration-2-style windows, rolled out for 100 learned steps on the
    1M-triangle configuration-2 workload with the FV judge every 8 steps
    (fallback above 1% stage L2 error).  value = FV-equivalent triangle-steps/s
    of the judged rollout (the FV steps the reference path needs for the same
    simulated time, per second)."""
    from anuga_cuda_amd import learned

    load_net = args.load_net
    if load_net is None and args.save_net is None:
        load_net = args.network  # the committed network; --save-net trains a new one
    res = learned.run_learned_workload(train_meshes=args.train_meshes, train_trajectories=args.train_trajectories,
                                       delta=args.delta, iters=args.train_iters,
                                       eval_points=args.points, save_net=args.save_net, load_net=load_net,
                                       judge_precision=args.judge_precision, merge_radius=args.merge_radius, log=log)
    out = {
        "metric": "triangle-steps/s at 1M tris, FV vs learned at stated error; % HBM peak",
        "value": res["learned_rate"],
        "unit": "triangle-steps/s",
        "n_gpus": 1,
        "steps": res["steps"],
        "warmup": 0,
        "ms_per_step": 1000.0 * res["learned_seconds"] / res["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+f32+f64",
        "data": "synthetic",
        "config": {"workload": "learned edge-GNN surrogate (3 message-passing layers, width 64, BF16 MFMA) on "
                               "the seeded 1M-triangle Delaunay workload; learned steps of %g s; FV judge "
                               "(%s step) every 8 steps, FP64 FV fallback above 1%% stage L2 error; error "
                               "against the FP64 FV path" % (args.delta, args.judge_precision),
                   "triangles": res["triangles"]},
        "learned": {k: v for k, v in res.items() if k not in ("stage_rel_l2", "judged")},
        "stage_rel_l2_at_window_ends": res["stage_rel_l2"],
        "fv_rate_same_run": res["fv_rate"],
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
