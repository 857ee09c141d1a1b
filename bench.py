"""Benchmark of the MI355X implicit cloth step.

Workload (default "batch64"): 64 independent square cloths whose vertices per
side are drawn uniformly from 17..129 (seeded), stiffness scaled
log-uniformly in [0.1, 10] per cloth, two top corners pinned, +-1 mm seeded
out-of-plane jitter, dt = 0.01 s, Jacobi-PCG to a 1e-6 relative residual.
All cloths step together (mixed precision: FP32 matrix and PCG recurrences,
FP64 dot products and solution, FP64 true-residual replacement).  A "step" is
one substep of all 64 cloths; the headline value is cloth-steps per second
(64 per step).

Reported beside it, in the same run:
  * step latency p50 / p99 (per batch step)
  * the plain FP64 form: every cloth stepped on its own, all-FP64, direct
    launches (the denominator of the batched/mixed speedup)
  * roofline of the dominant kernel (the PCG SpMV, with the direction update
    and the d.q partial fused in), timed with HIP events on the library's
    stream: algorithmic bytes per launch / mean launch time vs the 8 TB/s HBM
    peak
  * the CPU baseline: the reference's own CPU step (cBaraffCloth::UpdatePos,
    compiled from its sources -O3 -DNDEBUG into oracle/_ref/ref_cloth), one
    thread, on a bounded sample of the same cloths, each stepped from the
    state the GPU reached right after the timed steps and running the PCG
    iteration count the GPU needed there (the reference's solve runs a fixed
    count); without that binary, the oracle (the C restatement) in tolerance
    mode
  * the per-result error: the GPU's step from that state against the oracle
    solved to 1e-10 from the same state, over the same sample

Workload "irregular": one cloth of 1.5M seeded points (1 mm minimum
spacing) in a 2 m x 2 m square, Delaunay-triangulated, top side pinned,
dt = 0.002 s, mixed precision to a 1e-7 residual; its CPU baseline is
extrapolated from the oracle's assembly and per-iteration times on the same
mesh (a full CPU step takes minutes).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                       [--workload batch64|single257|single21|irregular]
       python bench.py --workload serve [--requests N] [--rate R]
       python bench.py --workload drape [--steps K] [--warmup W]
       python bench.py --roofline-only [--copy-calibration]
The second form only assembles the system and times the dominant kernel (and,
with --copy-calibration, first copies 256 MiB device-to-device three times):
the command profiled for the kernel's rocprofv3 duration and HBM counters.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from acousticsim_amd import capi  # noqa: E402
from acousticsim_amd import workloads as wl  # noqa: E402

METRIC = "cloth steps/s & p50/p99 step latency at 1e-6 PCG residual; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0
KERNEL = ("k_cg_spmv<float, %d, %s, %s, %d> (PCG SpMV: d = P r + beta d_prev, q = W d, partial "
          "d.q; directions %s; matrix loads %s)")


def batch_slices(ms):
    """the batch's 64-row slices: each cloth's, padded to a multiple of 4"""
    return sum(-(-(-(-(len(x) // 3) // 64)) // 4) * 4 for _, x in ms)


def kernel_name(lanes, staged, nt, slices):
    """the FP32 SpMV instance that a launch over every slice runs: the
    gathered product takes 8 / lanes slots per lane and chunk on grids of up
    to 2048 slices (BW_ILP_SLICES overrides), else 4 / lanes"""
    ilp = int(os.environ.get("BW_ILP_SLICES", "2048"))
    chunk = (8 if not staged and slices <= ilp else 4) // lanes
    return KERNEL % (lanes, "true" if staged else "false", "true" if nt else "false", chunk,
                     "staged in LDS per distinct slice column" if staged
                     else "gathered per slot",
                     "nontemporal" if nt else "cached")


# HBM bytes per launch of a k_cg_spmv variant (lanes per row, staged,
# nontemporal) on a workload, measured by rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE over the --roofline-only command with that variant fixed, scaled
# by the same counters' reading for a 256 MiB device-to-device copy; each
# entry names the profile it comes from.  A bench line reports traffic only
# when the autotuner picked a variant that has a measured entry here, and
# says which profile the figure comes from.
TRAFFIC = {
    "batch64": {(1, False, False): (211446746.0, "profiles/k_cg_spmv_batch64.md"),
                (2, False, False): (208552256.0, "profiles/k_cg_spmv_batch64.md")},
    "single257": {(1, False, False): (33980746.0, "profiles/k_cg_spmv_single257.md"),
                  (2, False, False): (33970307.0, "profiles/k_cg_spmv_single257.md")},
}


def traffic_for(workload, tuning):
    """(bytes, source) of the picked variant's counter profile, or (None, why)"""
    key = (int(tuning[0]), bool(tuning[1]), bool(tuning[2]))
    hit = TRAFFIC.get(workload, {}).get(key)
    if hit is None:
        return None, "no counter profile of the picked variant %s on this workload" % (key,)
    return hit[0], "%s (variant lanes=%d staged=%s nontemporal=%s)" % (hit[1], *key)


def members(workload, precision):
    if workload == "batch64":
        out = []
        for side, scale, seed in wl.batch64_members():
            c = wl.grid_config(side, dt=wl.BATCH64["dt"], stiffness_scale=scale)
            c["precision"] = precision
            out.append((c, wl.jittered_positions(c, seed)))
        return out
    if workload == "single257":
        c = wl.grid_config(257, dt=wl.SINGLE_LARGE["dt"])
        c["precision"] = precision
        return [(c, wl.jittered_positions(c, wl.SINGLE_LARGE["seed"]))]
    if workload == "single21":
        c = wl.grid_config(21, dt=wl.SINGLE_SMALL["dt"])
        c["precision"] = precision
        return [(c, wl.jittered_positions(c, wl.SINGLE_SMALL["seed"]))]
    if workload == "irregular":
        c = wl.irregular_config()
        c["precision"] = precision
        pos, _ = wl.delaunay_rest_positions(c)
        return [(c, wl.jitter(pos, wl.IRREGULAR["seed"]))]
    raise ValueError(workload)


def describe(workload, ms):
    verts = sum(len(x) // 3 for _, x in ms)
    names = {
        "batch64": "64 cloths, 17..129 vertices/side (seeded), stiffness x[0.1,10], dt 0.01 s",
        "single257": "one 257x257-vertex cloth, dt 0.005 s",
        "single21": "one 21x21-vertex cloth, dt 0.01 s",
        "irregular": "one Delaunay cloth of 1.5M seeded points (1 mm min spacing), 2 m x 2 m, "
                     "top side pinned, dt 0.002 s",
    }
    return {"workload": names[workload], "cloths": len(ms), "vertices": verts,
            "tolerance": ms[0][0].get("solver_tolerance", 1e-6),
            "precision": "mixed (FP32 matrix and recurrences; FP64 dots, solution, reliable updates)"}


def time_batch(ms, steps, warmup, graphs=True):
    b = capi.ClothBatch([c for c, _ in ms], [x for _, x in ms])
    b.set_graphs(graphs)
    b.tuning = b.autotune_all()
    for _ in range(warmup):
        b.step(1)
    lat = []
    iters = []
    worst = 0.0
    per_cloth_iters = np.zeros(b.n)
    critical = []
    fp64_matrix = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        b.step(1)
        lat.append(time.perf_counter() - s0)
        st = [b.stats(i) for i in range(b.n)]
        iters.append(np.mean([s["pcg_iterations"] for s in st]))
        critical.append(max(s["pcg_iterations"] + s["outer_iterations"] for s in st))
        per_cloth_iters += [s["pcg_iterations"] for s in st]
        fp64_matrix += sum(s["matrix_fp64"] for s in st)
        worst = max(worst, max(s["rel_residual"] for s in st))
    total = time.perf_counter() - t0
    b.per_cloth_iters = per_cloth_iters
    b.critical = np.array(critical, dtype=np.float64)
    b.fp64_matrix_fraction = fp64_matrix / float(steps * b.n)
    b.final_pos = [b.get_pos(i) for i in range(b.n)]
    return b, total, np.array(lat), float(np.mean(iters)), worst


def step_bytes(b, fp32=True):
    """algorithmic bytes of the timed steps' PCG iterations, per iteration of
    a cloth: its SpMV (matrix block + column index per true block; z, d_prev
    read, d, q written per vertex) and update (d, q, r, P read, dx32 read and
    written, r, z written): 40 nnzb + 144 V in the mixed form"""
    total = 0.0
    for i in range(b.n):
        total += b.per_cloth_iters[i] * (40.0 * b.num_blocks(i) + 144.0 * b.nv[i])
    return total


def sample_state(b):
    """two untimed steps after the timed ones: the state (x, x_pre) before
    the second and the GPU's result and PCG iteration count of it"""
    b.step(1)
    x_pre = [b.get_pos(i) for i in range(b.n)]
    b.step(1)
    x = [b.get_pos(i) for i in range(b.n)]
    b.step(1)
    x_new = [b.get_pos(i) for i in range(b.n)]
    iters = [b.stats(i)["pcg_iterations"] for i in range(b.n)]
    return x, x_pre, x_new, iters


def time_plain(ms, steps, warmup):
    """Every cloth on its own, FP64, direct launches.  Returns the time, the
    worst residual and the final positions."""
    objs = []
    for c, x in ms:
        c = dict(c, precision="fp64")
        b = capi.ClothBatch([c], [x])
        b.set_graphs(False)
        objs.append(b)
    for _ in range(warmup):
        for b in objs:
            b.step(1)
    worst = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        for b in objs:
            b.step(1)
            worst = max(worst, b.stats(0)["rel_residual"])
    t = time.perf_counter() - t0
    return t, worst, [b.get_pos(0) for b in objs]


def host_cpu():
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cpu


def cpu_baseline(ms, state, budget_s=20.0):
    """The CPU baseline: the oracle (the C restatement of the reference's
    plain method, bit-exact to the reference's compiled CPU step on every
    golden fixture; gcc -O3, no FMA contraction) in tolerance mode (1e-6),
    every 4th cloth until the budget is spent, each stepped once from the
    GPU's state after the timed steps -- one thread, and the same sample again
    with one thread per cloth on up to 16 host cores.
    Returns (cpu_baseline, all-cores baseline, per-result error): the error
    is the GPU's step from that state against the oracle solved to 1e-10
    from it, max over the sample of max |x_gpu - x_oracle| / max |x_oracle - x|."""
    import concurrent.futures as cf

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ffi as of

    x, x_pre, x_new, iters = state
    order = list(range(0, len(ms), 4)) if len(ms) > 4 else list(range(len(ms)))

    def one_step(i, tol):
        oc = of.OracleCloth(ms[i][0])
        oc.set_state(x[i], x_pre[i])
        t0 = time.perf_counter()
        oc.step(of.SOLVE_TOLERANCE, tol)
        return time.perf_counter() - t0, oc.pos()

    t_steps, done = 0.0, []
    start = time.perf_counter()
    for i in order:
        t_steps += one_step(i, 1e-6)[0]
        done.append(i)
        if time.perf_counter() - start > budget_s:
            break
    what = ("%d of %d cloths (every 4th until %.0f s), one step each from the GPU's state "
            "after the timed steps (GPU PCG iterations there %d..%d), tol 1e-6; the oracle "
            "(C restatement of the reference's plain method), gcc -O3; host %s"
            % (len(done), len(ms), budget_s, min(iters[i] for i in done),
               max(iters[i] for i in done), host_cpu()))
    base = {"value": len(done) / t_steps, "unit": "cloth steps/s", "cores": 1, "kind": "port",
            "sample": what + "; serial"}
    cores = min(16, os.cpu_count() or 1, len(done))
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        t0 = time.perf_counter()
        list(ex.map(lambda i: one_step(i, 1e-6), done))
        wall = time.perf_counter() - t0
    par = {"value": len(done) / wall, "unit": "cloth steps/s", "cores": cores, "kind": "port",
           "sample": what + "; one thread per cloth, %d threads" % cores}

    def check(i):
        xt = one_step(i, 1e-10)[1]
        return np.abs(x_new[i] - xt).max() / np.abs(xt - x[i]).max()

    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 4)) as ex:
        errs = list(ex.map(check, done))
    err = {"gpu_max_err_over_step": max(errs),
           "sample": "the CPU baseline's %d cloths: the GPU's step from the state after the "
                     "timed steps vs the oracle solved to 1e-10 from the same state" % len(done)}
    return base, par, err


def cpu_baseline_extrapolated(ms, gpu_iters):  # noqa: C901
    """For a cloth too large to step on the CPU within the budget: the
    oracle's time for a step with k PCG iterations is assembly + k * per
    iteration; both are measured on the same mesh (steps stopped after 6 and
    after 26 PCG iterations) and combined with the GPU's mean iteration count
    per step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ffi as of

    (c, x), = ms
    pos, tri = wl.delaunay_rest_positions(c)
    oc = of.OracleCloth(c, mesh=(pos, tri))
    oc.set_state(x, x)
    times = {}
    for k in (6, 26):
        t0 = time.perf_counter()
        oc.step(of.SOLVE_TOLERANCE, 0.0, k)
        times[k] = time.perf_counter() - t0
    per_iter = (times[26] - times[6]) / 20
    base = times[6] - 6 * per_iter
    step_s = base + gpu_iters * per_iter
    return {"value": 1.0 / step_s, "unit": "cloth steps/s", "cores": 1, "kind": "port",
            "sample": "extrapolated: oracle assembly %.2f s + %.1f PCG iterations (the GPU's "
                      "mean per step) x %.3f s per iteration, both timed on the same 1.5M-point "
                      "mesh; the oracle (C restatement), gcc -O3, serial; the reference itself "
                      "cannot build this mesh (its Delaunay generator is O(V^2)); host %s"
                      % (base, gpu_iters, per_iter, host_cpu())}


def trajectory_error(step, x_mixed, x_fp64):
    """The 257 x 257 cloth after `step` steps from x0 against the oracle's own
    FP64 trajectory (tests/golden/grid257_tol200.npz, a strided vertex
    subset), when the fixture holds that step: max deviation over the
    fixture's max displacement and the relative L2 deviation, for the mixed
    and the FP64 GPU form, beside the fixture's drift between two oracle
    trajectories started 1e-18 m apart (chaos_step)."""
    path = os.path.join(ROOT, "tests", "golden", "grid257_tol200.npz")
    if not os.path.exists(path):
        return {"note": "fixture grid257_tol200.npz not present"}
    g = np.load(path)
    if "x_step%d" % step not in g:
        return {"note": "the fixture keeps steps %s, not %d" % (list(g["steps"]), step)}
    sel = g["vertex_subset"]
    xr = g["x_step%d" % step]
    disp = np.abs(xr - g["x0_subset"]).max()
    out = {"step": step, "chaos_step": float(g["chaos_step%d" % step])}
    for name, x in (("mixed", x_mixed), ("fp64", x_fp64)):
        if x is None:
            continue
        xs = x.reshape(-1, 3)[sel].reshape(-1)
        out[name] = {"max_err_over_displacement": float(np.abs(xs - xr).max() / disp),
                     "rel_l2": float(np.linalg.norm(xs - xr) / np.linalg.norm(xr))}
    return out


def roofline_only(args, ms):
    if args.copy_calibration:
        import torch
        n = 64 << 20  # 256 MiB of float32
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("bench: copied %d bytes x3" % (4 * n), file=sys.stderr, flush=True)
    b = capi.ClothBatch([c for c, _ in ms], [x for _, x in ms])
    lanes, staged, nt, tune = b.autotune_all()
    us, nbytes = b.time_solver_kernel(reps=200)
    floor = b.time_launch_floor(reps=200)
    b.close()
    achieved = nbytes / us / 1e3
    print(json.dumps({"workload": args.workload, "kernel": kernel_name(lanes, staged, nt, batch_slices(ms)),
                      "tuning_us": {f: {str(k): float(v) for k, v in t.items()}
                                    for f, t in tune.items()},
                      "launches": 20 * sum(int(v > 0) for t in tune.values()
                                           for v in t.values()) + 201,
                      "launch_us": us, "alg_bytes": nbytes, "achieved_GBs": achieved,
                      "frac": achieved / HBM_PEAK_GBS, "launch_floor_us": floor,
                      "frac_if_at_floor": nbytes / floor / 1e3 / HBM_PEAK_GBS}), flush=True)


def serve(args):
    """Serving workload: requests arrive as a Poisson stream (seeded), each a
    unit-square cloth (regular grid or seeded Delaunay points) of log-uniform
    vertex count asking for 10 steps; the
    server admits them against a 200 GB device budget and solves whatever is
    resident together.  Reports throughput and request latency."""
    work = wl.serve_requests(args.requests, rate=args.rate)
    s = capi.Server(budget_bytes=200e9, max_resident=1024)
    print("bench: serving %d requests at %.0f req/s" % (len(work), args.rate), file=sys.stderr,
          flush=True)
    ids = []
    t0 = time.perf_counter()
    for t, c, x0 in work:
        delay = t0 + t - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(s.submit(c, x0, wl.SERVE["steps"]))
    for k, rid in enumerate(ids):
        if not s.wait(rid, 1200.0):
            raise RuntimeError("request %d did not finish" % k)
        if (k + 1) % 16 == 0:
            print("bench: %d requests done" % (k + 1), file=sys.stderr, flush=True)
    infos = [s.info(r) for r in ids]
    counters = s.counters()
    s.close()
    if args.dump_requests:
        with open(args.dump_requests, "w") as f:
            json.dump(infos, f)
    bad = [i for i in infos if i["status"] != 0]
    if bad:
        raise RuntimeError("%d requests failed; first: %s" % (len(bad), bad[0]))
    lat = np.array([i["t_done"] - i["t_submit"] for i in infos])
    # per step of a request once it runs: (done - first step) / steps
    step_lat = np.array([(i["t_done"] - i["t_first_step"]) / max(1, i["steps_done"]) for i in infos])
    queue = np.array([i["t_admit"] - i["t_submit"] for i in infos])
    t_first = min(i["t_submit"] for i in infos)
    makespan = max(i["t_done"] for i in infos) - t_first
    steps = sum(i["steps_done"] for i in infos)
    vsteps = sum(i["steps_done"] * i["num_vertices"] for i in infos)
    out = {
        "metric": METRIC,
        "value": steps / makespan,
        "unit": "cloth steps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "serving: Poisson %.0f req/s, %d requests, unit-square cloths "
                               "(regular grid or seeded Delaunay, even odds) with V log-uniform "
                               "in [1e3, 2e5], 10 steps each, dt 0.01 s, tol 1e-6, 200 GB budget"
                               % (args.rate, len(infos))},
        "delaunay_requests": sum(1 for _, c, _ in work if c["geometry_type"] == "delaunay"),
        "requests_per_s": len(infos) / makespan,
        "vertex_steps_per_s": vsteps / makespan,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "step_latency_p50_ms": 1000 * float(np.percentile(step_lat, 50)),
        "step_latency_p99_ms": 1000 * float(np.percentile(step_lat, 99)),
        "queue_p50_s": float(np.percentile(queue, 50)),
        "makespan_s": makespan,
        "max_resident": counters["max_resident_seen"],
        "pcg_iterations_mean": float(np.mean([i["pcg_iterations"] / i["steps_done"] for i in infos])),
        "max_rel_residual": max(i["max_rel_residual"] for i in infos),
        "solver_chunks": counters["chunks"],
    }
    print(json.dumps(out), flush=True)


def drape(args):
    """Draping workload: one 129x129 cloth (0.5 m, unpinned) released
    1.5 mm above a sphere of radius 0.15 m (9,024 triangles); before every
    step the device detects the cloth's contacts with the sphere (all
    vertex-triangle, triangle-vertex and edge-edge pairs below 3 mm), and the
    step applies the collision response (collision Hessian in W, repulsion
    impulses on the solution).  Reports steps/s including detection, and the
    detection and step times."""
    conf = wl.drape_config()
    conf["precision"] = "mixed"
    bp, bt = wl.drape_body()
    cl = capi.BaraffClothHip(conf)
    body = capi.KinematicBody(bp, bt)
    print("bench: drape, %d steps after %d warm-up" % (args.steps, args.warmup), file=sys.stderr,
          flush=True)
    det, stp, n_pt, n_ee, iters, worst = [], [], [], [], [], 0.0
    for k in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        ct = cl.detect_contacts(body, True)
        t1 = time.perf_counter()
        cl.UpdatePos()
        t2 = time.perf_counter()
        st = cl.last_stats()
        if k >= args.warmup:
            det.append(t1 - t0)
            stp.append(t2 - t1)
            n_pt.append(len(ct["pt_ids"]))
            n_ee.append(len(ct["ee_ids"]))
            iters.append(st["pcg_iterations"])
            worst = max(worst, st["rel_residual"])
    total = float(np.sum(det) + np.sum(stp))
    y = cl.GetPos().reshape(-1, 3)
    r = np.linalg.norm(y, axis=1)
    out = {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "cloth steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "drape: 129x129 cloth (0.5 m, unpinned) 1.5 mm above a sphere "
                               "(r 0.15 m, 4,514 vertices, 9,024 triangles), contacts detected "
                               "before every step, dt 0.01 s, tol 1e-6"},
        "detect_ms_mean": 1000 * float(np.mean(det)),
        "step_ms_mean": 1000 * float(np.mean(stp)),
        "point_triangle_contacts_mean": float(np.mean(n_pt)),
        "edge_edge_contacts_mean": float(np.mean(n_ee)),
        "pcg_iterations_mean": float(np.mean(iters)),
        "max_rel_residual": worst,
        "min_distance_to_centre": float(r.min()),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batch64",
                    choices=["batch64", "single257", "single21", "irregular", "serve", "drape"])
    ap.add_argument("--requests", type=int, default=64, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=wl.SERVE["rate"], help="serve: arrivals per s")
    ap.add_argument("--dump-requests", default=None, help="serve: write per-request records here")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--roofline-only", action="store_true",
                    help="assemble once and time the dominant kernel only")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="with --roofline-only: copy 256 MiB device-to-device first")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X; multi-GPU "
              "sharding (whole cloths per GPU, no per-iteration communication) is not part "
              "of it", file=sys.stderr)
        sys.exit(2)

    if args.workload == "serve":
        serve(args)
        return
    if args.workload == "drape":
        drape(args)
        return
    ms = members(args.workload, "mixed")
    if args.roofline_only:
        roofline_only(args, ms)
        return
    print("bench: %s, %d steps after %d warm-up" % (args.workload, args.steps, args.warmup),
          file=sys.stderr, flush=True)
    b, total, lat, iters, worst = time_batch(ms, args.steps, args.warmup)
    n = len(ms)
    value = n * args.steps / total
    whole_step_gbs = step_bytes(b) / total / 1e9
    state = None if args.no_cpu or args.workload == "irregular" else sample_state(b)
    us, nbytes = b.time_solver_kernel(reps=200)
    floor = b.time_launch_floor(reps=200)
    achieved = nbytes / us / 1e3  # GB/s
    traffic, traffic_src = traffic_for(args.workload, b.tuning)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cloth steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": describe(args.workload, ms),
        "p50_ms": 1000 * float(np.percentile(lat, 50)),
        "p99_ms": 1000 * float(np.percentile(lat, 99)),
        "pcg_iterations_mean": iters,
        # the slowest cloth's iterations (and true-residual checks) per step:
        # a step lasts as long as that cloth's chain of launches
        "critical_iterations_mean": float(np.mean(b.critical)),
        "us_per_critical_iteration": 1e6 * total / float(np.sum(b.critical)),
        # cloth-steps of the timed region whose solve read the FP64 matrix
        # (the mixed form's stall fallback, or its carry-over from an earlier
        # stalled solve of the same cloth)
        "fp64_matrix_fraction": b.fp64_matrix_fraction,
        "autotune": {"spmv_lanes_per_row": b.tuning[0], "spmv_staged": b.tuning[1],
                     "spmv_nontemporal": b.tuning[2],
                     "spmv_us": {f: {str(k): float(v) for k, v in t.items()}
                                 for f, t in b.tuning[3].items()}},
        "max_rel_residual": worst,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": kernel_name(*b.tuning[:3], batch_slices(ms)),
                     "launch_us": us, "alg_bytes": nbytes,
                     # an empty kernel of the same grid and workgroup, launched
                     # back to back: no launch of that shape is shorter, so
                     # alg_bytes over it bounds the fraction one launch can reach
                     "launch_floor_us": floor,
                     "frac_if_at_floor": nbytes / floor / 1e3 / HBM_PEAK_GBS},
        "whole_step": {"achieved": whole_step_gbs, "unit": "GB/s",
                       "frac": whole_step_gbs / HBM_PEAK_GBS,
                       "bytes": "per PCG iteration of a cloth 40 nnzb + 144 V (SpMV and update "
                                "launches), times its iterations over the timed steps, over "
                                "the timed wall time (assembly and host gaps included in the "
                                "time, not in the bytes)"},
    }
    b.close()
    x_plain = None
    if not args.no_plain:
        print("bench: plain FP64 form", file=sys.stderr, flush=True)
        tp, worst_p, x_plain = time_plain(ms, args.steps, args.warmup)
        plain = n * args.steps / tp
        out["plain_fp64"] = {"value": plain, "unit": "cloth steps/s",
                             "max_rel_residual": worst_p, "speedup": value / plain}
    if args.workload == "single257":
        out["trajectory_error"] = trajectory_error(args.warmup + args.steps, b.final_pos[0],
                                                   x_plain[0] if x_plain else None)
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        if args.workload == "irregular":
            out["cpu_baseline"] = cpu_baseline_extrapolated(ms, iters)
        else:
            (out["cpu_baseline"], out["cpu_baseline_all_cores"],
             out["per_result_error"]) = cpu_baseline(ms, state)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
