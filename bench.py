#!/usr/bin/env python
"""bench.py -- throughput of the voxelizer node pipeline on one MI355X.

One step = one pass of the hot path over synthetic meshes already resident in
HBM: grid planning (bounding box + dimensions), solid voxelization, surface
voxelization with material winners, erosion to the fixed point and boundary-id
node write, ending with the node grids resident in HBM
(Voxelizer<Node>::voxelizeToNodes with device-pointer output).

Workloads:
  hall (default) the 1.2 G-voxel large hall: 1536x1024x768 ShortNodes,
                 ~1.4M triangles, 8 material ids, voxelizeToNodes(1/64 m,
                 voxSplitRes=(1024,1024,768), matSplitRes=(512,512,512));
  batch          256 meshes (icospheres, rooms, boxes; 500..200k triangles,
                 64..512 voxels along the longest side), one batch call per step;
  room, fcc, icosphere  the shoebox room (LongNode), the FCC room (LongFCCNode)
                 and the 128^3 icosphere.

Schedules: "plain" processes surface bricks in index order; "learned" takes
the bricks the MLP cost model predicts to be expensive first (cut into
sub-bricks where there are few bricks per resident workgroup; the model is
trained first by the library's own GPU kernels on per-brick durations
recorded on a differently seeded mesh).  A batch runs
fused under both (every stage one launch over all meshes, bricks of all
meshes in one list), so learned / plain is the model's own share; the batch
also reports the meshes one after another.
Executions: "graph" replays a HIP graph of the pipeline captured once
(vox_graph_*; the step still plans the grid first), "eager" launches it call
by call.  Every (schedule, execution) is timed; `value` is the one named by
--schedule / --execution (default learned, graph).

Prints ONE JSON line.  `value` = allocated nodes per step / wall time per step
(Gvoxels/s), timed over --steps steps between two device synchronisations.
`roofline` is for the node-write kernel (k_nodes_cart / k_nodes_fcc): its
algorithmic bytes (one node written per node, N * sizeof(Node)) over its
average HIP-event duration on the library's stream, from a second run of the
same steps with events on (eager execution; events cost a synchronisation per
step, so the timed steps run without them); `pipeline_frac` puts the whole
step's algorithmic bytes (N * sizeof(Node) + 12 V + 12 T + T) over the step
time;
`traffic` is the HBM bytes per launch that rocprofv3 counted for it
(profiles/traffic_*.json: FETCH_SIZE and WRITE_SIZE scaled by a plain copy).
`cpu_baseline` times the CPU oracle (the restatement of the reference
algorithm) on the host cores: all cores on the full workload (`value`) and
one thread on a bounded sample of the same mesh.  `check_vs_plain` compares
the reported configuration's nodes with a plain eager run's.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
METRIC = "Gvoxels/sec solid+classify, learned vs plain schedule, bit-exact; % HBM roofline"

WORKLOADS = {
    # name: (mesh builder, plan mode, plan param, node kind, vox split, mat split, label)
    "hall": ("hall", 1, 1.0 / 64.0, 0, (1024, 1024, 768), (512, 512, 512),
             "large hall, ~1.4M triangles, 1536x1024x768 ShortNode, 8 material ids"),
    "room": ("room", 1, 0.01, 1, (1024, 512, 512), (512, 512, 512),
             "shoebox room 5x4x3 m + 64 boxes, 19212 triangles, 512x403x303 LongNode, ids 1..8"),
    "fcc": ("room", 1, 4.0 ** (1.0 / 3.0) * 0.01, 4, (1024, 512, 512), (512, 512, 512),
            "shoebox room on the FCC lattice, 640x255x384 LongFCCNode, ids 1..8"),
    "icosphere": ("icosphere", 0, 126, 0, (1024, 512, 512), (512, 512, 512),
                  "icosphere 1280 triangles, maxDimension 126 -> 128^3 ShortNode, 1 material"),
    "batch": ("batch", 1, None, None, (1024, 1024, 1024), (512, 512, 512),
              "256 meshes (icospheres, rooms, boxes), 500..200k triangles, 64..512 voxels "
              "on the longest side, ShortNode (ids <= 7) or LongNode"),
}


def build_mesh(name, seed=None):
    from voxelizer_amd import workloads as W
    if name == "hall":
        return W.hall() if seed is None else W.hall(seed=seed)
    if name == "room":
        return W.shoebox_room() if seed is None else W.shoebox_room(seed=seed)
    return W.icosphere() if seed is None else W.icosphere(seed=seed)


def node_bytes(kind):
    return {0: 1, 1: 2, 3: 2, 4: 4}[kind]


def host_cores():
    """Threads the CPU baseline uses: the host cores this process may run on,
    at most 16 (the GPU box's CPU share)."""
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return min(cores, 16)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(mesh, wl, budget_s):
    """The CPU oracle (the restatement of the reference algorithm) on the
    host: all cores (up to 16 threads) on the full workload, and one thread on
    the same mesh at a coarser cube length sized to the budget (the full
    1.2 G-node hall takes about a minute on one core).  value = all cores."""
    from oracle import oracle_py as O
    _, mode, param, kind, vs, ms, _ = wl
    cores = host_cores()
    fcc = 4 if kind in (3, 4) else 1

    def timed(g, threads):
        n = g.alloc[0] * g.alloc[1] * g.alloc[2] * fcc
        t0 = time.perf_counter()
        O.voxelize_nodes(mesh, g, kind, True, vox_split=vs, mat_split=ms, n_threads=threads)
        return n, time.perf_counter() - t0

    g = O.plan_grid(mesh.vertices, mode, param)
    n, dt = timed(g, cores)
    out = {"value": round(n / dt / 1e9, 4), "unit": "Gvoxels/s", "cores": cores, "kind": "port",
           "sample": f"full workload, {n} nodes in {dt:.2f} s on {cores} threads",
           "nproc": os.cpu_count(), "cpu_model": cpu_model()}
    # One thread: about 8x the all-core time per node on this box's share;
    # a coarser grid of the same mesh keeps it inside the budget.
    est1 = dt * max(cores, 1) * 0.7
    g1, sample1 = g, "full workload"
    if est1 > budget_s and mode == 1:
        scale = (budget_s / est1) ** (1.0 / 3.0)
        g1 = O.plan_grid(mesh.vertices, mode, param / scale)
        sample1 = (f"same mesh at cube length x{1 / scale:.2f} "
                   f"({g1.alloc[0]}x{g1.alloc[1]}x{g1.alloc[2]})")
    n1, dt1 = timed(g1, 1)
    out["single_thread"] = {"value": round(n1 / dt1 / 1e9, 4), "unit": "Gvoxels/s",
                            "sample": f"{sample1}, {n1} nodes in {dt1:.2f} s"}
    return out


def run_cpu_baseline_batch(meshes, kinds, budget_s):
    """The CPU oracle over the batch's meshes in order until the budget is used."""
    from oracle import oracle_py as O
    cores = host_cores()
    n = 0
    k = 0
    t0 = time.perf_counter()
    for m, kind in zip(meshes, kinds):
        g = O.plan_grid(m.vertices, 1, m.meta["cube_length"])
        O.voxelize_nodes(m, g, kind, True, vox_split=(1024, 1024, 1024), n_threads=cores)
        n += g.alloc[0] * g.alloc[1] * g.alloc[2]
        k += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": round(n / dt / 1e9, 4), "unit": "Gvoxels/s", "cores": cores, "kind": "port",
            "sample": f"first {k} of {len(meshes)} meshes; {n} nodes in {dt:.2f} s"}


def load_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def train_cost_model(ctx, va, mesh_name, mode, param, kind, vs, ms):
    """Records per-brick durations on a differently seeded mesh of the same
    family and trains the cost model on them with the library's kernels."""
    t0 = time.perf_counter()
    train = build_mesh(mesh_name, seed=5)
    ctx.set_mesh(train.vertices, train.indices)
    ctx.set_materials(train.materials, int(train.materials.max()) + 1)
    g = ctx.plan_grid(mode, param)
    buf = va.DeviceBuffer(ctx, ctx.node_count(g, kind) * node_bytes(kind))
    ctx.cost_record(True)
    ctx.voxelize_nodes_device(g, kind, buf.ptr.value, True, vs, ms)
    ctx.sync()
    f, c = ctx.cost_samples()
    ctx.cost_record(False)
    buf.free()
    mse = ctx.cost_train(f, c, epochs=20, batch=256, lr=2e-3, seed=0)
    return {"training_mesh": f"{train.name} seed 5", "samples": int(c.size),
            "mse_standardized": round(mse, 4), "seconds": round(time.perf_counter() - t0, 2)}


def bench_single(args, va, wl):
    mesh_name, mode, param, kind, vs, ms, label = wl
    t0 = time.perf_counter()
    mesh = build_mesh(mesh_name)
    print(f"# mesh {mesh.name}: {mesh.nv} vertices, {mesh.nt} triangles "
          f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    ctx = va.Context(0)
    scheds = [va.SCHEDULE_PLAIN, va.SCHEDULE_LEARNED]
    if args.only:
        scheds = [va.SCHEDULE_LEARNED if args.schedule == "learned" else va.SCHEDULE_PLAIN]
    training = None
    if va.SCHEDULE_LEARNED in scheds:
        training = train_cost_model(ctx, va, mesh_name, mode, param, kind, vs, ms)
        print(f"# cost model: {training}", file=sys.stderr, flush=True)
    vbuf = va.DeviceBuffer(ctx, mesh.vertices.nbytes)
    ibuf = va.DeviceBuffer(ctx, mesh.indices.nbytes)
    mbuf = va.DeviceBuffer(ctx, mesh.materials.nbytes)
    vbuf.upload(mesh.vertices)
    ibuf.upload(mesh.indices)
    mbuf.upload(mesh.materials)
    ctx.set_mesh_device(vbuf.ptr.value, mesh.nv, ibuf.ptr.value, mesh.nt)
    ctx.set_materials_device(mbuf.ptr.value, int(mesh.materials.max()) + 1)
    grid = ctx.plan_grid(mode, param)
    n_nodes = ctx.node_count(grid, kind)
    nb = node_bytes(kind)
    out = va.DeviceBuffer(ctx, n_nodes * nb)

    def step():
        g = ctx.plan_grid(mode, param)
        ctx.voxelize_nodes_device(g, kind, out.ptr.value, True, vs, ms)

    # Surface brick shape picked by the autotuner (before the timed steps).
    if args.autotune == "on":
        bdims, bms = ctx.autotune_bricks(grid, kind, out.ptr.value, 2, vs, ms)
        print(f"# autotuned brick {bdims} ({bms:.3f} ms)", file=sys.stderr, flush=True)
    else:
        bdims = ctx.brick_dims(kind in (va.NODE_SHORT_FCC, va.NODE_LONG_FCC))

    # Graph replay of the same step (the mesh as a one-item batch on the
    # grid planned above; the step still plans the grid first).
    item = (va.BatchItem * 1)()
    it = item[0]
    it.vertices_dev, it.nv = vbuf.ptr.value, mesh.nv
    it.indices_dev, it.nt = ibuf.ptr.value, mesh.nt
    it.materials_dev, it.n_unique = mbuf.ptr.value, int(mesh.materials.max()) + 1
    it.kind, it.grid = kind, grid
    it.vox_split[:] = vs
    it.mat_split[:] = ms
    it.nodes_dev = out.ptr.value

    def timed(run):
        for _ in range(args.warmup):
            run()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run()
        ctx.sync()
        return time.perf_counter() - t0

    execs = ["eager", "graph"] if not args.only else [args.execution]
    results = {}  # (schedule, execution) -> seconds for the timed steps
    for sched in scheds:
        ctx.set_schedule(sched)
        for ex in execs:
            if ex == "eager":
                results[(sched, ex)] = timed(step)
            else:
                gr = ctx.graph(item, sched)
                results[(sched, ex)] = timed(lambda: (ctx.plan_grid(mode, param), gr.launch()))
                gr.close()
    chosen = va.SCHEDULE_LEARNED if args.schedule == "learned" else va.SCHEDULE_PLAIN
    # HIP events on the library stream over the same number of steps (eager
    # execution, chosen schedule): per-stage times and the node kernel's
    # duration for the roofline.
    ctx.set_schedule(chosen)
    ctx.enable_timing(True)
    node_ms, stage = [], {}
    for _ in range(args.steps):
        step()
        st = ctx.stage_times()
        node_ms.append(st["nodes_ms"])
        for k, v in st.items():
            stage[k] = stage.get(k, 0.0) + v
    ctx.enable_timing(False)
    stage = {k: v / args.steps for k, v in stage.items()}
    ctx.set_schedule(va.SCHEDULE_PLAIN)

    # The nodes the reported configuration left in `out` against a plain,
    # eager run of the same step into a second buffer.
    check_plain = None
    if (chosen, args.execution) != (va.SCHEDULE_PLAIN, "eager"):
        ctx.set_schedule(chosen)
        if args.execution == "graph":
            gr = ctx.graph(item, chosen)
            gr.launch()
            gr.close()
        else:
            step()
        ctx.sync()
        got = out.download(np.uint8, n_nodes * nb)
        ctx.set_schedule(va.SCHEDULE_PLAIN)
        ref_buf = va.DeviceBuffer(ctx, n_nodes * nb)
        g0 = ctx.plan_grid(mode, param)
        ctx.voxelize_nodes_device(g0, kind, ref_buf.ptr.value, True, vs, ms)
        check_plain = bool(np.array_equal(got, ref_buf.download(np.uint8, n_nodes * nb)))
        ref_buf.free()
        del got
    elapsed = results[(chosen, args.execution)]
    ms_per_step = elapsed * 1e3 / args.steps
    value = n_nodes / (elapsed / args.steps) / 1e9
    gv = {f"{'learned' if s == va.SCHEDULE_LEARNED else 'plain'}_{ex}":
          n_nodes / (t / args.steps) / 1e9 for (s, ex), t in results.items()}

    kname = "k_nodes_fcc" if kind in (3, 4) else "k_nodes_cart"
    kern_ms = float(np.mean(node_ms))
    # Algorithmic bytes of the node kernel: one node written per node (N s);
    # the solid-bit read (1/8 B per node) is reported beside it.
    alg_bytes = n_nodes * nb
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    pipe_bytes = n_nodes * nb + 12 * mesh.nv + 12 * mesh.nt + mesh.nt
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(kname, args.workload),
            "kernel": kname, "kernel_ms": round(kern_ms, 4),
            "algorithmic_bytes": int(alg_bytes),
            "frac_with_bit_read": round(n_nodes * (nb + 0.125) / (kern_ms * 1e-3) / 1e9 /
                                        HBM_PEAK_GBS, 4),
            "pipeline_bytes": int(pipe_bytes),
            "pipeline_frac": round(pipe_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    cpu = run_cpu_baseline(mesh, wl, args.cpu_budget) if args.cpu_baseline == "on" else None
    check = None
    if args.check:
        from oracle import oracle_py as O
        go = O.plan_grid(mesh.vertices, mode, param)
        ref = O.voxelize_nodes(mesh, go, kind, True, vox_split=vs, mat_split=ms)
        check = bool(np.array_equal(ref, out.download(np.uint8, n_nodes * nb)))
    res = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gvoxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded procedural mesh)",
        "config": {"workload": label, "schedule": args.schedule, "execution": args.execution,
                   "nodes": n_nodes,
                   "dims": list(ctx.node_dims(grid, kind)), "triangles": mesh.nt,
                   "vox_split": list(vs), "mat_split": list(ms), "parallelism": "single GPU",
                   "surface_brick": list(bdims)},
        "schedules": schedules_block(gv, args.execution, training),
        "stages_ms": {k: round(v, 4) for k, v in stage.items()},
        "roofline": roof, "cpu_baseline": cpu,
    }
    if check_plain is not None:
        res["check_vs_plain"] = check_plain
    if check is not None:
        res["check_vs_oracle"] = check
    return res


def schedules_block(gv, execution, training):
    """Gvoxels/s of every (schedule, execution) timed, and learned / plain
    for the reported execution."""
    out = {"gvox_s": {k: round(v, 4) for k, v in gv.items()}}
    p, l = gv.get(f"plain_{execution}"), gv.get(f"learned_{execution}")
    if p:
        out["plain_gvox_s"] = round(p, 4)
    if l:
        out["learned_gvox_s"] = round(l, 4)
    if p and l:
        out["learned_over_plain"] = round(l / p, 4)
    out["cost_model"] = training
    return out


def bench_batch(args, va, wl):
    from voxelizer_amd import workloads as W
    _, _, _, _, vs, ms, label = wl
    t0 = time.perf_counter()
    meshes = W.batch_meshes()
    print(f"# batch: {len(meshes)} meshes, {sum(m.nt for m in meshes)} triangles "
          f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    kinds = [va.NODE_SHORT if int(m.materials.max()) <= 7 else va.NODE_LONG for m in meshes]
    ctx = va.Context(0)
    training = None
    if not (args.only and args.schedule == "plain"):
        training = train_cost_model(ctx, va, "room", 1, 0.01, va.NODE_LONG, (1024, 512, 512),
                                    (512, 512, 512))
    items = (va.BatchItem * len(meshes))()
    keep = []
    n_nodes = 0
    out_bytes = 0
    node_alg = 0.0
    for i, (m, kind) in enumerate(zip(meshes, kinds)):
        g = va.plan_grid_host(m.vertices, va.PLAN_CUBE_LENGTH, m.meta["cube_length"], vs[0])
        bufs = [va.DeviceBuffer(ctx, a.nbytes) for a in (m.vertices, m.indices, m.materials)]
        for b, a in zip(bufs, (m.vertices, m.indices, m.materials)):
            b.upload(np.ascontiguousarray(a))
        nn = ctx.node_count(g, kind)
        out = va.DeviceBuffer(ctx, nn * node_bytes(kind))
        keep.extend(bufs + [out])
        it = items[i]
        it.vertices_dev, it.nv = bufs[0].ptr.value, m.nv
        it.indices_dev, it.nt = bufs[1].ptr.value, m.nt
        it.materials_dev, it.n_unique = bufs[2].ptr.value, int(m.materials.max()) + 1
        it.kind = kind
        it.grid = g
        it.vox_split[:] = vs
        it.mat_split[:] = ms
        it.nodes_dev = out.ptr.value
        n_nodes += nn
        out_bytes += nn * node_bytes(kind)
        node_alg += nn * node_bytes(kind)  # node write
    # Modes: (brick order, streams, graph); streams 0 = fused (each stage one
    # launch over all meshes), 1 = one mesh after another, k = k worker
    # streams; graph = the fused batch captured once and replayed.  The
    # plain schedule the learned one is compared with is the same fused
    # execution with bricks in index order.
    modes = {"sequential_plain": (va.SCHEDULE_PLAIN, 1, False),
             "fused_plain": (va.SCHEDULE_PLAIN, 0, False),
             "learned": (va.SCHEDULE_LEARNED, 0, False),
             "fused_plain_graph": (va.SCHEDULE_PLAIN, 0, True),
             "learned_graph": (va.SCHEDULE_LEARNED, 0, True)}
    graph_x = args.execution == "graph"
    reported = ("learned_graph" if graph_x else "learned") if args.schedule == "learned" \
        else ("fused_plain_graph" if graph_x else "fused_plain")
    plain_mode = "fused_plain_graph" if graph_x else "fused_plain"
    learned_mode = "learned_graph" if graph_x else "learned"
    if args.only:  # profiling one mode
        modes = {reported: modes[reported]}
    results = {}
    for name, (sched, streams, graph) in modes.items():
        if graph:
            gr = ctx.graph(items, sched)
            run = gr.launch
        else:
            def run(sched=sched, streams=streams):
                ctx.voxelize_batch(items, sched, streams)
        for _ in range(args.warmup):
            run()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run()
        ctx.sync()
        results[name] = time.perf_counter() - t0
        if graph:
            gr.close()
        print(f"# batch {name}: {results[name] * 1e3 / args.steps:.2f} ms/step", file=sys.stderr,
              flush=True)
    # Per-stage HIP-event times of one fused run of the reported schedule.
    sched_r = va.SCHEDULE_LEARNED if args.schedule == "learned" else va.SCHEDULE_PLAIN
    ctx.enable_timing(True)
    ctx.voxelize_batch(items, sched_r, 0)
    stages = ctx.stage_times()
    ctx.enable_timing(False)
    elapsed = results[reported]
    gv = {k: n_nodes / (v / args.steps) / 1e9 for k, v in results.items()}
    plain_v, learned_v = gv.get(plain_mode), gv.get(learned_mode)

    def ratio(a, b):
        return round(a / b, 4) if a and b else None
    value = n_nodes / (elapsed / args.steps) / 1e9
    ms_per_step = elapsed * 1e3 / args.steps
    pipe_bytes = out_bytes + sum(12 * m.nv + 13 * m.nt for m in meshes)
    cpu = (run_cpu_baseline_batch(meshes, kinds, args.cpu_budget)
           if args.cpu_baseline == "on" else None)
    return {
        "metric": METRIC, "value": round(value, 4), "unit": "Gvoxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded procedural meshes)",
        "config": {"workload": label, "schedule": args.schedule, "meshes": len(meshes),
                   "nodes": n_nodes, "triangles": sum(m.nt for m in meshes),
                   "execution": "fused batch, replayed HIP graph" if graph_x else
                                "fused batch (each stage one launch over all meshes)",
                   "parallelism": "single GPU"},
        "schedules": {"plain_gvox_s": plain_v and round(plain_v, 4),
                      "learned_gvox_s": learned_v and round(learned_v, 4),
                      "learned_over_plain": ratio(learned_v, plain_v),
                      "plain_is": f"{plain_mode}: the same fused execution, bricks in index order",
                      "modes_gvox_s": {k: round(v, 4) for k, v in gv.items()},
                      "fused_over_sequential": ratio(gv.get("fused_plain"),
                                                     gv.get("sequential_plain")),
                      "stages_ms": stages, "cost_model": training},
        "roofline": {"bound": "hbm", "achieved": round(node_alg / (stages["nodes_ms"] * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(node_alg / (stages["nodes_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": "k_nodes_cart + k_nodes_fcc (node stage of the fused batch)",
                     "kernel_ms": round(stages["nodes_ms"], 4), "algorithmic_bytes": int(node_alg),
                     "pipeline_frac": round(pipe_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="hall", choices=sorted(WORKLOADS))
    ap.add_argument("--schedule", default="learned", choices=["learned", "plain"],
                    help="which schedule `value` reports (both are timed)")
    ap.add_argument("--only", action="store_true",
                    help="time only the schedule named by --schedule (plain: no cost-model "
                         "training; for profiling one schedule)")
    ap.add_argument("--execution", default="graph", choices=["graph", "eager"],
                    help="which execution `value` reports: a captured HIP graph replayed per "
                         "step, or the pipeline launched call by call (both are timed)")
    ap.add_argument("--autotune", default="on", choices=["on", "off"],
                    help="off: keep the default surface brick shape (profiling runs)")
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU work for the oracle baseline sample")
    ap.add_argument("--check", action="store_true",
                    help="compare the last output with the CPU oracle (slow for the hall)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (reference kernel of known "
                         "size for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 1 << 30
        src = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_each_way": n, "copies": 5}), flush=True)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1",
                          "n_gpus": args.gpus}))
        print("bench.py: only --gpus 1 is supported (the voxelizer runs the whole grid "
              "on one MI355X)", file=sys.stderr)
        return 2

    import voxelizer_amd as va
    wl = WORKLOADS[args.workload]
    res = bench_batch(args, va, wl) if args.workload == "batch" else bench_single(args, va, wl)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
