#!/usr/bin/env python
"""bench.py -- throughput of the voxelizer node pipeline on one MI355X.

One step = one pass of the hot path over one synthetic mesh already resident
in HBM: grid planning (bounding box + dimensions), solid voxelization,
surface voxelization with material winners, erosion to the fixed point and
boundary-id node write, ending with the node grid resident in HBM
(Voxelizer<Node>::voxelizeToNodes with device-pointer output).

Default workload: the 1.2 G-voxel large hall (1536x1024x768 ShortNodes,
~1.4M triangles, 8 material ids), voxelizeToNodes(1/64 m,
voxSplitRes=(1024,1024,768), matSplitRes=(512,512,512)).

Prints ONE JSON line.  `value` is Gvoxels/s = allocated nodes / wall time per
step, timed over --steps steps between two device synchronisations.
`roofline` is for the node-write kernel (k_nodes_cart / k_nodes_fcc): its
algorithmic bytes (one node write + one bit read per node) over its average
HIP-event duration on the library's stream.  `cpu_baseline` times the CPU
oracle (the restatement of the reference algorithm) on the host cores.
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
}


def build_mesh(name):
    from voxelizer_amd import workloads as W
    if name == "hall":
        return W.hall()
    if name == "room":
        return W.shoebox_room()
    return W.icosphere()


def node_bytes(kind):
    return {0: 1, 1: 2, 3: 2, 4: 4}[kind]


def run_cpu_baseline(mesh, wl, budget_s):
    """Times the CPU oracle on the host cores on this workload; if the full
    grid would exceed the budget, on a coarser grid of the same mesh (the
    sample is named in the output)."""
    from oracle import oracle_py as O
    _, mode, param, kind, vs, ms, _ = wl
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    cores = min(cores, 16)
    scale = 1.0
    g = O.plan_grid(mesh.vertices, mode, param)
    n_full = g.alloc[0] * g.alloc[1] * g.alloc[2] * (4 if kind in (3, 4) else 1)
    # About 60 ns per node per core on this oracle; keep inside the budget.
    est = n_full * 60e-9 * 8 / cores
    sample = "full workload"
    if est > budget_s and mode == 1:
        scale = (budget_s / est) ** (1.0 / 3.0)
        g = O.plan_grid(mesh.vertices, mode, param / scale)
        sample = f"same mesh at cube length x{1 / scale:.2f} ({g.alloc[0]}x{g.alloc[1]}x{g.alloc[2]})"
    n = g.alloc[0] * g.alloc[1] * g.alloc[2] * (4 if kind in (3, 4) else 1)
    t0 = time.perf_counter()
    O.voxelize_nodes(mesh, g, kind, True, vox_split=vs, mat_split=ms, n_threads=cores)
    dt = time.perf_counter() - t0
    return {"value": round(n / dt / 1e9, 4), "unit": "Gvoxels/s", "cores": cores,
            "kind": "port", "sample": f"{sample}; {n} nodes in {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="hall", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    ap.add_argument("--cpu-budget", type=float, default=25.0,
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
    mesh_name, mode, param, kind, vs, ms, label = wl
    t0 = time.perf_counter()
    mesh = build_mesh(mesh_name)
    gen_s = time.perf_counter() - t0
    print(f"# mesh {mesh.name}: {mesh.nv} vertices, {mesh.nt} triangles ({gen_s:.1f} s)",
          file=sys.stderr, flush=True)

    ctx = va.Context(0)
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

    for _ in range(args.warmup):
        step()
    ctx.sync()
    # Timed region: K steps between two synchronisations.
    ctx.enable_timing(True)
    node_ms = []
    stage = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        st = ctx.stage_times()
        node_ms.append(st["nodes_ms"])
        for k, v in st.items():
            stage[k] = stage.get(k, 0.0) + v
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ctx.enable_timing(False)
    ms_per_step = elapsed * 1e3 / args.steps
    value = n_nodes / (elapsed / args.steps) / 1e9
    for k in stage:
        stage[k] /= args.steps

    kern_ms = float(np.mean(node_ms))
    alg_bytes = n_nodes * (nb + 0.125)
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    pipe_bytes = n_nodes * nb + 12 * mesh.nv + 12 * mesh.nt + mesh.nt
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_nodes_fcc" if kind in (3, 4) else "k_nodes_cart",
            "kernel_ms": round(kern_ms, 4),
            "pipeline_frac": round(pipe_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    cpu = None
    if args.cpu_baseline == "on":
        cpu = run_cpu_baseline(mesh, wl, args.cpu_budget)
    check = None
    if args.check:
        from oracle import oracle_py as O
        go = O.plan_grid(mesh.vertices, mode, param)
        ref = O.voxelize_nodes(mesh, go, kind, True, vox_split=vs, mat_split=ms)
        got = out.download(np.uint8, n_nodes * nb)
        check = bool(np.array_equal(ref, got))
    res = {
        "metric": "Gvoxels/sec solid+classify (plain schedule), bit-exact; % HBM roofline",
        "value": round(value, 4), "unit": "Gvoxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded procedural mesh)",
        "config": {"workload": label, "nodes": n_nodes, "dims": list(ctx.node_dims(grid, kind)),
                   "triangles": mesh.nt, "vox_split": list(vs), "mat_split": list(ms),
                   "parallelism": "single GPU"},
        "stages_ms": {k: round(v, 4) for k, v in stage.items()},
        "roofline": roof, "cpu_baseline": cpu,
    }
    if check is not None:
        res["check_vs_oracle"] = check
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
