"""Per-voxel list of the nodes where the original CUDA build of the
reference could legitimately differ from this build (and from the oracle),
for the four single-mesh workloads.

Two causes (oracle/oracle.h, orc_fragile):
  * FMA contraction: nvcc's default build (--fmad=true) may contract the
    reference's multiply-adds; a test value within eps of its threshold
    (relative to the magnitude of its terms) can then fall the other way.
    eps64 = 1e-12 for the FP64 solid tests (about 2^13 ulps, far beyond what
    one contraction moves), eps32 = 1e-5 for the FP32 surface tests (about
    2^7 ulps).  A fragile crossing test marks the rest of its voxel row, whose
    parity it decides.
  * The racy material store of processVoxel (src/voxelizer.cu:2988): where
    the highest-class triangles writing a voxel carry different materials,
    any of them may survive there (this build keeps the largest
    (class, triangle) key).

Usage:
    python tests/golden/fragile_voxels.py            # rewrite fragile_counts.json
    python tests/golden/fragile_voxels.py --lists D  # also D/<workload>.npz:
        node indices per flag (solid, surface, multi_writer, material_race)
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle_py as O  # noqa: E402
from voxelizer_amd import workloads as W  # noqa: E402

EPS64, EPS32 = 1e-12, 1e-5

# name: (mesh, plan mode, param, node kind, voxSplitRes, matSplitRes)
WORKLOADS = {
    "icosphere_128_short": (lambda: W.icosphere(seed=0), 0, 126, O.NODE_SHORT,
                            (1024, 512, 512), (512, 512, 512)),
    "room_512x403x303_long": (lambda: W.shoebox_room(), 1, 0.01, O.NODE_LONG,
                              (1024, 512, 512), (512, 512, 512)),
    "hall_1536x1024x768_short": (lambda: W.hall(), 1, 1.0 / 64.0, O.NODE_SHORT,
                                 (1024, 1024, 768), (512, 512, 512)),
    "room_fcc_640x255x384_longfcc": (lambda: W.shoebox_room(), 1, 4.0 ** (1.0 / 3.0) * 0.01,
                                     O.NODE_LONG_FCC, (1024, 512, 512), (512, 512, 512)),
}


def analyse(name):
    mk, mode, param, kind, vs, ms = WORKLOADS[name]
    m = mk()
    g = O.plan_grid(m.vertices, mode, param, vs[0])
    flags, counts = O.fragile(m, g, kind, True, vs, ms, EPS64, EPS32)
    per_flag = {f: int(np.count_nonzero(flags & bit)) for f, bit in
                (("solid", O.FRAGILE_SOLID), ("surface", O.FRAGILE_SURF),
                 ("multi_writer", O.FRAGILE_MULTI), ("material_race", O.FRAGILE_RACE))}
    dims = list(O.node_dims(g, kind))
    out = {"mesh": m.name, "mesh_digest": m.digest(), "dims": dims, "nodes": int(flags.size),
           "vox_split": list(vs), "mat_split": list(ms), "eps64": EPS64, "eps32": EPS32,
           "tests": counts, "voxels": per_flag}
    return out, flags


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--lists", help="directory for the per-voxel lists (npz, not committed)")
    ap.add_argument("--only", nargs="*", help="workload names")
    args = ap.parse_args()
    res = {}
    for name in args.only or WORKLOADS:
        out, flags = analyse(name)
        res[name] = out
        print(name, json.dumps(out["voxels"]), flush=True)
        if args.lists:
            os.makedirs(args.lists, exist_ok=True)
            np.savez_compressed(os.path.join(args.lists, name + ".npz"),
                                solid=np.nonzero(flags & O.FRAGILE_SOLID)[0],
                                surface=np.nonzero(flags & O.FRAGILE_SURF)[0],
                                multi_writer=np.nonzero(flags & O.FRAGILE_MULTI)[0],
                                material_race=np.nonzero(flags & O.FRAGILE_RACE)[0])
    if not args.only:
        with open(os.path.join(HERE, "fragile_counts.json"), "w") as f:
            json.dump(res, f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main()
