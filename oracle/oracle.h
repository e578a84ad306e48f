/*
 * oracle.h -- CPU restatement of hakarlss/Voxelizer's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline); the product library never links or calls it.
 *
 * Every function follows the reference file:line given in oracle.cpp.  FP64 in
 * the solid stage, FP32 in the surface stage, no FMA contraction
 * (-ffp-contract=off), IEEE rounding, saturating float->int conversions (what
 * both CUDA's cvt.rzi and gfx950's v_cvt_* do).
 *
 * Material winners: the reference's processVoxel does a racy plain store
 * (src/voxelizer.cu:2988); this restatement uses the deterministic rule
 * "largest (bbType, triangle index) wins", which is one of the outcomes the
 * reference can produce (kernels run 1D, 2D, 3D in order, so the highest class
 * always wins there too).
 */
#ifndef VOX_ORACLE_H
#define VOX_ORACLE_H

#include <stdint.h>
#include "../include/voxelizer_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary id of a 6-bit neighbour permutation (src/voxelizer.cu:2142-2236). */
uint8_t orc_orientation(uint8_t permutation);

/* Bounding box + dims (src/host_side.cpp:2462-2634, 717-738).  mode 0:
 * maxDimension (param = R), mode 1: cubeLength.  Returns 0 on success. */
int orc_plan_grid(const float *vertices, uint32_t nv, int mode, double param,
                  uint32_t vox_split_x, vox_grid *out);

/* Solid bits exactly as performVoxelization + generateVoxelization compute
 * them for one device with no neighbours: ext_min is extMinVertex (minVertex
 * plus an optional FCC shift); split_y/split_z are voxSplitRes.y/z; the x
 * splits come from grid->x_splits.  bits: alloc words, zeroed by the callee. */
int orc_solid_bits(const float *vertices, const uint32_t *indices, uint32_t nt,
                   const vox_grid *grid, const double ext_min[3],
                   uint32_t split_y, uint32_t split_z, uint32_t *bits);

/* Whole node pipeline (voxelizeWorker / fccWorker for one device).
 * materials may be NULL (no surface stage).  nodes: node_count * node_size
 * bytes.  vox_split/mat_split: voxSplitRes/matSplitRes.  Returns 0 on
 * success, nonzero for an unsupported kind.  n_threads: OpenMP threads used
 * for the embarrassingly parallel loops (0 = all), results do not depend on it. */
int orc_voxelize_nodes(const float *vertices, uint32_t nv, const uint32_t *indices,
                       uint32_t nt, const uint8_t *materials, const vox_grid *grid,
                       int node_kind, const uint32_t vox_split[3],
                       const uint32_t mat_split[3], void *nodes, int n_threads);

/* Surface stage alone on a Cartesian grid: for each voxel of the alloc grid
 * the winning key (bbType<<30 | triangle) or 0 (src/voxelizer.cu:3307-3955). */
int orc_surface_keys(const float *vertices, const uint32_t *indices, uint32_t nt,
                     const vox_grid *grid, const double ext_min[3],
                     const uint32_t mat_split[3], uint32_t *keys);

/* Per-triangle surface classification (src/voxelizer.cu:3040-3221):
 * out[t] = bbType<<30 | domAxis<<28 | nrOfVoxCols. */
void orc_classify(const float *vertices, const uint32_t *indices, uint32_t nt,
                  const float model_min[3], float d, uint32_t *out);

/* Fill-stage test of one voxel centre against one triangle (src/voxelizer.cu:
 * 1375-1487): returns 1 and the x crossing voxelX (incl. the +1 padding) if the
 * centre passes; used by the hand-built tie-break tests. */
int orc_center_crossing(const float tri[9], const double ext_min[3], double d,
                        uint32_t cy, uint32_t cz, uint32_t *voxel_x);

/* Tile filter of the count pass (src/voxelizer.cu:840-1033): number of 4x4
 * tiles the triangle is queued on for the given ext y/z extent. */
uint32_t orc_tile_overlaps(const float tri[9], const double ext_min[3], double d,
                           uint32_t ext_y, uint32_t ext_z);

/* One device's partition of a Cartesian grid (voxelizeWorker,
 * src/host_side.cpp:901-1244): part_min/part_max bound the partition in
 * (y, z) ext coordinates of the plan; its neighbour flags, ext partition (one
 * row into each neighbour) and alloc layout follow the reference; nodes holds
 * res[0] x (py + 2) x (pz + 2) nodes, the border cleared when the partition
 * has a neighbour.  Returns 0 on success. */
int orc_voxelize_part(const float *vertices, uint32_t nv, const uint32_t *indices, uint32_t nt,
                      const uint8_t *materials, const vox_grid *grid, int node_kind,
                      const uint32_t vox_split[3], const uint32_t mat_split[3],
                      const uint32_t part_min[2], const uint32_t part_max[2], void *nodes,
                      int n_threads);
/* Solid bits of one partition (voxelize() with a devConfig: voxelizeWorker
 * up to its nodeOutput return, src/host_side.cpp:901-1055): alloc[0]/32 x
 * (py + 2) x (pz + 2) words, ext rows at their alloc positions, the rest zero.
 * Returns 0 on success. */
int orc_solid_bits_part(const float *vertices, const uint32_t *indices, uint32_t nt,
                        const vox_grid *grid, uint32_t split_y, uint32_t split_z,
                        const uint32_t part_min[2], const uint32_t part_max[2], uint32_t *bits);
/* One slice (voxelizeSlice / voxelizeSliceToRAM, src/host_side.cpp:
 * 1928-2262): plan by mode/param (0 maxDimension, 1 cubeLength), direction
 * 0 x (the model rotated, rotateVertices 3261-3290, and the nodes rotated
 * back, unRotateNodes src/voxelizer.cu:2061-2124), 1 y, 2 z; slice 1..res
 * along it (0 or past the grid: all zeros).  vox_split/mat_split are the
 * reference's (x, y) pairs.  dims receives the node array dimensions;
 * nodes may be NULL to query them.  Returns 0 on success. */
int orc_voxelize_slice(const float *vertices, uint32_t nv, const uint32_t *indices, uint32_t nt,
                       const uint8_t *materials, int mode, double param, int node_kind,
                       int direction, uint32_t slice, const uint32_t vox_split[2],
                       const uint32_t mat_split[2], void *nodes, uint32_t dims[3]);

/* calculateVoxelPlaneIntersectionVolume (src/voxelizer.cu:4220-4307 and its
 * helpers to 4785): the volume the reference computes for the box centred at
 * m + d * voxel cut by the plane through tri[0] with normal `normal`. */
float orc_cut_volume(const float tri[9], const int32_t voxel[3], const float normal[3],
                     const float m[3], float d);

/* Decisions where the original CUDA build could end with different nodes:
 * FP tests whose value lies within eps (relative to the magnitude of their
 * terms; eps64 for the FP64 solid tests, eps32 for the FP32 surface tests)
 * of the threshold, which FMA contraction in nvcc's default build can move,
 * and voxels several triangles write (the racy store of processVoxel,
 * src/voxelizer.cu:2988).  flags (one byte per node, node_count bytes):
 * 1 solid state rests on a fragile solid decision (a fragile crossing test
 * marks the rest of its row), 2 surface membership rests on a fragile
 * surface decision, 4 written by two or more triangles, 8 the writers of the
 * highest class carry two or more materials (the material may differ).
 * counts: [0] solid fill tests, [1] fragile ones, [2] fragile crossing
 * positions, [3] surface tests, [4] fragile ones, [5] voxels flagged 4,
 * [6] flagged 8, [7] flagged at all.  Returns 0 on success. */
int orc_fragile(const float *vertices, uint32_t nv, const uint32_t *indices, uint32_t nt,
                const uint8_t *materials, const vox_grid *grid, int node_kind,
                const uint32_t vox_split[3], const uint32_t mat_split[3], double eps64,
                double eps32, uint8_t *flags, uint64_t counts[8]);

/* Erosion + boundary ids on a raw Cartesian solid grid (0/1 per node),
 * per y/z subspace like the reference's host loop; literal = in-place
 * sequential passes, else snapshot passes (same fixed point). */
int orc_erode_nodes(const uint8_t *solid, uint32_t X, uint32_t Y, uint32_t Z, uint32_t split_y,
                    uint32_t split_z, int literal, uint8_t *bid_out);

/* 12-neighbour FCC masks on a raw FCC solid grid (one pass, in place). */
int orc_fcc_boundaries(const uint8_t *solid, uint32_t FX, uint32_t FY, uint32_t FZ,
                       uint16_t *bid_out);

#ifdef __cplusplus
}
#endif
#endif
