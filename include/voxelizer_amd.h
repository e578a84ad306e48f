/*
 * voxelizer_amd.h -- C ABI of the MI355X voxelizer (libvoxelizer_amd.so).
 *
 * Plain C: opaque context, plain pointers and sizes, status codes.  No C++
 * types and no exceptions cross this boundary.  The C++ class
 * vox::Voxelizer<Node> (include/vox/voxelizer.h) is a header-only wrapper over
 * these entry points, mirroring the reference API of
 * include/voxelizer.h:704-868 (hakarlss/Voxelizer).
 *
 * Every entry point below replaces one piece of the reference's host <-> device
 * boundary (include/host_device_interface.h:15-151 launchers, driven by
 * src/host_side.cpp); the replaced interface is named in each comment.
 *
 * Device pointers ("_dev") are HIP device addresses on the context's device.
 * All work is ordered on the context's stream; vox_sync() waits for it.  (A
 * call may also run work on a second stream of the context, forked from and
 * joined back into the context's stream inside the call.)
 * Calls on one context are not re-entrant (as in the reference, where one
 * Voxelizer instance mutates its options per call, src/host_side.cpp:1742).
 */
#ifndef VOXELIZER_AMD_H
#define VOXELIZER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vox_ctx vox_ctx;

typedef enum vox_status {
    VOX_OK = 0,
    VOX_ERR_INVALID = 1,     /* bad argument or unsupported configuration     */
    VOX_ERR_HIP = 2,         /* HIP runtime error (message in vox_last_error) */
    VOX_ERR_OOM = 3,         /* device allocation failed                      */
    VOX_ERR_UNSUPPORTED = 4, /* mode the reference rejects or that is not built */
    VOX_ERR_NO_DEVICE = 5    /* no HIP device (reference: initDevices throws)  */
} vox_status;

/* Node layouts, bit-identical to include/node_types.h of the reference. */
typedef enum vox_node_kind {
    VOX_NODE_SHORT = 0,     /* 1 B: mat<<5 | bid(5 bits)        node_types.h:19-177   */
    VOX_NODE_LONG = 1,      /* 2 B: {bid, mat}                  node_types.h:178-283  */
    VOX_NODE_PARTIAL = 2,   /* 8 B: {bid, mat, float ratio}     node_types.h:284-399  (not built) */
    VOX_NODE_SHORT_FCC = 3, /* 2 B: mat<<12 | bid(12 bits)      node_types.h:400-565  */
    VOX_NODE_LONG_FCC = 4   /* 4 B: {u16 bid, u8 mat, u8 pad=0} node_types.h:566-706  */
} vox_node_kind;

typedef enum vox_plan_mode {
    VOX_PLAN_MAX_DIMENSION = 0, /* param = voxels along the longest side (uint) */
    VOX_PLAN_CUBE_LENGTH = 1    /* param = distance between voxel centres       */
} vox_plan_mode;

/*
 * Grid plan: the reference's CommonHostData dimensioning (include/common.h:
 * 268-285) after determineDimensions + adjustResolution
 * (src/host_side.cpp:2497-2634).
 *   res[0]   = voxels along x incl. the +2 x padding, rounded up to a multiple
 *              of 32*x_splits;
 *   res[1..2]= voxels along y, z (no padding);
 *   alloc    = (res[0], res[1]+2, res[2]+2), the node/bit array dimensions
 *              (src/host_side.cpp:953-957), i.e. NodePointer.dim for
 *              Cartesian nodes.
 */
typedef struct vox_grid {
    double min_vertex[3];
    double max_vertex[3];
    double voxel_length;
    uint32_t res[3];
    uint32_t alloc[3];
    uint32_t x_splits;
    uint32_t pad_;
} vox_grid;

/* ---- context --------------------------------------------------------- */

/* Creates a context on HIP device `device` with its own stream and arena.
 * Replaces Voxelizer::initDevices (src/host_side.cpp:163-202). */
vox_status vox_create(int device, vox_ctx **out);
void vox_destroy(vox_ctx *ctx);
/* Last error message of this context ("" if none). */
const char *vox_last_error(const vox_ctx *ctx);
/* Number of HIP devices (reference: cudaGetDeviceCount in initVariables,
 * src/host_side.cpp:104-110). Returns 0 when the runtime finds none. */
int vox_device_count(void);
/* Waits for all work enqueued on the context's stream. */
vox_status vox_sync(vox_ctx *ctx);
/* Library/kernel build tag, e.g. "gfx950". */
const char *vox_build_info(void);

/* ---- device memory helpers (so C callers need no HIP headers) --------- */
vox_status vox_alloc(vox_ctx *ctx, size_t bytes, void **dev_out);
vox_status vox_free(vox_ctx *ctx, void *dev);
vox_status vox_memcpy_h2d(vox_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
vox_status vox_memcpy_d2h(vox_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);

/* ---- mesh ------------------------------------------------------------- */

/* Copies a host mesh to the device (reference: Voxelizer constructor copies
 * the arrays, src/host_side.cpp:44-47, and allocStaticMem uploads them,
 * src/host_side.cpp:447-476).  vertices: 3*nv floats, indices: 3*nt uint32. */
vox_status vox_set_mesh(vox_ctx *ctx, const float *vertices, uint32_t nv,
                        const uint32_t *indices, uint32_t nt);
/* Uses a mesh that is already resident on the device (no copy; the caller
 * keeps the buffers alive until the next vox_set_mesh*). */
vox_status vox_set_mesh_device(vox_ctx *ctx, const float *vertices_dev, uint32_t nv,
                               const uint32_t *indices_dev, uint32_t nt);
/* Per-triangle material ids; NULL disables materials.  Host pointer: copied.
 * Reference: Voxelizer::setMaterials (src/host_side.cpp:218-243). */
vox_status vox_set_materials(vox_ctx *ctx, const uint8_t *materials, uint32_t n_unique);
vox_status vox_set_materials_device(vox_ctx *ctx, const uint8_t *materials_dev, uint32_t n_unique);

/* ---- planning --------------------------------------------------------- */

/* Bounding box + dimensions of the current mesh, on the device.  Replaces
 * calculateBoundingBox + determineDimensions(d) / determineDimensions() +
 * splitResolutionByMaxDim(voxSplitRes.x) + adjustResolution
 * (src/host_side.cpp:2462-2634, 717-738). */
vox_status vox_plan_grid(vox_ctx *ctx, vox_plan_mode mode, double param,
                         uint32_t vox_split_x, vox_grid *out);
/* The same plan computed on the host from host vertices (no device needed);
 * backs Voxelizer::getArrayDimensions (src/host_side.cpp:3329-3470). */
vox_status vox_plan_grid_host(const float *vertices, uint32_t nv, vox_plan_mode mode,
                              double param, uint32_t vox_split_x, vox_grid *out);

/* ---- hot path --------------------------------------------------------- */

/* Solid (plain) voxelization into a bit grid: alloc[0]/32 words per row,
 * alloc[1]*alloc[2] rows, voxel x at bit 31-(x%32) (MSB first).  bits_dev
 * must hold alloc[0]*alloc[1]*alloc[2]/32 uint32 and is fully written.
 * shift (may be NULL) is added to min_vertex (the FCC sub-grid shifts,
 * src/host_side.cpp:1436-1463).  split_yz (may be NULL = one subspace) is
 * voxSplitRes.yz: the solid stage runs per y/z subspace as the reference does.
 * Replaces calcTileOverlap + prepareForConstructWorkQueue + calcWorkQueue +
 * sortWorkQueue + compactWorkQueue + calcVoxelization
 * (include/host_device_interface.h:15-60, src/host_side.cpp:275-397). */
vox_status vox_solid_bits(vox_ctx *ctx, const vox_grid *grid, const double shift[3],
                          const uint32_t split_yz[2], uint32_t *bits_dev);

/* Full node pipeline: solid voxelization, optional surface (materials)
 * voxelization with deterministic material winners, and boundary-id
 * classification (Cartesian 6-neighbour + erosion to the fixed point, or FCC
 * 12-neighbour mask).  nodes_dev must hold node_count(grid, kind) nodes.
 * vox_split: voxSplitRes (x,y,z), mat_split: matSplitRes (x,y,z); NULL means
 * the reference defaults (1024,512,512) and (512,512,512).
 * Replaces calcNodeList, launchConvertToFCCGrid, calcTriangleClassification,
 * calcOptSurfaceVoxelization, procNodeList (+ the host erosion loop) and
 * launchCalculateFCCBoundaries (include/host_device_interface.h:62-151,
 * src/host_side.cpp:901-1244 voxelizeWorker, 1262-1698 fccWorker). */
vox_status vox_voxelize_nodes(vox_ctx *ctx, const vox_grid *grid, vox_node_kind kind,
                              int use_materials, const uint32_t vox_split[3],
                              const uint32_t mat_split[3], void *nodes_dev);

/* One partition of a Cartesian grid, as one device of the reference's
 * multi-device mode voxelizes it (voxelizeWorker, src/host_side.cpp:901-1244,
 * the subspaces of voxelizeEntry's devConfig split, 782-813): part_min /
 * part_max bound the partition in the plan's ext (y, z) coordinates, x is
 * whole.  Where the partition stops short of the grid's edge it has a
 * neighbour and reaches one row or plane into it, so ids at the seam see the
 * neighbour's voxels (994-1011; src/voxelizer.cu:1487-1502, 3374-3383), and
 * the finished array's border is cleared (zeroPadding,
 * src/voxelizer.cu:4811-4844).  nodes_dev holds res[0] x (py + 2) x (pz + 2)
 * nodes.  ShortNode and LongNode (VOX_ERR_UNSUPPORTED for FCC). */
vox_status vox_voxelize_partition(vox_ctx *ctx, const vox_grid *grid, vox_node_kind kind,
                                  int use_materials, const uint32_t vox_split[3],
                                  const uint32_t mat_split[3], const uint32_t part_min[2],
                                  const uint32_t part_max[2], void *nodes_dev);
/* Plain (bit) voxelization of one partition, as voxelize() with a devConfig
 * leaves it on each device (voxelizeWorker up to its nodeOutput return,
 * src/host_side.cpp:901-1055): the solid bits of the ext partition (the
 * partition plus one row or plane into each neighbour) in an array of
 * alloc[0]/32 x (py + 2) x (pz + 2) words, ext row / plane e stored at
 * e + 1 less one where the partition has a neighbour below; rows and planes
 * with no ext row behind them are zero.  split_yz as in vox_solid_bits. */
vox_status vox_solid_bits_partition(vox_ctx *ctx, const vox_grid *grid, const uint32_t split_yz[2],
                                    const uint32_t part_min[2], const uint32_t part_max[2],
                                    uint32_t *bits_dev);
/* One slice of the current mesh (voxelizeSlice / voxelizeSliceToRAM,
 * src/host_side.cpp:1928-2262): the grid planned by mode/param, the slice
 * `slice` (1 .. resolution, 0 and past the grid give all zeros) along
 * direction 0 x, 1 y or 2 z, voxelized as a one-row partition with its two
 * neighbour rows.  Along x the model is rotated a quarter turn about z first
 * (rotateVertices, 3261-3290) and the nodes rotated back (unRotateNodes,
 * src/voxelizer.cu:2061-2124); that needs vox_split[1] at least the rotated
 * grid's x size (VOX_ERR_UNSUPPORTED otherwise).  vox_split / mat_split are
 * the reference's (x, y) pairs, y serving z too.  dims_out receives the node
 * array's dimensions: (X, 3, Z + 2) along y, (X, Y + 2, 3) along z,
 * (3, X', Z' + 2) along x; with nodes_dev NULL only the dimensions are
 * computed.  ShortNode and LongNode (VOX_ERR_UNSUPPORTED for FCC). */
vox_status vox_voxelize_slice(vox_ctx *ctx, vox_plan_mode mode, double param, vox_node_kind kind,
                              int use_materials, int direction, uint32_t slice,
                              const uint32_t vox_split[2], const uint32_t mat_split[2],
                              void *nodes_dev, uint32_t dims_out[3]);

/* vox_voxelize_nodes without allocating: no hipMalloc, hipFree or other
 * device or pinned allocation happens in the call.  No counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid it).  Every scratch buffer is
 * carved from a device array of 128 MiB that the library holds on each GPU it
 * runs on until the process ends (its address is looked up by vox_create);
 * grids that need more (about 1/8 B per voxel of solid bits, 1 B per voxel
 * with materials, ~100 B per triangle and 12 B per (triangle, tile) pair --
 * e.g. the 512x403x303 room fits, the 1.2 G-voxel hall does not) fail with
 * VOX_ERR_UNSUPPORTED before any work is queued.  All calls on one GPU share
 * that array and each leaves its contents behind (what must start at zero is
 * cleared on the stream); calls that overlap in time on one GPU (several
 * contexts, threads or streams) race on it and must be serialized by the
 * caller.  Cost recording (vox_cost_record) is not available here. */
vox_status vox_voxelize_nodes_noalloc(vox_ctx *ctx, const vox_grid *grid, vox_node_kind kind,
                                      int use_materials, const uint32_t vox_split[3],
                                      const uint32_t mat_split[3], void *nodes_dev);

/* Node array dimensions / count / node size for a plan and node kind
 * (FCC: 2*alloc[0] x alloc[1] x 2*alloc[2], src/host_side.cpp:1494-1499). */
void vox_node_dims(const vox_grid *grid, vox_node_kind kind, uint32_t dims_out[3]);
size_t vox_node_count(const vox_grid *grid, vox_node_kind kind);
size_t vox_node_size(vox_node_kind kind);

/* ---- instrumentation (no counterpart in the reference) ----------------- */

/* Per-stage device times (ms) of the last vox_voxelize_nodes/vox_solid_bits
 * call, measured with HIP events on the context stream when enabled. */
typedef struct vox_stage_times {
    float plan_ms, solid_bin_ms, solid_fill_ms, surface_ms, erode_ms, nodes_ms, total_ms;
    uint32_t erosion_passes;
    uint32_t solid_pairs;      /* (triangle, tile) work-queue entries */
    uint32_t surface_pairs;    /* (triangle, brick) surface work items */
    uint32_t surface_voxels;   /* voxels written by the surface stage */
    uint32_t learned_schedule; /* 1 if the surface bricks ran in learned order */
} vox_stage_times;
vox_status vox_enable_timing(vox_ctx *ctx, int enable);
vox_status vox_get_stage_times(const vox_ctx *ctx, vox_stage_times *out);

/* ---- brick schedule and its learned cost model (no counterpart in the
 * reference) --------------------------------------------------------------
 * The surface pass processes its non-empty bricks from a list with
 * persistent workgroups.  PLAIN lists them in index order; LEARNED moves the
 * bricks an MLP (16 features -> 64 -> 64 -> 1, ReLU, FP32) predicts to be
 * expensive to the front of the list and cuts the most expensive into
 * sub-bricks (see vox_set_heavy_threshold).  Outputs are identical under
 * both.  LEARNED falls back to PLAIN while no model is set
 * (vox_stage_times.learned_schedule says which ran). */
typedef enum vox_schedule { VOX_SCHEDULE_PLAIN = 0, VOX_SCHEDULE_LEARNED = 1 } vox_schedule;
#define VOX_COST_FEATURES 16
#define VOX_COST_PARAMS 5313 /* W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3 */
vox_status vox_set_schedule(vox_ctx *ctx, vox_schedule schedule);
/* LEARNED order: bricks with at least `screen ratio` times the mean number
 * of (triangle, brick) pairs (default 24) are scored by the model; those
 * whose predicted cost (standardized log cycles) exceeds `threshold` run
 * first, as they were predicted, one per grab; a brick predicted more than 1
 * above the threshold is cut into 2 sub-bricks along z (more than 2 above:
 * 4, at most its depth), so several workgroups share it.  All other bricks
 * keep index order (neighbouring bricks share triangles in cache).  Default
 * threshold 1.5. */
vox_status vox_set_heavy_threshold(vox_ctx *ctx, float threshold);
vox_status vox_set_screen_ratio(vox_ctx *ctx, float ratio);
/* At most `parts` (1, 2 or 4) sub-bricks per heavy brick; 0 (default) picks
 * it from the number of bricks per resident brick-pass workgroup (16 or more:
 * 1, 4 or more: 2, else 4): a part repeats the brick's triangle setup, which
 * pays only where one long brick could end the pass last. */
vox_status vox_set_split_parts(vox_ctx *ctx, uint32_t parts);
/* Model parameters (host arrays of VOX_COST_PARAMS floats). */
vox_status vox_cost_set_params(vox_ctx *ctx, const float *params);
vox_status vox_cost_get_params(vox_ctx *ctx, float *params);
/* While enabled, every surface pass appends, per non-empty brick, its
 * VOX_COST_FEATURES features and its measured duration (GPU clock cycles).
 * Enabling clears earlier samples. */
vox_status vox_cost_record(vox_ctx *ctx, int enable);
/* Copies up to cap samples (features row-major, cycles) to host arrays (may
 * be NULL to query the count) and reports the number recorded. */
vox_status vox_cost_samples(vox_ctx *ctx, float *features, float *cycles, uint32_t cap,
                            uint32_t *n_out);
/* Trains the model on host samples with the library's GPU kernels: targets
 * are standardized log(cycles); Adam, `epochs` passes of mini-batches of
 * `batch` rows (multiple of 32), deterministic for a given seed.  Installs
 * the result as the context's model and returns the final mean squared error
 * (standardized units) over all samples. */
vox_status vox_cost_train(vox_ctx *ctx, const float *features, const float *cycles, uint32_t n,
                          uint32_t epochs, uint32_t batch, float lr, uint32_t seed,
                          float *final_mse);
/* The MLP kernels on device arrays (training and tests): y[i] = f(x[i]) for
 * n rows of VOX_COST_FEATURES; and the gradient of
 * (1/n) sum_i (f(x[i]) - t[i])^2 with respect to the parameters (grad: device
 * array of VOX_COST_PARAMS) plus that loss (host). */
vox_status vox_mlp_forward(vox_ctx *ctx, const float *params_dev, const float *x_dev,
                           uint32_t n, float *y_dev);
vox_status vox_mlp_grad(vox_ctx *ctx, const float *params_dev, const float *x_dev,
                        const float *t_dev, uint32_t n, float *grad_dev, float *loss);

/* ---- surface brick shape and its autotuner (no counterpart in the
 * reference) -------------------------------------------------------------
 * The surface pass works on bricks of voxels (default 128x4x8 on Cartesian
 * grids, 64x4x4 on each FCC sub-grid).  Shapes are powers of two with x a
 * multiple of 64, y*z <= 32 rows and 256 to 4096 voxels; Cartesian shapes
 * need z a multiple of 4.  The shape changes speed only, never output. */
vox_status vox_set_brick_dims(vox_ctx *ctx, int fcc, const uint32_t dims[3]);
vox_status vox_get_brick_dims(const vox_ctx *ctx, int fcc, uint32_t dims[3]);
/* Times vox_voxelize_nodes (with materials) for this grid and node kind
 * with each candidate shape (one warm-up, then the best of `reps` runs),
 * keeps the fastest for the context and returns it and its time.  nodes_dev
 * ends up holding the nodes, as after a vox_voxelize_nodes call. */
vox_status vox_autotune_bricks(vox_ctx *ctx, const vox_grid *grid, vox_node_kind kind,
                               const uint32_t vox_split[3], const uint32_t mat_split[3],
                               void *nodes_dev, uint32_t reps, uint32_t best_dims[3],
                               float *best_ms);

/* ---- batches of meshes (no counterpart in the reference) ---------------
 * Many meshes voxelized in one call, each into its own node array.  Inputs
 * are device arrays; the grid of each item is planned by the caller
 * (vox_plan_grid / vox_plan_grid_host).  `streams` picks the execution:
 *   0  fused: every stage is one launch over all items (a per-item job table
 *      on the device; an FCC lattice counts as four grids), so a batch costs
 *      about as many launches and host round trips as one mesh.  Items whose
 *      Cartesian grid has several voxSplitRes y/z erosion subspaces run
 *      afterwards one by one;
 *   1  one item after another on the context's stream, as separate
 *      vox_voxelize_nodes calls would;
 *   k  k worker streams at once (one host thread and one internal context
 *      each, at most 64), items in decreasing order of estimated work (node
 *      bytes + 64 B per triangle: longest first).
 * `schedule` is the brick order of the surface pass (LEARNED: cost-model
 * order when a model is set).  Outputs are identical under every choice;
 * on return each item's status is set. */
typedef struct vox_batch_item {
    const float *vertices_dev;
    uint32_t nv;
    uint32_t nt;
    const uint32_t *indices_dev;
    const uint8_t *materials_dev; /* NULL: no materials */
    uint32_t n_unique;
    vox_node_kind kind;
    vox_grid grid;                /* planned grid of this mesh */
    uint32_t vox_split[3];        /* voxSplitRes */
    uint32_t mat_split[3];        /* matSplitRes */
    void *nodes_dev;              /* vox_node_count(&grid, kind) nodes */
    vox_status status;            /* out */
} vox_batch_item;
vox_status vox_voxelize_batch(vox_ctx *ctx, vox_batch_item *items, uint32_t n,
                              vox_schedule schedule, uint32_t streams);

/* ---- replayed batches (no counterpart in the reference) ----------------
 * A fused batch captured once as a HIP graph and replayed: every stage's
 * launches with no host round trip in between.  vox_graph_create runs the
 * batch once (fused, on private buffers of an internal context) to learn the
 * sizes of its pair lists and the number of erosion passes, then captures
 * the stages with those sizes (lists 1/8 longer, one pass more).  Each
 * vox_graph_launch replays the graph and then checks that the lists were
 * long enough and that erosion reached its fixed point; if not (the meshes
 * changed since), it runs the batch again without the graph and captures a
 * new one, so the nodes always equal vox_voxelize_batch's.  The items
 * (pointers, grids, kinds, splits) are fixed at creation; the contents of the
 * vertex, index and material arrays may change between launches.  Every item
 * must be fusable (see vox_voxelize_batch), else VOX_ERR_UNSUPPORTED.
 * vox_graph_launch returns once the nodes are written. */
typedef struct vox_graph vox_graph;
vox_status vox_graph_create(vox_ctx *ctx, const vox_batch_item *items, uint32_t n,
                            vox_schedule schedule, vox_graph **out);
vox_status vox_graph_launch(vox_graph *graph);
/* Launches that fell back to a run without the graph. */
uint32_t vox_graph_fallbacks(const vox_graph *graph);
void vox_graph_destroy(vox_graph *graph);

/* ---- mesh files and output dumps (host only, no context) -----------------
 * vox_mesh_load replaces the reference tester's VTK readers and its loadModel
 * copy loop (tests/src/main.cpp:359-404, 911-955): OBJ, PLY (ASCII and
 * binary), STL (ASCII and binary) and legacy VTK POLYDATA (ASCII and binary,
 * DataFile 2.0-5.1) by file extension, vertices as float in file order,
 * triangles in file order.  Polygons with more than three corners become a
 * fan (0, k, k+1); cells with fewer are skipped.  STL corners that coincide
 * exactly are merged into one vertex (first appearance order) and facets that
 * collapse under the merge are dropped, as VTK's STL reader does by default.
 * The arrays are malloc'ed; vox_mesh_free releases them.  On failure err
 * (may be NULL) receives a message. */
typedef struct vox_mesh {
    float *vertices;   /* 3 * nv */
    uint32_t nv;
    uint32_t *indices; /* 3 * nt */
    uint32_t nt;
} vox_mesh;
vox_status vox_mesh_load(const char *path, vox_mesh *out, char *err, size_t err_len);
void vox_mesh_free(vox_mesh *mesh);
/* A node array (dims x, y, z; x fastest) as a legacy VTK STRUCTURED_POINTS
 * file with the point scalars "bid" and "material" (and "ratio" for
 * PartialNode), ASCII or big-endian binary; origin may be NULL (zero). */
vox_status vox_write_nodes_vtk(const char *path, const void *nodes, vox_node_kind kind,
                               const uint32_t dims[3], const double origin[3], double spacing,
                               int binary);
/* A solid voxel bit array (dims x (a multiple of 32), y, z; voxel x of a
 * word at bit 31 - x % 32) as STRUCTURED_POINTS with the scalar "solid". */
vox_status vox_write_bits_vtk(const char *path, const uint32_t *bits, const uint32_t dims[3],
                              const double origin[3], double spacing, int binary);

#ifdef __cplusplus
}
#endif
#endif /* VOXELIZER_AMD_H */
