// vox/voxelizer.h -- header-only C++ class vox::Voxelizer<Node> over the C ABI
// of libvoxelizer_amd.so (include/voxelizer_amd.h).
//
// It keeps the public API of the reference class (include/voxelizer.h:704-868
// of hakarlss/Voxelizer): the same constructors, voxelize / voxelizeToRAM /
// voxelizeToNodes / voxelizeToNodesToRAM overloads with the same defaults,
// NodePointer, vox::Exception, the getters, setMaterials /
// setMaterialOutput / verboseOutput and getArrayDimensions.  Ownership follows
// the reference (src/host_side.cpp:3561-3625): device pointers are handed to
// the caller, who frees them with hipFree; host pointers come from new[] and
// the caller delete[]s them.  Unsupported requests print to std::cout and
// return an empty result exactly where the reference does (FCC nodes through
// maxDimension, plain voxelization with FCC nodes); library errors are thrown
// as vox::Exception.
//
// One device: a devConfig other than (1,1) runs inside simulateMultidevice,
// every partition on the default device as the reference's simulation does
// (src/host_side.cpp:163-202, 832-843); outside it, a request for more
// devices than exist throws the reference's "Not enough devices" exception.
// Slices (voxelizeSlice / voxelizeSliceToRAM) take devConfig 1 and Cartesian
// nodes; the model is rotated for x slices on every call, on a copy (the
// reference rotates its own vertex array once and keeps it rotated).
//
// Vector types are HIP's (uint2, uint3, double3, make_uint3, ...), the HIP
// counterparts of the CUDA vector types the reference API uses.  A host-only
// translation unit compiled with g++ needs -D__HIP_PLATFORM_AMD__ and
// -I/opt/rocm/include, as any HIP host program does.
#ifndef VOX_VOXELIZER_H_AMD
#define VOX_VOXELIZER_H_AMD

#include <hip/hip_vector_types.h>

#include <cstdint>
#include <functional>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "../voxelizer_amd.h"
#include "node_types.h"

namespace vox {

// include/common.h:173-197 of the reference.
class Exception : public std::exception {
public:
    Exception(std::string message) throw() : msg(message) {}
    Exception(char const *message) throw() : msg(std::string(message)) {}
    ~Exception() throw() {}
    const char *what() const throw() { return msg.c_str(); }

private:
    const std::string msg;
};

// include/voxelizer.h:674-684 of the reference.  dim is the allocated array
// size (NodePointer.dim), loc the position of the array in the whole grid
// (always (0,0,0) with one device).
template <class Node>
struct NodePointer {
    union {
        Node *ptr;
        VoxInt *vptr;
    };
    int dev;
    uint3 dim;
    uint3 loc;
};

template <class Node = ShortNode>
class Voxelizer {
public:
    // src/host_side.cpp:29-52: the arrays are copied, the bounding box is
    // computed at once; no device is touched until the first voxelize call.
    Voxelizer(float const *_vertices, uint const *_indices, uint _nrOfVertices,
              uint _nrOfTriangles)
        : vertices(_vertices, _vertices + 3 * (size_t)_nrOfVertices),
          indices(_indices, _indices + 3 * (size_t)_nrOfTriangles) {
        init(_nrOfTriangles);
    }
    // src/host_side.cpp:74-99.
    Voxelizer(float const *_vertices, uint const *_indices, uchar const *_materials,
              uint _nrOfVertices, uint _nrOfTriangles, uint _nrOfUniqueMaterials)
        : vertices(_vertices, _vertices + 3 * (size_t)_nrOfVertices),
          indices(_indices, _indices + 3 * (size_t)_nrOfTriangles) {
        init(_nrOfTriangles);
        setMaterials(_materials, _nrOfUniqueMaterials);
    }
    ~Voxelizer() {
        if (ctx) vox_destroy(ctx);
    }
    Voxelizer(const Voxelizer &) = delete;
    Voxelizer &operator=(const Voxelizer &) = delete;

    // ---- plain (bit) voxelization: src/host_side.cpp:562-650, 2275-2340 ----
    std::vector<NodePointer<Node> > voxelize(uint maxDimension = 256,
                                             uint2 devConfig = make_uint2(1, 1),
                                             uint3 voxSplitRes = make_uint3(1024, 512, 512)) {
        std::vector<NodePointer<Node> > result;
        if (Node::isFCCNode()) {
            std::cout << "Please use a standard Node-type when producing a "
                         "plain voxelization.\n";
            return result;
        }
        return plainParts(VOX_PLAN_MAX_DIMENSION, double(maxDimension), devConfig, voxSplitRes);
    }
    std::vector<NodePointer<Node> > voxelize(double cubeLength,
                                             uint2 devConfig = make_uint2(1, 1),
                                             uint3 voxSplitRes = make_uint3(1024, 512, 512)) {
        std::vector<NodePointer<Node> > result;
        if (Node::isFCCNode()) {
            std::cout << "Please use a standard Node-type when producing a "
                         "plain voxelization.\n";
            return result;
        }
        return plainParts(VOX_PLAN_CUBE_LENGTH, cubeLength, devConfig, voxSplitRes);
    }
    NodePointer<Node> voxelizeToRAM(uint maxDimension = 256,
                                    uint3 voxSplitRes = make_uint3(1024, 512, 512)) {
        if (Node::isFCCNode()) {
            std::cout << "Please use a standard Node-type when producing a "
                      << "plain voxelization" << std::endl;
            return NodePointer<Node>();
        }
        return plain(VOX_PLAN_MAX_DIMENSION, double(maxDimension), voxSplitRes, true);
    }
    NodePointer<Node> voxelizeToRAM(double cubeLength,
                                    uint3 voxSplitRes = make_uint3(1024, 512, 512)) {
        if (Node::isFCCNode()) {
            std::cout << "Please use a standard Node-type when producing a "
                      << "plain voxelization" << std::endl;
            return NodePointer<Node>();
        }
        return plain(VOX_PLAN_CUBE_LENGTH, cubeLength, voxSplitRes, true);
    }

    // ---- node voxelization: src/host_side.cpp:1734-1895 ---------------------
    std::vector<NodePointer<Node> > voxelizeToNodes(uint maxDimension = 256,
                                                    uint2 devConfig = make_uint2(1, 1),
                                                    uint3 voxSplitRes = make_uint3(1024, 512, 512),
                                                    uint3 matSplitRes = make_uint3(512, 512, 512)) {
        std::vector<NodePointer<Node> > result;
        if (Node::isFCCNode()) {
            std::cout << "FCC Nodes can currently only be defined through the "
                         "side length of their cubic cells";
            return result;
        }
        return nodeParts(VOX_PLAN_MAX_DIMENSION, double(maxDimension), devConfig, voxSplitRes,
                         matSplitRes);
    }
    std::vector<NodePointer<Node> > voxelizeToNodes(double cubeLength,
                                                    uint2 devConfig = make_uint2(1, 1),
                                                    uint3 voxSplitRes = make_uint3(1024, 512, 512),
                                                    uint3 matSplitRes = make_uint3(512, 512, 512)) {
        return nodeParts(VOX_PLAN_CUBE_LENGTH, cubeLength, devConfig, voxSplitRes, matSplitRes);
    }
    NodePointer<Node> voxelizeToNodesToRAM(uint maxDimension = 256,
                                           uint3 voxSplitRes = make_uint3(1024, 512, 512),
                                           uint3 matSplitRes = make_uint3(512, 512, 512)) {
        if (Node::isFCCNode()) {
            std::cout << "FCC Nodes can currently only be defined through the "
                      << "length of their cubic cells";
            return NodePointer<Node>();
        }
        return nodes(VOX_PLAN_MAX_DIMENSION, double(maxDimension), voxSplitRes, matSplitRes, true);
    }
    NodePointer<Node> voxelizeToNodesToRAM(double cubeLength,
                                           uint3 voxSplitRes = make_uint3(1024, 512, 512),
                                           uint3 matSplitRes = make_uint3(512, 512, 512)) {
        return nodes(VOX_PLAN_CUBE_LENGTH, cubeLength, voxSplitRes, matSplitRes, true);
    }

    // ---- slicing: src/host_side.cpp:1927-2262 -------------------------------
    // One slice (1 .. resolution along `direction`, 0 x, 1 y, 2 z; 0 and
    // slices past the grid give an all-zero array) as a one-row partition
    // with its neighbour rows: dim (X, 3, Z + 2) along y, (X, Y + 2, 3) along
    // z, (3, X', Z' + 2) along x (X', Z' of the rotated model's grid).
    std::vector<NodePointer<Node> > voxelizeSlice(uint maxDimension, int direction, uint slice,
                                                  uint devConfig = 1,
                                                  uint2 voxSplitRes = make_uint2(1024, 512),
                                                  uint2 matSplitRes = make_uint2(512, 512)) {
        std::vector<NodePointer<Node> > result;
        if (Node::isFCCNode()) {
            std::cout << "FCC Nodes can currently only be defined through the "
                      << "length of their cubic cells";
            return result;
        }
        checkSliceDevConfig(devConfig);
        result.push_back(sliceNodes(VOX_PLAN_MAX_DIMENSION, double(maxDimension), direction, slice,
                                    voxSplitRes, matSplitRes, false));
        return result;
    }
    std::vector<NodePointer<Node> > voxelizeSlice(double cubeLength, int direction, uint slice,
                                                  uint devConfig = 1,
                                                  uint2 voxSplitRes = make_uint2(1024, 512),
                                                  uint2 matSplitRes = make_uint2(512, 512)) {
        std::vector<NodePointer<Node> > result;
        checkSliceDevConfig(devConfig);
        result.push_back(sliceNodes(VOX_PLAN_CUBE_LENGTH, cubeLength, direction, slice,
                                    voxSplitRes, matSplitRes, false));
        return result;
    }
    NodePointer<Node> voxelizeSliceToRAM(uint maxDimension, int direction, uint slice,
                                         uint2 voxSplitRes = make_uint2(1024, 512),
                                         uint2 matSplitRes = make_uint2(512, 512)) {
        if (Node::isFCCNode()) {
            std::cout << "FCC Nodes can currently only be defined through the "
                         "length of their cubic cells";
            return NodePointer<Node>();
        }
        return sliceNodes(VOX_PLAN_MAX_DIMENSION, double(maxDimension), direction, slice,
                          voxSplitRes, matSplitRes, true);
    }
    NodePointer<Node> voxelizeSliceToRAM(double cubeLength, int direction, uint slice,
                                         uint2 voxSplitRes = make_uint2(1024, 512),
                                         uint2 matSplitRes = make_uint2(512, 512)) {
        return sliceNodes(VOX_PLAN_CUBE_LENGTH, cubeLength, direction, slice, voxSplitRes,
                          matSplitRes, true);
    }

    // ---- getters: include/voxelizer.h:791-846 ------------------------------
    double getVoxelLength() const throw() { return voxelLength; }
    double3 getMinVertex() const throw() { return minVertex; }
    double3 getMaxVertex() const throw() { return maxVertex; }
    uint3 getResolution() const throw() { return resolution; }
    std::vector<float> getVertices() const throw() { return vertices; }
    std::vector<uint> getIndices() const throw() { return indices; }
    uint getNrOfTriangles() const throw() { return nrOfTriangles; }

    // src/host_side.cpp:218-243: same checks and messages; enables materials.
    void setMaterials(uchar const *_materials, uint _nrOfUniqueMaterials) {
        if (_materials == NULL)
            throw Exception("No valid materials provided (Received null pointer)"
                            " @ setMaterials.");
        if (_nrOfUniqueMaterials > uint(Node::maxMat() + 1))
            throw Exception("The number of unique materials exceeds the "
                            "capabilities of the node type @ setMaterials.");
        nrOfUniqueMaterials = _nrOfUniqueMaterials;
        materials.assign(_materials, _materials + nrOfTriangles);
        materialsDirty = true;
        useMaterials = true;
    }
    // src/host_side.cpp:254-258.
    void setMaterialOutput(bool _materials) throw() { useMaterials = _materials; }
    // src/host_side.cpp:3247-3250: when on, each call prints its stage times.
    void verboseOutput(bool verbose) throw() { this->verbose = verbose; }

    // src/host_side.cpp:3329-3507 (the non-slicing branch): the array size a
    // voxelization with these parameters produces; FCC nodes double x and z.
    uint3 getArrayDimensions(uint longestSizeInVoxels, uint maxInternalXSize, bool sliceAlongX) {
        return arrayDims(VOX_PLAN_MAX_DIMENSION, double(longestSizeInVoxels), maxInternalXSize,
                         sliceAlongX);
    }
    uint3 getArrayDimensions(double cubeLength, uint maxInternalXSize, bool sliceAlongX) {
        return arrayDims(VOX_PLAN_CUBE_LENGTH, cubeLength, maxInternalXSize, sliceAlongX);
    }

    // src/host_side.cpp:3651-3659: from here on every partition of a
    // devConfig runs on the default device (the flag stays set, as in the
    // reference); returns what func returns.
    std::vector<NodePointer<Node> > simulateMultidevice(
        std::function<std::vector<NodePointer<Node> >()> func) {
        simulate = true;
        return func();
    }

    // Per-stage device times of the last call (no counterpart in the reference).
    vox_stage_times lastStageTimes() const { return stageTimes; }

private:
    std::vector<float> vertices;
    std::vector<uint> indices;
    std::vector<uchar> materials;
    uint nrOfTriangles = 0;
    uint nrOfUniqueMaterials = 0;
    bool useMaterials = false;
    bool materialsDirty = false;
    bool verbose = false;
    bool simulate = false;
    vox_ctx *ctx = nullptr;
    double3 minVertex = make_double3(0, 0, 0), maxVertex = make_double3(0, 0, 0);
    double voxelLength = 0.0;
    uint3 resolution = make_uint3(256, 0, 0);  // initVariables, src/host_side.cpp:112-113
    vox_stage_times stageTimes = vox_stage_times();

    void init(uint nT) {
        nrOfTriangles = nT;
        // calculateBoundingBox (src/host_side.cpp:2462-2488).
        const size_t nv = vertices.size() / 3;
        if (nv == 0) return;
        double mn[3] = {vertices[0], vertices[1], vertices[2]};
        double mx[3] = {mn[0], mn[1], mn[2]};
        for (size_t i = 1; i < nv; i++)
            for (int c = 0; c < 3; c++) {
                const double x = vertices[3 * i + c];
                if (x > mx[c]) mx[c] = x;
                if (x < mn[c]) mn[c] = x;
            }
        minVertex = make_double3(mn[0], mn[1], mn[2]);
        maxVertex = make_double3(mx[0], mx[1], mx[2]);
    }

    void check(vox_status st, const char *where) {
        if (st == VOX_OK) return;
        std::string m = ctx ? vox_last_error(ctx) : "";
        if (m.empty()) m = "vox status " + std::to_string(int(st));
        throw Exception(m + " @ " + where + ".");
    }

    // initDevices (src/host_side.cpp:163-202): the partitions of devConfig,
    // ordered z outer and y inner, each with its place (0, i % x, i / x) in
    // the array of partitions (voxelizeEntry, 782-813).
    struct Part {
        uint32_t min[2], max[2];
        uint3 loc;
    };
    // splitResolutionByNrOfParts (src/host_side.cpp:3046-3133) along one
    // axis: from the top down, part i takes the remaining length / i.
    static std::vector<std::pair<uint32_t, uint32_t> > splitParts(uint32_t len, uint32_t n) {
        std::vector<std::pair<uint32_t, uint32_t> > r(n);
        uint32_t rest = len;
        for (uint32_t i = n; i > 0; --i) {
            r[i - 1] = std::make_pair(rest - rest / i, rest);
            rest = r[i - 1].first;
        }
        return r;
    }
    std::vector<Part> partitions(uint2 devConfig, const vox_grid &g) {
        const uint32_t n = devConfig.x * devConfig.y;
        if (n == 0) throw Exception("devConfig must name at least one device.");
        if (n > 1 && !simulate) {
            if (n > uint32_t(vox_device_count()))
                throw Exception("Not enough devices to satisfy request.\n");
            throw Exception("Placing partitions on several devices is not built; run the call "
                            "inside simulateMultidevice to voxelize every partition on one "
                            "device.");
        }
        const std::vector<std::pair<uint32_t, uint32_t> > ys = splitParts(g.res[1], devConfig.x),
                                                          zs = splitParts(g.res[2], devConfig.y);
        std::vector<Part> parts;
        for (uint32_t i = 0; i < n; i++) {
            Part p;
            const uint32_t y = i % devConfig.x, z = i / devConfig.x;
            p.min[0] = ys[y].first;
            p.max[0] = ys[y].second;
            p.min[1] = zs[z].first;
            p.max[1] = zs[z].second;
            p.loc = make_uint3(0, y, z);
            if (p.min[0] == p.max[0] || p.min[1] == p.max[1])
                throw Exception("devConfig splits the grid into empty partitions.");
            parts.push_back(p);
        }
        return parts;
    }
    void checkSliceDevConfig(uint devConfig) {
        if (devConfig != 1)
            throw Exception("Slices are voxelized on one device: devConfig must be 1.");
    }
    void ensureContext() {
        if (!ctx) {
            if (vox_device_count() <= 0)
                throw Exception("Couldn't find any HIP capable devices.\n");
            check(vox_create(0, &ctx), "initDevices");
            check(vox_set_mesh(ctx, vertices.data(), uint32_t(vertices.size() / 3), indices.data(),
                               nrOfTriangles),
                  "initDevices");
            if (verbose) check(vox_enable_timing(ctx, 1), "initDevices");
        }
        if (materialsDirty) {
            check(vox_set_materials(ctx, materials.data(), nrOfUniqueMaterials), "setMaterials");
            materialsDirty = false;
        }
        check(vox_enable_timing(ctx, verbose ? 1 : 0), "verboseOutput");
    }

    vox_grid plan(vox_plan_mode mode, double param, uint xSplit) {
        vox_grid g;
        check(vox_plan_grid(ctx, mode, param, xSplit, &g), "determineDimensions");
        voxelLength = g.voxel_length;
        resolution = make_uint3(g.res[0], g.res[1], g.res[2]);
        return g;
    }

    void report(const char *what) {
        if (!verbose) return;
        check(vox_get_stage_times(ctx, &stageTimes), what);
        std::cout << what << ": " << stageTimes.total_ms << " ms (solid "
                  << stageTimes.solid_bin_ms + stageTimes.solid_fill_ms << ", surface "
                  << stageTimes.surface_ms << ", erosion " << stageTimes.erode_ms << " in "
                  << stageTimes.erosion_passes << " passes, nodes " << stageTimes.nodes_ms
                  << ")\n";
    }

    NodePointer<Node> plain(vox_plan_mode mode, double param, uint3 vsr, bool toHost) {
        ensureContext();
        const vox_grid g = plan(mode, param, vsr.x);
        const size_t words = (size_t)g.alloc[0] / 32 * g.alloc[1] * g.alloc[2];
        void *dev = nullptr;
        check(vox_alloc(ctx, words * sizeof(VoxInt), &dev), "voxelize");
        const uint32_t split[2] = {vsr.y, vsr.z};
        vox_status st = vox_solid_bits(ctx, &g, nullptr, split, (uint32_t *)dev);
        if (st == VOX_OK) st = vox_sync(ctx);
        if (st != VOX_OK) {
            vox_free(ctx, dev);
            check(st, "voxelize");
        }
        report("voxelize");
        NodePointer<Node> r = NodePointer<Node>();
        r.dev = 0;
        r.dim = make_uint3(g.alloc[0], g.alloc[1], g.alloc[2]);
        r.loc = make_uint3(0, 0, 0);
        if (!toHost) {
            r.vptr = (VoxInt *)dev;
            return r;
        }
        r.vptr = new VoxInt[words];
        st = vox_memcpy_d2h(ctx, r.vptr, dev, words * sizeof(VoxInt));
        vox_free(ctx, dev);
        if (st != VOX_OK) {
            delete[] r.vptr;
            check(st, "voxelizeToRAM");
        }
        return r;
    }

    NodePointer<Node> nodes(vox_plan_mode mode, double param, uint3 vsr, uint3 msr, bool toHost) {
        ensureContext();
        const vox_grid g = plan(mode, param, vsr.x);
        const vox_node_kind kind = vox_node_kind(Node::kind);
        const size_t n = vox_node_count(&g, kind);
        void *dev = nullptr;
        check(vox_alloc(ctx, n * sizeof(Node), &dev), "voxelizeToNodes");
        const uint32_t vs[3] = {vsr.x, vsr.y, vsr.z}, ms[3] = {msr.x, msr.y, msr.z};
        vox_status st = vox_voxelize_nodes(ctx, &g, kind, useMaterials ? 1 : 0, vs, ms, dev);
        if (st == VOX_OK) st = vox_sync(ctx);
        if (st != VOX_OK) {
            std::string m = vox_last_error(ctx);
            vox_free(ctx, dev);
            throw Exception((m.empty() ? "vox status " + std::to_string(int(st)) : m) +
                            " @ voxelizeToNodes.");
        }
        report("voxelizeToNodes");
        uint32_t d[3];
        vox_node_dims(&g, kind, d);
        NodePointer<Node> r = NodePointer<Node>();
        r.dev = 0;
        r.dim = make_uint3(d[0], d[1], d[2]);
        r.loc = make_uint3(0, 0, 0);
        if (!toHost) {
            r.ptr = (Node *)dev;
            return r;
        }
        r.ptr = new Node[n];
        st = vox_memcpy_d2h(ctx, r.ptr, dev, n * sizeof(Node));
        vox_free(ctx, dev);
        if (st != VOX_OK) {
            delete[] r.ptr;
            check(st, "voxelizeToNodesToRAM");
        }
        return r;
    }

    std::vector<NodePointer<Node> > plainParts(vox_plan_mode mode, double param, uint2 devConfig,
                                               uint3 vsr) {
        std::vector<NodePointer<Node> > result;
        if (devConfig.x == 1 && devConfig.y == 1) {
            result.push_back(plain(mode, param, vsr, false));
            return result;
        }
        ensureContext();
        const vox_grid g = plan(mode, param, vsr.x);
        const std::vector<Part> parts = partitions(devConfig, g);
        const uint32_t split[2] = {vsr.y, vsr.z};
        try {
            for (const Part &p : parts) {
                const uint3 dim =
                    make_uint3(g.alloc[0], p.max[0] - p.min[0] + 2, p.max[1] - p.min[1] + 2);
                const size_t words = (size_t)dim.x / 32 * dim.y * dim.z;
                void *dev = nullptr;
                check(vox_alloc(ctx, words * sizeof(VoxInt), &dev), "voxelize");
                vox_status st =
                    vox_solid_bits_partition(ctx, &g, split, p.min, p.max, (uint32_t *)dev);
                if (st == VOX_OK) st = vox_sync(ctx);
                if (st != VOX_OK) {
                    vox_free(ctx, dev);
                    check(st, "voxelize");
                }
                NodePointer<Node> r = NodePointer<Node>();
                r.vptr = (VoxInt *)dev;
                r.dev = 0;
                r.dim = dim;
                r.loc = p.loc;
                result.push_back(r);
            }
        } catch (...) {
            for (NodePointer<Node> &r : result) vox_free(ctx, r.vptr);
            throw;
        }
        report("voxelize");
        return result;
    }

    std::vector<NodePointer<Node> > nodeParts(vox_plan_mode mode, double param, uint2 devConfig,
                                              uint3 vsr, uint3 msr) {
        std::vector<NodePointer<Node> > result;
        if (devConfig.x == 1 && devConfig.y == 1) {
            result.push_back(nodes(mode, param, vsr, msr, false));
            return result;
        }
        ensureContext();
        const vox_grid g = plan(mode, param, vsr.x);
        const std::vector<Part> parts = partitions(devConfig, g);
        const vox_node_kind kind = vox_node_kind(Node::kind);
        const uint32_t vs[3] = {vsr.x, vsr.y, vsr.z}, ms[3] = {msr.x, msr.y, msr.z};
        try {
            for (const Part &p : parts) {
                const uint3 dim =
                    make_uint3(g.alloc[0], p.max[0] - p.min[0] + 2, p.max[1] - p.min[1] + 2);
                const size_t n = (size_t)dim.x * dim.y * dim.z;
                void *dev = nullptr;
                check(vox_alloc(ctx, n * sizeof(Node), &dev), "voxelizeToNodes");
                vox_status st = vox_voxelize_partition(ctx, &g, kind, useMaterials ? 1 : 0, vs,
                                                       ms, p.min, p.max, dev);
                if (st == VOX_OK) st = vox_sync(ctx);
                if (st != VOX_OK) {
                    vox_free(ctx, dev);
                    check(st, "voxelizeToNodes");
                }
                NodePointer<Node> r = NodePointer<Node>();
                r.ptr = (Node *)dev;
                r.dev = 0;
                r.dim = dim;
                r.loc = p.loc;
                result.push_back(r);
            }
        } catch (...) {
            for (NodePointer<Node> &r : result) vox_free(ctx, r.ptr);
            throw;
        }
        report("voxelizeToNodes");
        return result;
    }

    NodePointer<Node> sliceNodes(vox_plan_mode mode, double param, int direction, uint slice,
                                 uint2 vsr, uint2 msr, bool toHost) {
        ensureContext();
        const vox_node_kind kind = vox_node_kind(Node::kind);
        const uint32_t vs[2] = {vsr.x, vsr.y}, ms[2] = {msr.x, msr.y};
        const char *where = toHost ? "voxelizeSliceToRAM" : "voxelizeSlice";
        uint32_t d[3];
        check(vox_voxelize_slice(ctx, mode, param, kind, useMaterials ? 1 : 0, direction, slice,
                                 vs, ms, nullptr, d),
              where);
        const size_t n = (size_t)d[0] * d[1] * d[2];
        void *dev = nullptr;
        check(vox_alloc(ctx, n * sizeof(Node), &dev), where);
        vox_status st = vox_voxelize_slice(ctx, mode, param, kind, useMaterials ? 1 : 0,
                                           direction, slice, vs, ms, dev, d);
        if (st == VOX_OK) st = vox_sync(ctx);
        if (st != VOX_OK) {
            std::string m = vox_last_error(ctx);
            vox_free(ctx, dev);
            throw Exception((m.empty() ? "vox status " + std::to_string(int(st)) : m) + " @ " +
                            where + ".");
        }
        report(where);
        NodePointer<Node> r = NodePointer<Node>();
        r.dev = 0;
        r.dim = make_uint3(d[0], d[1], d[2]);
        r.loc = make_uint3(0, 0, 0);
        if (!toHost) {
            r.ptr = (Node *)dev;
            return r;
        }
        r.ptr = new Node[n];
        st = vox_memcpy_d2h(ctx, r.ptr, dev, n * sizeof(Node));
        vox_free(ctx, dev);
        if (st != VOX_OK) {
            delete[] r.ptr;
            check(st, where);
        }
        return r;
    }

    // getArrayDimensions (src/host_side.cpp:3329-3396): along x, the grid of
    // the model rotated a quarter turn about z (rotateVertices, 3261-3290:
    // (x, y) -> (float(a - y), float(b + x)), a = cy + cx, b = cy - cx with c
    // half the bounding box's extent), its dims un-rotated (x and y swap).
    uint3 arrayDims(vox_plan_mode mode, double param, uint maxInternalXSize, bool sliceAlongX) {
        vox_grid g;
        vox_status st;
        const uint32_t nv = uint32_t(vertices.size() / 3);
        if (sliceAlongX) {
            const double cx = (maxVertex.x - minVertex.x) / 2.0,
                         cy = (maxVertex.y - minVertex.y) / 2.0;
            const double a = cy + cx, b = cy - cx;
            std::vector<float> r(vertices);
            for (uint32_t i = 0; i < nv; i++) {
                r[3 * i] = float(a - double(vertices[3 * i + 1]));
                r[3 * i + 1] = float(b + double(vertices[3 * i]));
            }
            st = vox_plan_grid_host(r.data(), nv, mode, param, maxInternalXSize, &g);
        } else {
            st = vox_plan_grid_host(vertices.data(), nv, mode, param, maxInternalXSize, &g);
        }
        if (st != VOX_OK)
            throw Exception("Invalid arguments @ getArrayDimensions.");
        uint3 r = sliceAlongX ? make_uint3(g.alloc[1], g.alloc[0], g.alloc[2])
                              : make_uint3(g.alloc[0], g.alloc[1], g.alloc[2]);
        if (Node::isFCCNode()) {
            r.x *= 2;
            r.z *= 2;
        }
        return r;
    }
};

}  // namespace vox

#endif
