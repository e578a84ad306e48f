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
// Built here for one device: devConfig must be (1,1), and the slicing entry
// points and simulateMultidevice throw vox::Exception.
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
        checkDevConfig(devConfig);
        result.push_back(plain(VOX_PLAN_MAX_DIMENSION, double(maxDimension), voxSplitRes, false));
        return result;
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
        checkDevConfig(devConfig);
        result.push_back(plain(VOX_PLAN_CUBE_LENGTH, cubeLength, voxSplitRes, false));
        return result;
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
        checkDevConfig(devConfig);
        result.push_back(nodes(VOX_PLAN_MAX_DIMENSION, double(maxDimension), voxSplitRes,
                               matSplitRes, false));
        return result;
    }
    std::vector<NodePointer<Node> > voxelizeToNodes(double cubeLength,
                                                    uint2 devConfig = make_uint2(1, 1),
                                                    uint3 voxSplitRes = make_uint3(1024, 512, 512),
                                                    uint3 matSplitRes = make_uint3(512, 512, 512)) {
        std::vector<NodePointer<Node> > result;
        checkDevConfig(devConfig);
        result.push_back(nodes(VOX_PLAN_CUBE_LENGTH, cubeLength, voxSplitRes, matSplitRes, false));
        return result;
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

    // ---- slicing (src/host_side.cpp:1927-2262): not built ------------------
    std::vector<NodePointer<Node> > voxelizeSlice(uint, int, uint, uint = 1,
                                                  uint2 = make_uint2(1024, 512),
                                                  uint2 = make_uint2(512, 512)) {
        throw Exception("Slice voxelization is not built by this library @ voxelizeSlice.");
    }
    std::vector<NodePointer<Node> > voxelizeSlice(double, int, uint, uint = 1,
                                                  uint2 = make_uint2(1024, 512),
                                                  uint2 = make_uint2(512, 512)) {
        throw Exception("Slice voxelization is not built by this library @ voxelizeSlice.");
    }
    NodePointer<Node> voxelizeSliceToRAM(uint, int, uint, uint2 = make_uint2(1024, 512),
                                         uint2 = make_uint2(512, 512)) {
        throw Exception("Slice voxelization is not built by this library @ voxelizeSliceToRAM.");
    }
    NodePointer<Node> voxelizeSliceToRAM(double, int, uint, uint2 = make_uint2(1024, 512),
                                         uint2 = make_uint2(512, 512)) {
        throw Exception("Slice voxelization is not built by this library @ voxelizeSliceToRAM.");
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

    // src/host_side.cpp:3651-3659: not built (one device).
    std::vector<NodePointer<Node> > simulateMultidevice(
        std::function<std::vector<NodePointer<Node> >()>) {
        throw Exception("simulateMultidevice is not built by this library.");
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

    // initDevices (src/host_side.cpp:163-202) for one device.
    void checkDevConfig(uint2 devConfig) {
        if (devConfig.x * devConfig.y != 1)
            throw Exception("Multi-device voxelization is not built: devConfig must be (1,1).");
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

    uint3 arrayDims(vox_plan_mode mode, double param, uint maxInternalXSize, bool sliceAlongX) {
        if (sliceAlongX)
            throw Exception("Slicing along x is not built by this library @ getArrayDimensions.");
        vox_grid g;
        const vox_status st = vox_plan_grid_host(vertices.data(), uint32_t(vertices.size() / 3),
                                                 mode, param, maxInternalXSize, &g);
        if (st != VOX_OK)
            throw Exception("Invalid arguments @ getArrayDimensions.");
        uint3 r = make_uint3(g.alloc[0], g.alloc[1], g.alloc[2]);
        if (Node::isFCCNode()) {
            r.x *= 2;
            r.z *= 2;
        }
        return r;
    }
};

}  // namespace vox

#endif
