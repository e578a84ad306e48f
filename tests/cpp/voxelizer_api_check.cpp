// Exercises the C++ class vox::Voxelizer<Node> (include/vox/voxelizer.h) the
// way a program written against the reference class would.
//
//   voxelizer_api_check host
//       Checks that need no device: getters, getArrayDimensions (also along
//       x), the setMaterials checks, the printed-and-empty results for FCC
//       nodes in the maxDimension / plain / slice calls, the devConfig
//       exceptions, and (when no device exists) the "no device" exception.
//       Prints "host ok".
//   voxelizer_api_check dims <mesh.bin> <kind> <mode> <param> <maxX> <alongX>
//       Prints getArrayDimensions as "dims X Y Z".
//   voxelizer_api_check parts <mesh.bin> <kind> <plain> <mode> <param>
//                             <materials> <dcx> <dcy> <vsx> <vsy> <vsz>
//                             <msx> <msy> <msz> <out.bin>
//       voxelize / voxelizeToNodes with devConfig (dcx, dcy) inside
//       simulateMultidevice; writes the number of parts, then per part
//       dim[3], loc[3] (uint32) and the raw array.
//   voxelizer_api_check slice <mesh.bin> <kind> <mode> <param> <materials>
//                             <direction> <slice> <vsx> <vsy> <msx> <msy>
//                             <to_ram 0|1> <out.bin>
//       voxelizeSlice / voxelizeSliceToRAM; writes dim[3] and the array.
//   voxelizer_api_check run <mesh.bin> <node kind 0..4> <plain 0|1>
//                           <mode 0=maxDimension|1=cubeLength> <param>
//                           <materials 0|1> <vsx> <vsy> <vsz> <msx> <msy> <msz>
//                           <to_ram 0|1> <out.bin>
//       Voxelizes through voxelize{,ToRAM} (plain=1) or
//       voxelizeToNodes{,ToRAM} and writes dim[3] (uint32) followed by the
//       raw array.  mesh.bin: uint32 nv, nt, has_materials, n_unique, then
//       float32 vertices[3nv], uint32 indices[3nt], uint8 materials[nt].
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vox/voxelizer.h"

using namespace vox;

static int failures = 0;
#define EXPECT(c)                                                              \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            failures++;                                                        \
        }                                                                      \
    } while (0)

template <class F>
static bool throws(F f, const char *needle) {
    try {
        f();
    } catch (const Exception &e) {
        return std::strstr(e.what(), needle) != nullptr;
    }
    return false;
}

// A closed axis-aligned box [0,2]x[0,1]x[0,0.5] as 12 triangles.
static void box_mesh(std::vector<float> &v, std::vector<uint> &t) {
    const float X = 2.f, Y = 1.f, Z = 0.5f;
    v = {0, 0, 0, X, 0, 0, X, Y, 0, 0, Y, 0, 0, 0, Z, X, 0, Z, X, Y, Z, 0, Y, Z};
    t = {0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
         3, 7, 6, 3, 6, 2, 0, 4, 7, 0, 7, 3, 1, 2, 6, 1, 6, 5};
}

static int host_checks() {
    std::vector<float> v;
    std::vector<uint> t;
    box_mesh(v, t);
    Voxelizer<ShortNode> vx(v.data(), t.data(), 8, 12);
    EXPECT(vx.getNrOfTriangles() == 12);
    EXPECT(vx.getVertices() == v);
    EXPECT(vx.getIndices() == t);
    const double3 mn = vx.getMinVertex(), mx = vx.getMaxVertex();
    EXPECT(mn.x == 0.0 && mn.y == 0.0 && mn.z == 0.0);
    EXPECT(mx.x == 2.0 && mx.y == 1.0 && mx.z == 0.5);
    EXPECT(vx.getResolution().x == 256);  // initVariables default

    // maxDimension 101 on the x-dominant box: d = 2/100, x = 101 + 2 -> 128
    // (one x split of 1024, rounded to 32); y = ceil(1/0.02)+1 = 51 (+2);
    // z = ceil(0.5/0.02)+1 = 26 (+2).
    uint3 a = vx.getArrayDimensions(101u, 1024u, false);
    EXPECT(a.x == 128 && a.y == 53 && a.z == 28);
    // cubeLength 0.05: x = 40+1+2 = 43 -> 64, y = 21 (+2), z = 11 (+2).
    a = vx.getArrayDimensions(0.05, 1024u, false);
    EXPECT(a.x == 64 && a.y == 23 && a.z == 13);
    // Two x splits (maxInternalXSize 32) round x to a multiple of 64.
    a = vx.getArrayDimensions(0.05, 32u, false);
    EXPECT(a.x == 64);
    // Along x the box is turned a quarter about z: (x, y) -> (1.5 - y,
    // x - 0.5), extents 1 x 2 x 0.5, rotated grid 32 x 43 x 13 (x 21 + 2 ->
    // 32; y 41 + 2; z 11 + 2), reported un-rotated.
    a = vx.getArrayDimensions(0.05, 1024u, true);
    EXPECT(a.x == 43 && a.y == 32 && a.z == 13);

    Voxelizer<LongFCCNode> fx(v.data(), t.data(), 8, 12);
    a = fx.getArrayDimensions(0.05, 1024u, false);
    EXPECT(a.x == 128 && a.y == 23 && a.z == 26);
    a = fx.getArrayDimensions(0.05, 1024u, true);
    EXPECT(a.x == 86 && a.y == 32 && a.z == 26);
    // Requests the reference answers with a message and an empty result.
    EXPECT(fx.voxelizeToNodes(128u).empty());
    EXPECT(fx.voxelize(128u).empty());
    EXPECT(fx.voxelizeToRAM(0.05).ptr == nullptr);
    EXPECT(fx.voxelizeToNodesToRAM(128u).ptr == nullptr);
    EXPECT(fx.voxelizeSlice(128u, 0, 3u).empty());
    EXPECT(fx.voxelizeSliceToRAM(128u, 2, 3u).ptr == nullptr);

    std::vector<uchar> m(12, 3);
    EXPECT(throws([&] { vx.setMaterials(nullptr, 2); }, "null pointer"));
    EXPECT(throws([&] { vx.setMaterials(m.data(), 9); }, "exceeds"));
    vx.setMaterials(m.data(), 8);  // ShortNode holds 8 materials
    EXPECT(throws([&] { vx.voxelizeSlice(64u, 1, 3u, 2u); }, "devConfig must be 1"));

    if (vox_device_count() == 0) {
        EXPECT(throws([&] { vx.voxelizeToNodes(64u); }, "Couldn't find any HIP"));
        EXPECT(throws([&] { vx.voxelizeToNodes(64u, make_uint2(1, 2)); }, "Couldn't find any HIP"));
    } else if (vox_device_count() == 1) {
        EXPECT(throws([&] { vx.voxelizeToNodes(64u, make_uint2(1, 2)); }, "Not enough devices"));
    }
    std::printf("\n%s\n", failures ? "host FAILED" : "host ok");
    return failures ? 1 : 0;
}

template <class Node>
static int run(const std::vector<float> &v, const std::vector<uint> &t,
               const std::vector<uchar> &m, uint nUnique, bool plainOnly, int mode,
               double param, bool mats, uint3 vsr, uint3 msr, bool toRam, const char *out) {
    Voxelizer<Node> vx(v.data(), t.data(), uint(v.size() / 3), uint(t.size() / 3));
    if (!m.empty()) vx.setMaterials(m.data(), nUnique);
    vx.setMaterialOutput(mats);
    NodePointer<Node> p = NodePointer<Node>();
    size_t bytes = 0;
    if (plainOnly) {
        if (toRam)
            p = mode == 0 ? vx.voxelizeToRAM(uint(param), vsr) : vx.voxelizeToRAM(param, vsr);
        else
            p = (mode == 0 ? vx.voxelize(uint(param), make_uint2(1, 1), vsr)
                           : vx.voxelize(param, make_uint2(1, 1), vsr))
                    .at(0);
        bytes = (size_t)p.dim.x / 32 * p.dim.y * p.dim.z * sizeof(VoxInt);
    } else {
        if (toRam)
            p = mode == 0 ? vx.voxelizeToNodesToRAM(uint(param), vsr, msr)
                          : vx.voxelizeToNodesToRAM(param, vsr, msr);
        else
            p = (mode == 0 ? vx.voxelizeToNodes(uint(param), make_uint2(1, 1), vsr, msr)
                           : vx.voxelizeToNodes(param, make_uint2(1, 1), vsr, msr))
                    .at(0);
        bytes = (size_t)p.dim.x * p.dim.y * p.dim.z * sizeof(Node);
    }
    std::vector<unsigned char> host(bytes);
    if (toRam) {
        std::memcpy(host.data(), p.ptr, bytes);
        if (plainOnly)
            delete[] p.vptr;
        else
            delete[] p.ptr;
    } else {
        if (hipMemcpy(host.data(), p.ptr, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
            hipFree(p.ptr) != hipSuccess) {
            std::fprintf(stderr, "hipMemcpy/hipFree failed\n");
            return 1;
        }
    }
    FILE *f = std::fopen(out, "wb");
    if (!f) return 1;
    const uint32_t dim[3] = {p.dim.x, p.dim.y, p.dim.z};
    std::fwrite(dim, 4, 3, f);
    std::fwrite(host.data(), 1, bytes, f);
    std::fclose(f);
    std::printf("run ok: dim %u %u %u, voxel length %.17g\n", dim[0], dim[1], dim[2],
                vx.getVoxelLength());
    return 0;
}

struct Mesh {
    std::vector<float> v;
    std::vector<uint> t;
    std::vector<uchar> m;
    uint nUnique = 0;
};

static bool read_mesh(const char *path, Mesh &M) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    uint32_t hdr[4];
    if (std::fread(hdr, 4, 4, f) != 4) {
        std::fclose(f);
        return false;
    }
    M.v.resize(3 * (size_t)hdr[0]);
    M.t.resize(3 * (size_t)hdr[1]);
    M.m.resize(hdr[2] ? hdr[1] : 0);
    M.nUnique = hdr[3];
    bool ok = std::fread(M.v.data(), 4, M.v.size(), f) == M.v.size() &&
              std::fread(M.t.data(), 4, M.t.size(), f) == M.t.size() &&
              std::fread(M.m.data(), 1, M.m.size(), f) == M.m.size();
    std::fclose(f);
    return ok;
}

template <class Node>
static int dims(const Mesh &M, int mode, double param, uint maxX, bool alongX) {
    Voxelizer<Node> vx(M.v.data(), M.t.data(), uint(M.v.size() / 3), uint(M.t.size() / 3));
    const uint3 a = mode == 0 ? vx.getArrayDimensions(uint(param), maxX, alongX)
                              : vx.getArrayDimensions(param, maxX, alongX);
    std::printf("dims %u %u %u\n", a.x, a.y, a.z);
    return 0;
}

template <class Node>
static int parts(const Mesh &M, bool plainOnly, int mode, double param, bool mats, uint2 dc,
                 uint3 vsr, uint3 msr, const char *out) {
    Voxelizer<Node> vx(M.v.data(), M.t.data(), uint(M.v.size() / 3), uint(M.t.size() / 3));
    if (!M.m.empty()) vx.setMaterials(M.m.data(), M.nUnique);
    vx.setMaterialOutput(mats);
    std::vector<NodePointer<Node> > ps = vx.simulateMultidevice([&] {
        if (plainOnly)
            return mode == 0 ? vx.voxelize(uint(param), dc, vsr) : vx.voxelize(param, dc, vsr);
        return mode == 0 ? vx.voxelizeToNodes(uint(param), dc, vsr, msr)
                         : vx.voxelizeToNodes(param, dc, vsr, msr);
    });
    FILE *f = std::fopen(out, "wb");
    if (!f) return 1;
    const uint32_t n = uint32_t(ps.size());
    std::fwrite(&n, 4, 1, f);
    int rc = 0;
    for (NodePointer<Node> &p : ps) {
        const size_t bytes = plainOnly ? (size_t)p.dim.x / 32 * p.dim.y * p.dim.z * sizeof(VoxInt)
                                       : (size_t)p.dim.x * p.dim.y * p.dim.z * sizeof(Node);
        std::vector<unsigned char> host(bytes);
        if (hipMemcpy(host.data(), p.ptr, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
            hipFree(p.ptr) != hipSuccess)
            rc = 1;
        const uint32_t hdr[6] = {p.dim.x, p.dim.y, p.dim.z, p.loc.x, p.loc.y, p.loc.z};
        std::fwrite(hdr, 4, 6, f);
        std::fwrite(host.data(), 1, bytes, f);
    }
    std::fclose(f);
    std::printf("parts ok: %u parts\n", n);
    return rc;
}

template <class Node>
static int slice(const Mesh &M, int mode, double param, bool mats, int direction, uint sl,
                 uint2 vsr, uint2 msr, bool toRam, const char *out) {
    Voxelizer<Node> vx(M.v.data(), M.t.data(), uint(M.v.size() / 3), uint(M.t.size() / 3));
    if (!M.m.empty()) vx.setMaterials(M.m.data(), M.nUnique);
    vx.setMaterialOutput(mats);
    NodePointer<Node> p;
    if (toRam)
        p = mode == 0 ? vx.voxelizeSliceToRAM(uint(param), direction, sl, vsr, msr)
                      : vx.voxelizeSliceToRAM(param, direction, sl, vsr, msr);
    else
        p = (mode == 0 ? vx.voxelizeSlice(uint(param), direction, sl, 1u, vsr, msr)
                       : vx.voxelizeSlice(param, direction, sl, 1u, vsr, msr))
                .at(0);
    const size_t bytes = (size_t)p.dim.x * p.dim.y * p.dim.z * sizeof(Node);
    std::vector<unsigned char> host(bytes);
    if (toRam) {
        std::memcpy(host.data(), p.ptr, bytes);
        delete[] p.ptr;
    } else if (hipMemcpy(host.data(), p.ptr, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
               hipFree(p.ptr) != hipSuccess) {
        std::fprintf(stderr, "hipMemcpy/hipFree failed\n");
        return 1;
    }
    FILE *f = std::fopen(out, "wb");
    if (!f) return 1;
    const uint32_t dim[3] = {p.dim.x, p.dim.y, p.dim.z};
    std::fwrite(dim, 4, 3, f);
    std::fwrite(host.data(), 1, bytes, f);
    std::fclose(f);
    std::printf("slice ok: dim %u %u %u\n", dim[0], dim[1], dim[2]);
    return 0;
}

static int sub_command(int argc, char **argv) {
    const std::string cmd = argv[1];
    Mesh M;
    if (argc < 3 || !read_mesh(argv[2], M)) return 2;
    try {
        if (cmd == "dims" && argc == 8) {
            const int kind = std::atoi(argv[3]);
            const int mode = std::atoi(argv[4]);
            const double param = std::strtod(argv[5], nullptr);
            const uint maxX = uint(std::atoi(argv[6]));
            const bool alongX = std::atoi(argv[7]) != 0;
            switch (kind) {
            case 0: return dims<ShortNode>(M, mode, param, maxX, alongX);
            case 1: return dims<LongNode>(M, mode, param, maxX, alongX);
            case 3: return dims<ShortFCCNode>(M, mode, param, maxX, alongX);
            case 4: return dims<LongFCCNode>(M, mode, param, maxX, alongX);
            }
            return 2;
        }
        if (cmd == "parts" && argc == 17) {
            const int kind = std::atoi(argv[3]);
            const bool plainOnly = std::atoi(argv[4]) != 0;
            const int mode = std::atoi(argv[5]);
            const double param = std::strtod(argv[6], nullptr);
            const bool mats = std::atoi(argv[7]) != 0;
            const uint2 dc = make_uint2(std::atoi(argv[8]), std::atoi(argv[9]));
            const uint3 vsr = make_uint3(std::atoi(argv[10]), std::atoi(argv[11]), std::atoi(argv[12]));
            const uint3 msr = make_uint3(std::atoi(argv[13]), std::atoi(argv[14]), std::atoi(argv[15]));
            switch (kind) {
            case 0: return parts<ShortNode>(M, plainOnly, mode, param, mats, dc, vsr, msr, argv[16]);
            case 1: return parts<LongNode>(M, plainOnly, mode, param, mats, dc, vsr, msr, argv[16]);
            case 2: return parts<PartialNode>(M, plainOnly, mode, param, mats, dc, vsr, msr, argv[16]);
            case 3: return parts<ShortFCCNode>(M, plainOnly, mode, param, mats, dc, vsr, msr, argv[16]);
            }
            return 2;
        }
        if (cmd == "slice" && argc == 15) {
            const int kind = std::atoi(argv[3]);
            const int mode = std::atoi(argv[4]);
            const double param = std::strtod(argv[5], nullptr);
            const bool mats = std::atoi(argv[6]) != 0;
            const int direction = std::atoi(argv[7]);
            const uint sl = uint(std::atoi(argv[8]));
            const uint2 vsr = make_uint2(std::atoi(argv[9]), std::atoi(argv[10]));
            const uint2 msr = make_uint2(std::atoi(argv[11]), std::atoi(argv[12]));
            const bool toRam = std::atoi(argv[13]) != 0;
            switch (kind) {
            case 0: return slice<ShortNode>(M, mode, param, mats, direction, sl, vsr, msr, toRam, argv[14]);
            case 1: return slice<LongNode>(M, mode, param, mats, direction, sl, vsr, msr, toRam, argv[14]);
            case 2: return slice<PartialNode>(M, mode, param, mats, direction, sl, vsr, msr, toRam, argv[14]);
            case 4: return slice<LongFCCNode>(M, mode, param, mats, direction, sl, vsr, msr, toRam, argv[14]);
            }
            return 2;
        }
    } catch (const Exception &e) {
        std::fprintf(stderr, "vox::Exception: %s\n", e.what());
        return 3;
    }
    return 2;
}

int main(int argc, char **argv) {
    if (argc >= 2 && std::string(argv[1]) == "host") return host_checks();
    if (argc >= 2 && (std::string(argv[1]) == "dims" || std::string(argv[1]) == "parts" ||
                      std::string(argv[1]) == "slice"))
        return sub_command(argc, argv);
    if (argc != 16 || std::string(argv[1]) != "run") {
        std::fprintf(stderr, "usage: %s host | run <mesh.bin> <kind> <plain> <mode> <param> "
                             "<mats> <vsx> <vsy> <vsz> <msx> <msy> <msz> <to_ram> <out.bin>\n",
                     argv[0]);
        return 2;
    }
    FILE *f = std::fopen(argv[2], "rb");
    if (!f) return 2;
    uint32_t hdr[4];
    if (std::fread(hdr, 4, 4, f) != 4) return 2;
    std::vector<float> v(3 * (size_t)hdr[0]);
    std::vector<uint> t(3 * (size_t)hdr[1]);
    std::vector<uchar> m(hdr[2] ? hdr[1] : 0);
    bool ok = std::fread(v.data(), 4, v.size(), f) == v.size() &&
              std::fread(t.data(), 4, t.size(), f) == t.size() &&
              std::fread(m.data(), 1, m.size(), f) == m.size();
    std::fclose(f);
    if (!ok) return 2;
    const int kind = std::atoi(argv[3]);
    const bool plainOnly = std::atoi(argv[4]) != 0;
    const int mode = std::atoi(argv[5]);
    const double param = std::strtod(argv[6], nullptr);
    const bool mats = std::atoi(argv[7]) != 0;
    const uint3 vsr = make_uint3(std::atoi(argv[8]), std::atoi(argv[9]), std::atoi(argv[10]));
    const uint3 msr = make_uint3(std::atoi(argv[11]), std::atoi(argv[12]), std::atoi(argv[13]));
    const bool toRam = std::atoi(argv[14]) != 0;
    try {
        switch (kind) {
        case 0: return run<ShortNode>(v, t, m, hdr[3], plainOnly, mode, param, mats, vsr, msr, toRam, argv[15]);
        case 1: return run<LongNode>(v, t, m, hdr[3], plainOnly, mode, param, mats, vsr, msr, toRam, argv[15]);
        case 2: return run<PartialNode>(v, t, m, hdr[3], plainOnly, mode, param, mats, vsr, msr, toRam, argv[15]);
        case 3: return run<ShortFCCNode>(v, t, m, hdr[3], plainOnly, mode, param, mats, vsr, msr, toRam, argv[15]);
        case 4: return run<LongFCCNode>(v, t, m, hdr[3], plainOnly, mode, param, mats, vsr, msr, toRam, argv[15]);
        }
    } catch (const Exception &e) {
        std::fprintf(stderr, "vox::Exception: %s\n", e.what());
        return 3;
    }
    return 2;
}
