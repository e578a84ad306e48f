// vox/node_types.h -- node structs with the bit layouts of the reference's
// include/node_types.h (hakarlss/Voxelizer), so arrays produced here can be
// read by code written against the reference:
//   ShortNode     1 B : bim = mat<<5 | bid&31            (node_types.h:19-177)
//   LongNode      2 B : {uchar bid, uchar mat}           (node_types.h:178-283)
//   PartialNode   8 B : {uchar bid, uchar mat, float r}  (node_types.h:284-399)
//   ShortFCCNode  2 B : bim = mat<<12 | bid&4095         (node_types.h:400-565)
//   LongFCCNode   4 B : {ushort bid, uchar mat, pad}     (node_types.h:566-706)
// The accessors keep the reference's names and semantics (bid(), mat(),
// setters, maxMat(), isFCCNode(), ...).  Host and device (HIP) usable.
#ifndef VOX_NODE_TYPES_H_AMD
#define VOX_NODE_TYPES_H_AMD

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define VOX_HD __host__ __device__
#else
#define VOX_HD
#endif

namespace vox {

typedef unsigned int uint;
typedef uint16_t ushort;
typedef uint8_t uchar;
typedef uint32_t VoxInt;  // 32-bit voxel words, voxel x at bit 31 - x%32

class ShortNode {
public:
    VOX_HD ShortNode() : bim(0) {}
    VOX_HD ShortNode(uchar b) : bim(b) {}
    VOX_HD ShortNode(uchar bid_, uchar mat_) { set(mat_, bid_); }
    VOX_HD uchar mat() const { return bim >> SHIFT; }
    VOX_HD uchar bid() const { return bim & MASK; }
    VOX_HD float r() const { return 0.0f; }
    VOX_HD void mat(uchar m) { bim = (uchar)((m << SHIFT) | (bim & MASK)); }
    VOX_HD void bid(uchar b) { bim = (uchar)((b & MASK) | (bim & MASK_INV)); }
    VOX_HD void r(float) {}
    VOX_HD void set(uchar m, uchar b) { bim = (uchar)((m << SHIFT) | (b & MASK)); }
    static VOX_HD uchar maxMat() { return 7; }
    static VOX_HD uchar maxBid() { return 31; }
    static VOX_HD bool hasRatio() { return false; }
    static VOX_HD bool isFCCNode() { return false; }
    static constexpr int kind = 0;

private:
    uchar bim;
    static const uchar MASK = 31, MASK_INV = 224, SHIFT = 5;
};

class LongNode {
public:
    VOX_HD LongNode() : _bid(0), _mat(0) {}
    VOX_HD LongNode(uchar b) : _bid(b), _mat(0) {}
    VOX_HD LongNode(uchar b, uchar m) : _bid(b), _mat(m) {}
    VOX_HD uchar mat() const { return _mat; }
    VOX_HD uchar bid() const { return _bid; }
    VOX_HD float r() const { return 0.0f; }
    VOX_HD void mat(uchar m) { _mat = m; }
    VOX_HD void bid(uchar b) { _bid = b; }
    VOX_HD void r(float) {}
    VOX_HD void set(uchar m, uchar b) { _mat = m; _bid = b; }
    static VOX_HD uchar maxMat() { return 255; }
    static VOX_HD uchar maxBid() { return 255; }
    static VOX_HD bool hasRatio() { return false; }
    static VOX_HD bool isFCCNode() { return false; }
    static constexpr int kind = 1;

private:
    uchar _bid, _mat;
};

class PartialNode {
public:
    VOX_HD PartialNode() : _bid(0), _mat(0), _r(0.0f) {}
    VOX_HD PartialNode(uchar b) : _bid(b), _mat(0), _r(0.0f) {}
    VOX_HD PartialNode(uchar b, uchar m) : _bid(b), _mat(m), _r(0.0f) {}
    VOX_HD PartialNode(uchar b, uchar m, float ratio) : _bid(b), _mat(m), _r(ratio) {}
    VOX_HD uchar mat() const { return _mat; }
    VOX_HD uchar bid() const { return _bid; }
    VOX_HD float r() const { return _r; }
    VOX_HD void mat(uchar m) { _mat = m; }
    VOX_HD void bid(uchar b) { _bid = b; }
    VOX_HD void r(float ratio) { _r = ratio; }
    VOX_HD void set(uchar m, uchar b) { _mat = m; _bid = b; }
    VOX_HD void set(uchar m, uchar b, float ratio) { _mat = m; _bid = b; _r = ratio; }
    static VOX_HD uchar maxMat() { return 255; }
    static VOX_HD uchar maxBid() { return 255; }
    static VOX_HD bool hasRatio() { return true; }
    static VOX_HD bool isFCCNode() { return false; }
    static constexpr int kind = 2;

private:
    uchar _bid, _mat;
    float _r;
};

class ShortFCCNode {
public:
    VOX_HD ShortFCCNode() : bim(0) {}
    VOX_HD ShortFCCNode(ushort b) : bim(b) {}
    VOX_HD ShortFCCNode(ushort bid_, ushort mat_) { set(mat_, bid_); }
    VOX_HD ushort mat() const { return bim >> SHIFT; }
    VOX_HD ushort bid() const { return bim & MASK; }
    VOX_HD float r() const { return 0.0f; }
    VOX_HD void mat(ushort m) { bim = (ushort)((m << SHIFT) | (bim & MASK)); }
    VOX_HD void bid(ushort b) { bim = (ushort)((b & MASK) | (bim & MASK_INV)); }
    VOX_HD void r(float) {}
    VOX_HD void set(ushort m, ushort b) { bim = (ushort)((m << SHIFT) | (b & MASK)); }
    static VOX_HD ushort maxMat() { return 15; }
    static VOX_HD bool hasRatio() { return false; }
    static VOX_HD bool isFCCNode() { return true; }
    static constexpr int kind = 3;

private:
    ushort bim;
    static const ushort MASK = 4095, MASK_INV = 61440, SHIFT = 12;
};

class LongFCCNode {
public:
    VOX_HD LongFCCNode() : _bid(0), _mat(0), _pad(0) {}
    VOX_HD LongFCCNode(ushort b) : _bid(b), _mat(0), _pad(0) {}
    VOX_HD LongFCCNode(ushort b, uchar m) : _bid(b), _mat(m), _pad(0) {}
    VOX_HD uchar mat() const { return _mat; }
    VOX_HD ushort bid() const { return _bid; }
    VOX_HD float r() const { return 0.0f; }
    VOX_HD void mat(uchar m) { _mat = m; }
    VOX_HD void bid(ushort b) { _bid = b; }
    VOX_HD void r(float) {}
    VOX_HD void set(uchar m, ushort b) { _bid = b; _mat = m; }
    static VOX_HD ushort maxMat() { return 255; }
    static VOX_HD bool hasRatio() { return false; }
    static VOX_HD bool isFCCNode() { return true; }
    static constexpr int kind = 4;

private:
    ushort _bid;
    uchar _mat;
    uchar _pad;  // written as 0 by this library
};

static_assert(sizeof(ShortNode) == 1 && sizeof(LongNode) == 2 && sizeof(PartialNode) == 8 &&
                  sizeof(ShortFCCNode) == 2 && sizeof(LongFCCNode) == 4,
              "node sizes must match the reference layouts");

}  // namespace vox

#endif
