// LBVH build launchers (lbvh.hip) and the 4-wide BVH built from it.
#pragma once

#include "common.h"

#include <vector>

namespace mray
{

int    LaunchIota(uint32_t* dOut, uint32_t n, hipStream_t s);
int    LaunchPrimCentersAabbs(const float* dPos, const uint32_t* dIdx, uint32_t n, float* dCenters,
                              float* dAabbs, hipStream_t s);
int    LaunchReduceAabb(const float* dAabbs, uint32_t n, float* dOut, float* dPartial,
                        hipStream_t s);
size_t ReduceAabbTempBytes();
int    LaunchMorton63(const float* dCenters, uint32_t n, const float* dBox, uint64_t* dCodes,
                      hipStream_t s);
int    LaunchMorton30(const float* dCenters, uint32_t n, const float* dBox, uint32_t* dCodes,
                      hipStream_t s);
size_t RadixSortTempBytes(uint32_t n);
int    LaunchRadixSort(uint64_t* dKeys, uint32_t* dVals, uint32_t n, void* dTemp,
                       size_t tempBytes, hipStream_t s);
size_t LbvhTempBytes(uint32_t n);
int    LaunchLbvhBuild(const uint64_t* dSortedCodes, const uint32_t* dLeafOrder,
                       const float* dLeafAabbs, uint32_t n, uint32_t* dNodes, float* dBoxes,
                       uint32_t* dNodeRange, void* dTemp, size_t tempBytes, hipStream_t s,
                       uint32_t strictTie = 0);
// True when every internal node and leaf is reached exactly once from the
// root and each node's two children split its leaf range at one point.
bool ValidateBinaryTree(const std::vector<uint32_t>& nodes, const std::vector<uint32_t>& range,
                        uint32_t n);

// ---------------------------------------------------------------------------
// 4-wide BVH node: one 128-byte line, child boxes as SoA lanes so a node is
// eight dwordx4 loads.  child[k]: internal node index, or kLeafFlag |
// (count-1) << 26 | first, indexing the triangle soup (leaf order).  Unused
// slots carry an empty box (lo = +inf, hi = -inf).
// ---------------------------------------------------------------------------
struct alignas(128) BNode4
{
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(BNode4) == 128, "BNode4 must be one 128-byte line");

static constexpr uint32_t kLeafCountShift = 26;
static constexpr uint32_t kLeafStartMask = (1u << kLeafCountShift) - 1u;
// Triangles per leaf slot.  One: every triangle gets its own box in its
// parent node.  Measured on the 2.25 M-triangle terrain at 1080p x 64 spp:
// 4 -> 481, 3 -> 503, 2 -> 525, 1 -> 543 Mpaths/s; the box tests that cull a
// triangle cost less than the divergent intersection tests of a wider leaf.
static constexpr uint32_t kLeafMax = 1;

// Soup triangle: v0.xyz + triangle id bits, v1.xyz + cull flag, v2.xyz + 0.
struct alignas(16) SoupTri
{
    float v0[3];
    uint32_t prim;
    float v1[3];
    uint32_t cull;
    float v2[3];
    uint32_t pad;
};
static_assert(sizeof(SoupTri) == 48, "SoupTri layout");

struct DeviceArena;
// collapse.hip: the 4-wide tree of an n >= 2 leaf binary tree built on the
// device (nodes allocated from `ar`), and the triangle soup in leaf order.
int CollapseToBvh4Device(const uint32_t* dBinNodes, const float* dBinBoxes, const uint32_t* dBinRange,
                         const float* dTriAabbs, uint32_t n, DeviceArena& ar, BNode4** dNodesOut,
                         uint32_t* nodeCount, uint32_t* leafSlots, uint32_t* maxDepth, hipStream_t s);
int LaunchGatherSoup(const float* dPos, const uint32_t* dIdx, const uint8_t* dCull,
                     const uint32_t* dLeafOrder, uint32_t n, SoupTri* dSoup, hipStream_t s);

// Compressed 8-wide node (one 128-byte line): child boxes quantised to 8
// bits per plane on a per-node grid, decoded as fmaf(q, 2^e, origin) -- the
// builder checks every decoded box contains the exact one, so traversal stays
// conservative.  child[k] as in BNode4; empty slots have q_lo = 255 > q_hi = 0.
struct alignas(128) CNode8
{
    float ox, oy, oz;
    int8_t ex, ey, ez;
    uint8_t nChild;
    uint8_t qlox[8], qhix[8];
    uint8_t qloy[8], qhiy[8];
    uint8_t qloz[8], qhiz[8];
    uint32_t child[8];
    uint32_t pad[8];
};
static_assert(sizeof(CNode8) == 128, "CNode8 must be one 128-byte line");

struct CBvh8
{
    std::vector<CNode8> nodes;
    uint32_t maxDepth = 0;
    uint32_t leafSlots = 0;
};

struct WideBvh
{
    std::vector<BNode4> nodes;
    uint32_t maxDepth = 0;     // root = depth 1
    uint32_t leafSlots = 0;
};

// Binned-SAH binary tree over triangle boxes aabbs [n][6], in the LBVH's
// binary layout (nodes [n-1][3], boxes [n-1][6], leaf-order ranges [n-1][2])
// plus its leaf order; false if the result fails ValidateBinaryTree.
bool BuildSahBinary(const std::vector<float>& aabbs, uint32_t n, std::vector<uint32_t>& nodes,
                    std::vector<float>& boxes, std::vector<uint32_t>& range,
                    std::vector<uint32_t>& order);

// Collapse the binary LBVH (reference layout) into BNode4s, top-down,
// expanding the child with the largest surface area until four children;
// subtrees of <= leafMax primitives become leaves over their leaf-order range.
WideBvh CollapseToBvh4(const std::vector<uint32_t>& binNodes, const std::vector<float>& binBoxes,
                       const std::vector<uint32_t>& binRange, const std::vector<float>& leafAabbs,
                       const std::vector<uint32_t>& leafOrder,
                       uint32_t leafMax = kLeafMax);

// Same collapse with eight children per node, then quantised into CNode8s;
// false if a box cannot be represented conservatively.
bool CollapseToCBvh8(const std::vector<uint32_t>& binNodes, const std::vector<float>& binBoxes,
                     const std::vector<uint32_t>& binRange, const std::vector<float>& leafAabbs,
                     const std::vector<uint32_t>& leafOrder, CBvh8& out);
bool QuantizeCNode8(const float box[8][6], const uint32_t child[8], int count, CNode8& out);

} // namespace mray
