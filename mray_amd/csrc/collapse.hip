// GPU collapse of the binary LBVH into the 4-wide traversal BVH, and the
// triangle soup in leaf order: the device half of the LBVH builder, so a
// scene goes from triangles to traversal tree without leaving HBM.
//
// The rule is the host collapse's (bvh4.cpp, CollapseWide): a wide node
// starts from its binary node's two children and repeatedly splits the
// child of largest surface area that holds more than kLeafMax triangles,
// until it has four children or none can be split; children are ordered by
// leaf-order position; a child of <= kLeafMax triangles is a leaf slot over
// its leaf-order range.  Levels are built one at a time: one thread per wide
// node of the level counts its internal children, an exclusive scan numbers
// them, and a second pass writes the nodes and the next level.  Nodes are
// therefore numbered level by level (the root 0, then its children, ...),
// where the host collapse numbers them depth first; the tree is the same
// tree, so hits and node visits are the same.
#include "lbvh.h"
#include "scene.h"

#include <hipcub/hipcub.hpp>

#include <string>

namespace mray
{
namespace
{
constexpr uint32_t kCollapseTPB = 256;

struct Kid
{
    uint32_t idx;        // binary node, or triangle for a binary leaf
    uint32_t lo, hi;     // leaf-order range
    bool     leaf;       // a binary leaf (one triangle)
    float    box[6];
};

__device__ __forceinline__ float Area(const float* b)
{
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    if(dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ Kid ChildOf(const uint32_t* __restrict__ binNodes,
                                       const float* __restrict__ binBoxes,
                                       const uint32_t* __restrict__ binRange,
                                       const float* __restrict__ triAabbs, uint32_t parent, int side)
{
    Kid c;
    const uint32_t v = binNodes[size_t(parent) * 3 + side];
    if(v & kLeafFlag)
    {
        c.leaf = true;
        c.idx = v & ~kLeafFlag;
        c.lo = c.hi = binRange[2 * size_t(parent) + side];
        for(int k = 0; k < 6; k++) c.box[k] = triAabbs[size_t(c.idx) * 6 + k];
    }
    else
    {
        c.leaf = false;
        c.idx = v;
        c.lo = binRange[2 * size_t(v)];
        c.hi = binRange[2 * size_t(v) + 1];
        for(int k = 0; k < 6; k++) c.box[k] = binBoxes[size_t(v) * 6 + k];
    }
    return c;
}

// The children of the wide node over binary node `bin`, in leaf order.
__device__ __forceinline__ int Expand(const uint32_t* __restrict__ binNodes,
                                      const float* __restrict__ binBoxes,
                                      const uint32_t* __restrict__ binRange,
                                      const float* __restrict__ triAabbs, uint32_t bin, Kid kids[4])
{
    int count = 2;
    kids[0] = ChildOf(binNodes, binBoxes, binRange, triAabbs, bin, 0);
    kids[1] = ChildOf(binNodes, binBoxes, binRange, triAabbs, bin, 1);
    while(count < 4)
    {
        int best = -1;
        float bestArea = -1.0f;
        for(int k = 0; k < count; k++)
        {
            const uint32_t cnt = kids[k].hi - kids[k].lo + 1u;
            if(kids[k].leaf || cnt <= kLeafMax) continue;
            const float ar = Area(kids[k].box);
            if(ar > bestArea)
            {
                bestArea = ar;
                best = k;
            }
        }
        if(best < 0) break;
        const uint32_t b = kids[best].idx;
        kids[best] = ChildOf(binNodes, binBoxes, binRange, triAabbs, b, 0);
        kids[count++] = ChildOf(binNodes, binBoxes, binRange, triAabbs, b, 1);
    }
    // insertion sort by leaf-order position (the ranges are disjoint)
    for(int i = 1; i < count; i++)
    {
        Kid x = kids[i];
        int j = i - 1;
        while(j >= 0 && kids[j].lo > x.lo)
        {
            kids[j + 1] = kids[j];
            j--;
        }
        kids[j + 1] = x;
    }
    return count;
}

__device__ __forceinline__ bool Internal(const Kid& c) { return c.hi - c.lo + 1u > kLeafMax; }

// Internal children of every wide node of the level.
__global__ void __launch_bounds__(kCollapseTPB)
KCollapseCount(const uint32_t* __restrict__ levelBin, uint32_t levelCount,
               const uint32_t* __restrict__ binNodes, const float* __restrict__ binBoxes,
               const uint32_t* __restrict__ binRange, const float* __restrict__ triAabbs,
               uint32_t* __restrict__ counts)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= levelCount) return;
    Kid kids[4];
    const int n = Expand(binNodes, binBoxes, binRange, triAabbs, levelBin[i], kids);
    uint32_t c = 0;
    for(int k = 0; k < n; k++) c += Internal(kids[k]) ? 1u : 0u;
    counts[i] = c;
}

// Writes the level's nodes: wide node levelBase + i over levelBin[i]; its
// internal children become wide nodes nextBase + offsets[i] + j and the next
// level's entries offsets[i] + j.
__global__ void __launch_bounds__(kCollapseTPB)
KCollapseEmit(const uint32_t* __restrict__ levelBin, uint32_t levelCount, uint32_t levelBase,
              const uint32_t* __restrict__ offsets, uint32_t nextBase,
              const uint32_t* __restrict__ binNodes, const float* __restrict__ binBoxes,
              const uint32_t* __restrict__ binRange, const float* __restrict__ triAabbs,
              BNode4* __restrict__ nodes, uint32_t* __restrict__ nextBin,
              unsigned long long* __restrict__ leafSlots)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= levelCount) return;
    Kid kids[4];
    const int n = Expand(binNodes, binBoxes, binRange, triAabbs, levelBin[i], kids);
    BNode4 nd;
    uint32_t next = offsets[i], leaves = 0;
    for(int k = 0; k < 4; k++)
    {
        if(k < n)
        {
            const Kid& c = kids[k];
            nd.lox[k] = c.box[0]; nd.loy[k] = c.box[1]; nd.loz[k] = c.box[2];
            nd.hix[k] = c.box[3]; nd.hiy[k] = c.box[4]; nd.hiz[k] = c.box[5];
            if(Internal(c))
            {
                nd.child[k] = nextBase + next;
                nextBin[next] = c.idx;
                next++;
            }
            else
            {
                nd.child[k] = kLeafFlag | ((c.hi - c.lo) << kLeafCountShift) | c.lo;
                leaves++;
            }
        }
        else
        {
            nd.lox[k] = nd.loy[k] = nd.loz[k] = INFINITY;
            nd.hix[k] = nd.hiy[k] = nd.hiz[k] = -INFINITY;
            nd.child[k] = kInvalid;
        }
        nd.pad[k] = 0;
    }
    nodes[levelBase + i] = nd;
    if(leaves) atomicAdd(leafSlots, (unsigned long long)leaves);
}

// soup[p] = triangle leafOrder[p]: its three positions, id and cull flag.
__global__ void __launch_bounds__(kCollapseTPB)
KGatherSoup(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
            const uint8_t* __restrict__ cull, const uint32_t* __restrict__ leafOrder, uint32_t n,
            SoupTri* __restrict__ soup)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= n) return;
    const uint32_t t = leafOrder[p];
    SoupTri o;
    float* dst[3] = {o.v0, o.v1, o.v2};
    for(int v = 0; v < 3; v++)
    {
        const float* src = pos + 3 * size_t(idx[3 * size_t(t) + v]);
        dst[v][0] = src[0];
        dst[v][1] = src[1];
        dst[v][2] = src[2];
    }
    o.prim = t;
    o.cull = cull ? uint32_t(cull[t]) : 0u;
    o.pad = 0;
    soup[p] = o;
}
}  // namespace

// The 4-wide tree of a binary tree of n >= 2 leaves, on the device.
// dNodesOut receives a device array of *nodeCount nodes allocated from `ar`.
int CollapseToBvh4Device(const uint32_t* dBinNodes, const float* dBinBoxes, const uint32_t* dBinRange,
                         const float* dTriAabbs, uint32_t n, DeviceArena& ar, BNode4** dNodesOut,
                         uint32_t* nodeCount, uint32_t* leafSlots, uint32_t* maxDepth, hipStream_t s)
{
    if(n < 2)
    {
        SetError("device collapse needs at least two triangles");
        return MRAY_ERR_INVALID;
    }
    // at most n - 1 wide nodes (each takes at least one binary node) and at
    // most n - 1 entries on a level
    const uint32_t cap = n - 1u;
    BNode4* dWork = nullptr;
    uint32_t *dLevel = nullptr, *dNext = nullptr, *dCounts = nullptr, *dOffsets = nullptr;
    unsigned long long* dLeafSlots = nullptr;
    void* dScan = nullptr;
    size_t scanBytes = 0;
    MRAY_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, (const uint32_t*)nullptr,
                                                    (uint32_t*)nullptr, int(cap + 1u), s));
    int rc = MRAY_OK;
    auto fail = [&](hipError_t e) { SetError(std::string("collapse: ") + hipGetErrorString(e)); rc = MRAY_ERR_HIP; };
    hipError_t e = hipSuccess;
    if((e = hipMalloc((void**)&dWork, size_t(cap) * sizeof(BNode4))) != hipSuccess ||
       (e = hipMalloc((void**)&dLevel, size_t(cap) * 4)) != hipSuccess ||
       (e = hipMalloc((void**)&dNext, size_t(cap) * 4)) != hipSuccess ||
       (e = hipMalloc((void**)&dCounts, size_t(cap + 1u) * 4)) != hipSuccess ||
       (e = hipMalloc((void**)&dOffsets, size_t(cap + 1u) * 4)) != hipSuccess ||
       (e = hipMalloc((void**)&dLeafSlots, sizeof(unsigned long long))) != hipSuccess ||
       (e = hipMalloc(&dScan, std::max<size_t>(scanBytes, 4))) != hipSuccess)
        fail(e);
    uint32_t total = 1, levelCount = 1, levelBase = 0, depth = 0;
    const uint32_t root = 0;
    const unsigned long long zero = 0;
    if(rc == MRAY_OK)
    {
        if((e = hipMemcpyAsync(dLevel, &root, 4, hipMemcpyHostToDevice, s)) != hipSuccess ||
           (e = hipMemcpyAsync(dLeafSlots, &zero, sizeof(zero), hipMemcpyHostToDevice, s)) != hipSuccess ||
           (e = hipStreamSynchronize(s)) != hipSuccess)
            fail(e);
    }
    while(rc == MRAY_OK && levelCount > 0)
    {
        depth++;
        const uint32_t blocks = DivUp(levelCount, kCollapseTPB);
        KCollapseCount<<<blocks, kCollapseTPB, 0, s>>>(dLevel, levelCount, dBinNodes, dBinBoxes,
                                                        dBinRange, dTriAabbs, dCounts);
        // counts[levelCount] = 0 makes offsets[levelCount] the level's total
        if((e = hipMemcpyAsync(dCounts + levelCount, &root, 4, hipMemcpyHostToDevice, s)) != hipSuccess) { fail(e); break; }
        if((e = hipcub::DeviceScan::ExclusiveSum(dScan, scanBytes, dCounts, dOffsets,
                                                 int(levelCount + 1u), s)) != hipSuccess) { fail(e); break; }
        uint32_t nextCount = 0;
        if((e = hipMemcpyAsync(&nextCount, dOffsets + levelCount, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
           (e = hipStreamSynchronize(s)) != hipSuccess) { fail(e); break; }
        if(total + nextCount > cap)
        {
            SetError("collapse: more wide nodes than binary nodes (malformed tree)");
            rc = MRAY_ERR_INVALID;
            break;
        }
        KCollapseEmit<<<blocks, kCollapseTPB, 0, s>>>(dLevel, levelCount, levelBase, dOffsets, total,
                                                       dBinNodes, dBinBoxes, dBinRange, dTriAabbs,
                                                       dWork, dNext, dLeafSlots);
        if((e = hipGetLastError()) != hipSuccess) { fail(e); break; }
        levelBase = total;
        total += nextCount;
        levelCount = nextCount;
        std::swap(dLevel, dNext);
    }
    if(rc == MRAY_OK)
    {
        unsigned long long slots = 0;
        BNode4* dNodes = nullptr;
        rc = DevAlloc(ar, (void**)&dNodes, size_t(total) * sizeof(BNode4));
        if(rc == MRAY_OK &&
           ((e = hipMemcpyAsync(dNodes, dWork, size_t(total) * sizeof(BNode4), hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(&slots, dLeafSlots, sizeof(slots), hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess))
            fail(e);
        if(rc == MRAY_OK)
        {
            *dNodesOut = dNodes;
            *nodeCount = total;
            *leafSlots = uint32_t(slots);
            *maxDepth = depth;
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(dWork);
    (void)hipFree(dLevel);
    (void)hipFree(dNext);
    (void)hipFree(dCounts);
    (void)hipFree(dOffsets);
    (void)hipFree(dLeafSlots);
    (void)hipFree(dScan);
    return rc;
}

int LaunchGatherSoup(const float* dPos, const uint32_t* dIdx, const uint8_t* dCull,
                     const uint32_t* dLeafOrder, uint32_t n, SoupTri* dSoup, hipStream_t s)
{
    if(n == 0) return MRAY_OK;
    KGatherSoup<<<DivUp(n, kCollapseTPB), kCollapseTPB, 0, s>>>(dPos, dIdx, dCull, dLeafOrder, n, dSoup);
    MRAY_HIP_CHECK(hipGetLastError());
    return MRAY_OK;
}

}  // namespace mray
