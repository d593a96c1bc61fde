// Binned-SAH binary BVH over triangle bounds (host, multi-threaded), in the
// binary-tree layout the LBVH build produces (children / parent per internal
// node, leaf-order range, node box), so the same 4-wide collapse consumes it.
//
// The reference traverses its Karras LBVH directly (AcceleratorLBVH.hpp:14-107).
// Morton-order splits give boxes that overlap heavily on terrain-like meshes;
// the surface-area heuristic (MacDonald & Booth 1990; binned evaluation as in
// Wald 2007, "On fast construction of SAH-based bounding volume hierarchies")
// cuts the nodes and triangles each ray visits.  The LBVH is still built on
// the GPU for the reference-layout products; this tree only feeds traversal.
//
// Internal nodes are numbered by their split position (the last leaf-order
// position of the left child): every internal node of a binary tree over
// ordered leaves has a distinct split position in [0, n-2], so the numbering
// is a bijection, independent of the build order (deterministic with any
// number of threads).  The root's number is swapped with 0.
#include "lbvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

namespace mray
{

namespace
{

constexpr int kBins = 16;

struct Box
{
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void Grow(const float* b)
    {
        for(int d = 0; d < 3; d++)
        {
            lo[d] = std::min(lo[d], b[d]);
            hi[d] = std::max(hi[d], b[3 + d]);
        }
    }
    void Grow(const Box& b)
    {
        for(int d = 0; d < 3; d++)
        {
            lo[d] = std::min(lo[d], b.lo[d]);
            hi[d] = std::max(hi[d], b.hi[d]);
        }
    }
    float Area() const
    {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if(!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Builder
{
    const std::vector<float>& aabbs;
    std::vector<float> centers;
    std::vector<uint32_t>& order;
    std::vector<uint32_t>& nodes;
    std::vector<float>& boxes;
    std::vector<uint32_t>& range;
    std::atomic<int> threadsLeft;
    uint32_t n;

    Builder(const std::vector<float>& a, uint32_t count, std::vector<uint32_t>& o,
            std::vector<uint32_t>& nd, std::vector<float>& bx, std::vector<uint32_t>& rg, int threads)
        : aabbs(a), order(o), nodes(nd), boxes(bx), range(rg), threadsLeft(threads), n(count)
    {
        centers.resize(size_t(n) * 3);
        for(uint32_t i = 0; i < n; i++)
            for(int d = 0; d < 3; d++)
                centers[size_t(i) * 3 + d] = 0.5f * (aabbs[size_t(i) * 6 + d] + aabbs[size_t(i) * 6 + 3 + d]);
    }

    // Split [lo, hi] (hi > lo); returns the split position (last of left).
    uint32_t Split(uint32_t lo, uint32_t hi)
    {
        float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        for(uint32_t p = lo; p <= hi; p++)
        {
            const float* c = &centers[size_t(order[p]) * 3];
            for(int d = 0; d < 3; d++)
            {
                cmin[d] = std::min(cmin[d], c[d]);
                cmax[d] = std::max(cmax[d], c[d]);
            }
        }
        float bestCost = INFINITY;
        int bestAxis = -1, bestBin = -1;
        for(int d = 0; d < 3; d++)
        {
            float ext = cmax[d] - cmin[d];
            if(!(ext > 0.0f)) continue;
            float inv = float(kBins) / ext;
            Box bb[kBins];
            uint32_t cnt[kBins] = {};
            for(uint32_t p = lo; p <= hi; p++)
            {
                uint32_t t = order[p];
                int b = std::min(kBins - 1, int((centers[size_t(t) * 3 + d] - cmin[d]) * inv));
                cnt[b]++;
                bb[b].Grow(&aabbs[size_t(t) * 6]);
            }
            float rightArea[kBins];
            uint32_t rightCnt[kBins];
            Box acc;
            uint32_t c = 0;
            for(int b = kBins - 1; b > 0; b--)
            {
                acc.Grow(bb[b]);
                c += cnt[b];
                rightArea[b] = acc.Area();
                rightCnt[b] = c;
            }
            Box left;
            uint32_t lc = 0;
            for(int b = 0; b < kBins - 1; b++)
            {
                left.Grow(bb[b]);
                lc += cnt[b];
                uint32_t rc = rightCnt[b + 1];
                if(lc == 0 || rc == 0) continue;
                float cost = left.Area() * float(lc) + rightArea[b + 1] * float(rc);
                if(cost < bestCost)
                {
                    bestCost = cost;
                    bestAxis = d;
                    bestBin = b;
                }
            }
        }
        if(bestAxis < 0)
        {
            // all centroids coincide: split the range in the middle
            return lo + (hi - lo) / 2u;
        }
        const float inv = float(kBins) / (cmax[bestAxis] - cmin[bestAxis]);
        auto mid = std::stable_partition(order.begin() + lo, order.begin() + hi + 1, [&](uint32_t t)
        {
            int b = std::min(kBins - 1, int((centers[size_t(t) * 3 + bestAxis] - cmin[bestAxis]) * inv));
            return b <= bestBin;
        });
        uint32_t m = uint32_t(mid - order.begin());
        if(m <= lo || m > hi) return lo + (hi - lo) / 2u;
        return m - 1u;
    }

    // Builds the subtree over [lo, hi]; returns its child reference and box.
    uint32_t Build(uint32_t lo, uint32_t hi, uint32_t parent, Box& boxOut)
    {
        if(lo == hi)
        {
            uint32_t t = order[lo];
            boxOut = Box{};
            boxOut.Grow(&aabbs[size_t(t) * 6]);
            return kLeafFlag | t;
        }
        uint32_t mid = Split(lo, hi);
        uint32_t self = mid;
        Box lb, rb;
        uint32_t l, r;
        const uint32_t size = hi - lo + 1u;
        bool spawned = false;
        if(size > 32768u && threadsLeft.fetch_sub(1) > 0)
        {
            std::thread th([&] { l = Build(lo, mid, self, lb); });
            r = Build(mid + 1u, hi, self, rb);
            th.join();
            spawned = true;
        }
        if(!spawned)
        {
            l = Build(lo, mid, self, lb);
            r = Build(mid + 1u, hi, self, rb);
        }
        nodes[size_t(self) * 3 + 0] = l;
        nodes[size_t(self) * 3 + 1] = r;
        nodes[size_t(self) * 3 + 2] = parent;
        range[size_t(self) * 2 + 0] = lo;
        range[size_t(self) * 2 + 1] = hi;
        boxOut = lb;
        boxOut.Grow(rb);
        for(int d = 0; d < 3; d++)
        {
            boxes[size_t(self) * 6 + d] = boxOut.lo[d];
            boxes[size_t(self) * 6 + 3 + d] = boxOut.hi[d];
        }
        return self;
    }
};

} // namespace

bool BuildSahBinary(const std::vector<float>& aabbs, uint32_t n, std::vector<uint32_t>& nodes,
                    std::vector<float>& boxes, std::vector<uint32_t>& range,
                    std::vector<uint32_t>& order)
{
    if(n == 0) return false;
    order.resize(n);
    for(uint32_t i = 0; i < n; i++) order[i] = i;
    const size_t internal = std::max<uint32_t>(1u, n - 1u);
    nodes.assign(internal * 3, kInvalid);
    boxes.assign(internal * 6, 0.0f);
    range.assign(internal * 2, 0u);
    if(n == 1)
    {
        nodes[0] = kLeafFlag | 0u;
        range[0] = 0;
        range[1] = 0;
        for(int k = 0; k < 6; k++) boxes[k] = aabbs[k];
        return true;
    }
    int threads = int(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    Builder b(aabbs, n, order, nodes, boxes, range, threads - 1);
    Box root;
    uint32_t rootId = b.Build(0, n - 1u, kInvalid, root);
    if(rootId != 0u)
    {
        // swap node numbers rootId <-> 0 and fix the references to both
        auto ref = [&](uint32_t v) -> uint32_t
        {
            if(v == kInvalid || (v & kLeafFlag)) return v;
            if(v == rootId) return 0u;
            if(v == 0u) return rootId;
            return v;
        };
        for(int k = 0; k < 3; k++) std::swap(nodes[size_t(rootId) * 3 + k], nodes[k]);
        for(int k = 0; k < 2; k++) std::swap(range[size_t(rootId) * 2 + k], range[k]);
        for(int k = 0; k < 6; k++) std::swap(boxes[size_t(rootId) * 6 + k], boxes[k]);
        for(size_t v = 0; v < internal; v++)
            for(int k = 0; k < 3; k++) nodes[v * 3 + k] = ref(nodes[v * 3 + k]);
    }
    return ValidateBinaryTree(nodes, range, n);
}

} // namespace mray
