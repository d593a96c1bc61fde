// Host-side collapse of the binary LBVH into the 4-wide BNode4 layout.
// The binary tree is the reference's (AcceleratorLBVH.cu:170-435); the wide
// layout is this build's own (MI355X has no ray-tracing units, so fewer,
// wider node fetches per ray are what bounds traversal).
#include "lbvh.h"

#include <algorithm>
#include <cmath>

namespace mray
{

namespace
{
struct BChild
{
    bool leaf;
    uint32_t idx;     // binary node (internal) or triangle (leaf)
    uint32_t lo, hi;  // leaf-order range
    float box[6];
};

float Area(const float* b)
{
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    if(dx < 0 || dy < 0 || dz < 0) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}
}

bool ValidateBinaryTree(const std::vector<uint32_t>& nodes, const std::vector<uint32_t>& range,
                        uint32_t n)
{
    if(n == 0) return false;
    if(n == 1) return nodes.size() >= 3 && nodes[0] == (kLeafFlag | 0u);
    const uint32_t nodeCount = n - 1;
    if(range[0] != 0 || range[1] != n - 1) return false;
    std::vector<uint8_t> seenNode(nodeCount, 0), seenLeaf(n, 0);
    std::vector<uint32_t> stack{0u};
    uint32_t visitedNodes = 0, visitedLeaves = 0;
    while(!stack.empty())
    {
        uint32_t v = stack.back();
        stack.pop_back();
        if(v >= nodeCount || seenNode[v]) return false;
        seenNode[v] = 1;
        visitedNodes++;
        uint32_t lo = range[2 * v], hi = range[2 * v + 1];
        if(lo >= hi || hi >= n) return false;
        uint32_t l = nodes[3 * size_t(v)], r = nodes[3 * size_t(v) + 1];
        uint32_t mid;
        if(l & kLeafFlag)
        {
            uint32_t t = l & ~kLeafFlag;
            if(t >= n || seenLeaf[t]) return false;
            seenLeaf[t] = 1;
            visitedLeaves++;
            mid = lo;
        }
        else
        {
            if(l >= nodeCount || range[2 * l] != lo) return false;
            mid = range[2 * l + 1];
            stack.push_back(l);
        }
        if(r & kLeafFlag)
        {
            uint32_t t = r & ~kLeafFlag;
            if(t >= n || seenLeaf[t] || hi != mid + 1) return false;
            seenLeaf[t] = 1;
            visitedLeaves++;
        }
        else
        {
            if(r >= nodeCount || range[2 * r] != mid + 1 || range[2 * r + 1] != hi) return false;
            stack.push_back(r);
        }
    }
    return visitedNodes == nodeCount && visitedLeaves == n;
}

WideBvh CollapseToBvh4(const std::vector<uint32_t>& binNodes, const std::vector<float>& binBoxes,
                       const std::vector<uint32_t>& binRange, const std::vector<float>& leafAabbs,
                       const std::vector<uint32_t>& leafOrder)
{
    WideBvh out;
    const uint32_t n = uint32_t(leafOrder.size());
    auto childOf = [&](uint32_t parent, int side) -> BChild
    {
        BChild c;
        uint32_t v = binNodes[size_t(parent) * 3 + side];
        if(v & kLeafFlag)
        {
            c.leaf = true;
            c.idx = v & ~kLeafFlag;
            uint32_t pos = (side == 0) ? binRange[2 * parent] : binRange[2 * parent + 1];
            c.lo = c.hi = pos;
            std::copy_n(&leafAabbs[size_t(c.idx) * 6], 6, c.box);
        }
        else
        {
            c.leaf = false;
            c.idx = v;
            c.lo = binRange[2 * v];
            c.hi = binRange[2 * v + 1];
            std::copy_n(&binBoxes[size_t(v) * 6], 6, c.box);
        }
        return c;
    };
    auto setEmpty = [](BNode4& nd, int k)
    {
        nd.lox[k] = nd.loy[k] = nd.loz[k] = INFINITY;
        nd.hix[k] = nd.hiy[k] = nd.hiz[k] = -INFINITY;
        nd.child[k] = kInvalid;
    };
    auto setBox = [](BNode4& nd, int k, const float* b)
    {
        nd.lox[k] = b[0]; nd.loy[k] = b[1]; nd.loz[k] = b[2];
        nd.hix[k] = b[3]; nd.hiy[k] = b[4]; nd.hiz[k] = b[5];
    };

    out.nodes.reserve(std::max<size_t>(1, n / 2));
    out.nodes.emplace_back();
    BNode4& root = out.nodes[0];
    for(int k = 0; k < 4; k++) setEmpty(root, k);
    root.pad[0] = root.pad[1] = root.pad[2] = root.pad[3] = 0;
    if(n == 0) return out;
    if(n == 1)
    {
        setBox(root, 0, &leafAabbs[size_t(leafOrder[0]) * 6]);
        root.child[0] = kLeafFlag | 0u;
        out.maxDepth = 1;
        out.leafSlots = 1;
        return out;
    }
    struct Work { uint32_t wide, bin, depth; };
    std::vector<Work> stack;
    stack.push_back({0u, 0u, 1u});
    while(!stack.empty())
    {
        Work w = stack.back();
        stack.pop_back();
        out.maxDepth = std::max(out.maxDepth, w.depth);
        BChild kids[4];
        int count = 2;
        kids[0] = childOf(w.bin, 0);
        kids[1] = childOf(w.bin, 1);
        while(count < 4)
        {
            int best = -1;
            float bestArea = -1.0f;
            for(int k = 0; k < count; k++)
            {
                uint32_t cnt = kids[k].hi - kids[k].lo + 1;
                if(kids[k].leaf || cnt <= kLeafMax) continue;
                float a = Area(kids[k].box);
                if(a > bestArea) { bestArea = a; best = k; }
            }
            if(best < 0) break;
            BChild c = kids[best];
            kids[best] = childOf(c.idx, 0);
            kids[count++] = childOf(c.idx, 1);
        }
        // children ordered by leaf-order position (keeps memory order coherent)
        std::sort(kids, kids + count, [](const BChild& a, const BChild& b) { return a.lo < b.lo; });
        BNode4 nd;
        for(int k = 0; k < 4; k++) setEmpty(nd, k);
        nd.pad[0] = nd.pad[1] = nd.pad[2] = nd.pad[3] = 0;
        for(int k = 0; k < count; k++)
        {
            const BChild& c = kids[k];
            setBox(nd, k, c.box);
            uint32_t cnt = c.hi - c.lo + 1;
            if(cnt <= kLeafMax)
            {
                nd.child[k] = kLeafFlag | ((cnt - 1) << kLeafCountShift) | c.lo;
                out.leafSlots++;
            }
            else
            {
                uint32_t wi = uint32_t(out.nodes.size());
                out.nodes.emplace_back();
                nd.child[k] = wi;
                stack.push_back({wi, c.idx, w.depth + 1});
            }
        }
        out.nodes[w.wide] = nd;
    }
    return out;
}

} // namespace mray
