// Host-side collapse of the binary LBVH into the 4-wide BNode4 layout.
// The binary tree is the reference's (AcceleratorLBVH.cu:170-435); the wide
// layout is this build's own (MI355X has no ray-tracing units, so fewer,
// wider node fetches per ray are what bounds traversal).
#include "lbvh.h"

#include <algorithm>
#include <cmath>
#include <cstring>

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

namespace
{
template<int W>
struct WideF
{
    float box[W][6];
    uint32_t child[W];
    int count;
};

// Top-down collapse of a binary tree into W-wide nodes: expand the child with
// the largest surface area until W children; subtrees of <= leafMax
// primitives become leaves over their leaf-order range.  Nodes are numbered in
// depth-first order (root 0).
template<int W>
void CollapseWide(const std::vector<uint32_t>& binNodes, const std::vector<float>& binBoxes,
                  const std::vector<uint32_t>& binRange, const std::vector<float>& leafAabbs,
                  const std::vector<uint32_t>& leafOrder, std::vector<WideF<W>>& out,
                  uint32_t& maxDepth, uint32_t& leafSlots, uint32_t leafMax = kLeafMax)
{
    const uint32_t n = uint32_t(leafOrder.size());
    maxDepth = 0;
    leafSlots = 0;
    out.clear();
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
    auto empty = []()
    {
        WideF<W> nd;
        for(int k = 0; k < W; k++)
        {
            nd.box[k][0] = nd.box[k][1] = nd.box[k][2] = INFINITY;
            nd.box[k][3] = nd.box[k][4] = nd.box[k][5] = -INFINITY;
            nd.child[k] = kInvalid;
        }
        nd.count = 0;
        return nd;
    };
    out.reserve(std::max<size_t>(1, n / (W / 2)));
    out.push_back(empty());
    if(n == 0) return;
    if(n == 1)
    {
        std::copy_n(&leafAabbs[size_t(leafOrder[0]) * 6], 6, out[0].box[0]);
        out[0].child[0] = kLeafFlag | 0u;
        out[0].count = 1;
        maxDepth = 1;
        leafSlots = 1;
        return;
    }
    struct Work { uint32_t wide, bin, depth; };
    std::vector<Work> stack;
    stack.push_back({0u, 0u, 1u});
    while(!stack.empty())
    {
        Work w = stack.back();
        stack.pop_back();
        maxDepth = std::max(maxDepth, w.depth);
        BChild kids[W];
        int count = 2;
        kids[0] = childOf(w.bin, 0);
        kids[1] = childOf(w.bin, 1);
        while(count < W)
        {
            int best = -1;
            float bestArea = -1.0f;
            for(int k = 0; k < count; k++)
            {
                uint32_t cnt = kids[k].hi - kids[k].lo + 1;
                if(kids[k].leaf || cnt <= leafMax) continue;
                float ar = Area(kids[k].box);
                if(ar > bestArea) { bestArea = ar; best = k; }
            }
            if(best < 0) break;
            BChild c = kids[best];
            kids[best] = childOf(c.idx, 0);
            kids[count++] = childOf(c.idx, 1);
        }
        // children ordered by leaf-order position (keeps memory order coherent)
        std::sort(kids, kids + count, [](const BChild& x, const BChild& y) { return x.lo < y.lo; });
        WideF<W> nd = empty();
        nd.count = count;
        for(int k = 0; k < count; k++)
        {
            const BChild& c = kids[k];
            std::copy_n(c.box, 6, nd.box[k]);
            uint32_t cnt = c.hi - c.lo + 1;
            if(cnt <= leafMax)
            {
                nd.child[k] = kLeafFlag | ((cnt - 1) << kLeafCountShift) | c.lo;
                leafSlots++;
            }
            else
            {
                uint32_t wi = uint32_t(out.size());
                out.push_back(empty());
                nd.child[k] = wi;
                stack.push_back({wi, c.idx, w.depth + 1});
            }
        }
        out[w.wide] = nd;
    }
}

// Smallest e with fmaf(255, 2^e, lo) >= hi (conservative decode of the top).
int QuantExponent(float lo, float hi)
{
    float ext = hi - lo;
    int e = -126;
    if(ext > 0.0f)
    {
        e = int(std::ceil(std::log2(double(ext) / 255.0)));
        e = std::max(e, -126);
    }
    while(e < 127 && std::fma(255.0f, std::ldexp(1.0f, e), lo) < hi) e++;
    return e;
}
} // namespace

WideBvh CollapseToBvh4(const std::vector<uint32_t>& binNodes, const std::vector<float>& binBoxes,
                       const std::vector<uint32_t>& binRange, const std::vector<float>& leafAabbs,
                       const std::vector<uint32_t>& leafOrder, uint32_t leafMax)
{
    WideBvh out;
    std::vector<WideF<4>> wide;
    CollapseWide<4>(binNodes, binBoxes, binRange, leafAabbs, leafOrder, wide, out.maxDepth,
                    out.leafSlots, leafMax);
    out.nodes.resize(wide.size());
    for(size_t i = 0; i < wide.size(); i++)
    {
        BNode4& nd = out.nodes[i];
        const WideF<4>& w = wide[i];
        for(int k = 0; k < 4; k++)
        {
            nd.lox[k] = w.box[k][0]; nd.loy[k] = w.box[k][1]; nd.loz[k] = w.box[k][2];
            nd.hix[k] = w.box[k][3]; nd.hiy[k] = w.box[k][4]; nd.hiz[k] = w.box[k][5];
            nd.child[k] = w.child[k];
            nd.pad[k] = 0;
        }
    }
    return out;
}

bool QuantizeCNode8(const float box[8][6], const uint32_t child[8], int count, CNode8& out)
{
    std::memset(&out, 0, sizeof(out));
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for(int k = 0; k < count; k++)
        for(int d = 0; d < 3; d++)
        {
            lo[d] = std::min(lo[d], box[k][d]);
            hi[d] = std::max(hi[d], box[k][3 + d]);
        }
    uint8_t* q[6] = {out.qlox, out.qloy, out.qloz, out.qhix, out.qhiy, out.qhiz};
    float org[3] = {lo[0], lo[1], lo[2]};
    int ex[3] = {0, 0, 0};
    if(count > 0)
        for(int d = 0; d < 3; d++) ex[d] = QuantExponent(lo[d], hi[d]);
    out.ox = org[0]; out.oy = org[1]; out.oz = org[2];
    out.ex = int8_t(ex[0]); out.ey = int8_t(ex[1]); out.ez = int8_t(ex[2]);
    out.nChild = uint8_t(count);
    for(int k = 0; k < 8; k++)
    {
        out.child[k] = (k < count) ? child[k] : kInvalid;
        if(k >= count)
        {
            // empty slot: kInvalid child (traversal masks on it), bytes 255 / 0
            for(int d = 0; d < 3; d++) { q[d][k] = 255; q[3 + d][k] = 0; }
            continue;
        }
        for(int d = 0; d < 3; d++)
        {
            float s = std::ldexp(1.0f, ex[d]);
            double fl = std::floor(double(box[k][d] - org[d]) / double(s));
            int ql = int(std::max(0.0, std::min(255.0, fl)));
            while(ql > 0 && std::fma(float(ql), s, org[d]) > box[k][d]) ql--;
            double fh = std::ceil(double(box[k][3 + d] - org[d]) / double(s));
            int qh = int(std::max(0.0, std::min(255.0, fh)));
            while(qh < 255 && std::fma(float(qh), s, org[d]) < box[k][3 + d]) qh++;
            if(std::fma(float(ql), s, org[d]) > box[k][d] || std::fma(float(qh), s, org[d]) < box[k][3 + d])
                return false;
            q[d][k] = uint8_t(ql);
            q[3 + d][k] = uint8_t(qh);
        }
    }
    return true;
}

bool CollapseToCBvh8(const std::vector<uint32_t>& binNodes, const std::vector<float>& binBoxes,
                     const std::vector<uint32_t>& binRange, const std::vector<float>& leafAabbs,
                     const std::vector<uint32_t>& leafOrder, CBvh8& out)
{
    std::vector<WideF<8>> wide;
    CollapseWide<8>(binNodes, binBoxes, binRange, leafAabbs, leafOrder, wide, out.maxDepth,
                    out.leafSlots);
    out.nodes.resize(wide.size());
    for(size_t i = 0; i < wide.size(); i++)
        if(!QuantizeCNode8(wide[i].box, wide[i].child, wide[i].count, out.nodes[i])) return false;
    return true;
}

} // namespace mray
