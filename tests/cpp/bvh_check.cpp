// Host checks of the traversal trees built by libmray_amd's host code
// (sah.cpp, bvh4.cpp), compiled into this test from the same sources:
//   * the binned-SAH binary tree passes ValidateBinaryTree;
//   * BNode4 and CNode8 trees reference every triangle exactly once through
//     their leaf ranges;
//   * every CNode8 child box, decoded as fmaf(q, 2^e, origin) exactly as the
//     kernels do, contains the exact bounds of everything below it;
//   * a host walk of the CNode8 tree with the decoded boxes finds the same
//     closest hits as brute force for random rays.
// Input: u32 n, then float[n][9] triangle vertices.  Prints "ok" on success.
#include "../../mray_amd/csrc/lbvh.h"

#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

using namespace mray;

namespace
{
int fails = 0;
void Check(bool c, const char* what)
{
    if(!c && fails++ < 20) std::printf("FAIL: %s\n", what);
}

struct Tri { float v[9]; };

float Pow2(int e) { return std::ldexp(1.0f, e); }

void DecodeBox(const CNode8& n, int k, float lo[3], float hi[3])
{
    const uint8_t* ql[3] = {n.qlox, n.qloy, n.qloz};
    const uint8_t* qh[3] = {n.qhix, n.qhiy, n.qhiz};
    const float o[3] = {n.ox, n.oy, n.oz};
    const int e[3] = {n.ex, n.ey, n.ez};
    for(int d = 0; d < 3; d++)
    {
        lo[d] = std::fma(float(ql[d][k]), Pow2(e[d]), o[d]);
        hi[d] = std::fma(float(qh[d][k]), Pow2(e[d]), o[d]);
    }
}

// exact bounds of a child subtree; marks its triangles (by leaf position)
void SubtreeBounds(const std::vector<CNode8>& nodes, uint32_t c, const std::vector<Tri>& soup,
                   std::vector<int>& seen, float lo[3], float hi[3])
{
    for(int d = 0; d < 3; d++) { lo[d] = INFINITY; hi[d] = -INFINITY; }
    if(c & kLeafFlag)
    {
        uint32_t first = c & kLeafStartMask, cnt = ((c & ~kLeafFlag) >> kLeafCountShift) + 1;
        for(uint32_t i = first; i < first + cnt; i++)
        {
            seen[i]++;
            for(int v = 0; v < 3; v++)
                for(int d = 0; d < 3; d++)
                {
                    lo[d] = std::min(lo[d], soup[i].v[3 * v + d]);
                    hi[d] = std::max(hi[d], soup[i].v[3 * v + d]);
                }
        }
        return;
    }
    const CNode8& n = nodes[c];
    for(int k = 0; k < n.nChild; k++)
    {
        float cl[3], ch[3], dl[3], dh[3];
        SubtreeBounds(nodes, n.child[k], soup, seen, cl, ch);
        DecodeBox(n, k, dl, dh);
        for(int d = 0; d < 3; d++)
        {
            Check(dl[d] <= cl[d] && dh[d] >= ch[d], "decoded CNode8 box contains its subtree");
            lo[d] = std::min(lo[d], cl[d]);
            hi[d] = std::max(hi[d], ch[d]);
        }
    }
    for(int k = n.nChild; k < 8; k++)
        Check(n.child[k] == kInvalid, "empty CNode8 slot holds kInvalid");
}

bool HitTri(const float* o, const float* dir, const Tri& t, float& tOut)
{
    // plain Moller-Trumbore (only used to compare the two walks on equal terms)
    float e0[3], e1[3], p[3], tv[3], q[3];
    for(int d = 0; d < 3; d++) { e0[d] = t.v[3 + d] - t.v[d]; e1[d] = t.v[6 + d] - t.v[d]; tv[d] = o[d] - t.v[d]; }
    p[0] = dir[1] * e1[2] - dir[2] * e1[1]; p[1] = dir[2] * e1[0] - dir[0] * e1[2]; p[2] = dir[0] * e1[1] - dir[1] * e1[0];
    float det = e0[0] * p[0] + e0[1] * p[1] + e0[2] * p[2];
    if(std::fabs(det) < 1e-7f) return false;
    float inv = 1.0f / det;
    float u = (tv[0] * p[0] + tv[1] * p[1] + tv[2] * p[2]) * inv;
    if(u < 0 || u > 1) return false;
    q[0] = tv[1] * e0[2] - tv[2] * e0[1]; q[1] = tv[2] * e0[0] - tv[0] * e0[2]; q[2] = tv[0] * e0[1] - tv[1] * e0[0];
    float v = (dir[0] * q[0] + dir[1] * q[1] + dir[2] * q[2]) * inv;
    if(v < 0 || u + v > 1) return false;
    float tt = (e1[0] * q[0] + e1[1] * q[1] + e1[2] * q[2]) * inv;
    if(tt <= 1e-7f) return false;
    tOut = tt;
    return true;
}

bool BoxHit(const float* o, const float* inv, const float lo[3], const float hi[3], float tMax)
{
    float t0 = 0.0f, t1 = tMax;
    for(int d = 0; d < 3; d++)
    {
        bool neg = std::signbit(inv[d]);
        float n = ((neg ? hi[d] : lo[d]) - o[d]) * inv[d];
        float f = ((neg ? lo[d] : hi[d]) - o[d]) * inv[d];
        t0 = std::fmax(t0, n);
        t1 = std::fmin(t1, f);
    }
    return t1 * (1.0f + 4.8e-7f) >= t0;
}
} // namespace

int main(int argc, char** argv)
{
    if(argc < 2) { std::printf("usage: bvh_check soup.bin\n"); return 2; }
    FILE* f = std::fopen(argv[1], "rb");
    if(!f) return 2;
    uint32_t n = 0;
    if(std::fread(&n, 4, 1, f) != 1) return 2;
    std::vector<Tri> tris(n);
    if(std::fread(tris.data(), sizeof(Tri), n, f) != n) return 2;
    std::fclose(f);
    std::vector<float> aabbs(size_t(n) * 6);
    for(uint32_t i = 0; i < n; i++)
        for(int d = 0; d < 3; d++)
        {
            float lo = INFINITY, hi = -INFINITY;
            for(int v = 0; v < 3; v++) { lo = std::min(lo, tris[i].v[3 * v + d]); hi = std::max(hi, tris[i].v[3 * v + d]); }
            aabbs[size_t(i) * 6 + d] = lo;
            aabbs[size_t(i) * 6 + 3 + d] = hi;
        }
    std::vector<uint32_t> nodes, range, order;
    std::vector<float> boxes;
    Check(BuildSahBinary(aabbs, n, nodes, boxes, range, order), "SAH tree valid");
    std::vector<Tri> soup(n);
    for(uint32_t p = 0; p < n; p++) soup[p] = tris[order[p]];
    // BNode4: every leaf position exactly once
    WideBvh w4 = CollapseToBvh4(nodes, boxes, range, aabbs, order);
    std::vector<int> seen4(n, 0);
    for(const BNode4& nd : w4.nodes)
        for(int k = 0; k < 4; k++)
        {
            uint32_t c = nd.child[k];
            if(c == kInvalid || !(c & kLeafFlag)) continue;
            uint32_t first = c & kLeafStartMask, cnt = ((c & ~kLeafFlag) >> kLeafCountShift) + 1;
            for(uint32_t i = first; i < first + cnt && i < n; i++) seen4[i]++;
        }
    for(uint32_t i = 0; i < n; i++) Check(seen4[i] == 1, "BNode4 leaves cover each triangle once");
    // CNode8: containment + coverage
    CBvh8 w8;
    Check(CollapseToCBvh8(nodes, boxes, range, aabbs, order, w8), "CNode8 quantisation");
    std::vector<int> seen8(n, 0);
    float lo[3], hi[3];
    if(n > 0) SubtreeBounds(w8.nodes, 0u, soup, seen8, lo, hi);
    for(uint32_t i = 0; i < n; i++) Check(seen8[i] == 1, "CNode8 leaves cover each triangle once");
    // host walk of the CNode8 tree vs brute force
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> u(-1.5f, 1.5f), g(-1.0f, 1.0f);
    int hits = 0;
    for(int r = 0; r < 4000; r++)
    {
        float o[3] = {u(rng), u(rng), u(rng)}, d[3] = {g(rng), g(rng), g(rng)};
        if(r % 7 == 0) { d[0] = 0.0f; d[1] = (r % 2) ? -0.0f : 0.0f; }
        float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if(len < 1e-3f) continue;
        for(float& x : d) x /= len;
        float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
        float bt = INFINITY;
        uint32_t bp = kInvalid;
        for(uint32_t i = 0; i < n; i++)
        {
            float t;
            if(HitTri(o, d, soup[i], t) && (t < bt || (t == bt && i < bp))) { bt = t; bp = i; }
        }
        float wt = INFINITY;
        uint32_t wp = kInvalid;
        std::vector<uint32_t> st{0u};
        while(!st.empty())
        {
            uint32_t c = st.back();
            st.pop_back();
            const CNode8& nd = w8.nodes[c];
            for(int k = 0; k < nd.nChild; k++)
            {
                float dl[3], dh[3];
                DecodeBox(nd, k, dl, dh);
                if(!BoxHit(o, inv, dl, dh, wt)) continue;
                uint32_t ch = nd.child[k];
                if(ch & kLeafFlag)
                {
                    uint32_t first = ch & kLeafStartMask, cnt = ((ch & ~kLeafFlag) >> kLeafCountShift) + 1;
                    for(uint32_t i = first; i < first + cnt; i++)
                    {
                        float t;
                        if(HitTri(o, d, soup[i], t) && (t < wt || (t == wt && i < wp))) { wt = t; wp = i; }
                    }
                }
                else st.push_back(ch);
            }
        }
        Check(wp == bp && (bp == kInvalid || wt == bt), "CNode8 walk == brute force");
        hits += (bp != kInvalid);
    }
    std::printf("n=%u wide4=%zu wide8=%zu depth8=%u hits=%d fails=%d\n", n, w4.nodes.size(),
                w8.nodes.size(), w8.maxDepth, hits, fails);
    if(fails == 0) std::printf("ok\n");
    return fails == 0 ? 0 : 1;
}
