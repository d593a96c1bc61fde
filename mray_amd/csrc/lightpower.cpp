// Light selection by emitted power: an alias table over a scene's emissive
// triangles (Walker 1977; construction after Vose 1991, "A linear algorithm
// for generating random numbers with a given distribution").
//
// The reference selects NEE lights uniformly (DirectLightSamplerUniform,
// LightSampler.hpp:30-119); its StratifiedIntegerAliasTable
// (Core/DataStructures.h:10-72) is an integer range lookup with no caller.
// This sampler is the build's own and is opt-in (MRAY_RENDER_LIGHT_POWER).
// It keeps the uniform sampler's handling of the boundary light: the last
// of the n + 1 equal strata of the selection number picks the boundary
// light with probability 1 / (n + 1), exactly as the uniform sampler does.
// The other n strata are the table's buckets; a triangle light is chosen
// with probability n / (n + 1) * w_i / sum(w), where
// w_i = Y(radiance) * area * (2 if two-sided else 1) and Y is the Rec.709
// luminance; if every weight is zero, the weights become equal.
#include "scene.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace mray
{

void LightPowerTable(const float* pos9, const float4* rad, uint32_t n, float* prob, uint32_t* alias,
                     float* pmf)
{
    if(n == 0) return;
    const double yR = 0.2126390058715103, yG = 0.7151686787677559, yB = 0.0721923153607337;
    std::vector<double> w(n);
    double W = 0.0;
    for(uint32_t i = 0; i < n; i++)
    {
        const float* p = pos9 + size_t(i) * 9;
        const double e1[3] = {double(p[3]) - p[0], double(p[4]) - p[1], double(p[5]) - p[2]};
        const double e2[3] = {double(p[6]) - p[0], double(p[7]) - p[1], double(p[8]) - p[2]};
        const double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                             e1[0] * e2[1] - e1[1] * e2[0]};
        const double area = 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        const double lum = std::fabs(yR * rad[i].x + yG * rad[i].y + yB * rad[i].z);
        double wi = lum * area * (rad[i].w != 0.0f ? 2.0 : 1.0);
        if(!std::isfinite(wi)) wi = 0.0;
        w[i] = wi;
        W += wi;
    }
    if(!(W > 0.0) || !std::isfinite(W))
    {
        for(uint32_t i = 0; i < n; i++) w[i] = 1.0;
        W = double(n);
    }
    // bucket masses scaled to mean 1, split into under- and over-full lists
    std::vector<double> q(n);
    std::vector<uint32_t> small, large;
    for(uint32_t i = 0; i < n; i++)
    {
        q[i] = w[i] * double(n) / W;
        (q[i] < 1.0 ? small : large).push_back(i);
    }
    while(!small.empty() && !large.empty())
    {
        const uint32_t s = small.back();
        small.pop_back();
        const uint32_t l = large.back();
        prob[s] = float(q[s]);
        alias[s] = l;
        q[l] = (q[l] + q[s]) - 1.0;
        if(q[l] < 1.0)
        {
            large.pop_back();
            small.push_back(l);
        }
    }
    for(uint32_t i : large) { prob[i] = 1.0f; alias[i] = i; }
    for(uint32_t i : small) { prob[i] = 1.0f; alias[i] = i; }   // rounding leftovers
    const double share = double(n) / double(n + 1);
    for(uint32_t i = 0; i < n; i++) pmf[i] = float(share * w[i] / W);
}

int BuildLightPower(DeviceArena& ar, const SceneView& v, const std::vector<uint32_t>& rangeOff,
                    const std::vector<uint32_t>& rangeCount, uint32_t total, uint2** outAlias,
                    float** outPmf)
{
    *outAlias = nullptr;
    *outPmf = nullptr;
    std::vector<float> pos(size_t(total) * 9);
    std::vector<float4> rad(total);
    std::vector<float> prob(total, 1.0f), pmf(total, 0.0f);
    std::vector<uint32_t> alias(total, 0u);
    if(total)
    {
        MRAY_HIP_CHECK(hipMemcpy(pos.data(), v.lightPos, pos.size() * 4, hipMemcpyDeviceToHost));
        MRAY_HIP_CHECK(hipMemcpy(rad.data(), v.lightRad, rad.size() * sizeof(float4), hipMemcpyDeviceToHost));
    }
    for(size_t r = 0; r < rangeOff.size(); r++)
    {
        const uint32_t o = rangeOff[r], c = rangeCount[r];
        if(o + c > total)
        {
            SetError("light power table: light range outside the light table");
            return MRAY_ERR_INVALID;
        }
        LightPowerTable(&pos[size_t(o) * 9], &rad[o], c, &prob[o], &alias[o], &pmf[o]);
    }
    std::vector<uint2> packed(std::max<uint32_t>(total, 1u));
    for(uint32_t i = 0; i < total; i++)
    {
        uint32_t bits;
        std::memcpy(&bits, &prob[i], 4);
        packed[i] = make_uint2(bits, alias[i]);
    }
    uint2* dAlias = nullptr;
    float* dPmf = nullptr;
    MRAY_TRY(DevAlloc(ar, (void**)&dAlias, packed.size() * sizeof(uint2)));
    MRAY_TRY(DevAlloc(ar, (void**)&dPmf, std::max<size_t>(total, 1) * sizeof(float)));
    if(total)
    {
        MRAY_HIP_CHECK(hipMemcpy(dAlias, packed.data(), total * sizeof(uint2), hipMemcpyHostToDevice));
        MRAY_HIP_CHECK(hipMemcpy(dPmf, pmf.data(), total * sizeof(float), hipMemcpyHostToDevice));
    }
    *outAlias = dAlias;
    *outPmf = dPmf;
    return MRAY_OK;
}

} // namespace mray
