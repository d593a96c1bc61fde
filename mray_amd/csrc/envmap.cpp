// Textured Skysphere: the radiance texture and its luminance distribution
// (LightGroupSkysphere::Finalize, LightsDefault.hpp:860-907, and
// DistributionGroupPwC2D::Construct, Distributions.cu:465-538).
//
// The distribution is built on the host once per scene: per texel the
// Rec.709 luminance Y of the texel's RGB (ColorConverter.cu:450-470 reads
// texel centres with DIRECT mode), a row-wise inclusive scan of |Y| in
// double rounded to float per element (KCSegmentedScanPrecise, the
// reference's CPU kernel at Distributions.cu:262-281), the marginal as a
// float running sum of the row totals (KCCopyScanY, :241-255), then each
// row and the marginal multiplied by the float reciprocal of its last
// element (KCNormalizeXY, :73-119).  One deviation: a row whose luminance
// is all zero keeps a uniform CDF (the reference divides by its zero total,
// which makes the row's pdf NaN for rays that escape through it); the
// marginal gives such a row zero probability either way.
#include "scene.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace mray
{

// Y row of the Rec.709 (D65) RGB -> XYZ matrix, from the primaries
// R (0.64, 0.33), G (0.30, 0.60), B (0.15, 0.06) and white (0.3127, 0.3290).
static void LuminanceRow(float out[3])
{
    const double xr = 0.64, yr = 0.33, xg = 0.30, yg = 0.60, xb = 0.15, yb = 0.06;
    const double xw = 0.3127, yw = 0.3290;
    // columns: XYZ of each primary at Y = 1
    const double P[3][3] = {{xr / yr, xg / yg, xb / yb},
                            {1.0, 1.0, 1.0},
                            {(1 - xr - yr) / yr, (1 - xg - yg) / yg, (1 - xb - yb) / yb}};
    const double W[3] = {xw / yw, 1.0, (1 - xw - yw) / yw};
    // solve P s = W (Cramer)
    auto det3 = [](const double m[3][3])
    {
        return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
               m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    };
    const double d = det3(P);
    for(int c = 0; c < 3; c++)
    {
        double m[3][3];
        std::memcpy(m, P, sizeof(m));
        for(int r = 0; r < 3; r++) m[r][c] = W[r];
        out[c] = float(det3(m) / d);   // Y row entry = 1 * s_c
    }
}

int BuildEnvironment(DeviceArena& ar, SceneView& v, uint32_t W, uint32_t H, const float* hRgb)
{
    if(v.boundaryType != MRAY_BOUNDARY_SKYSPHERE_COOCTA &&
       v.boundaryType != MRAY_BOUNDARY_SKYSPHERE_SPHERICAL)
    {
        SetError("set_environment: the scene's boundary is not a Skysphere");
        return MRAY_ERR_INVALID;
    }
    if(!hRgb || W == 0 || H == 0 || W > 8192 || H > 8192)
    {
        SetError("set_environment: NULL texels or size outside 1..8192");
        return MRAY_ERR_INVALID;
    }
    const size_t P = size_t(W) * H;
    float yRow[3];
    LuminanceRow(yRow);
    std::vector<float4> tex(P);
    std::vector<float> cdfX(P), cdfY(H);
    for(size_t i = 0; i < P; i++)
    {
        const float r = hRgb[3 * i], g = hRgb[3 * i + 1], b = hRgb[3 * i + 2];
        if(!std::isfinite(r) || !std::isfinite(g) || !std::isfinite(b))
        {
            SetError("set_environment: non-finite texel " + std::to_string(i));
            return MRAY_ERR_INVALID;
        }
        tex[i] = make_float4(r, g, b, 0.0f);
    }
    float ySum = 0.0f;
    for(uint32_t y = 0; y < H; y++)
    {
        float* row = &cdfX[size_t(y) * W];
        double sum = 0.0;
        for(uint32_t x = 0; x < W; x++)
        {
            const float4 t = tex[size_t(y) * W + x];
            const float lum = yRow[0] * t.x + yRow[1] * t.y + yRow[2] * t.z;
            sum += double(std::fabs(lum));
            row[x] = float(sum);
        }
        ySum += row[W - 1];
        cdfY[y] = ySum;
    }
    if(!(ySum > 0.0f))
    {
        SetError("set_environment: the texture has zero luminance everywhere");
        return MRAY_ERR_INVALID;
    }
    for(uint32_t y = 0; y < H; y++)
    {
        float* row = &cdfX[size_t(y) * W];
        if(row[W - 1] > 0.0f)
        {
            const float recip = 1.0f / row[W - 1];
            for(uint32_t x = 0; x < W; x++) row[x] *= recip;
        }
        else
        {
            for(uint32_t x = 0; x < W; x++) row[x] = float(x + 1) / float(W);
        }
    }
    {
        const float recip = 1.0f / cdfY[H - 1];
        for(uint32_t y = 0; y < H; y++) cdfY[y] *= recip;
    }
    float4* dTex = nullptr;
    float *dX = nullptr, *dY = nullptr;
    MRAY_TRY(DevAlloc(ar, (void**)&dTex, P * sizeof(float4)));
    MRAY_TRY(DevAlloc(ar, (void**)&dX, P * sizeof(float)));
    MRAY_TRY(DevAlloc(ar, (void**)&dY, size_t(H) * sizeof(float)));
    MRAY_HIP_CHECK(hipMemcpy(dTex, tex.data(), P * sizeof(float4), hipMemcpyHostToDevice));
    MRAY_HIP_CHECK(hipMemcpy(dX, cdfX.data(), P * sizeof(float), hipMemcpyHostToDevice));
    MRAY_HIP_CHECK(hipMemcpy(dY, cdfY.data(), size_t(H) * sizeof(float), hipMemcpyHostToDevice));
    v.envTex = dTex;
    v.envCdfX = dX;
    v.envCdfY = dY;
    v.envW = W;
    v.envH = H;
    return MRAY_OK;
}

int CopyEnvironment(DeviceArena& ar, const SceneView& from, SceneView& to)
{
    to.envTex = nullptr;
    to.envCdfX = to.envCdfY = nullptr;
    to.envW = to.envH = 0;
    if(from.envW == 0) return MRAY_OK;
    const size_t P = size_t(from.envW) * from.envH;
    float4* dTex = nullptr;
    float *dX = nullptr, *dY = nullptr;
    MRAY_TRY(DevAlloc(ar, (void**)&dTex, P * sizeof(float4)));
    MRAY_TRY(DevAlloc(ar, (void**)&dX, P * sizeof(float)));
    MRAY_TRY(DevAlloc(ar, (void**)&dY, size_t(from.envH) * sizeof(float)));
    MRAY_HIP_CHECK(hipMemcpy(dTex, from.envTex, P * sizeof(float4), hipMemcpyDeviceToDevice));
    MRAY_HIP_CHECK(hipMemcpy(dX, from.envCdfX, P * sizeof(float), hipMemcpyDeviceToDevice));
    MRAY_HIP_CHECK(hipMemcpy(dY, from.envCdfY, size_t(from.envH) * sizeof(float),
                             hipMemcpyDeviceToDevice));
    to.envTex = dTex;
    to.envCdfX = dX;
    to.envCdfY = dY;
    to.envW = from.envW;
    to.envH = from.envH;
    return MRAY_OK;
}

} // namespace mray

using namespace mray;

extern "C" {

int mray_accel_set_environment(MRayAccel* a, uint32_t width, uint32_t height, const float* hRgb)
{
    if(!a) { SetError("accel_set_environment: NULL accelerator"); return MRAY_ERR_INVALID; }
    return BuildEnvironment(a->arena, a->view, width, height, hRgb);
}

int mray_accel_environment_cdfs(const MRayAccel* a, float* hCdfX, float* hCdfY)
{
    if(!a || !hCdfX || !hCdfY) { SetError("accel_environment_cdfs: NULL argument"); return MRAY_ERR_INVALID; }
    const SceneView& v = a->view;
    if(v.envW == 0) { SetError("accel_environment_cdfs: no environment texture"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipMemcpy(hCdfX, v.envCdfX, size_t(v.envW) * v.envH * sizeof(float),
                             hipMemcpyDeviceToHost));
    MRAY_HIP_CHECK(hipMemcpy(hCdfY, v.envCdfY, size_t(v.envH) * sizeof(float), hipMemcpyDeviceToHost));
    return MRAY_OK;
}

int mray_instanced_scene_set_environment(MRayInstScene* s, uint32_t width, uint32_t height,
                                         const float* hRgb)
{
    if(!s) { SetError("instanced_scene_set_environment: NULL scene"); return MRAY_ERR_INVALID; }
    return BuildEnvironment(s->arena, s->view, width, height, hRgb);
}

} // extern "C"
