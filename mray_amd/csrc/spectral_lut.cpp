// RGB -> spectrum coefficient tables (the capability of the reference's
// SpectraLUTGen, Source/SpectraLUTGen/main.cpp, loaded at
// Tracer/SpectrumContext.cu:298-472), built by the published method:
//
//   Jakob & Hanika 2019, "A Low-Dimensional Function Space for Efficient
//   Spectral Upsampling": a reflectance is S(c0 l^2 + c1 l + c2) with
//   S(x) = 1/2 + x / (2 sqrt(1 + x^2)); per target RGB the three coefficients
//   are found by Gauss-Newton on the CIELAB difference of the reflectance's
//   colour under the working illuminant.  Three tables (largest channel) of
//   N^3 cells; x, y = the other two channels over the largest, z = largest
//   value on a twice-smoothstepped axis; each (x, y) column is solved along z
//   from N/5 outwards with warm starts.
//
// The fit follows the reference generator's choices step for step, so the
// table matches the one the reference renders with up to rounding
// (SpectraLUTGen/main.cpp:80-436):
//   quadrature   Simpson's 3/8 rule over the 471 1-nm samples (:99-110)
//   integrand    XYZ->RGB(observer) * D65 / 10566.864 * weight (:121-131)
//   Lab white    the illuminant's XYZ under the same quadrature (:130-135)
//   variable     normalised wavelength i / 471 (:160-168)
//   solver       15 passes of Gauss-Newton, central differences with step
//                1e-4, LU solve, coefficients scaled down when the largest
//                exceeds 200, early exit when |residual|^2 < 1e-7 (:186-260)
//   storage      coefficients in nanometres through l' = (l - 360) / 470
//                (:350-356: the stored polynomial is the fitted one
//                stretched by 471/470 in wavelength, as the reference's is)
//
// Colour data: the CIE 1931 2-degree observer and the D65 SPD
// (colour_data.cpp), Rec.709 primaries with the D65 white point.
// Stored coefficients evaluate in nanometres, as Converter::ConvertAlbedo does.
#include "spectral.h"
#include "colour_data.h"

#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <dlfcn.h>
#include <sys/stat.h>

namespace mray
{

namespace
{
constexpr int kLutN = 64;
constexpr int kFitPasses = 15;

using M3 = std::array<double, 9>;
M3 Inverse(const M3& m)
{
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
    double A = e * i - f * h, B = -(d * i - f * g), Cc = d * h - e * g;
    double det = a * A + b * B + c * Cc;
    M3 r = {A, -(b * i - c * h), b * f - c * e, B, a * i - c * g, -(a * f - c * d), Cc,
            -(a * h - b * g), a * e - b * d};
    for(double& v : r) v /= det;
    return r;
}
std::array<double, 3> Mul(const M3& m, const double v[3])
{
    return {m[0] * v[0] + m[1] * v[1] + m[2] * v[2], m[3] * v[0] + m[4] * v[1] + m[5] * v[2],
            m[6] * v[0] + m[7] * v[1] + m[8] * v[2]};
}

// RGB -> XYZ of a colour space from its primaries and white point
// (Color::GenRGBToXYZ, Core/ColorFunctions.h:347-361, after Lindbloom): the
// primaries' XYZ at Y = 1 as columns, scaled so RGB (1, 1, 1) is the white.
// The reference evaluates it in float; the entries are rounded to float here.
M3 RgbToXyz(const double p[8])
{
    auto xyz = [](double x, double y) { return std::array<double, 3>{x / y, 1.0, (1.0 - x - y) / y}; };
    auto r = xyz(p[0], p[1]), g = xyz(p[2], p[3]), b = xyz(p[4], p[5]), w = xyz(p[6], p[7]);
    M3 P = {r[0], g[0], b[0], r[1], g[1], b[1], r[2], g[2], b[2]};
    M3 Pi = Inverse(P);
    auto S = Mul(Pi, w.data());
    M3 out;
    for(int row = 0; row < 3; row++)
        for(int k = 0; k < 3; k++) out[3 * row + k] = double(float(P[3 * row + k] * S[k]));
    return out;
}

struct Colour
{
    std::vector<double> toRgb;   // [471][3] fit integrand: XYZ->RGB(observer) * D65 * weight
    double white[3];             // Lab white point: the illuminant's XYZ, same quadrature
    M3 rgbToXyz, xyzToRgb;       // Rec.709 with the D65 white
    std::vector<float> cmf;      // [471][3] device observer: each channel over its integral
    std::vector<float> illum;    // [471] device illuminant: D65 * integral(ybar) / 10566.864
};

// Simpson's 3/8 weight of sample i of n (n - 1 divisible by 3): 1 3 3 2 3 3 2 ... 3 3 1
double Simpson38(int i, int n)
{
    const double W = 3.0 / 8.0;
    if(i == 0 || i == n - 1) return W;
    return ((i - 1) % 3 == 2) ? 2.0 * W : 3.0 * W;
}

Colour MakeColour()
{
    Colour c;
    c.rgbToXyz = RgbToXyz(kRec709Primaries);
    M3 inv = Inverse(c.rgbToXyz);
    for(double& v : inv) v = double(float(v));
    c.xyzToRgb = inv;
    c.toRgb.resize(size_t(kCieRows) * 3);
    c.white[0] = c.white[1] = c.white[2] = 0.0;
    c.cmf.resize(size_t(kCieRows) * 3);
    c.illum.resize(kCieRows);
    // device tables, in float as the reference computes them at load time
    // (SpectrumContext.cu:409-426)
    const float wX = 1.0f / kCieIntegralX, wY = 1.0f / kCieIntegralY, wZ = 1.0f / kCieIntegralZ;
    float illumFactor = kCieIntegralY;
    illumFactor /= kD65NormFactor;
    for(int i = 0; i < kCieRows; i++)
    {
        const float* row = kCieObserverD65[i];
        c.cmf[3 * i] = row[0] * wX;
        c.cmf[3 * i + 1] = row[1] * wY;
        c.cmf[3 * i + 2] = row[2] * wZ;
        c.illum[i] = row[3] * illumFactor;
        // fit integrand
        const double xyz[3] = {double(row[0]), double(row[1]), double(row[2])};
        const double I = double(row[3]) / double(kD65NormFactor);
        const double w = Simpson38(i, kCieRows);
        auto rgb = Mul(c.xyzToRgb, xyz);
        for(int k = 0; k < 3; k++)
        {
            c.toRgb[3 * i + k] = rgb[k] * I * w;
            c.white[k] += xyz[k] * I * w;
        }
    }
    return c;
}

// XYZ -> CIELAB against a white point (Color::XYZToCIELab, ColorFunctions.h:259-284)
void Lab(const double white[3], const double xyz[3], double out[3])
{
    auto f = [](double t)
    {
        const double d = 6.0 / 29.0;
        if(t > d * d * d) return std::cbrt(t);
        const double k = 1.0 / (d * d * 3.0);
        return t * k + 4.0 / 29.0;
    };
    double fx = f(xyz[0] / white[0]), fy = f(xyz[1] / white[1]), fz = f(xyz[2] / white[2]);
    out[0] = 116.0 * fy - 16.0;
    out[1] = 500.0 * (fx - fy);
    out[2] = 200.0 * (fy - fz);
}

// target Lab minus the Lab of the reflectance of normalised-wavelength
// coefficients
void Residual(const Colour& c, const double coef[3], const double targetLab[3], double r[3])
{
    double rgb[3] = {0, 0, 0};
    const double norm = 1.0 / double(kCieRows);
    for(int i = 0; i < kCieRows; i++)
    {
        double l = double(i) * norm;
        double x = coef[0];
        x = x * l + coef[1];
        x = x * l + coef[2];
        double s = 0.5 * x;
        s /= std::sqrt(1.0 + x * x);
        s += 0.5;
        rgb[0] += c.toRgb[3 * i] * s;
        rgb[1] += c.toRgb[3 * i + 1] * s;
        rgb[2] += c.toRgb[3 * i + 2] * s;
    }
    auto xyz = Mul(c.rgbToXyz, rgb);
    double lab[3];
    Lab(c.white, xyz.data(), lab);
    for(int k = 0; k < 3; k++) r[k] = targetLab[k] - lab[k];
}

// A x = b by LU with partial pivoting; false when A is singular
bool Solve3(const double Ain[9], const double b[3], double x[3])
{
    double A[3][3];
    int perm[3] = {0, 1, 2};
    for(int r = 0; r < 3; r++)
        for(int k = 0; k < 3; k++) A[r][k] = Ain[3 * r + k];
    for(int col = 0; col < 3; col++)
    {
        int piv = col;
        for(int r = col + 1; r < 3; r++)
            if(std::fabs(A[r][col]) > std::fabs(A[piv][col])) piv = r;
        if(A[piv][col] == 0.0) return false;
        if(piv != col)
        {
            for(int k = 0; k < 3; k++) std::swap(A[col][k], A[piv][k]);
            std::swap(perm[col], perm[piv]);
        }
        for(int r = col + 1; r < 3; r++)
        {
            A[r][col] /= A[col][col];
            for(int k = col + 1; k < 3; k++) A[r][k] -= A[r][col] * A[col][k];
        }
    }
    double y[3];
    for(int r = 0; r < 3; r++)
    {
        y[r] = b[perm[r]];
        for(int k = 0; k < r; k++) y[r] -= A[r][k] * y[k];
    }
    for(int r = 2; r >= 0; r--)
    {
        x[r] = y[r];
        for(int k = r + 1; k < 3; k++) x[r] -= A[r][k] * x[k];
        x[r] /= A[r][r];
    }
    return true;
}

// Gauss-Newton from coef (in place), normalised-wavelength coefficients
void GaussNewton(const Colour& c, const double rgb[3], double coef[3])
{
    auto xyz = Mul(c.rgbToXyz, rgb);
    double target[3];
    Lab(c.white, xyz.data(), target);
    const double eps = 1e-4;
    for(int pass = 0; pass < kFitPasses; pass++)
    {
        double r[3];
        Residual(c, coef, target, r);
        double J[9];
        for(int k = 0; k < 3; k++)
        {
            double lo[3] = {coef[0], coef[1], coef[2]}, hi[3] = {coef[0], coef[1], coef[2]};
            lo[k] -= eps;
            hi[k] += eps;
            double r0[3], r1[3];
            Residual(c, lo, target, r0);
            Residual(c, hi, target, r1);
            for(int q = 0; q < 3; q++) J[3 * q + k] = (r1[q] - r0[q]) * (0.5 / eps);
        }
        double step[3];
        if(!Solve3(J, r, step)) { coef[0] = coef[1] = coef[2] = NAN; return; }
        for(int k = 0; k < 3; k++) coef[k] -= step[k];
        int mi = 0;
        if(coef[1] > coef[mi]) mi = 1;
        if(coef[2] > coef[mi]) mi = 2;
        if(coef[mi] > 200.0)
        {
            const double f = 200.0 / coef[mi];
            for(int k = 0; k < 3; k++) coef[k] *= f;
        }
        if(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < 1e-7) break;
    }
}

double Smoothstep(double x) { return x * x * (3.0 - 2.0 * x); }

struct HostTables
{
    std::vector<float> lut;   // [3][N][N][N][4]
    Colour colour;
    bool built = false;
    bool failed = false;      // the fit produced a non-finite coefficient
};

// The table file lives where the reference looks for it, relative to the
// binary: <library dir>/SpectraLUT/REC_709.mrspectra (LUT_FOLDER_NAME and
// the colour-space name, Core/ColorFunctions.h:112-116, Definitions.h:615).
std::string LibDir()
{
    Dl_info info;
    if(dladdr(reinterpret_cast<void*>(&LibDir), &info) && info.dli_fname)
    {
        std::string p(info.dli_fname);
        size_t s = p.find_last_of('/');
        return (s == std::string::npos) ? std::string(".") : p.substr(0, s);
    }
    return ".";
}
std::string CachePath() { return LibDir() + "/SpectraLUT/REC_709.mrspectra"; }

// .mrspectra (SpectraLUTGen/main.cpp:590-609, read back by
// ReadMRSpectraFileHeader, SpectrumContext.cu:298-352): the ten characters
// "MR_SPECTRA", u32 resolution (64), u32 mode (1 = float), then nine
// N^3 float volumes, x fastest, ordered [largest channel][c0, c1, c2].
constexpr char kCC[10] = {'M', 'R', '_', 'S', 'P', 'E', 'C', 'T', 'R', 'A'};

bool WriteMrSpectra(const std::vector<float>& lut, const std::string& path, std::string& err)
{
    const size_t N3 = size_t(kLutN) * kLutN * kLutN;
    std::vector<float> vol(N3);
    FILE* f = std::fopen(path.c_str(), "wb");
    if(!f) { err = "cannot open " + path + " for writing"; return false; }
    uint32_t n = kLutN, mode = 1;
    bool ok = std::fwrite(kCC, 1, sizeof(kCC), f) == sizeof(kCC) &&
              std::fwrite(&n, 4, 1, f) == 1 && std::fwrite(&mode, 4, 1, f) == 1;
    for(int l = 0; l < 3 && ok; l++)
        for(int c = 0; c < 3 && ok; c++)
        {
            for(size_t v = 0; v < N3; v++) vol[v] = lut[(size_t(l) * N3 + v) * 4 + c];
            ok = std::fwrite(vol.data(), 4, N3, f) == N3;
        }
    ok = (std::fclose(f) == 0) && ok;
    if(!ok) err = "short write to " + path;
    return ok;
}

bool ReadMrSpectra(const std::string& path, std::vector<float>& lut, std::string& err)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if(!f) { err = "Unable to open spectra lut file " + path; return false; }
    char cc[10] = {};
    uint32_t n = 0, mode = 0;
    bool ok = std::fread(cc, 1, sizeof(cc), f) == sizeof(cc);
    const char rev[4] = {kCC[3], kCC[2], kCC[1], kCC[0]};
    if(ok && std::memcmp(cc, rev, 4) == 0)
        err = "Wrong character code in .mrspectra file! Endianness issue maybe";
    else if(!ok || std::memcmp(cc, kCC, sizeof(kCC)) != 0)
        err = "Wrong character code in .mrspectra file!";
    else if(std::fread(&n, 4, 1, f) != 1 || n != uint32_t(kLutN))
        err = "Wrong size (" + std::to_string(n) + "), spectra lut size must be " +
              std::to_string(kLutN) + "!";
    else if(std::fread(&mode, 4, 1, f) != 1 || mode != 1)
        err = "Wrong mode (" + std::to_string(mode) + "), spectra lut must have mode \"1\" (aka. Float)!";
    if(!err.empty()) { std::fclose(f); return false; }
    const size_t N3 = size_t(kLutN) * kLutN * kLutN;
    std::vector<float> vol(N3);
    lut.assign(3 * N3 * 4, 0.0f);
    for(int l = 0; l < 3 && err.empty(); l++)
        for(int c = 0; c < 3 && err.empty(); c++)
        {
            if(std::fread(vol.data(), 4, N3, f) != N3) { err = "Unable to read sprectum lut file!"; break; }
            for(size_t v = 0; v < N3; v++)
            {
                if(!std::isfinite(vol[v])) { err = "non-finite coefficient in " + path; break; }
                lut[(size_t(l) * N3 + v) * 4 + c] = vol[v];
            }
        }
    std::fclose(f);
    return err.empty();
}

// Fills the table; false if a cell's fit is not finite (the reference's
// generator fails the same way, SpectraLUTGen/main.cpp:312-334).
bool BuildLut(HostTables& h)
{
    const int N = kLutN;
    const Colour& c = h.colour;
    h.lut.assign(size_t(3) * N * N * N * 4, 0.0f);
    std::vector<double> scale(N);
    for(int k = 0; k < N; k++) scale[k] = Smoothstep(Smoothstep(double(k) / double(N - 1)));
    // normalised wavelength i / 471 -> nanometres through (l - 360) / 470
    const double p0 = double(kCieFirstNm), p1 = 1.0 / double(kCieRows - 1);
    std::atomic<int> next{0};
    std::atomic<bool> ok{true};
    auto worker = [&]()
    {
        for(;;)
        {
            int job = next.fetch_add(1);
            if(job >= 3 * N * N) break;
            int ch = job / (N * N);
            int yi = (job / N) % N;
            int xi = job % N;
            double x = double(xi) / double(N - 1), y = double(yi) / double(N - 1);
            auto cell = [&](int k, double coef[3])
            {
                double z = scale[k];
                double rgb[3];
                rgb[ch] = z;
                rgb[(ch + 1) % 3] = x * z;
                rgb[(ch + 2) % 3] = y * z;
                GaussNewton(c, rgb, coef);
                if(!std::isfinite(coef[0]) || !std::isfinite(coef[1]) || !std::isfinite(coef[2]))
                    ok = false;
                const double A = coef[0], B = coef[1], Cc = coef[2];
                size_t o = ((((size_t(ch) * N + k) * N + yi) * N + xi) * 4);
                h.lut[o] = float(A * p1 * p1);
                h.lut[o + 1] = float(B * p1 - 2.0 * A * p0 * p1 * p1);
                h.lut[o + 2] = float(Cc - B * p0 * p1 + A * p0 * p1 * p0 * p1);
                h.lut[o + 3] = 0.0f;
            };
            // from N/5 upwards with warm starts, then from N/5 - 1 down,
            // warm-started from the N/5 cell
            const int mid = N / 5;
            double coef[3] = {0, 0, 0}, midCoef[3] = {0, 0, 0};
            for(int k = mid; k < N; k++)
            {
                cell(k, coef);
                if(k == mid) std::memcpy(midCoef, coef, sizeof(coef));
            }
            std::memcpy(coef, midCoef, sizeof(coef));
            for(int k = mid - 1; k >= 0; k--) cell(k, coef);
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for(unsigned t = 0; t < nt; t++) pool.emplace_back(worker);
    for(auto& t : pool) t.join();
    return ok;
}

HostTables& Host()
{
    static HostTables h;
    static std::mutex m;
    std::lock_guard<std::mutex> lk(m);
    if(h.built) return h;
    h.colour = MakeColour();
    std::string err;
    if(!ReadMrSpectra(CachePath(), h.lut, err))
    {
        h.failed = !BuildLut(h);
        if(!h.failed)
        {
            (void)::mkdir((LibDir() + "/SpectraLUT").c_str(), 0755);
            std::string werr;
            (void)WriteMrSpectra(h.lut, CachePath(), werr);   // cache only: failure is harmless
        }
    }
    h.built = true;
    return h;
}

// Tables read from a file replace the built ones, before first use only.
std::mutex gDevMutex;
bool gDevUploaded = false;

} // namespace

int GetSpectralTables(SpectralTables& out)
{
    static SpectralTables dev;
    std::lock_guard<std::mutex> lk(gDevMutex);
    if(!dev.ready)
    {
        HostTables& h = Host();
        if(h.failed)
        {
            SetError("spectral tables: the RGB->spectrum fit produced a non-finite coefficient");
            return MRAY_ERR_INVALID;
        }
        gDevUploaded = true;
        float4* dLut = nullptr;
        float4* dCmf = nullptr;
        float* dIllum = nullptr;
        const size_t lutN = size_t(3) * kLutN * kLutN * kLutN;
        MRAY_HIP_CHECK(hipMalloc((void**)&dLut, lutN * sizeof(float4)));
        MRAY_HIP_CHECK(hipMemcpy(dLut, h.lut.data(), lutN * sizeof(float4), hipMemcpyHostToDevice));
        std::vector<float4> cmf(kCieCount);
        for(uint32_t i = 0; i < kCieCount; i++)
            cmf[i] = make_float4(h.colour.cmf[3 * i], h.colour.cmf[3 * i + 1],
                                 h.colour.cmf[3 * i + 2], 0.0f);
        MRAY_HIP_CHECK(hipMalloc((void**)&dCmf, kCieCount * sizeof(float4)));
        MRAY_HIP_CHECK(hipMemcpy(dCmf, cmf.data(), kCieCount * sizeof(float4), hipMemcpyHostToDevice));
        MRAY_HIP_CHECK(hipMalloc((void**)&dIllum, kCieCount * sizeof(float)));
        MRAY_HIP_CHECK(hipMemcpy(dIllum, h.colour.illum.data(), kCieCount * sizeof(float),
                                 hipMemcpyHostToDevice));
        dev.view.lut = dLut;
        dev.view.cmf = dCmf;
        dev.view.illum = dIllum;
        for(int k = 0; k < 9; k++) dev.view.xyzToRgb[k] = float(h.colour.xyzToRgb[k]);
        dev.view.N = kLutN;
        dev.ready = true;   // kept for the process lifetime (shared by renderers)
    }
    out = dev;
    return MRAY_OK;
}

} // namespace mray

extern "C" {

// Host copies of the spectral tables (builds them on first use):
// lut [3][N][N][N][4] (c0, c1, c2, 0; x fastest), cmf [471][3], illum [471],
// xyzToRgb [9].  Any pointer may be NULL; *hN receives N.
int mray_spectral_tables(int* hN, float* hLut, float* hCmf, float* hIllum, float* hXyzToRgb)
{
    using namespace mray;
    HostTables& h = Host();
    if(h.failed)
    {
        SetError("spectral tables: the RGB->spectrum fit produced a non-finite coefficient");
        return MRAY_ERR_INVALID;
    }
    if(hN) *hN = kLutN;
    if(hLut) std::memcpy(hLut, h.lut.data(), h.lut.size() * sizeof(float));
    if(hCmf) std::memcpy(hCmf, h.colour.cmf.data(), h.colour.cmf.size() * sizeof(float));
    if(hIllum) std::memcpy(hIllum, h.colour.illum.data(), h.colour.illum.size() * sizeof(float));
    if(hXyzToRgb)
        for(int k = 0; k < 9; k++) hXyzToRgb[k] = float(h.colour.xyzToRgb[k]);
    return MRAY_OK;
}

int mray_spectral_write_mrspectra(const char* path)
{
    using namespace mray;
    if(!path) { SetError("spectral_write_mrspectra: NULL path"); return MRAY_ERR_INVALID; }
    std::string err;
    if(!WriteMrSpectra(Host().lut, path, err)) { SetError(err); return MRAY_ERR_INVALID; }
    return MRAY_OK;
}

int mray_spectral_read_mrspectra(const char* path, int* hN, float* hLut)
{
    using namespace mray;
    if(!path) { SetError("spectral_read_mrspectra: NULL path"); return MRAY_ERR_INVALID; }
    std::vector<float> lut;
    std::string err;
    if(!ReadMrSpectra(path, lut, err)) { SetError(err); return MRAY_ERR_INVALID; }
    if(hN) *hN = kLutN;
    if(hLut) std::memcpy(hLut, lut.data(), lut.size() * sizeof(float));
    return MRAY_OK;
}

int mray_spectral_use_mrspectra(const char* path)
{
    using namespace mray;
    if(!path) { SetError("spectral_use_mrspectra: NULL path"); return MRAY_ERR_INVALID; }
    std::vector<float> lut;
    std::string err;
    if(!ReadMrSpectra(path, lut, err)) { SetError(err); return MRAY_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(gDevMutex);
    if(gDevUploaded)
    {
        SetError("spectral_use_mrspectra: the tables are already in use by a spectral renderer");
        return MRAY_ERR_INVALID;
    }
    Host().lut.swap(lut);
    return MRAY_OK;
}

} // extern "C"
