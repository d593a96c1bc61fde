// RGB -> spectrum coefficient tables (the capability of the reference's
// SpectraLUTGen, Source/SpectraLUTGen/main.cpp, loaded at
// Tracer/SpectrumContext.cu:298-472), built from the published method:
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
// Colour data (this build's own, no tables are copied):
//   observer    CIE 1931 2-degree colour matching functions, multi-lobe
//               piecewise-Gaussian fit of Wyman, Sloan & Shirley 2013 (JCGT 2(2))
//   illuminant  Planck radiator at 6504 K, normalised to Y = 1
//   RGB space   Rec.709 primaries with the illuminant as white point
// Stored coefficients evaluate in nanometres, as Converter::ConvertAlbedo does.
#include "spectral.h"

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
constexpr double kLmin = 360.0, kLmax = 830.0;

double Lobe(double x, double mu, double s1, double s2)
{
    double t = (x - mu) / ((x < mu) ? s1 : s2);
    return std::exp(-0.5 * t * t);
}
double XBar(double l)
{
    return 1.056 * Lobe(l, 599.8, 37.9, 31.0) + 0.362 * Lobe(l, 442.0, 16.0, 26.7) -
           0.065 * Lobe(l, 501.1, 20.4, 26.2);
}
double YBar(double l) { return 0.821 * Lobe(l, 568.8, 46.9, 40.5) + 0.286 * Lobe(l, 530.9, 16.3, 31.1); }
double ZBar(double l) { return 1.217 * Lobe(l, 437.0, 11.8, 36.0) + 0.681 * Lobe(l, 459.0, 26.0, 13.8); }
double Planck(double lnm, double T)
{
    const double c1 = 3.741771852e-16, c2 = 1.438776877e-2;
    double l = lnm * 1e-9;
    return c1 / (l * l * l * l * l * (std::exp(c2 / (l * T)) - 1.0));
}

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

struct Colour
{
    // fit quadrature (1 nm)
    std::vector<double> lam, wx, wy, wz;   // CMF * illuminant * dl, normalised
    double white[3];
    M3 rgbToXyz, xyzToRgb;
    std::vector<float> cmf;                // [471][3]
    std::vector<float> illum;              // [471]
};

// Observer and illuminant normalised as the reference does at load time
// (SpectrumContext.cu:409-426): each CMF channel over its own integral, the
// illuminant scaled by integral(ybar) / integral(illum * ybar), so the
// illuminant's Y is 1 and a flat unit spectrum has XYZ = (1, 1, 1).
Colour MakeColour()
{
    Colour c;
    const double T = 6504.0;
    double ix = 0, iy = 0, iz = 0, illumY = 0;
    for(int i = 0; i <= 470; i++)
    {
        double l = kLmin + i;
        ix += XBar(l);
        iy += YBar(l);
        iz += ZBar(l);
        illumY += Planck(l, T) * YBar(l);
    }
    const double illumScale = iy / illumY;
    c.cmf.resize(471 * 3);
    c.illum.resize(471);
    c.white[0] = c.white[1] = c.white[2] = 0.0;
    for(int i = 0; i <= 470; i++)
    {
        double l = kLmin + i;
        double cx = XBar(l) / ix, cy = YBar(l) / iy, cz = ZBar(l) / iz;
        double il = Planck(l, T) * illumScale;
        c.cmf[3 * i] = float(cx);
        c.cmf[3 * i + 1] = float(cy);
        c.cmf[3 * i + 2] = float(cz);
        c.illum[i] = float(il);
        c.lam.push_back(l);
        c.wx.push_back(cx * il);
        c.wy.push_back(cy * il);
        c.wz.push_back(cz * il);
        c.white[0] += cx * il;
        c.white[1] += cy * il;
        c.white[2] += cz * il;
    }
    // Rec.709 primaries, white = the illuminant's XYZ
    const double xr = 0.64, yr = 0.33, xg = 0.30, yg = 0.60, xb = 0.15, yb = 0.06;
    M3 P = {xr / yr, xg / yg, xb / yb, 1.0, 1.0, 1.0, (1 - xr - yr) / yr, (1 - xg - yg) / yg,
            (1 - xb - yb) / yb};
    M3 Pi = Inverse(P);
    double S[3];
    for(int r = 0; r < 3; r++)
        S[r] = Pi[3 * r] * c.white[0] + Pi[3 * r + 1] * c.white[1] + Pi[3 * r + 2] * c.white[2];
    for(int r = 0; r < 3; r++)
        for(int k = 0; k < 3; k++) c.rgbToXyz[3 * r + k] = P[3 * r + k] * S[k];
    c.xyzToRgb = Inverse(c.rgbToXyz);
    return c;
}

double Sigmoid(double x) { return 0.5 + x / (2.0 * std::sqrt(1.0 + x * x)); }

void Lab(const Colour& c, const double xyz[3], double out[3])
{
    auto f = [](double t)
    {
        const double d = 6.0 / 29.0;
        return (t > d * d * d) ? std::cbrt(t) : t / (3 * d * d) + 4.0 / 29.0;
    };
    double fx = f(xyz[0] / c.white[0]), fy = f(xyz[1] / c.white[1]), fz = f(xyz[2] / c.white[2]);
    out[0] = 116.0 * fy - 16.0;
    out[1] = 500.0 * (fx - fy);
    out[2] = 200.0 * (fy - fz);
}

// residual in Lab of normalised-wavelength coefficients
void Residual(const Colour& c, const double coef[3], const double targetLab[3], double r[3])
{
    double xyz[3] = {0, 0, 0};
    const size_t n = c.lam.size();
    for(size_t i = 0; i < n; i++)
    {
        double t = (c.lam[i] - kLmin) / (kLmax - kLmin);
        double s = Sigmoid((coef[0] * t + coef[1]) * t + coef[2]);
        xyz[0] += s * c.wx[i];
        xyz[1] += s * c.wy[i];
        xyz[2] += s * c.wz[i];
    }
    double lab[3];
    Lab(c, xyz, lab);
    for(int k = 0; k < 3; k++) r[k] = lab[k] - targetLab[k];
}

bool Solve3(double A[9], double b[3], double x[3])
{
    // Gaussian elimination with partial pivoting
    double M[3][4];
    for(int r = 0; r < 3; r++)
    {
        for(int k = 0; k < 3; k++) M[r][k] = A[3 * r + k];
        M[r][3] = b[r];
    }
    for(int col = 0; col < 3; col++)
    {
        int piv = col;
        for(int r = col + 1; r < 3; r++)
            if(std::fabs(M[r][col]) > std::fabs(M[piv][col])) piv = r;
        if(std::fabs(M[piv][col]) < 1e-30) return false;
        std::swap(M[col], M[piv]);
        for(int r = 0; r < 3; r++)
        {
            if(r == col) continue;
            double f = M[r][col] / M[col][col];
            for(int k = col; k < 4; k++) M[r][k] -= f * M[col][k];
        }
    }
    for(int r = 0; r < 3; r++) x[r] = M[r][3] / M[r][r];
    return true;
}

void GaussNewton(const Colour& c, const double rgb[3], double coef[3])
{
    double xyz[3];
    for(int r = 0; r < 3; r++)
        xyz[r] = c.rgbToXyz[3 * r] * rgb[0] + c.rgbToXyz[3 * r + 1] * rgb[1] +
                 c.rgbToXyz[3 * r + 2] * rgb[2];
    double target[3];
    Lab(c, xyz, target);
    auto norm = [](const double r[3]) { return std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]); };
    double r[3];
    Residual(c, coef, target, r);
    for(int it = 0; it < 30; it++)
    {
        double rn = norm(r);
        if(rn < 1e-6) break;
        double J[9];
        const double eps = 1e-5;
        for(int k = 0; k < 3; k++)
        {
            double cp[3] = {coef[0], coef[1], coef[2]};
            cp[k] += eps;
            double r1[3];
            Residual(c, cp, target, r1);
            for(int q = 0; q < 3; q++) J[3 * q + k] = (r1[q] - r[q]) / eps;
        }
        double delta[3];
        double rhs[3] = {r[0], r[1], r[2]};
        if(!Solve3(J, rhs, delta)) break;
        // backtracking: accept the first step length that lowers the residual
        double step = 1.0;
        bool improved = false;
        for(int ls = 0; ls < 12; ls++, step *= 0.5)
        {
            double cand[3] = {coef[0] - step * delta[0], coef[1] - step * delta[1],
                              coef[2] - step * delta[2]};
            double rc[3];
            Residual(c, cand, target, rc);
            if(norm(rc) < rn)
            {
                coef[0] = cand[0]; coef[1] = cand[1]; coef[2] = cand[2];
                r[0] = rc[0]; r[1] = rc[1]; r[2] = rc[2];
                improved = true;
                break;
            }
        }
        if(!improved) break;
    }
}

double Smoothstep(double x) { return x * x * (3.0 - 2.0 * x); }

struct HostTables
{
    std::vector<float> lut;   // [3][N][N][N][4]
    Colour colour;
    bool built = false;
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

void BuildLut(HostTables& h)
{
    const int N = kLutN;
    const Colour& c = h.colour;
    h.lut.assign(size_t(3) * N * N * N * 4, 0.0f);
    std::vector<double> scale(N);
    for(int k = 0; k < N; k++) scale[k] = Smoothstep(Smoothstep(double(k) / double(N - 1)));
    const double range = kLmax - kLmin;
    std::atomic<int> next{0};
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
            auto solveRange = [&](int k0, int k1, int step)
            {
                double coef[3] = {0, 0, 0};
                for(int k = k0; k != k1; k += step)
                {
                    double z = scale[k];
                    double rgb[3];
                    rgb[ch] = z;
                    rgb[(ch + 1) % 3] = x * z;
                    rgb[(ch + 2) % 3] = y * z;
                    if(k == 0)
                    {
                        // black: the sigmoid only reaches 0 asymptotically;
                        // S(-50) ~ 1e-4 everywhere
                        coef[0] = 0.0; coef[1] = 0.0; coef[2] = -50.0;
                    }
                    else GaussNewton(c, rgb, coef);
                    // normalised wavelength -> nanometres
                    double A = coef[0], B = coef[1], Cc = coef[2];
                    double c0 = A / (range * range);
                    double c1 = B / range - 2.0 * A * kLmin / (range * range);
                    double c2 = A * kLmin * kLmin / (range * range) - B * kLmin / range + Cc;
                    size_t o = ((((size_t(ch) * N + k) * N + yi) * N + xi) * 4);
                    h.lut[o] = float(c0);
                    h.lut[o + 1] = float(c1);
                    h.lut[o + 2] = float(c2);
                    h.lut[o + 3] = 0.0f;
                }
            };
            int start = N / 5;
            solveRange(start, N, 1);
            solveRange(start - 1, -1, -1);
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for(unsigned t = 0; t < nt; t++) pool.emplace_back(worker);
    for(auto& t : pool) t.join();
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
        BuildLut(h);
        (void)::mkdir((LibDir() + "/SpectraLUT").c_str(), 0755);
        std::string werr;
        (void)WriteMrSpectra(h.lut, CachePath(), werr);   // cache only: failure is harmless
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
        gDevUploaded = true;
        HostTables& h = Host();
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
