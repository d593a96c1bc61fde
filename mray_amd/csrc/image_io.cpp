// Film -> OpenEXR, the image the reference's run command saves when a
// renderer reports triggerSave (MRay/RunCommand.cpp:873-930): per-pixel
// weighted mean R, G, B as 32-bit float channels, top row first (the
// reference's writer flips the film, whose row 0 is the bottom of the camera
// plane, ImageLoader.cpp:118-123).
//
// The file follows the OpenEXR 2 single-part scanline layout, written here
// from the published file-format description: magic 20000630, version 2,
// a header of (name, type, size, value) attributes ending in a zero byte,
// a table of one 64-bit offset per scanline, then per scanline the y
// coordinate, the byte count and the channels' pixels in channel-name
// order (B, G, R), uncompressed.
#include "common.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace mray
{
namespace
{
struct Bytes
{
    std::vector<unsigned char> b;
    void Raw(const void* p, size_t n)
    {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        b.insert(b.end(), c, c + n);
    }
    void Str(const char* s) { Raw(s, std::strlen(s) + 1); }
    void U8(uint8_t v) { Raw(&v, 1); }
    void I32(int32_t v) { Raw(&v, 4); }
    void F32(float v) { Raw(&v, 4); }
    void U64(uint64_t v) { Raw(&v, 8); }
    void Attr(const char* name, const char* type, int32_t size)
    {
        Str(name);
        Str(type);
        I32(size);
    }
};
} // namespace

int WriteFilmExr(const float* film, uint32_t W, uint32_t H, const char* path)
{
    const size_t P = size_t(W) * H;
    Bytes h;
    const uint32_t magic = 20000630u;
    h.Raw(&magic, 4);
    h.I32(2);                                   // version 2, single-part scanline
    const char* chans[3] = {"B", "G", "R"};
    h.Attr("channels", "chlist", 3 * (2 + 16) + 1);
    for(const char* c : chans)
    {
        h.Str(c);
        h.I32(2);                               // FLOAT
        h.U8(0);                                // pLinear
        h.U8(0); h.U8(0); h.U8(0);              // reserved
        h.I32(1);                               // xSampling
        h.I32(1);                               // ySampling
    }
    h.U8(0);
    h.Attr("compression", "compression", 1);
    h.U8(0);                                    // NO_COMPRESSION
    h.Attr("dataWindow", "box2i", 16);
    h.I32(0); h.I32(0); h.I32(int32_t(W) - 1); h.I32(int32_t(H) - 1);
    h.Attr("displayWindow", "box2i", 16);
    h.I32(0); h.I32(0); h.I32(int32_t(W) - 1); h.I32(int32_t(H) - 1);
    h.Attr("lineOrder", "lineOrder", 1);
    h.U8(0);                                    // INCREASING_Y
    h.Attr("pixelAspectRatio", "float", 4);
    h.F32(1.0f);
    h.Attr("screenWindowCenter", "v2f", 8);
    h.F32(0.0f); h.F32(0.0f);
    h.Attr("screenWindowWidth", "float", 4);
    h.F32(1.0f);
    h.U8(0);                                    // end of header
    const uint64_t lineBytes = uint64_t(W) * 3 * 4;
    const uint64_t first = h.b.size() + uint64_t(H) * 8;
    for(uint32_t y = 0; y < H; y++) h.U64(first + uint64_t(y) * (8 + lineBytes));
    std::vector<float> line(size_t(W) * 3);
    FILE* f = std::fopen(path, "wb");
    if(!f)
    {
        SetError(std::string("film_write_exr: cannot open ") + path);
        return MRAY_ERR_INVALID;
    }
    bool ok = std::fwrite(h.b.data(), 1, h.b.size(), f) == h.b.size();
    for(uint32_t y = 0; y < H && ok; y++)
    {
        const size_t row = size_t(H - 1 - y) * W;   // top row first
        for(uint32_t x = 0; x < W; x++)
        {
            const size_t i = row + x;
            const double w = film[3 * P + i];
            for(int c = 0; c < 3; c++)
            {
                // the run command's accumulation: sum / weight, 0 where no weight
                const double v = (w == 0.0) ? 0.0 : double(film[size_t(c) * P + i]) / w;
                line[size_t(2 - c) * W + x] = float(v);   // B, G, R planes
            }
        }
        int32_t hdr[2] = {int32_t(y), int32_t(lineBytes)};
        ok = std::fwrite(hdr, 4, 2, f) == 2 &&
             std::fwrite(line.data(), 4, line.size(), f) == line.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    if(!ok)
    {
        SetError(std::string("film_write_exr: short write to ") + path);
        return MRAY_ERR_INVALID;
    }
    return MRAY_OK;
}

} // namespace mray

extern "C" int mray_film_write_exr(const float* hFilm, uint32_t width, uint32_t height,
                                   const char* path)
{
    if(!hFilm || !path || width == 0 || height == 0)
    {
        mray::SetError("film_write_exr: NULL argument or empty image");
        return MRAY_ERR_INVALID;
    }
    return mray::WriteFilmExr(hFilm, width, height, path);
}
