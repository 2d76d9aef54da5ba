// TIFF export of reconstructed slices, host part: baseline 16-bit TIFF files
// written without libtiff (TIFF 6.0: little-endian header, one IFD, one row
// per strip, uncompressed), and the export of a volume in host memory.
#include "tomokvas/fdk.h"
#include "tomokvas/fdk_tiff.h"
#include "fdk_internal.hpp"
#include "fdk_tiff_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#pragma clang fp contract(off)

using tkv::fail;
using tkv::scale_u16;

namespace {

void put16(std::vector<unsigned char> &b, size_t at, unsigned v)
{
    b[at] = (unsigned char)(v & 0xff);
    b[at + 1] = (unsigned char)(v >> 8);
}

void put32(std::vector<unsigned char> &b, size_t at, unsigned v)
{
    for (int i = 0; i < 4; ++i) b[at + i] = (unsigned char)(v >> (8 * i));
}

std::string slice_name(const char *dir, int index)
{
    char name[32];
    std::snprintf(name, sizeof name, "%08d.tif", index);
    std::string d = dir && *dir ? dir : ".";
    if (d.back() != '/') d += '/';
    return d + name;
}

}  // namespace

namespace tkv {

int write_slices(const uint16_t *img, int x, int y, int count, const char *dir, int first)
{
    for (int k = 0; k < count; ++k) {
        const std::string p = slice_name(dir, first + k);
        int rc = fdk_tiff_write_u16(p.c_str(), img + (size_t)k * x * y, x, y);
        if (rc) return rc;
    }
    return FDK_OK;
}

int writer_threads(int count)
{
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    return std::max(1, std::min({8, hw, count}));
}

}  // namespace tkv

extern "C" int fdk_tiff_write_u16(const char *path, const uint16_t *img, int width, int height)
{
    if (!path || !img || width < 1 || height < 1)
        return fail(FDK_ERR_INVALID_ARGUMENT, "tiff: null argument or empty image");
    const unsigned row = 2u * (unsigned)width;
    if ((unsigned long long)row * height > 0xffff0000ull)
        return fail(FDK_ERR_UNSUPPORTED, "tiff: image above 4 GiB");
    // little-endian baseline TIFF: header, one IFD of 10 entries, the strip
    // offset and byte-count arrays (inline when one row), then the rows
    const int nent = 10;
    const size_t ifd = 8, ifd_bytes = 2 + 12 * nent + 4;
    const bool inline_arrays = height == 1;
    const size_t offs_at = ifd + ifd_bytes;
    const size_t counts_at = offs_at + (inline_arrays ? 0 : 4 * (size_t)height);
    const size_t data_at = counts_at + (inline_arrays ? 0 : 4 * (size_t)height);
    std::vector<unsigned char> h(data_at, 0);
    h[0] = 'I';
    h[1] = 'I';
    put16(h, 2, 42);
    put32(h, 4, (unsigned)ifd);
    put16(h, ifd, nent);
    size_t e = ifd + 2;
    auto entry = [&](unsigned tag, unsigned type, unsigned count, unsigned value) {
        put16(h, e, tag);
        put16(h, e + 2, type);
        put32(h, e + 4, count);
        if (type == 3 && count == 1) put16(h, e + 8, value);
        else put32(h, e + 8, value);
        e += 12;
    };
    const unsigned SHORT = 3, LONG = 4;
    entry(256, LONG, 1, (unsigned)width);                       // ImageWidth
    entry(257, LONG, 1, (unsigned)height);                      // ImageLength
    entry(258, SHORT, 1, 16);                                   // BitsPerSample
    entry(259, SHORT, 1, 1);                                    // Compression: none
    entry(262, SHORT, 1, 1);                                    // Photometric: min-is-black
    entry(273, LONG, (unsigned)height,                          // StripOffsets
          inline_arrays ? (unsigned)data_at : (unsigned)offs_at);
    entry(277, SHORT, 1, 1);                                    // SamplesPerPixel
    entry(278, LONG, 1, 1);                                     // RowsPerStrip
    entry(279, LONG, (unsigned)height, inline_arrays ? row : (unsigned)counts_at);  // StripByteCounts
    entry(284, SHORT, 1, 1);                                    // PlanarConfiguration
    put32(h, e, 0);                                             // no next IFD
    if (!inline_arrays) {
        for (int r = 0; r < height; ++r) {
            put32(h, offs_at + 4 * (size_t)r, (unsigned)(data_at + (size_t)r * row));
            put32(h, counts_at + 4 * (size_t)r, row);
        }
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(FDK_ERR_INVALID_ARGUMENT, std::string("tiff: cannot write ") + path);
    bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    // samples are little-endian on this host, as the header says
    ok = ok && std::fwrite(img, 2, (size_t)width * height, f) == (size_t)width * height;
    ok = std::fclose(f) == 0 && ok;
    return ok ? FDK_OK : fail(FDK_ERR_INVALID_ARGUMENT, std::string("tiff: short write to ") + path);
}

extern "C" int fdk_tiff_export_host(const float *vol, int x, int y, int z, double vmin,
                                    double vmax, const char *dir, int first_index)
{
    if (!vol || x < 1 || y < 1 || z < 1)
        return fail(FDK_ERR_INVALID_ARGUMENT, "tiff export: empty volume");
    const size_t plane = (size_t)x * y;
    const int threads = tkv::writer_threads(z);
    std::vector<std::vector<uint16_t>> img(threads, std::vector<uint16_t>(plane));
    std::string err;
    const int rc = tkv::write_slices_parallel(z, threads, err, [&](int t, int k) {
        const float *s = vol + (size_t)k * plane;
        uint16_t *d = img[t].data();
        for (size_t i = 0; i < plane; ++i) d[i] = scale_u16(s[i], vmin, vmax);
        return tkv::write_slices(d, x, y, 1, dir, first_index + k);
    });
    if (rc) return fail(rc, err);
    return FDK_OK;
}
