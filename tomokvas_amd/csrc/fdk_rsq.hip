// Scanco RSQ raw projections: header on the host, preprocessing on the GPU.
//
// The raw sinograms stay 16-bit until they are on the device (a quarter of the
// bytes of the reference's FP64 sinograms cross the host link); one kernel then
// forms every output sample from at most two raw samples: flat/dark
// correction and the beam-hardening line on each, and for the full-fan
// rebinning the linear interpolation between the two angle rows of the
// redundant ray.  All arithmetic is FP64 in the reference's order of
// operations (no contraction into FMAs); atan(t / distance) comes from a
// per-column table computed on the host, as the reference does (atan_vect,
// rsq_load.cpp:329-333); log is the device's FP64 function.
#include "tomokvas/fdk.h"
#include "tomokvas/fdk_rsq.h"
#include "fdk_internal.hpp"

#include <hip/hip_runtime.h>
#include <sys/types.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#pragma clang fp contract(off)

using tkv::fail;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kHeaderBytes = 512;     // sizeof(ima_data_type)
constexpr int kBhOffset = 3432, kBhBytes = 152;

struct RsqArgs {
    const uint16_t *raw;   // [n][sin_y + 2][sin_x]
    void *out;             // [n][sy][sx]
    const double *at;      // atan(t / distance) for t = -1 .. -(sin_x-1-axis_move), at[-t]
    int n, sin_x, sin_y, sx, sy, axis_move, start, flip, invert, complete;
    double A, B, distance, delta_fi;
};

// line integral of raw sample (row y of the projections, column x) of one
// sinogram: rows 0 and 1 are dark and flat (FlatFieldCorrection,
// rsq_load.cpp:299-325)
__device__ double line_integral(const RsqArgs &a, const uint16_t *blk, int y, int x)
{
    const int d = blk[x], f = blk[a.sin_x + x], i = blk[(size_t)(y + 2) * a.sin_x + x];
    double v = (double)(i - d) / (double)(f - d);
    v = a.A * log(v) + a.B;
    return a.invert ? -v : v;
}

template <typename T>
__global__ void __launch_bounds__(256) rsq_kernel(RsqArgs a)
{
    const size_t per = (size_t)a.sy * a.sx;
    const size_t total = per * a.n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(i / per);
        const size_t rem = i - (size_t)k * per;
        const int r = (int)(rem / a.sx), c = (int)(rem - (size_t)r * a.sx);
        const uint16_t *blk = a.raw + (size_t)k * (a.sin_y + 2) * a.sin_x;
        double v;
        if (!a.complete) {
            // CutHalfSin (rsq_load.cpp:393-410)
            v = line_integral(a, blk, a.flip ? a.sin_y - 1 - r : r, a.axis_move + c);
        } else {
            // Half2FullFan (rsq_load.cpp:327-391): output row r is angle f
            const int f = a.flip ? a.sy - 1 - r + a.start : r;
            const int c0 = a.sin_x - 1 - a.axis_move;   // column of the central ray
            if (c >= c0) {
                v = line_integral(a, blk, f, a.axis_move + c - c0);
            } else {
                const int t = c - c0;                    // < 0
                const double at = a.at[-t];              // atan(t / distance), host table
                const double fi = f * a.delta_fi;
                double nf = (fi - 2 * at - kPi) / a.delta_fi;
                if (nf > a.sin_y - 1 || nf < 0) nf = (fi - 2 * at + kPi) / a.delta_fi;
                const double ca = ceil(nf), fb = floor(nf);
                int ia = (int)ca, ib = (int)fb;
                // inside the scan for delta_fi = 2 pi / sin_y (ia <= sin_y wraps
                // to row 0 as in the reference); any other angle wraps too
                ia = ((ia % a.sin_y) + a.sin_y) % a.sin_y;
                ib = ((ib % a.sin_y) + a.sin_y) % a.sin_y;
                const int col = -t + a.axis_move;
                const double A = line_integral(a, blk, ia, col);
                if (ca != fb) {
                    const double B = line_integral(a, blk, ib, col);
                    v = A + (B - A) * (-nf + ca);
                } else {
                    v = A;
                }
            }
        }
        static_cast<T *>(a.out)[i] = (T)v;
    }
}

int read_i32(const unsigned char *b, int off)
{
    int v;
    std::memcpy(&v, b + off, 4);   // little-endian file, little-endian host
    return v;
}

// the file layout: sinogram size and data offset
int check_layout(const fdk_rsq_params *p)
{
    if (!p) return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: null parameters");
    if (p->sin_x < 1 || p->sin_y < 1 || p->header_offset < 0)
        return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: sizes must be positive");
    return FDK_OK;
}

// the layout and the processing options
int check_params(const fdk_rsq_params *p)
{
    if (int rc = check_layout(p)) return rc;
    if (p->axis_move < 0 || p->axis_move >= p->sin_x)
        return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: axis_move outside the detector");
    if (p->complete && (!(p->distance > 0) || !(p->delta_fi > 0)))
        return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: distance and delta_fi must be positive");
    return FDK_OK;
}

// CalculateSizeOfSinogram (rsq_load.cpp:82-105), parallel branch not provided
void out_size(const fdk_rsq_params *p, int *sx, int *sy)
{
    if (p->complete) {
        *sx = 2 * p->sin_x - 1 - 2 * p->axis_move;
        *sy = p->minimal ? (int)std::ceil((kPi + 2 * std::atan(*sx / 2. / p->distance)) / p->delta_fi)
                         : p->sin_y;
        if (p->minimal && *sy > p->sin_y) *sy = p->sin_y;
    } else {
        *sx = p->sin_x - p->axis_move;
        *sy = p->sin_y;
    }
}

}  // namespace

struct fdk_rsq {
    std::string path;
    fdk_rsq_header h{};
};

extern "C" int fdk_rsq_default_params(fdk_rsq_params *p)
{
    if (!p) return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: null parameters");
    p->sin_x = 1024;
    p->sin_y = 2000;
    p->header_offset = 3584;
    p->axis_move = 21;
    p->distance = 125810. / 24.;
    p->delta_fi = 2 * kPi / p->sin_y;
    p->flip = p->invert = p->complete = p->minimal = 0;
    return FDK_OK;
}

extern "C" int fdk_rsq_open(const char *path, fdk_rsq **out)
{
    if (!path || !out) return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: null argument");
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(FDK_ERR_INVALID_ARGUMENT, std::string("rsq: cannot open ") + path);
    unsigned char b[kBhOffset + kBhBytes] = {};
    const size_t got = std::fread(b, 1, sizeof b, f);
    fseeko(f, 0, SEEK_END);
    const long long size = (long long)ftello(f);
    std::fclose(f);
    if (got < sizeof b) return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: file shorter than its header");
    auto *r = new fdk_rsq;
    r->path = path;
    fdk_rsq_header &h = r->h;
    // ima_data_type (rsq_load.h:13-49): check[16], then 4-byte ints; name[40]
    // at 128, data_offset the last int of the 512 bytes
    std::memcpy(h.check, b, 16);
    h.check[16] = 0;
    h.data_type = read_i32(b, 16);
    h.nr_of_bytes = read_i32(b, 20);
    h.nr_of_blocks = read_i32(b, 24);
    h.patient_index = read_i32(b, 28);
    h.scanner_id = read_i32(b, 32);
    h.dimx_p = read_i32(b, 44);
    h.dimy_p = read_i32(b, 48);
    h.dimz_p = read_i32(b, 52);
    h.dimx_um = read_i32(b, 56);
    h.dimy_um = read_i32(b, 60);
    h.dimz_um = read_i32(b, 64);
    h.mu_scaling = read_i32(b, 88);
    h.nr_of_samples = read_i32(b, 92);
    h.nr_of_projections = read_i32(b, 96);
    h.scandist_um = read_i32(b, 100);
    h.scanner_type = read_i32(b, 104);
    h.reference_line_um = read_i32(b, 120);
    h.energy = read_i32(b, 168);
    h.intensity = read_i32(b, 172);
    h.data_offset = read_i32(b, kHeaderBytes - 4);
    h.file_size = size;
    char bh[kBhBytes + 1];
    std::memcpy(bh, b + kBhOffset, kBhBytes);
    bh[kBhBytes] = 0;
    if (std::sscanf(bh, "Corr at %d %d %d: %lf %lf %lf", &h.bh_x[0], &h.bh_x[1], &h.bh_x[2],
                    &h.bh_y[0], &h.bh_y[1], &h.bh_y[2]) != 6 ||
        h.bh_x[2] == h.bh_x[0]) {
        delete r;
        return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: no beam-hardening line at byte 3432");
    }
    h.bh_a = (h.bh_y[2] - h.bh_y[0]) / (h.bh_x[2] - h.bh_x[0]);
    h.bh_b = h.bh_y[0] - h.bh_a * h.bh_x[0];
    *out = r;
    return FDK_OK;
}

extern "C" void fdk_rsq_close(fdk_rsq *r) { delete r; }

extern "C" int fdk_rsq_get_header(const fdk_rsq *r, fdk_rsq_header *h)
{
    if (!r || !h) return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: null argument");
    *h = r->h;
    return FDK_OK;
}

namespace {
int count_sinograms(const fdk_rsq *r, const fdk_rsq_params *p)
{
    const long long per = 2LL * p->sin_x * (p->sin_y + 2);
    const long long data = r->h.file_size - p->header_offset;
    return data > 0 ? (int)(data / per) : 0;
}
}  // namespace

extern "C" int fdk_rsq_sizes(const fdk_rsq *r, const fdk_rsq_params *p, int *num, int *sx,
                             int *sy)
{
    if (!r || !num || !sx || !sy) return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: null argument");
    int rc = check_params(p);
    if (rc) return rc;
    *num = count_sinograms(r, p);
    out_size(p, sx, sy);
    return FDK_OK;
}

extern "C" int fdk_rsq_load(const fdk_rsq *r, const fdk_rsq_params *p, int first, int last,
                            uint16_t *dst)
{
    if (!r || !dst) return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: null argument");
    int rc = check_layout(p);
    if (rc) return rc;
    const int num = count_sinograms(r, p);
    if (first < 0 || last < first || last >= num)
        return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: sinograms " + std::to_string(first) + ".." +
                                                  std::to_string(last) + " outside 0.." +
                                                  std::to_string(num - 1));
    FILE *f = std::fopen(r->path.c_str(), "rb");
    if (!f) return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: cannot open " + r->path);
    const size_t per = (size_t)p->sin_x * (p->sin_y + 2);
    const size_t count = per * (size_t)(last + 1 - first);
    const off_t off = (off_t)(p->header_offset + (long long)per * 2 * first);
    const bool ok = fseeko(f, off, SEEK_SET) == 0 && std::fread(dst, 2, count, f) == count;
    std::fclose(f);
    return ok ? FDK_OK : fail(FDK_ERR_INVALID_ARGUMENT, "rsq: short read from " + r->path);
}

extern "C" int fdk_rsq_process(const fdk_rsq_params *p, double bh_a, double bh_b,
                               const uint16_t *d_raw, int nsin, int out_dtype, void *d_out,
                               void *stream)
{
    int rc = check_params(p);
    if (rc) return rc;
    if (!d_raw || !d_out || nsin < 0)
        return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: null buffer or negative count");
    if (out_dtype != FDK_F32 && out_dtype != FDK_F64)
        return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: output must be FDK_F32 or FDK_F64");
    RsqArgs a;
    a.raw = d_raw;
    a.out = d_out;
    a.n = nsin;
    a.sin_x = p->sin_x;
    a.sin_y = p->sin_y;
    out_size(p, &a.sx, &a.sy);
    a.axis_move = p->axis_move;
    a.start = p->flip && p->complete ? p->sin_y - a.sy : 0;
    a.flip = p->flip;
    a.invert = p->invert;
    a.complete = p->complete;
    a.A = bh_a;
    a.B = bh_b;
    a.distance = p->distance;
    a.delta_fi = p->delta_fi;
    const size_t total = (size_t)a.sx * a.sy * nsin;
    if (total == 0) return FDK_OK;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
    hipStream_t st = (hipStream_t)stream;
    // atan table, index -t for t = -1 .. -(sin_x - 1 - axis_move) (entry 0 unused)
    const int nt = p->sin_x - p->axis_move;
    std::vector<double> at(nt, 0.0);
    for (int t = -p->sin_x + 1 + p->axis_move; t < 0; ++t) at[-t] = std::atan(t / p->distance);
    double *d_at = nullptr;
    TKV_HIP(hipMalloc((void **)&d_at, nt * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d_at, at.data(), nt * sizeof(double), hipMemcpyHostToDevice, st);
    a.at = d_at;
    if (e == hipSuccess) {
        if (out_dtype == FDK_F64)
            hipLaunchKernelGGL(rsq_kernel<double>, dim3(blocks), dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL(rsq_kernel<float>, dim3(blocks), dim3(256), 0, st, a);
        e = hipGetLastError();
    }
    // the table lives until the kernel is done
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_at);
    if (e != hipSuccess) return tkv::hip_fail(e, "rsq");
    return FDK_OK;
}
