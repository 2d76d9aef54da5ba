// Scanco RSQ raw projections: header on the host, preprocessing on the GPU.
//
// The raw sinograms stay 16-bit until they are on the device (a quarter of the
// bytes of the reference's FP64 sinograms cross the host link).  Every output
// sample comes from at most two raw samples: flat/dark correction and the
// beam-hardening line on each, and for the full-fan rebinning the linear
// interpolation between the two angle rows of the redundant ray.  The FP64
// log dominates the cost, so the rebinning runs in two passes over groups of
// sinograms: the line integrals of the raw samples it reads, once each, into
// an FP64 scratch small enough to stay in the last-level cache, then the
// rebinned rows from them (the same FP64 values the single-pass form would
// compute, so the result does not depend on the grouping).  All arithmetic is
// FP64 in the reference's order of operations (no contraction into FMAs);
// atan(t / distance) comes from a per-column table computed on the host, as
// the reference does (atan_vect, rsq_load.cpp:329-333); log is the device's
// FP64 function.
#include "tomokvas/fdk.h"
#include "tomokvas/fdk_rsq.h"
#include "fdk_internal.hpp"

#include <hip/hip_runtime.h>
#include <sys/types.h>

#include <algorithm>
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
// line-integral scratch of one group of sinograms in the full-fan rebinning
// (one sinogram when a single one is larger)
constexpr size_t kGroupBytes = size_t(64) << 20;

struct RsqArgs {
    const uint16_t *raw;   // [n][sin_y + 2][sin_x]
    void *out;             // [n][sy][sx]
    const double *at;      // atan(t / distance) for t = -1 .. -(sin_x-1-axis_move), at[-t]
    int sin_x, sin_y, sx, sy, axis_move, start, flip, invert, complete;
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

// CutHalfSin (rsq_load.cpp:393-410): one raw sample per output sample;
// grid (column tiles, output rows, sinograms)
template <typename T>
__global__ void __launch_bounds__(256) rsq_cut_kernel(RsqArgs a)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, k = blockIdx.z;
    if (c >= a.sx) return;
    const uint16_t *blk = a.raw + (size_t)k * (a.sin_y + 2) * a.sin_x;
    const double v = line_integral(a, blk, a.flip ? a.sin_y - 1 - r : r, a.axis_move + c);
    static_cast<T *>(a.out)[((size_t)k * a.sy + r) * a.sx + c] = (T)v;
}

// First pass of Half2FullFan (rsq_load.cpp:327-391): the line integrals of
// columns axis_move .. sin_x-1 (the only ones the rebinning reads) of
// sinograms k0 .. k0+gridDim.z-1 into lines [gridDim.z][sin_y][sin_x -
// axis_move], and the measured half of the output (columns from the central
// ray c0 on are these values of angle row f, written to output row r).
// Grid (column tiles, rows, sinograms).
template <typename T>
__global__ void __launch_bounds__(256) rsq_lines_kernel(RsqArgs a, int k0, double *lines)
{
    const int w = a.sin_x - a.axis_move, c0 = a.sin_x - 1 - a.axis_move;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, f = blockIdx.y, k = blockIdx.z;
    if (x >= w) return;
    const uint16_t *blk = a.raw + (size_t)(k0 + k) * (a.sin_y + 2) * a.sin_x;
    const double v = line_integral(a, blk, f, a.axis_move + x);
    lines[((size_t)k * a.sin_y + f) * w + x] = v;
    const int r = a.flip ? a.sy - 1 - f + a.start : f;
    if (r >= 0 && r < a.sy)
        static_cast<T *>(a.out)[((size_t)(k0 + k) * a.sy + r) * a.sx + c0 + x] = (T)v;
}

// Second pass of Half2FullFan: the columns left of the central ray, the
// redundant rays, interpolated between two angle rows of the line integrals.
// The angle rows and the weight depend on (r, c) only, so each thread finds
// them once and applies them to the nk sinograms of the group.  Grid (column
// tiles over 0 .. c0-1, output rows).
template <typename T>
__global__ void __launch_bounds__(256) rsq_rebin_kernel(RsqArgs a, int k0, int nk,
                                                        const double *lines)
{
    const int w = a.sin_x - a.axis_move, c0 = a.sin_x - 1 - a.axis_move;
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (c >= c0) return;
    const size_t lstride = (size_t)a.sin_y * w, ostride = (size_t)a.sy * a.sx;
    T *out = static_cast<T *>(a.out) + (size_t)k0 * ostride + (size_t)r * a.sx + c;
    const int f = a.flip ? a.sy - 1 - r + a.start : r;   // angle row of output row r
    const int t = c - c0;                        // < 0
    const double at = a.at[-t];                  // atan(t / distance), host table
    const double fi = f * a.delta_fi;
    double nf = (fi - 2 * at - kPi) / a.delta_fi;
    if (nf > a.sin_y - 1 || nf < 0) nf = (fi - 2 * at + kPi) / a.delta_fi;
    const double ca = ceil(nf), fb = floor(nf);
    int ia = (int)ca, ib = (int)fb;
    // inside the scan for delta_fi = 2 pi / sin_y (ia <= sin_y wraps to row 0
    // as in the reference); any other angle wraps too
    if ((unsigned)ia >= (unsigned)a.sin_y) ia = ((ia % a.sin_y) + a.sin_y) % a.sin_y;
    if ((unsigned)ib >= (unsigned)a.sin_y) ib = ((ib % a.sin_y) + a.sin_y) % a.sin_y;
    const double *LA = lines + (size_t)ia * w + (-t);   // column axis_move - t
    const double *LB = lines + (size_t)ib * w + (-t);
    const double wgt = -nf + ca;
    const bool two = ca != fb;
    for (int k = 0; k < nk; ++k) {
        const double A = LA[k * lstride];
        double v = A;
        if (two) {
            const double B = LB[k * lstride];
            v = A + (B - A) * wgt;
        }
        out[k * ostride] = (T)v;
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
    if ((size_t)a.sx * a.sy * nsin == 0) return FDK_OK;
    // grids are (column tiles, rows, sinograms)
    if (a.sin_y > 65535)
        return fail(FDK_ERR_INVALID_ARGUMENT, "rsq: more than 65535 projections per sinogram");
    hipStream_t st = (hipStream_t)stream;
    const bool f64 = out_dtype == FDK_F64;
    hipError_t e = hipSuccess;
    if (!a.complete) {
        a.at = nullptr;
        for (int k0 = 0; k0 < nsin && e == hipSuccess; k0 += 65535) {
            RsqArgs b = a;
            const int nk = std::min(nsin - k0, 65535);
            b.raw = a.raw + (size_t)k0 * (a.sin_y + 2) * a.sin_x;
            b.out = (char *)a.out + (size_t)k0 * a.sy * a.sx * (f64 ? 8 : 4);
            const dim3 grid((a.sx + 255) / 256, a.sy, nk);
            if (f64) hipLaunchKernelGGL(rsq_cut_kernel<double>, grid, dim3(256), 0, st, b);
            else hipLaunchKernelGGL(rsq_cut_kernel<float>, grid, dim3(256), 0, st, b);
            e = hipGetLastError();
        }
        if (e != hipSuccess) return tkv::hip_fail(e, "rsq");
        return FDK_OK;
    }
    // atan table, index -t for t = -1 .. -(sin_x - 1 - axis_move) (entry 0 unused)
    const int w = p->sin_x - p->axis_move;
    std::vector<double> at(w, 0.0);
    for (int t = -p->sin_x + 1 + p->axis_move; t < 0; ++t) at[-t] = std::atan(t / p->distance);
    const size_t per = (size_t)a.sin_y * w * sizeof(double);
    const int group = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nsin, kGroupBytes / per,
                                                                 (size_t)65535}));
    double *d_at = nullptr, *d_lines = nullptr;
    TKV_HIP(hipMalloc((void **)&d_at, w * sizeof(double)));
    e = hipMalloc((void **)&d_lines, per * group);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_at, at.data(), w * sizeof(double), hipMemcpyHostToDevice, st);
    a.at = d_at;
    const int c0 = p->sin_x - 1 - p->axis_move;
    for (int k0 = 0; k0 < nsin && e == hipSuccess; k0 += group) {
        const int nk = std::min(group, nsin - k0);
        const dim3 g1((w + 255) / 256, a.sin_y, nk), g2((c0 + 255) / 256, a.sy);
        if (f64) hipLaunchKernelGGL(rsq_lines_kernel<double>, g1, dim3(256), 0, st, a, k0, d_lines);
        else hipLaunchKernelGGL(rsq_lines_kernel<float>, g1, dim3(256), 0, st, a, k0, d_lines);
        if (c0 > 0) {
            if (f64) hipLaunchKernelGGL(rsq_rebin_kernel<double>, g2, dim3(256), 0, st, a, k0,
                                        nk, (const double *)d_lines);
            else hipLaunchKernelGGL(rsq_rebin_kernel<float>, g2, dim3(256), 0, st, a, k0, nk,
                                    (const double *)d_lines);
        }
        e = hipGetLastError();
    }
    // the table and the scratch live until the kernels are done
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_lines);
    (void)hipFree(d_at);
    if (e != hipSuccess) return tkv::hip_fail(e, "rsq");
    return FDK_OK;
}
