// TIFF export of reconstructed slices, device part: range and 16-bit scaling
// on the GPU, files written by the host writer of fdk_tiff_host.cpp.
//
// The scaling u = (unsigned short)(65535 * (x - min) / (max - min)) is
// evaluated in FP64 in that order (no FMA contraction), on the device and on
// the host alike, so both paths write the same bytes.  Values outside
// [min, max] clamp to 0 / 65535 and NaN gives 0 (the reference's conversion
// of such values to unsigned short is undefined).
#include "tomokvas/fdk.h"
#include "tomokvas/fdk_tiff.h"
#include "fdk_internal.hpp"
#include "fdk_tiff_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#pragma clang fp contract(off)

using tkv::fail;
using tkv::scale_u16;
using tkv::write_slices;
using tkv::write_slices_parallel;
using tkv::writer_threads;

namespace {

// float -> key whose unsigned order is the float order
__device__ inline unsigned order_key(float x)
{
    const unsigned b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

float from_key(unsigned k)
{
    const unsigned b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float x;
    std::memcpy(&x, &b, 4);
    return x;
}

__global__ void minmax_init_kernel(unsigned *keys)
{
    if (threadIdx.x == 0) {
        keys[0] = 0xffffffffu;   // running minimum
        keys[1] = 0u;            // running maximum
    }
}

// VEC: v is 16-byte aligned; the body reads float4 and the last n % 4
// values one by one
template <bool VEC>
__global__ void __launch_bounds__(256) minmax_kernel(const float *v, size_t n, unsigned *keys)
{
    unsigned lo = 0xffffffffu, hi = 0u;
    auto take = [&](float x) {
        if (x == x) {
            const unsigned k = order_key(x);
            lo = k < lo ? k : lo;
            hi = k > hi ? k : hi;
        }
    };
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    size_t done = 0;
    if (VEC) {
        const float4 *v4 = reinterpret_cast<const float4 *>(v);
        const size_t n4 = n / 4;
        for (size_t i = tid; i < n4; i += nth) {
            const float4 q = v4[i];
            take(q.x);
            take(q.y);
            take(q.z);
            take(q.w);
        }
        done = 4 * n4;
    }
    for (size_t i = done + tid; i < n; i += nth) take(v[i]);
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    __shared__ unsigned slo[4], shi[4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        slo[w] = lo;
        shi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; ++k) {
            lo = slo[k] < lo ? slo[k] : lo;
            hi = shi[k] > hi ? shi[k] : hi;
        }
        atomicMin(keys, lo);
        atomicMax(keys + 1, hi);
    }
}

// VEC: v is 16-byte and out 8-byte aligned; the body converts four values per
// step, the last n % 4 one by one
template <bool VEC>
__global__ void __launch_bounds__(256) to_u16_kernel(const float *v, size_t n, double vmin,
                                                     double vmax, uint16_t *out)
{
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    size_t done = 0;
    if (VEC) {
        const float4 *v4 = reinterpret_cast<const float4 *>(v);
        uint2 *o4 = reinterpret_cast<uint2 *>(out);
        const size_t n4 = n / 4;
        for (size_t i = tid; i < n4; i += nth) {
            const float4 q = v4[i];
            const unsigned a = scale_u16(q.x, vmin, vmax), b = scale_u16(q.y, vmin, vmax);
            const unsigned c = scale_u16(q.z, vmin, vmax), d = scale_u16(q.w, vmin, vmax);
            o4[i] = make_uint2(a | (b << 16), c | (d << 16));
        }
        done = 4 * n4;
    }
    for (size_t i = done + tid; i < n; i += nth) out[i] = scale_u16(v[i], vmin, vmax);
}

unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 8192); }

}  // namespace

extern "C" int fdk_volume_minmax(const float *d_vol, size_t n, double *minmax, void *stream)
{
    if (!d_vol || !minmax || n == 0) return fail(FDK_ERR_INVALID_ARGUMENT, "minmax: empty input");
    hipStream_t st = (hipStream_t)stream;
    unsigned *keys = nullptr;
    TKV_HIP(hipMalloc((void **)&keys, 2 * sizeof(unsigned)));
    unsigned hk[2] = {0, 0};
    hipLaunchKernelGGL(minmax_init_kernel, dim3(1), dim3(64), 0, st, keys);
    if (((uintptr_t)d_vol & 15) == 0)
        hipLaunchKernelGGL(minmax_kernel<true>, dim3(grid_for(n)), dim3(256), 0, st, d_vol, n,
                           keys);
    else
        hipLaunchKernelGGL(minmax_kernel<false>, dim3(grid_for(n)), dim3(256), 0, st, d_vol, n,
                           keys);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hk, keys, sizeof hk, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(keys);
    if (e != hipSuccess) return tkv::hip_fail(e, "minmax");
    if (hk[0] == 0xffffffffu && hk[1] == 0u)
        return fail(FDK_ERR_INVALID_ARGUMENT, "minmax: no value that is not NaN");
    minmax[0] = from_key(hk[0]);
    minmax[1] = from_key(hk[1]);
    return FDK_OK;
}

extern "C" int fdk_volume_to_u16(const float *d_vol, size_t n, double vmin, double vmax,
                                 uint16_t *d_out, void *stream)
{
    if (!d_vol || !d_out) return fail(FDK_ERR_INVALID_ARGUMENT, "to_u16: null buffer");
    if (n == 0) return FDK_OK;
    if (((uintptr_t)d_vol & 15) == 0 && ((uintptr_t)d_out & 7) == 0)
        hipLaunchKernelGGL(to_u16_kernel<true>, dim3(grid_for(n)), dim3(256), 0,
                           (hipStream_t)stream, d_vol, n, vmin, vmax, d_out);
    else
        hipLaunchKernelGGL(to_u16_kernel<false>, dim3(grid_for(n)), dim3(256), 0,
                           (hipStream_t)stream, d_vol, n, vmin, vmax, d_out);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

extern "C" int fdk_tiff_export_device(const float *d_vol, int x, int y, int z, const char *dir,
                                      int first_index, double *minmax_out)
{
    if (!d_vol || x < 1 || y < 1 || z < 1)
        return fail(FDK_ERR_INVALID_ARGUMENT, "tiff export: empty volume");
    const size_t plane = (size_t)x * y;
    double mm[2];
    int rc = fdk_volume_minmax(d_vol, plane * z, mm, nullptr);
    if (rc) return rc;
    if (minmax_out) {
        minmax_out[0] = mm[0];
        minmax_out[1] = mm[1];
    }
    // slabs of about 64 MiB of 16-bit slices: slab k is written on a thread
    // while slab k+1 is converted and copied
    const int slab = (int)std::max<size_t>(1, std::min<size_t>(z, (64u << 20) / (2 * plane)));
    uint16_t *d_u16 = nullptr, *h_u16[2] = {nullptr, nullptr};
    hipError_t e = hipMalloc((void **)&d_u16, (size_t)slab * plane * 2);
    for (int i = 0; i < 2 && e == hipSuccess; ++i)
        e = hipHostMalloc((void **)&h_u16[i], (size_t)slab * plane * 2, hipHostMallocDefault);
    std::thread writer;
    int wrc = FDK_OK;
    std::string werr;
    auto join = [&]() {
        if (writer.joinable()) writer.join();
    };
    for (int z0 = 0; z0 < z && e == hipSuccess && !rc; z0 += slab) {
        const int nz = std::min(slab, z - z0);
        uint16_t *h = h_u16[(z0 / slab) & 1];
        rc = fdk_volume_to_u16(d_vol + (size_t)z0 * plane, (size_t)nz * plane, mm[0], mm[1], d_u16,
                               nullptr);
        if (rc) break;
        join();   // the buffer written two slabs ago is free again
        if (wrc) break;
        e = hipMemcpy(h, d_u16, (size_t)nz * plane * 2, hipMemcpyDeviceToHost);
        if (e != hipSuccess) break;
        auto write = [&, h, nz, z0]() {
            wrc = write_slices_parallel(nz, writer_threads(nz), werr, [&](int, int k) {
                return write_slices(h + (size_t)k * plane, x, y, 1, dir, first_index + z0 + k);
            });
        };
        try {
            writer = std::thread(write);
        } catch (const std::system_error &) {
            write();   // no thread available: write now
        }
    }
    join();
    (void)hipFree(d_u16);
    for (auto p : h_u16)
        if (p) (void)hipHostFree(p);
    if (rc) return rc;
    if (e != hipSuccess) return tkv::hip_fail(e, "tiff export");
    if (wrc) return fail(wrc, werr);
    return FDK_OK;
}
