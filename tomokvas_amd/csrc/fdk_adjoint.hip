// fdk_adjoint.hip -- adjoint of the backprojector: voxel-driven transposed-bilinear
// forward projection with atomic detector splats (gfx950).
//
// The backprojector (fdk_backproject.hip, reference FBPConeKernel/fQ3D,
// FBPG/kernel.cu:68-177) computes x = s * sum_beta U^2 * bilinear(Q_beta; u, v).
// Its transpose scatters every voxel value, weighted by s * U^2 and the same
// four bilinear weights, onto the detector samples that the voxel reads:
//   y(iv,   iu)   += s U^2 (1-fu)(1-fv) x      y(iv,   iu+1) += s U^2 fu (1-fv) x
//   y(iv+1, iu)   += s U^2 (1-fu) fv   x      y(iv+1, iu+1) += s U^2 fu fv     x
// with the same zero-outside-detector rule.  The thread layout matches the
// backprojector (one (x,y) column and a z-run per thread, u fixed along the
// run).
//
// FP32 (fp_lds_kernel): per angle, a workgroup's splats land in an LDS copy of
// its detector footprint box (the box the backprojector stages, fdk_box.hpp)
// through LDS float atomics; the box is then added to the detector with one
// coalesced global atomic per touched sample and cleared, which cuts global
// atomics from four per voxel-update to about a quarter.  FP64 (fp_kernel)
// splats straight into HBM with hardware double atomics.
#include "fdk_internal.hpp"
#include "fdk_box.hpp"

namespace tkv {

template <typename R>
struct FPArgs {
    const R *vol;
    R *sino;                     // [Nv][nang][Nu]
    const R *ang, *ax, *ay, *az;
    int Nu, Nv, nang, Nx, Ny, Nz;
    R cX, cZ, D, umax, vmax, out_scale;
    double maskR2;
    int tiles_x, tiles_xy;
};

template <typename R>
__device__ __forceinline__ R rfl(R x);
template <>
__device__ __forceinline__ float rfl(float x) { return floorf(x); }
template <>
__device__ __forceinline__ double rfl(double x) { return floor(x); }

template <typename R>
__device__ __forceinline__ R rfm(R a, R b, R c);
template <>
__device__ __forceinline__ float rfm(float a, float b, float c) { return fmaf(a, b, c); }
template <>
__device__ __forceinline__ double rfm(double a, double b, double c) { return fma(a, b, c); }

template <typename R>
__device__ __forceinline__ void splat(R *p, R v)
{
    unsafeAtomicAdd(p, v);
}

template <typename R, int ZR>
__global__ void __launch_bounds__(256) fp_kernel(FPArgs<R> a)
{
    const int lb = blockIdx.x;
    const int zr = lb / a.tiles_xy;
    const int tile = lb - zr * a.tiles_xy;
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = tx * 16 + (wave & 1) * 8 + (lane & 7);
    const int y = ty * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int z0 = zr * ZR;
    if (x >= a.Nx || y >= a.Ny) return;
    const R aX = a.ax[x], aY = a.ay[y];
    if (!((double)aX * (double)aX + (double)aY * (double)aY < a.maskR2)) return;

    R val[ZR], az[ZR];
    const size_t plane = (size_t)a.Nx * a.Ny;
    bool any = false;
#pragma unroll
    for (int k = 0; k < ZR; ++k) {
        int z = z0 + k;
        val[k] = z < a.Nz ? a.vol[(size_t)z * plane + (size_t)y * a.Nx + x] * a.out_scale : (R)0;
        az[k] = a.az[z0 + k];
        any = any || val[k] != (R)0;
    }
    if (!any) return;

    const size_t rstride = (size_t)a.nang * a.Nu;   // one detector row to the next
    for (int b = 0; b < a.nang; ++b) {
        const R c = a.ang[2 * b], s = a.ang[2 * b + 1];
        R S = rfm(aX, c, aY * s);
        R T = rfm(aY, c, -(aX * s));
        R U = a.D / (a.D - S);
        R u = rfm(T, U, a.cX);
        if (!(u >= (R)0 && u <= a.umax)) continue;
        R fl = rfl(u);
        const int iu = (int)fl;
        const R fu = u - fl;
        const R W = U * U;
        const R w1 = W * fu, w0 = W - W * fu;
        const bool has_u1 = iu + 1 < a.Nu;
        R *col = a.sino + (size_t)b * a.Nu + iu;
#pragma unroll
        for (int k = 0; k < ZR; ++k) {
            if (val[k] == (R)0) continue;
            R v = rfm(az[k], U, a.cZ);
            if (!(v >= (R)0 && v <= a.vmax)) continue;
            R fv0 = rfl(v);
            const int iv = (int)fv0;
            const R fv = v - fv0;
            const R x0 = val[k] * (1 - fv), x1 = val[k] * fv;
            R *r0 = col + (size_t)iv * rstride;
            splat(r0, w0 * x0);
            if (has_u1) splat(r0 + 1, w1 * x0);
            if (iv + 1 < a.Nv) {
                splat(r0 + rstride, w0 * x1);
                if (has_u1) splat(r0 + rstride + 1, w1 * x1);
            }
        }
    }
}

constexpr int FP_ZR = 32;                       // slices per thread (FP32)
constexpr int FP_BOX_H = box_rows<FP_ZR>();

__global__ void __launch_bounds__(256) fp_lds_kernel(FPArgs<float> a)
{
    __shared__ float box[BOX_W * FP_BOX_H];
    __shared__ int4 boxes[64];
    const int lb = blockIdx.x;
    const int zr = lb / a.tiles_xy;
    const int tile = lb - zr * a.tiles_xy;
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = tx * 16 + (wave & 1) * 8 + (lane & 7);
    const int y = ty * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int z0 = zr * FP_ZR;
    const bool valid = x < a.Nx && y < a.Ny;
    const float aX = valid ? a.ax[x] : 0.f, aY = valid ? a.ay[y] : 0.f;
    const bool inside =
        valid && ((double)aX * (double)aX + (double)aY * (double)aY < a.maskR2);

    float val[FP_ZR], az[FP_ZR];
    const size_t plane = (size_t)a.Nx * a.Ny;
    bool any = false;
#pragma unroll
    for (int k = 0; k < FP_ZR; ++k) {
        const int z = z0 + k;
        val[k] = inside && z < a.Nz
                     ? a.vol[(size_t)z * plane + (size_t)y * a.Nx + x] * a.out_scale
                     : 0.f;
        az[k] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(a.az[z0 + k])));
        any = any || val[k] != 0.f;
    }
    // the barriers below are workgroup-wide: skip only whole workgroups
    if (!__syncthreads_or(any)) return;
    const bool live = __any(any);   // this wave has something to splat

    for (int i = threadIdx.x; i < BOX_W * FP_BOX_H; i += 256) box[i] = 0.f;
    const float azlo = fminf(az[0], az[FP_ZR - 1]), azhi = fmaxf(az[0], az[FP_ZR - 1]);
    const int xa = tx * 16, xb = min(xa + 15, a.Nx - 1);
    const int ya = ty * 16, yb = min(ya + 15, a.Ny - 1);
    const float cx0 = a.ax[xa], cx1 = a.ax[xb], cy0 = a.ay[ya], cy1 = a.ay[yb];
    const size_t rstride = (size_t)a.nang * a.Nu;   // one detector row to the next

    for (int b = 0; b < a.nang; ++b) {
        if ((b & 63) == 0) {
            __syncthreads();
            if (wave == 0 && b + lane < a.nang)
                boxes[lane] = footprint(a.ang[2 * (b + lane)], a.ang[2 * (b + lane) + 1], cx0,
                                        cx1, cy0, cy1, azlo, azhi, a.cX, a.cZ, a.D, a.Nu, a.Nv,
                                        FP_BOX_H);
            __syncthreads();
        }
        const int bw = __builtin_amdgcn_readfirstlane(boxes[b & 63].z);
        if (bw == 0) continue;   // the block misses the detector at this angle
        const float c = a.ang[2 * b], s = a.ang[2 * b + 1];
        const float S = fmaf(aX, c, aY * s);
        const float T = fmaf(aY, c, -(aX * s));
        const float U = a.D / (a.D - S);
        const float u = fmaf(T, U, a.cX);
        const bool act = any && u >= 0.f && u <= a.umax;
        const float fl = floorf(u);
        const int iu = (int)fl;
        const float fu = u - fl;
        const float W = U * U;
        const float w1 = W * fu, w0 = W - W * fu;

        if (bw < 0) {
            // box beyond the LDS tile: splat straight into HBM
            if (!act) continue;
            const bool has_u1 = iu + 1 < a.Nu;
            float *col = a.sino + (size_t)b * a.Nu + iu;
#pragma unroll
            for (int k = 0; k < FP_ZR; ++k) {
                if (val[k] == 0.f) continue;
                const float v = fmaf(az[k], U, a.cZ);
                if (!(v >= 0.f && v <= a.vmax)) continue;
                const float fv0 = floorf(v);
                const int iv = (int)fv0;
                const float fv = v - fv0;
                const float x0 = val[k] * (1 - fv), x1 = val[k] * fv;
                float *r0 = col + (size_t)iv * rstride;
                splat(r0, w0 * x0);
                if (has_u1) splat(r0 + 1, w1 * x0);
                if (iv + 1 < a.Nv) {
                    splat(r0 + rstride, w0 * x1);
                    if (has_u1) splat(r0 + rstride + 1, w1 * x1);
                }
            }
            continue;
        }

        // splat into the LDS box; taps with zero weight past the detector's
        // last row or column land inside the box and are dropped at the flush
        const int4 bx = boxes[b & 63];
        const int bu0 = __builtin_amdgcn_readfirstlane(bx.x);
        const int bv0 = __builtin_amdgcn_readfirstlane(bx.y);
        const int bh = __builtin_amdgcn_readfirstlane(bx.w);
        const bool inner = bw & 0x10000;
        if (live && act) {
            const int colb = (iu - bu0) * 4 - bv0 * (4 * BOX_W);
            char *lds = (char *)box;
#pragma unroll
            for (int k = 0; k < FP_ZR; ++k) {
                const float v = fmaf(az[k], U, a.cZ);
                if (!inner && !(v >= 0.f && v <= a.vmax)) continue;
                const float fv = __builtin_amdgcn_fractf(v);
                float *p0 = (float *)(lds + (int)v * (4 * BOX_W) + colb);
                const float x1 = val[k] * fv, x0 = val[k] - x1;
                atomicAdd(p0, w0 * x0);
                atomicAdd(p0 + 1, w1 * x0);
                atomicAdd(p0 + BOX_W, w0 * x1);
                atomicAdd(p0 + BOX_W + 1, w1 * x1);
            }
        }
        __syncthreads();
        // flush: one global atomic per touched sample, rows of 64 lanes
        for (int r = threadIdx.x >> 6; r <= bh; r += 4) {
            float *cell = box + r * BOX_W + lane;
            const float t = *cell;
            if (t != 0.f) {
                *cell = 0.f;
                const int vr = bv0 + r, uc = bu0 + lane;
                if (vr < a.Nv && uc < a.Nu)
                    splat(a.sino + (size_t)vr * rstride + (size_t)b * a.Nu + uc, t);
            }
        }
        __syncthreads();
    }
}

template <typename R>
static int launch_fp_t(const FPLaunch &L, hipStream_t st)
{
    constexpr int ZR = 8;
    FPArgs<R> a;
    a.vol = (const R *)L.vol;
    a.sino = (R *)L.sino;
    a.ang = (const R *)L.ang;
    a.ax = (const R *)L.ax;
    a.ay = (const R *)L.ay;
    a.az = (const R *)L.az;
    a.Nu = L.Nu;
    a.Nv = L.Nv;
    a.nang = L.nang;
    a.Nx = L.Nx;
    a.Ny = L.Ny;
    a.Nz = L.Nz;
    a.cX = (R)L.cX;
    a.cZ = (R)L.cZ;
    a.D = (R)L.D;
    a.umax = (R)(L.Nu - 1);
    a.vmax = (R)(L.Nv - 1);
    a.out_scale = (R)L.out_scale;
    a.maskR2 = L.maskR2;
    a.tiles_x = (L.Nx + 15) / 16;
    a.tiles_xy = a.tiles_x * ((L.Ny + 15) / 16);
    const int zrun = sizeof(R) == 4 ? FP_ZR : ZR;
    long long nb = (long long)a.tiles_xy * ((L.Nz + zrun - 1) / zrun);
    if (nb == 0 || L.nang == 0) return FDK_OK;
    if (nb > 0x7fffffffLL) return fail(FDK_ERR_UNSUPPORTED, "forward_project: volume too large");
    if constexpr (sizeof(R) == 4)
        hipLaunchKernelGGL(fp_lds_kernel, dim3((unsigned)nb), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((fp_kernel<R, ZR>), dim3((unsigned)nb), dim3(256), 0, st, a);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

int launch_forward_project(int real_bytes, const FPLaunch &a, hipStream_t st)
{
    return real_bytes == 8 ? launch_fp_t<double>(a, st) : launch_fp_t<float>(a, st);
}

}  // namespace tkv
