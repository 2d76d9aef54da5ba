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
// run); the splats are hardware float/double atomics in L2.
#include "fdk_internal.hpp"

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
    long long nb = (long long)a.tiles_xy * ((L.Nz + ZR - 1) / ZR);
    if (nb == 0 || L.nang == 0) return FDK_OK;
    if (nb > 0x7fffffffLL) return fail(FDK_ERR_UNSUPPORTED, "forward_project: volume too large");
    hipLaunchKernelGGL((fp_kernel<R, ZR>), dim3((unsigned)nb), dim3(256), 0, st, a);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

int launch_forward_project(int real_bytes, const FPLaunch &a, hipStream_t st)
{
    return real_bytes == 8 ? launch_fp_t<double>(a, st) : launch_fp_t<float>(a, st);
}

}  // namespace tkv
