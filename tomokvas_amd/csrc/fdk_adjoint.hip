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
// through LDS float atomics, with the lanes of a wave along z; the box is then
// added to the detector with one coalesced global atomic per touched sample
// and cleared, which cuts global atomics from four per voxel-update to about
// a tenth.  FP64 (fp_kernel)
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
    float fx_bound;              // FP32: see FPLaunch::fx_bound
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

constexpr int FP_ZR = 64;                       // slices per workgroup: one per lane
constexpr int FP_BOX_H = box_rows<FP_ZR>();
constexpr int FP_P = BOX_W + 1;                 // odd pitch: the rows of one column
                                                // fall in distinct LDS banks

// Lanes run along z: in each splat instruction the 64 lanes of a wave are the
// 64 slices of one (x, y) column, so their taps share the column and sit on
// distinct detector rows.  The per-column work (S, T, U, u, weights) is done
// once per angle by one thread per column and handed over through LDS.
//
// The box accumulates in 32-bit fixed point with integer LDS atomics: float
// LDS atomics (ds_add_f32) measured 20x slower than ds_add_u32 or plain stores
// on gfx950 (918 against 45 ms for 50 projections of a 1024^3 volume).  The
// workgroup's scale fxs is a power of two with fxs * fx_bound * max|x| <= 2^30,
// so no sample of the box can overflow; each splat is rounded to the nearest
// multiple of 1/fxs, i.e. to within 2^-31 of the box's bound (the largest single
// splat keeps about 21 significant bits), and the sums are exact.  The flush
// converts back to float.
__global__ void __launch_bounds__(256) fp_lds_kernel(FPArgs<float> a)
{
    __shared__ int box[FP_P * FP_BOX_H];  // fixed point, units of 1/fxs
    __shared__ float4 colrec[256];        // per column: U, box byte column (-1: none), w0, w1 (x fxs)
    __shared__ unsigned wmax[4];
    __shared__ int4 boxes[64];
    __shared__ unsigned char colin[256];  // column inside the reconstruction cylinder
    const int lb = blockIdx.x;
    const int zr = lb / a.tiles_xy;
    const int tile = lb - zr * a.tiles_xy;
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int z0 = zr * FP_ZR;

    // column role: thread t owns column (t & 15, t >> 4) of the 16 x 16 tile
    const int x = tx * 16 + (threadIdx.x & 15), y = ty * 16 + (threadIdx.x >> 4);
    const bool valid = x < a.Nx && y < a.Ny;
    const float aX = valid ? a.ax[x] : 0.f, aY = valid ? a.ay[y] : 0.f;
    const bool inside =
        valid && ((double)aX * (double)aX + (double)aY * (double)aY < a.maskR2);
    colin[threadIdx.x] = inside ? 1 : 0;
    __syncthreads();

    // slice role: lane = slice z0 + lane; wave w walks columns 64w .. 64w+63
    const int z = z0 + lane;
    const float azl = a.az[z];            // the z table is padded past Nz
    const size_t plane = (size_t)a.Nx * a.Ny;
    float val[64];
    bool any = false;
#pragma unroll
    for (int j = 0; j < 64; ++j) {
        const int c = 64 * wave + j;
        const int xc = tx * 16 + (c & 15), yc = ty * 16 + (c >> 4);
        val[j] = colin[c] && z < a.Nz
                     ? a.vol[(size_t)z * plane + (size_t)yc * a.Nx + xc] * a.out_scale
                     : 0.f;
        any = any || val[j] != 0.f;
    }
    // the barriers below are workgroup-wide: skip only whole workgroups
    if (!__syncthreads_or(any)) return;

    // the workgroup's largest |x| sets the fixed-point scale
    float m = 0.f;
#pragma unroll
    for (int j = 0; j < 64; ++j) m = fmaxf(m, fabsf(val[j]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (lane == 0) wmax[wave] = __float_as_uint(m);
    for (int i = threadIdx.x; i < FP_P * FP_BOX_H; i += 256) box[i] = 0;
    __syncthreads();
    const float gmax = fmaxf(fmaxf(__uint_as_float(wmax[0]), __uint_as_float(wmax[1])),
                             fmaxf(__uint_as_float(wmax[2]), __uint_as_float(wmax[3])));
    // fixed-point scale fxs = 2^(30 - ceil(log2(bound))): fxs * bound <= 2^30
    // (exponent capped so fxs stays a finite float; a smaller scale only
    // coarsens the rounding)
    const float fxs = exp2f(fminf(30.f - ceilf(log2f(a.fx_bound * gmax)), 126.f));
    const float fxinv = 1.f / fxs;
    const float az0 = a.az[z0], az1 = a.az[z0 + FP_ZR - 1];
    const float azlo = fminf(az0, az1), azhi = fmaxf(az0, az1);
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
        const int4 bx = boxes[b & 63];
        const int bw = __builtin_amdgcn_readfirstlane(bx.z);
        if (bw == 0) continue;   // the block misses the detector at this angle
        const int bu0 = __builtin_amdgcn_readfirstlane(bx.x);
        const int bv0 = __builtin_amdgcn_readfirstlane(bx.y);
        const int bh = __builtin_amdgcn_readfirstlane(bx.w);

        // column role: the reference's S, T, U, u and the two column weights
        const float c = a.ang[2 * b], s = a.ang[2 * b + 1];
        const float S = fmaf(aX, c, aY * s);
        const float T = fmaf(aY, c, -(aX * s));
        const float U = a.D / (a.D - S);
        const float u = fmaf(T, U, a.cX);
        const bool act = inside && u >= 0.f && u <= a.umax;
        const float fl = floorf(u);
        const int iu = (int)fl;
        const float fu = u - fl;
        const float W = U * U;
        const float w1 = W * fu, w0 = W - W * fu;
        // box byte column (LDS splats) or detector column (HBM splats)
        const int cb = !act ? -1 : bw > 0 ? (iu - bu0) * 4 : iu;
        // lane t of wave w holds column t here and splats columns 64w..64w+63
        // below: whether all of them are active is known to the wave
        const bool wave_active = __all(act);
        __syncthreads();   // every wave is done with the previous angle's records
        // LDS splats take the weights in units of 1/fxs, HBM splats as they are
        const float ws = bw > 0 ? fxs : 1.f;
        colrec[threadIdx.x] = make_float4(U, __int_as_float(cb), w0 * ws, w1 * ws);
        __syncthreads();

        if (bw > 0) {
            // splat into the LDS box; taps with zero weight past the
            // detector's last row or column land inside the box and are
            // dropped at the flush
            const bool inner = bw & 0x10000;
            char *lds = (char *)box;
            const int rowb = -bv0 * (4 * FP_P);
            // The next column's record is read before this column's splats, and
            // the column test stays a lane predicate (the record is uniform over
            // the wave): read into a scalar register, it made every column wait
            // for its record.  Zero voxels splat zeros rather than carry a
            // per-column mask (64 masks spilled): 44 -> 37 ms for the adjoint of
            // a 256^3 crop of 1024^3 over 1200 projections of 1536^2.
            float4 rn = colrec[64 * wave];
            if (inner && wave_active) {
                // every column of the wave active and every slice's taps on
                // the detector (the box covers the whole z-run, padding slices
                // included): no lane tests, so no branch splits the waits
                // (33.8 against 36.7 ms for the crop adjoint above)
#pragma unroll
                for (int j = 0; j < 64; ++j) {
                    const float4 r = rn;
                    if (j + 1 < 64) rn = colrec[64 * wave + j + 1];
                    const float v = fmaf(azl, r.x, a.cZ);
                    const float g = val[j];
                    const float fv = __builtin_amdgcn_fractf(v);
                    int *p0 = (int *)(lds + (int)v * (4 * FP_P) + rowb + __float_as_int(r.y));
                    const float x1 = g * fv, x0 = g - x1;
                    atomicAdd(p0, __float2int_rn(r.z * x0));
                    atomicAdd(p0 + 1, __float2int_rn(r.w * x0));
                    atomicAdd(p0 + FP_P, __float2int_rn(r.z * x1));
                    atomicAdd(p0 + FP_P + 1, __float2int_rn(r.w * x1));
                }
            } else {
#pragma unroll
                for (int j = 0; j < 64; ++j) {
                    const float4 r = rn;
                    if (j + 1 < 64) rn = colrec[64 * wave + j + 1];
                    const int cbj = __float_as_int(r.y);
                    const float v = fmaf(azl, r.x, a.cZ);
                    const float g = val[j];
                    // cbj < 0: column inactive at this angle
                    if (cbj < 0 || (!inner && !(v >= 0.f && v <= a.vmax))) continue;
                    const float fv = __builtin_amdgcn_fractf(v);
                    int *p0 = (int *)(lds + (int)v * (4 * FP_P) + rowb + cbj);
                    const float x1 = g * fv, x0 = g - x1;
                    // rounded to nearest: zero-mean errors, which partly cancel
                    // in the sums (truncation biased every sample toward zero)
                    atomicAdd(p0, __float2int_rn(r.z * x0));
                    atomicAdd(p0 + 1, __float2int_rn(r.w * x0));
                    atomicAdd(p0 + FP_P, __float2int_rn(r.z * x1));
                    atomicAdd(p0 + FP_P + 1, __float2int_rn(r.w * x1));
                }
            }
            __syncthreads();
            // flush: one global atomic per touched sample, rows of 64 lanes
            for (int rr = threadIdx.x >> 6; rr <= bh; rr += 4) {
                int *cell = box + rr * FP_P + lane;
                const int t = *cell;
                if (t != 0) {
                    *cell = 0;
                    const int vr = bv0 + rr, uc = bu0 + lane;
                    if (vr < a.Nv && uc < a.Nu)
                        splat(a.sino + (size_t)vr * rstride + (size_t)b * a.Nu + uc,
                              (float)t * fxinv);
                }
            }
        } else {
            // box beyond the LDS tile: splat straight into HBM
#pragma unroll
            for (int j = 0; j < 64; ++j) {
                const float4 r = colrec[64 * wave + j];
                const int iuj = __builtin_amdgcn_readfirstlane(__float_as_int(r.y));
                if (iuj < 0) continue;
                const float v = fmaf(azl, r.x, a.cZ);
                const float g = val[j];
                if (g == 0.f || !(v >= 0.f && v <= a.vmax)) continue;
                const float fv0 = floorf(v);
                const int iv = (int)fv0;
                const float fv = v - fv0;
                const float x1 = g * fv, x0 = g - x1;
                float *r0 = a.sino + (size_t)b * a.Nu + iuj + (size_t)iv * rstride;
                const bool has_u1 = iuj + 1 < a.Nu;
                splat(r0, r.z * x0);
                if (has_u1) splat(r0 + 1, r.w * x0);
                if (iv + 1 < a.Nv) {
                    splat(r0 + rstride, r.z * x1);
                    if (has_u1) splat(r0 + rstride + 1, r.w * x1);
                }
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
    a.fx_bound = (float)L.fx_bound;
    a.tiles_x = (L.Nx + 15) / 16;
    a.tiles_xy = a.tiles_x * ((L.Ny + 15) / 16);
    const int zrun = sizeof(R) == 4 ? FP_ZR : ZR;
    long long nb = (long long)a.tiles_xy * ((L.Nz + zrun - 1) / zrun);
    if (nb == 0 || L.nang == 0) return FDK_OK;
    if (!grid_fits(nb, 256)) return fail(FDK_ERR_UNSUPPORTED, "forward_project: volume too large");
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
