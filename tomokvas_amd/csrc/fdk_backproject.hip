// fdk_backproject.hip -- voxel-driven cone-beam backprojection for gfx950.
//
// Reference: FBPConeKernel + fQ3D (FBPG/kernel.cu:68-177): per voxel
//   a = scale*(idx - centre) (sign flips), inside iff |(aX,aY)| < Nu/2,
//   sum over beta of U^2 * bilinear(Q_beta; T*U + cX, aZ*U + cZ),
//   S = aX cos + aY sin, T = aY cos - aX sin, U = D/(D - S), times delta_f/p;
//   samples outside [0,Nu-1] x [0,Nv-1] contribute 0.
//
// MI355X design:
//  * filtered projections are stored as "quad cells": for every detector
//    sample (v,u) the four bilinear coefficients
//        A = Q(v,u)              B = Q(v+1,u) - Q(v,u)
//        C = Q(v,u+1) - Q(v,u)   D = Q(v+1,u+1) - Q(v+1,u) - Q(v,u+1) + Q(v,u)
//    so bilinear(v+fv, u+fu) = (A + fu C) + fv (B + fu D): ONE aligned 16-byte
//    (FP32) or 8-byte (FP16) load and three FMAs per voxel-update;
//  * a thread owns one (x,y) column and a run of ZR consecutive slices in
//    registers: S, T, U, u = T*U + cX, fu and U^2 depend only on (x, y, beta)
//    and are paid once per ZR updates;
//  * per update: v = aZ*U + cZ, row = trunc(v) (v >= 0 there), fv = fract(v),
//    one address, one load, 4 FMAs;
//  * a wave covers an 8x8 column tile (compact detector footprint), a
//    workgroup 16x16 columns; all lanes walk the angles in lockstep, so the
//    per-angle cos/sin/scale and the projection base address are scalar;
//  * z-runs that project entirely inside the detector for every lane of the
//    wave (the common case) skip the per-slice range test;
//  * workgroup ids are remapped so the workgroups an XCD runs together cover a
//    compact block of the volume and share detector rows in that XCD's L2;
//  * FP16 quads feed v_fma_mix_f32 directly: FP16 coefficients, FP32 weights
//    and FP32 accumulators, half the gather bytes of FP32 quads.
#include "fdk_internal.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

// Tuning knobs (defaults chosen by measurement on MI355X):
//   TKV_BP_ZR            slices per thread (register z-run) for FP32/FP16/BF16
//   TKV_BP_LB            gathers per lane issued together
//   TKV_BP_BATCH_BARRIER 1 = no load is scheduled ahead of the previous batch
#ifndef TKV_BP_ZR
#define TKV_BP_ZR 32
#endif
#ifndef TKV_BP_LB
#define TKV_BP_LB 8
#endif
#ifndef TKV_BP_BATCH_BARRIER
#define TKV_BP_BATCH_BARRIER 1
#endif

namespace tkv {

template <typename R>
struct BPArgs {
    const char *cells;
    size_t angle_bytes;           // bytes per projection in the cell stack
    uint32_t row_bytes;           // bytes per detector row of cells
    const R *ang;                 // (cos, sin) per angle
    const R *scales;              // per angle dequantisation factor
    const R *ax, *ay, *az;
    R *vol;
    int Nx, Ny, Nz, nang;
    R cX, cZ, D, umax, vmax, out_scale;
    double maskR2;
    int accumulate;
    int tiles_x, tiles_xy, nzrun, nblocks;
};

// ---- quad cells: load one, evaluate (A + fu C) + fv (B + fu D) ------------------------
struct quad_f32;
struct quad_f16;
struct quad_bf16;
struct quad_f64;

template <typename Q>
struct Quad;

template <>
struct Quad<quad_f32> {
    using R = float;
    using L = float4;   // (A, B, C, D)
    static constexpr int BYTES = 16;
    static __device__ __forceinline__ float eval(const L &q, float fu, float fv)
    {
        float t0 = fmaf(fu, q.z, q.x);
        float t1 = fmaf(fu, q.w, q.y);
        return fmaf(fv, t1, t0);
    }
};

template <>
struct Quad<quad_f64> {
    using R = double;
    using L = double4;
    static constexpr int BYTES = 32;
    static __device__ __forceinline__ double eval(const L &q, double fu, double fv)
    {
        double t0 = fma(fu, q.z, q.x);
        double t1 = fma(fu, q.w, q.y);
        return fma(fv, t1, t0);
    }
};

template <>
struct Quad<quad_f16> {
    using R = float;
    using L = uint2;    // x = (A, B) halves, y = (C, D) halves
    static constexpr int BYTES = 8;
    // FP16 operands fold into v_fma_mix_f32: FP32 arithmetic, no conversions
    static __device__ __forceinline__ float eval(const L &q, float fu, float fv)
    {
        __half2 ab = *reinterpret_cast<const __half2 *>(&q.x);
        __half2 cd = *reinterpret_cast<const __half2 *>(&q.y);
        float t0 = fmaf(fu, __low2float(cd), __low2float(ab));
        float t1 = fmaf(fu, __high2float(cd), __high2float(ab));
        return fmaf(fv, t1, t0);
    }
};

template <>
struct Quad<quad_bf16> {
    using R = float;
    using L = uint2;
    static constexpr int BYTES = 8;
    static __device__ __forceinline__ float eval(const L &q, float fu, float fv)
    {
        float A = __uint_as_float(q.x << 16), B = __uint_as_float(q.x & 0xffff0000u);
        float C = __uint_as_float(q.y << 16), D = __uint_as_float(q.y & 0xffff0000u);
        float t0 = fmaf(fu, C, A);
        float t1 = fmaf(fu, D, B);
        return fmaf(fv, t1, t0);
    }
};

template <typename R>
__device__ __forceinline__ R rfloor(R x);
template <>
__device__ __forceinline__ float rfloor(float x) { return floorf(x); }
template <>
__device__ __forceinline__ double rfloor(double x) { return floor(x); }
template <typename R>
__device__ __forceinline__ R rfma(R a, R b, R c);
template <>
__device__ __forceinline__ float rfma(float a, float b, float c) { return fmaf(a, b, c); }
template <>
__device__ __forceinline__ double rfma(double a, double b, double c) { return fma(a, b, c); }
// x - floor(x) in one instruction (v_fract); exact for the finite x used here
template <typename R>
__device__ __forceinline__ R rfract(R x);
template <>
__device__ __forceinline__ float rfract(float x) { return __builtin_amdgcn_fractf(x); }
template <>
__device__ __forceinline__ double rfract(double x) { return __builtin_amdgcn_fract(x); }

template <typename Q, typename L>
__device__ __forceinline__ L load_quad(const char *base, uint32_t off)
{
    return *reinterpret_cast<const L *>(base + off);
}

template <typename Q, int ZR, int LB>
__global__ void __launch_bounds__(256) bp_kernel(BPArgs<typename Quad<Q>::R> a)
{
    using R = typename Quad<Q>::R;
    using QT = Quad<Q>;
    using L = typename QT::L;

    // XCD-aware block order: hardware dispatches workgroup i to XCD i % 8; the
    // remap gives each XCD a contiguous range of logical blocks.
    const int nb = a.nblocks;
    const int bid = blockIdx.x;
    const int per = (nb + 7) / 8;
    const int xcd = bid & 7;
    const int full = nb - (per - 1) * 8;   // XCDs that receive `per` blocks
    int lb = (xcd < full) ? xcd * per + (bid >> 3)
                          : full * per + (xcd - full) * (per - 1) + (bid >> 3);
    const int zr = lb / a.tiles_xy;        // z-run major: neighbours share rows
    const int tile = lb - zr * a.tiles_xy;
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = tx * 16 + (wave & 1) * 8 + (lane & 7);
    const int y = ty * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int z0 = zr * ZR;
    const bool valid_xy = x < a.Nx && y < a.Ny;

    const R aX = valid_xy ? a.ax[x] : (R)0;
    const R aY = valid_xy ? a.ay[y] : (R)0;
    const bool inside =
        valid_xy && ((double)aX * (double)aX + (double)aY * (double)aY < a.maskR2);

    R acc[ZR];
#pragma unroll
    for (int k = 0; k < ZR; ++k) acc[k] = 0;

    if (__any(inside)) {
        R az[ZR];
#pragma unroll
        for (int k = 0; k < ZR; ++k) az[k] = a.az[z0 + k];
        const char *proj = a.cells;
        for (int b = 0; b < a.nang; ++b, proj += a.angle_bytes) {
            const R c = a.ang[2 * b], s = a.ang[2 * b + 1];
            const R sc = a.scales[b];
            // the reference's S, T, U and u with nvcc's contractions
            R S = rfma(aX, c, aY * s);
            R T = rfma(aY, c, -(aX * s));
            R U = a.D / (a.D - S);
            R u = rfma(T, U, a.cX);
            const bool act = inside && u >= (R)0 && u <= a.umax;
            const R fu = rfract(u);
            const int iu = act ? (int)u : 0;            // u >= 0: trunc = floor
            const R W = act ? U * U * sc : (R)0;
            const R Uv = act ? U : (R)0;
            const R cz = act ? a.cZ : (R)0;
            const uint32_t colb = (uint32_t)iu * (uint32_t)QT::BYTES;
            const R v_lo = rfma(az[0], Uv, cz), v_hi = rfma(az[ZR - 1], Uv, cz);
            const R vmin = v_lo < v_hi ? v_lo : v_hi, vmax = v_lo < v_hi ? v_hi : v_lo;
            // (a 0.01-row margin keeps float rounding of the slice table away
            // from the unchecked path's row bounds)
            const bool fast = __all(vmin >= (R)0.01 && vmax <= a.vmax - (R)0.01);
#pragma unroll
            for (int k0 = 0; k0 < ZR; k0 += LB) {
                R val[LB];
                if (fast) {
                    // whole run inside the detector for every lane: no range tests
                    uint32_t off[LB];
                    R fv[LB];
#pragma unroll
                    for (int k = 0; k < LB; ++k) {
                        R v = rfma(az[k0 + k], Uv, cz);
                        fv[k] = rfract(v);
                        off[k] = __umul24((uint32_t)(int)v, a.row_bytes) + colb;
                    }
                    L q[LB];
#pragma unroll
                    for (int k = 0; k < LB; ++k) q[k] = load_quad<Q, L>(proj, off[k]);
#pragma unroll
                    for (int k = 0; k < LB; ++k) val[k] = QT::eval(q[k], fu, fv[k]);
                } else {
                    uint32_t off[LB];
                    R fv[LB];
                    bool ok[LB];
#pragma unroll
                    for (int k = 0; k < LB; ++k) {
                        R v = rfma(az[k0 + k], Uv, cz);
                        ok[k] = v >= (R)0 && v <= a.vmax;
                        fv[k] = rfract(v);
                        off[k] = __umul24(ok[k] ? (uint32_t)(int)v : 0u, a.row_bytes) + colb;
                    }
                    L q[LB];
#pragma unroll
                    for (int k = 0; k < LB; ++k) q[k] = load_quad<Q, L>(proj, off[k]);
#pragma unroll
                    for (int k = 0; k < LB; ++k)
                        val[k] = ok[k] ? QT::eval(q[k], fu, fv[k]) : (R)0;
                }
#pragma unroll
                for (int k = 0; k < LB; ++k) acc[k0 + k] = rfma(W, val[k], acc[k0 + k]);
#if TKV_BP_BATCH_BARRIER
                // keep the next batch's loads behind this batch's arithmetic so
                // at most LB gathers per lane are in flight (bounds the VGPRs)
                __builtin_amdgcn_sched_barrier(0);
#endif
            }
        }
    }

    if (!valid_xy) return;
    const size_t plane = (size_t)a.Nx * a.Ny;
    R *out = a.vol + (size_t)y * a.Nx + x;
#pragma unroll
    for (int k = 0; k < ZR; ++k) {
        int z = z0 + k;
        if (z < a.Nz) {
            R r = inside ? acc[k] * a.out_scale : (R)0;
            if (a.accumulate)
                out[(size_t)z * plane] += r;
            else
                out[(size_t)z * plane] = r;
        }
    }
}

template <typename Q>
constexpr int zrun_of()
{
    return sizeof(typename Quad<Q>::R) == 8 ? 8 : TKV_BP_ZR;
}

int bp_zrun(int cell_dtype)
{
    switch (cell_dtype) {
    case FDK_F64: return zrun_of<quad_f64>();
    case FDK_F16: return zrun_of<quad_f16>();
    case FDK_BF16: return zrun_of<quad_bf16>();
    default: return zrun_of<quad_f32>();
    }
}

int cell_bytes_of(int cell_dtype)
{
    switch (cell_dtype) {
    case FDK_F32: return Quad<quad_f32>::BYTES;
    case FDK_F16: return Quad<quad_f16>::BYTES;
    case FDK_BF16: return Quad<quad_bf16>::BYTES;
    case FDK_F64: return Quad<quad_f64>::BYTES;
    default: return 0;
    }
}

template <typename Q>
static int launch_bp_t(const BPLaunch &L, hipStream_t st)
{
    using R = typename Quad<Q>::R;
    constexpr int ZR = zrun_of<Q>();
    BPArgs<R> a;
    a.cells = (const char *)L.cells;
    a.row_bytes = (uint32_t)((size_t)L.pitch * Quad<Q>::BYTES);
    a.angle_bytes = (size_t)a.row_bytes * L.Nv;
    if (a.angle_bytes > 0xffffffffull || a.row_bytes >= (1u << 24))
        return fail(FDK_ERR_UNSUPPORTED, "backproject: projection larger than 4 GiB");
    a.ang = (const R *)L.ang;
    a.scales = (const R *)L.scales;
    a.ax = (const R *)L.ax;
    a.ay = (const R *)L.ay;
    a.az = (const R *)L.az;
    a.vol = (R *)L.vol;
    a.Nx = L.Nx;
    a.Ny = L.Ny;
    a.Nz = L.Nz;
    a.nang = L.nang;
    a.cX = (R)L.cX;
    a.cZ = (R)L.cZ;
    a.D = (R)L.D;
    a.umax = (R)(L.Nu - 1);
    a.vmax = (R)(L.Nv - 1);
    a.out_scale = (R)L.out_scale;
    a.maskR2 = L.maskR2;
    a.accumulate = L.accumulate;
    a.tiles_x = (L.Nx + 15) / 16;
    a.tiles_xy = a.tiles_x * ((L.Ny + 15) / 16);
    a.nzrun = (L.Nz + ZR - 1) / ZR;
    long long nb = (long long)a.tiles_xy * a.nzrun;
    if (nb == 0) return FDK_OK;
    if (nb > 0x7fffffffLL) return fail(FDK_ERR_UNSUPPORTED, "backproject: volume too large");
    a.nblocks = (int)nb;
    hipLaunchKernelGGL((bp_kernel<Q, ZR, (ZR < TKV_BP_LB ? ZR : TKV_BP_LB)>), dim3((unsigned)nb),
                       dim3(256), 0, st, a);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

int launch_backproject(int cell_dtype, const BPLaunch &a, hipStream_t st)
{
    switch (cell_dtype) {
    case FDK_F32: return launch_bp_t<quad_f32>(a, st);
    case FDK_F16: return launch_bp_t<quad_f16>(a, st);
    case FDK_BF16: return launch_bp_t<quad_bf16>(a, st);
    case FDK_F64: return launch_bp_t<quad_f64>(a, st);
    default: return fail(FDK_ERR_INVALID_ARGUMENT, "backproject: unknown cell type");
    }
}

}  // namespace tkv
