// fdk_backproject.hip -- voxel-driven cone-beam backprojection for gfx950.
//
// Reference: FBPConeKernel + fQ3D (FBPG/kernel.cu:68-177): per voxel
//   a = scale*(idx - centre) (sign flips), inside iff |(aX,aY)| < Nu/2,
//   sum over beta of U^2 * bilinear(Q_beta; T*U + cX, aZ*U + cZ),
//   S = aX cos + aY sin, T = aY cos - aX sin, U = D/(D - S), times delta_f/p;
//   samples outside [0,Nu-1] x [0,Nv-1] contribute 0.
//
// MI355X design:
//  * a thread owns one (x,y) column and a run of ZR consecutive slices kept in
//    registers: S, T, U and the detector column u = T*U + cX depend only on
//    (x, y, beta), so the per-angle geometry is paid once per ZR updates;
//  * per update: v = aZ*U + cZ, floor/fract, one address, ONE load of two
//    adjacent cells = the whole 2x2 bilinear neighbourhood, 6 FMA-class ops;
//  * a wave covers an 8x8 column tile (a compact detector footprint per
//    angle), a workgroup 16x16 columns; all lanes walk the angles in lockstep,
//    so the per-angle cos/sin/scale and the base address are scalar (SGPR);
//  * z-runs that project entirely inside the detector for every lane of the
//    wave (the common case) skip the per-slice range test;
//  * workgroup ids are remapped so the workgroups an XCD runs together cover a
//    compact block of the volume and share detector rows in that XCD's L2;
//  * FP16 cells feed v_fma_mix_f32 directly: FP16 taps, FP32 weights and FP32
//    accumulators, half the gather bytes of FP32 cells.
#include "fdk_internal.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

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

// ---- tap access ------------------------------------------------------------------
// load(): the two adjacent cells at byte offset `off` (the 2x2 neighbourhood);
// lerp(): c0 = (1-fv) Q(v,u) + fv Q(v+1,u), c1 likewise at u+1.  Loads of a whole
// z-run are issued before any is consumed, so each wave keeps ZR gathers in
// flight.
template <typename Cell>
struct Taps;

template <>
struct Taps<float2> {
    using R = float;
    using L = float4;
    static constexpr int BYTES = 8;
    static __device__ __forceinline__ L load(const char *base, uint32_t off)
    {
        L q;
        __builtin_memcpy(&q, base + off, sizeof(L));
        return q;
    }
    static __device__ __forceinline__ void lerp(const L &q, float fv, float gv, float &c0,
                                                float &c1)
    {
        c0 = fmaf(fv, q.y - q.x, q.x);
        c1 = fmaf(fv, q.w - q.z, q.z);
    }
};

template <>
struct Taps<double2> {
    using R = double;
    using L = double4;
    static constexpr int BYTES = 16;
    static __device__ __forceinline__ L load(const char *base, uint32_t off)
    {
        L q;
        __builtin_memcpy(&q, base + off, sizeof(L));
        return q;
    }
    static __device__ __forceinline__ void lerp(const L &q, double fv, double gv, double &c0,
                                                double &c1)
    {
        c0 = fma(fv, q.y - q.x, q.x);
        c1 = fma(fv, q.w - q.z, q.z);
    }
};

struct half2u {
    unsigned v;
};
struct bf162u {
    unsigned v;
};

template <>
struct Taps<half2u> {
    using R = float;
    using L = uint2;
    static constexpr int BYTES = 4;
    static __device__ __forceinline__ L load(const char *base, uint32_t off)
    {
        L q;
        __builtin_memcpy(&q, base + off, sizeof(L));
        return q;
    }
    // every FP16 operand sits in an FMA with FP32 weights: v_fma_mix_f32,
    // no separate conversions
    static __device__ __forceinline__ void lerp(const L &q, float fv, float gv, float &c0,
                                                float &c1)
    {
        __half2 h0 = *reinterpret_cast<const __half2 *>(&q.x);
        __half2 h1 = *reinterpret_cast<const __half2 *>(&q.y);
        c0 = fmaf(fv, __high2float(h0), fmaf(gv, __low2float(h0), 0.0f));
        c1 = fmaf(fv, __high2float(h1), fmaf(gv, __low2float(h1), 0.0f));
    }
};

template <>
struct Taps<bf162u> {
    using R = float;
    using L = uint2;
    static constexpr int BYTES = 4;
    static __device__ __forceinline__ L load(const char *base, uint32_t off)
    {
        L q;
        __builtin_memcpy(&q, base + off, sizeof(L));
        return q;
    }
    static __device__ __forceinline__ void lerp(const L &q, float fv, float gv, float &c0,
                                                float &c1)
    {
        float a00 = __uint_as_float(q.x << 16), a10 = __uint_as_float(q.x & 0xffff0000u);
        float a01 = __uint_as_float(q.y << 16), a11 = __uint_as_float(q.y & 0xffff0000u);
        c0 = fmaf(fv, a10 - a00, a00);
        c1 = fmaf(fv, a11 - a01, a01);
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

template <typename Cell, int ZR, int LB>
__global__ void __launch_bounds__(256)
    bp_kernel(BPArgs<typename Taps<Cell>::R> a)
{
    using R = typename Taps<Cell>::R;
    using TP = Taps<Cell>;

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
            R S = aX * c + aY * s;
            R T = aY * c - aX * s;
            R U = a.D / (a.D - S);
            R u = T * U + a.cX;
            const bool act = inside && u >= (R)0 && u <= a.umax;
            R fl = rfloor(u);
            int iu = act ? (int)fl : 0;
            R fu = u - fl;
            R W = U * U * sc;
            R wu1 = act ? W * fu : (R)0;
            R wu0 = act ? W - W * fu : (R)0;
            const R Uv = act ? U : (R)0;
            const R cz = act ? a.cZ : (R)0;
            const uint32_t colb = (uint32_t)iu * (uint32_t)TP::BYTES;
            const R v_lo = rfma(az[0], Uv, cz), v_hi = rfma(az[ZR - 1], Uv, cz);
            const R vmin = v_lo < v_hi ? v_lo : v_hi, vmax = v_lo < v_hi ? v_hi : v_lo;
            // (a 0.01-row margin keeps float rounding of the slice table away
            // from the unchecked path's row bounds)
            if (__all(vmin >= (R)0.01 && vmax <= a.vmax - (R)0.01)) {
                // whole run inside the detector for every lane: no range tests
#pragma unroll
                for (int k0 = 0; k0 < ZR; k0 += LB) {
                    uint32_t off[LB];
                    R fv[LB];
#pragma unroll
                    for (int k = 0; k < LB; ++k) {
                        R v = rfma(az[k0 + k], Uv, cz);
                        R fl = rfloor(v);
                        fv[k] = v - fl;
                        off[k] = __umul24((uint32_t)(int)fl, a.row_bytes) + colb;
                    }
                    typename TP::L q[LB];
#pragma unroll
                    for (int k = 0; k < LB; ++k) q[k] = TP::load(proj, off[k]);
#pragma unroll
                    for (int k = 0; k < LB; ++k) {
                        R c0, c1;
                        TP::lerp(q[k], fv[k], (R)1 - fv[k], c0, c1);
                        acc[k0 + k] = rfma(wu1, c1, rfma(wu0, c0, acc[k0 + k]));
                    }
                }
            } else {
#pragma unroll
                for (int k0 = 0; k0 < ZR; k0 += LB) {
                    uint32_t off[LB];
                    R fv[LB];
                    bool ok[LB];
#pragma unroll
                    for (int k = 0; k < LB; ++k) {
                        R v = rfma(az[k0 + k], Uv, cz);
                        ok[k] = v >= (R)0 && v <= a.vmax;
                        R fl = rfloor(v);
                        fv[k] = v - fl;
                        off[k] = __umul24(ok[k] ? (uint32_t)(int)fl : 0u, a.row_bytes) + colb;
                    }
                    typename TP::L q[LB];
#pragma unroll
                    for (int k = 0; k < LB; ++k) q[k] = TP::load(proj, off[k]);
#pragma unroll
                    for (int k = 0; k < LB; ++k) {
                        R c0, c1;
                        TP::lerp(q[k], fv[k], (R)1 - fv[k], c0, c1);
                        R w0 = ok[k] ? wu0 : (R)0, w1 = ok[k] ? wu1 : (R)0;
                        acc[k0 + k] = rfma(w1, c1, rfma(w0, c0, acc[k0 + k]));
                    }
                }
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

template <typename Cell>
constexpr int zrun_of()
{
    return sizeof(typename Taps<Cell>::R) == 8 ? 8 : 16;
}

int bp_zrun(int cell_dtype)
{
    switch (cell_dtype) {
    case FDK_F64: return zrun_of<double2>();
    case FDK_F16: return zrun_of<half2u>();
    case FDK_BF16: return zrun_of<bf162u>();
    default: return zrun_of<float2>();
    }
}

template <typename Cell>
static int launch_bp_t(const BPLaunch &L, hipStream_t st)
{
    using R = typename Taps<Cell>::R;
    constexpr int ZR = zrun_of<Cell>();
    BPArgs<R> a;
    a.cells = (const char *)L.cells;
    a.row_bytes = (uint32_t)((size_t)L.pitch * sizeof(Cell));
    a.angle_bytes = (size_t)a.row_bytes * L.Nv;
    if ((size_t)a.row_bytes * L.Nv > 0xffffffffull)
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
    hipLaunchKernelGGL((bp_kernel<Cell, ZR, (ZR < 8 ? ZR : 8)>), dim3((unsigned)nb), dim3(256), 0, st, a);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

int launch_backproject(int cell_dtype, const BPLaunch &a, hipStream_t st)
{
    switch (cell_dtype) {
    case FDK_F32: return launch_bp_t<float2>(a, st);
    case FDK_F16: return launch_bp_t<half2u>(a, st);
    case FDK_BF16: return launch_bp_t<bf162u>(a, st);
    case FDK_F64: return launch_bp_t<double2>(a, st);
    default: return fail(FDK_ERR_INVALID_ARGUMENT, "backproject: unknown cell type");
    }
}

}  // namespace tkv
