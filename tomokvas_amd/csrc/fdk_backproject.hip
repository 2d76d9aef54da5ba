// fdk_backproject.hip -- voxel-driven cone-beam backprojection for gfx950.
//
// Reference: FBPConeKernel + fQ3D (FBPG/kernel.cu:68-177): per voxel
//   a = scale*(idx - centre) (sign flips), inside iff |(aX,aY)| < Nu/2,
//   sum over beta of U^2 * bilinear(Q_beta; T*U + cX, aZ*U + cZ),
//   S = aX cos + aY sin, T = aY cos - aX sin, U = D/(D - S), times delta_f/p;
//   samples outside [0,Nu-1] x [0,Nv-1] contribute 0.
//
// MI355X design:
//  * a workgroup owns a block of 16x16 columns x ZR slices; a thread owns one
//    (x,y) column and its ZR slices in registers: S, T, U, u = T*U + cX and U^2
//    depend only on (x, y, beta) and are paid once per ZR updates;
//  * per angle, the block's detector footprint -- a box of samples that the
//    workgroup derives from its corner columns (u and U are linear-fractional
//    in (aX, aY), so their extremes over the block sit at the corners) -- is
//    staged from HBM into LDS with coalesced loads; every bilinear gather of
//    the angle then reads LDS, so L2/HBM carry about one stored sample per two
//    voxel-updates instead of four taps per update;
//  * the stored stack is plain and angle-major ([beta][v][pitch]) in the
//    storage type: FP32, FP16 with a per-projection scale, or BF16; FP16 and
//    BF16 halve the staging bytes.  FP32 and BF16 boxes hold FP32 samples.
//    FP16 boxes keep FP16 sample pairs (q[c], q[c+1]) per 32-bit word, which
//    halves the LDS bytes of a gather; the FP16 path (FDK_F16) widens them
//    inside mixed-precision FMAs (v_fma_mix_f32: FP16 samples, FP32 fraction,
//    differences and weights, FP32 result), so for every storage type the
//    interpolation weights, lerps and accumulation are FP32 and the FP16
//    volume is the FP32 arithmetic on the decoded samples.  FDK_F16_PACKED
//    (opt-in) instead lerps in packed FP16 and weights with v_dot2 (FP16
//    fraction and column weights): faster, outside that precision form;
//  * z-runs that project entirely inside the detector for every active lane of
//    a wave (the common case) skip the per-slice range test;
//  * workgroup ids are remapped so the workgroups an XCD runs together cover a
//    compact part of the volume and share detector rows in that XCD's L2;
//  * boxes larger than the LDS tile (extreme magnification, blocks near the
//    source) and the FP64 instantiation gather straight from HBM with the same
//    arithmetic.
#include "fdk_internal.hpp"
#include "fdk_box.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <type_traits>

// Tuning knobs (defaults chosen by measurement on MI355X, 1024^3 from
// 1800 x 2048^2):
//   TKV_BP_ZR_F32 / _F16 / _BF16 / _F16P  slices per thread (register z-run)
//                                  per storage type (FP64 uses 8)
//   TKV_BP_LB                      LDS gathers per lane issued together
//   TKV_BP_GROUP_FENCE             1: no instruction moves between groups of
//                                  LB gathers
//   TKV_BP_BOXP_F32 / _F16 / _BF16 LDS row pitch of the staged box in 32-bit
//                                  words (>= 64, multiple of 4): rows of one
//                                  column fall in different banks (the FP16
//                                  pitch serves FDK_F16 and FDK_F16_PACKED)
#ifndef TKV_BP_ZR_F32
#define TKV_BP_ZR_F32 24
#endif
#ifndef TKV_BP_ZR_F16
#define TKV_BP_ZR_F16 32
#endif
#ifndef TKV_BP_ZR_BF16
#define TKV_BP_ZR_BF16 32
#endif
#ifndef TKV_BP_ZR_F16P
#define TKV_BP_ZR_F16P 32
#endif
#ifndef TKV_BP_LB
#define TKV_BP_LB 4
#endif
// TKV_BP_GROUP_FENCE: the compiler schedules each group of LB gathers on its own
#ifndef TKV_BP_GROUP_FENCE
#define TKV_BP_GROUP_FENCE 1
#endif
#ifndef TKV_BP_BOXP_F16
#define TKV_BP_BOXP_F16 72
#endif
#ifndef TKV_BP_BOXP_F32
#define TKV_BP_BOXP_F32 80
#endif
#ifndef TKV_BP_BOXP_BF16
#define TKV_BP_BOXP_BF16 80
#endif

namespace tkv {

static_assert(box_rows<64>() <= STACK_TAIL_ROWS, "stored stack tail shorter than a box");

// Row pitch of the stored stack in samples: a multiple of 64 with at least
// BOX_W + 4 zero samples after column Nu-1, so a staged box row that starts
// inside the detector never leaves its own row.
int stack_pitch(int Nu) { return ((Nu + BOX_W + 4 + 63) / 64) * 64; }

template <typename R>
struct BPArgs {
    const char *stack;
    size_t angle_bytes;           // bytes per projection of the stack
    uint32_t row_bytes;           // bytes per detector row of the stack
    const R *ang;                 // (cos, sin) per angle
    const R *scales;              // per angle dequantisation factor
    const R *ax, *ay, *az;
    R *vol;
    int Nx, Ny, Nz, Nu, Nv, nang;
    R cX, cZ, D, umax, vmax, out_scale;
    float one;                    // 1.0 (a run-time value: see gather_mix)
    double maskR2;
    int accumulate;
    int tiles_x, tiles_xy, nblocks;
};

// ---- storage types ----------------------------------------------------------------
// PAIRS: the LDS box holds FP16 sample pairs; MIX: they are interpolated with
// FP32 fractions and weights in mixed-precision FMAs (else packed FP16)
struct st_f32;
struct st_f16;
struct st_bf16;
struct st_f16p;
struct st_f64;
template <typename S>
struct Store;
template <>
struct Store<st_f32> {
    using R = float;
    static constexpr int BYTES = 4;
    static constexpr bool STAGED = true;
    static constexpr int ZR = TKV_BP_ZR_F32;
    static constexpr int P = TKV_BP_BOXP_F32;
    static constexpr bool PAIRS = false;
    static constexpr bool MIX = false;
    static __device__ __forceinline__ float get(const char *p) { return *(const float *)p; }
    static __device__ __forceinline__ float half(unsigned, int) { return 0.f; }
};
template <>
struct Store<st_f16> {
    using R = float;
    static constexpr int BYTES = 2;
    static constexpr bool STAGED = true;
    static constexpr int ZR = TKV_BP_ZR_F16;
    static constexpr int P = TKV_BP_BOXP_F16;
    static constexpr bool PAIRS = true;
    static constexpr bool MIX = true;
    static __device__ __forceinline__ float get(const char *p)
    {
        return __half2float(__ushort_as_half(*(const unsigned short *)p));
    }
    static __device__ __forceinline__ float half(unsigned w, int hi)
    {
        return __half2float(__ushort_as_half((unsigned short)(hi ? w >> 16 : w & 0xffffu)));
    }
};
template <>
struct Store<st_bf16> {
    using R = float;
    static constexpr int BYTES = 2;
    static constexpr bool STAGED = true;
    static constexpr int ZR = TKV_BP_ZR_BF16;
    static constexpr int P = TKV_BP_BOXP_BF16;
    static constexpr bool PAIRS = false;
    static constexpr bool MIX = false;
    static __device__ __forceinline__ float get(const char *p)
    {
        return __uint_as_float((unsigned)(*(const unsigned short *)p) << 16);
    }
    static __device__ __forceinline__ float half(unsigned w, int hi)
    {
        return __uint_as_float(hi ? (w & 0xffff0000u) : (w << 16));
    }
};
template <>
struct Store<st_f16p> : Store<st_f16> {
    static constexpr int ZR = TKV_BP_ZR_F16P;
    static constexpr bool MIX = false;
};
template <>
struct Store<st_f64> {
    using R = double;
    static constexpr int BYTES = 8;
    static constexpr bool STAGED = false;
    static constexpr int ZR = 8;
    static constexpr int P = 64;   // unused: FP64 is gathered from HBM
    static constexpr bool PAIRS = false;
    static constexpr bool MIX = false;
    static __device__ __forceinline__ double get(const char *p) { return *(const double *)p; }
    static __device__ __forceinline__ float half(unsigned, int) { return 0.f; }
};

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

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

// ---- LDS gathers of one angle for LB slices of a thread's z-run ----------------------
// acc[k] += W * bilinear(box; u, v_k) with v_k = az[k]*Uv + cz; colb is the
// byte offset of the thread's column minus that of the box's first row.
// FAST: every v_k of every lane of the wave lies inside [0.01, Nv-1.01].
template <bool FAST, int LB, int P>
__device__ __forceinline__ void gather_lds(float *acc, const float *az, const float *box,
                                           float Uv, float cz, int colb, float wu0, float wu1,
                                           float vmax)
{
    const char *lds = (const char *)box;
    float c0[LB], c1[LB];
    bool okk[LB];
#pragma unroll
    for (int k = 0; k < LB; ++k) {
        const float v = fmaf(az[k], Uv, cz);
        const float fv = __builtin_amdgcn_fractf(v);
        bool ok = true;
        int off = (int)__umul24((unsigned)(int)v, 4u * P) + colb;   // v >= 0 where used
        if (!FAST) {
            ok = v >= 0.f && v <= vmax;
            off = ok ? off : 0;
        }
        const float *p0 = (const float *)(lds + off);
        const float q00 = p0[0], q01 = p0[1];
        const float q10 = p0[P], q11 = p0[P + 1];
        c0[k] = fmaf(fv, q10 - q00, q00);
        c1[k] = fmaf(fv, q11 - q01, q01);
        okk[k] = ok;
    }
    // taps outside the detector get zero weight (the staged values are finite)
#pragma unroll
    for (int k = 0; k < LB; ++k)
        acc[k] = fmaf(okk[k] ? wu1 : 0.f, c1[k], fmaf(okk[k] ? wu0 : 0.f, c0[k], acc[k]));
}

// FP16 pairs in LDS, FP32 interpolation (FDK_F16): word (r, c) holds samples
// c and c+1 of row r; one ds_read2 fetches both rows of the 2x2
// neighbourhood.  Each column's difference along v and its lerp are
// mixed-precision FMAs reading the FP16 halves directly (v_fma_mix_f32: the
// product is exact and the result rounds once to FP32, so d = q10 - q00 and
// c = q00 + fv*d equal the FP32 operations on the widened samples); the
// column weights (FP32, including the projection's scale) accumulate in FP32.
template <bool FAST, int LB, int P>
__device__ __forceinline__ void gather_mix(float *acc, const float *az, const float *box,
                                           float Uv, float cz, int colb, float wu0, float wu1,
                                           float vmax, float one)
{
    const char *lds = (const char *)box;
    float c0[LB], c1[LB];
    bool okk[LB];
#pragma unroll
    for (int k = 0; k < LB; ++k) {
        const float v = fmaf(az[k], Uv, cz);
        const float fv = __builtin_amdgcn_fractf(v);
        bool ok = true;
        int off = (int)__umul24((unsigned)(int)v, 4u * P) + colb;   // v >= 0 where used
        if (!FAST) {
            ok = v >= 0.f && v <= vmax;
            off = ok ? off : 0;
        }
        const half2_t w0 = *(const half2_t *)(lds + off);          // (q00, q01)
        const half2_t w1 = *(const half2_t *)(lds + off + 4 * P);  // (q10, q11)
        // d = q1*one - q0 (one = 1 at run time: an FMA of the widened halves,
        // which the compiler forms as v_fma_mix_f32), then c = q0 + fv*d
        const float d0 = fmaf((float)w1.x, one, -(float)w0.x);
        const float d1 = fmaf((float)w1.y, one, -(float)w0.y);
        c0[k] = fmaf(fv, d0, (float)w0.x);
        c1[k] = fmaf(fv, d1, (float)w0.y);
        okk[k] = ok;
    }
    // taps outside the detector get zero weight (the staged values are finite)
#pragma unroll
    for (int k = 0; k < LB; ++k)
        acc[k] = fmaf(okk[k] ? wu1 : 0.f, c1[k], fmaf(okk[k] ? wu0 : 0.f, c0[k], acc[k]));
}

// FP16 pairs in LDS, packed-FP16 interpolation (FDK_F16_PACKED): the lerp
// along v runs in packed FP16; the two column weights (FP16, normalised to at
// most U^2) and the two lerped columns meet in one v_dot2_f32_f16, summed in
// FP32.
template <bool FAST, int LB, int P>
__device__ __forceinline__ void gather_pairs(float *acc, const float *az, const float *box,
                                             float Uv, float cz, int colb, half2_t wuh,
                                             float vmax)
{
    static_assert(LB % 2 == 0, "fractions are converted in pairs");
    const char *lds = (const char *)box;
    half2_t c[LB];
    bool okk[LB];
#pragma unroll
    for (int k = 0; k < LB; k += 2) {
        const float va = fmaf(az[k], Uv, cz), vb = fmaf(az[k + 1], Uv, cz);
        // one conversion serves two slices; each lerp broadcasts its half
        const half2_t fab = __builtin_bit_cast(
            half2_t, __builtin_amdgcn_cvt_pkrtz(__builtin_amdgcn_fractf(va),
                                                __builtin_amdgcn_fractf(vb)));
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float v = j ? vb : va;
            bool ok = true;
            int off = (int)__umul24((unsigned)(int)v, 4u * P) + colb;   // v >= 0 where used
            if (!FAST) {
                ok = v >= 0.f && v <= vmax;
                off = ok ? off : 0;
            }
            const half2_t w0 = *(const half2_t *)(lds + off);
            const half2_t w1 = *(const half2_t *)(lds + off + 4 * P);
            const half2_t f = j ? half2_t{fab.y, fab.y} : half2_t{fab.x, fab.x};
            c[k + j] = f * (w1 - w0) + w0;
            okk[k + j] = ok;
        }
    }
    // the three-address VOP3P form (the builtin selects the two-address
    // v_dot2c form, whose tied accumulator costs register copies here)
#pragma unroll
    for (int k = 0; k < LB; ++k) {
        const half2_t w = okk[k] ? wuh : half2_t{0, 0};
        float r;
        asm("v_dot2_f32_f16 %0, %1, %2, %3" : "=v"(r) : "v"(c[k]), "v"(w), "v"(acc[k]));
        acc[k] = r;
    }
}

// all ZR slices of a thread for one staged angle, LB gathers at a time
template <bool FAST, bool PAIRS, bool MIX, int ZR, int LB, int P>
__device__ __forceinline__ void gather_run(float (&acc)[ZR], const float (&az)[ZR],
                                           const float *box, float Uv, float cz, int colb,
                                           float wu0, float wu1, half2_t wuh, float vmax,
                                           float one)
{
#pragma unroll
    for (int k0 = 0; k0 < ZR; k0 += LB) {
        if constexpr (PAIRS && MIX)
            gather_mix<FAST, LB, P>(acc + k0, az + k0, box, Uv, cz, colb, wu0, wu1, vmax, one);
        else if constexpr (PAIRS)
            gather_pairs<FAST, LB, P>(acc + k0, az + k0, box, Uv, cz, colb, wuh, vmax);
        else
            gather_lds<FAST, LB, P>(acc + k0, az + k0, box, Uv, cz, colb, wu0, wu1, vmax);
#if TKV_BP_GROUP_FENCE
        __builtin_amdgcn_sched_barrier(0);
#endif
    }
}

// Box staging in two steps, so the loads of the next angle's box fly while
// the current angle is gathered: issue() loads rows [bv0, bv0+bh] of a
// projection (the last is the +1 neighbour row; rows past the detector are
// zero) x columns [bu0, bu0+64) into registers, 4 samples per lane and row,
// 16 lanes per row (columns past the row end read the stack's zero padding);
// commit() writes them into the LDS box.  FP32 boxes hold the samples as
// floats; FP16 pair boxes hold words (q[c], q[c+1]), the 4th word of a lane
// taking the next lane's first sample (word 63 is never read).
template <typename S, int ROWS>
struct BoxStage {
    using ST = Store<S>;
    static constexpr int PER = (ROWS + 15) / 16;   // rows per thread
    using Raw = typename std::conditional<ST::BYTES == 4, float4, uint2>::type;
    Raw raw[PER];

    __device__ __forceinline__ void issue(const char *proj, int bu0, int bv0, int bh,
                                          uint32_t row_bytes, int Nv)
    {
        // buffer loads: a scalar descriptor at the box origin whose range ends
        // with the last row inside the detector, so rows past it read as zero
        // (raw buffer range check), and one 32-bit offset per row
        // (a projection is below 2 GiB, checked at launch: 32-bit offsets)
        const char *base = proj + ((uint32_t)bv0 * row_bytes + (uint32_t)bu0 * ST::BYTES);
        const int rows = min(bh, Nv - 1 - bv0);   // last row inside the detector
        const int nrec = (rows + 1) * (int)row_bytes - bu0 * ST::BYTES;
        const __amdgpu_buffer_rsrc_t rsrc =
            __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, nrec, 0x00020000);
        const uint32_t off0 = (threadIdx.x >> 4) * row_bytes + (threadIdx.x & 15) * 4 * ST::BYTES;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int off = (int)(off0 + (uint32_t)(16 * i) * row_bytes);
            if constexpr (ST::BYTES == 4) {
                const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, 0);
                raw[i] = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]),
                                     __uint_as_float(v[2]), __uint_as_float(v[3]));
            } else {
                const auto v = __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, 0);
                raw[i] = make_uint2(v[0], v[1]);
            }
        }
    }

    // rows are written in blocks of 16 (one per load) while the block starts
    // inside the box (bh uniform: a scalar branch); rows past the box in the
    // last block were loaded as zeros (outside the descriptor range), so the
    // tile stays finite
    __device__ __forceinline__ void commit(float *box, int bh) const
    {
        const int r0 = threadIdx.x >> 4, col = (threadIdx.x & 15) * 4;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            if (16 * i > bh) break;
            const int r = r0 + 16 * i;
            float *dst = box + r * ST::P + col;
            if constexpr (ST::PAIRS) {
                const uint2 w = raw[i];
                // the next lane's first two samples (DPP row shift by one
                // within the 16 lanes of a row; the last lane's 4th word, column
                // 63, is never read)
                const unsigned nx = (unsigned)__builtin_amdgcn_mov_dpp((int)w.x, 0x101, 0xf, 0xf,
                                                                       true);
                *(uint4 *)dst = make_uint4(w.x, __builtin_amdgcn_alignbyte(w.y, w.x, 2), w.y,
                                           __builtin_amdgcn_alignbyte(nx, w.y, 2));
            } else if constexpr (ST::BYTES == 4) {
                *(float4 *)dst = raw[i];
            } else {
                const uint2 w = raw[i];
                *(float4 *)dst = make_float4(ST::half(w.x, 0), ST::half(w.x, 1),
                                             ST::half(w.y, 0), ST::half(w.y, 1));
            }
        }
    }
};

// a workgroup-uniform value, kept in a scalar register
template <typename R>
__device__ __forceinline__ R uniform_load(const R *p) { return *p; }
template <>
__device__ __forceinline__ float uniform_load(const float *p)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(*p)));
}

// ---- the backprojector --------------------------------------------------------------
template <typename S, int ZR, int LB>
__global__ void __launch_bounds__(256) bp_kernel(BPArgs<typename Store<S>::R> a)
{
    using R = typename Store<S>::R;
    using ST = Store<S>;
    constexpr bool STAGED = ST::STAGED;
    constexpr int BOX_H = box_rows<ZR>();
    constexpr int P = ST::P;
    static_assert(P >= BOX_W && P % 4 == 0, "box pitch");
    __shared__ float boxmem[STAGED ? P * BOX_H : 1];
    // per group of 64 angles, in two slots (groups g and g+1): the footprints,
    // the group's angle order (inner angles from the front, edge angles from
    // the back) and the two counts; and whether any angle is left to pass 2
    __shared__ int4 boxes[STAGED ? 128 : 1];
    __shared__ unsigned char order[STAGED ? 128 : 1];
    __shared__ int counts[4];
    __shared__ int hbm_flag;

    // XCD-aware block order: hardware dispatches workgroup i to XCD i % 8; the
    // remap gives each XCD a contiguous range of logical blocks.
    const int nb = a.nblocks;
    const int bid = blockIdx.x;
    const int per = (nb + 7) / 8;
    const int xcd = bid & 7;
    const int full = nb - (per - 1) * 8;   // XCDs that receive `per` blocks
    const int lb = (xcd < full) ? xcd * per + (bid >> 3)
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

    // packed FP16 carries the column weights in FP16: they are normalised by
    // the largest projection scale (rounded up to a power of two, g), so they
    // stay within [0, U^2], and the sums are multiplied by g at the end
    constexpr bool PACKED = ST::PAIRS && !ST::MIX;
    R wnorm = 1, wback = 1;
    if constexpr (PACKED) {
        float m = 0.f;
        for (int i = lane; i < a.nang; i += 64) m = fmaxf(m, (float)a.scales[i]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        const float g = m > 0.f ? exp2f(ceilf(log2f(m))) : 1.f;
        wnorm = 1.f / g;
        wback = g;
    }

    // every wave of the workgroup joins the staging barriers, so the decision to
    // skip the angle loop is workgroup-wide
    if (__syncthreads_or(inside)) {
        R az[ZR];
#pragma unroll
        for (int k = 0; k < ZR; ++k) az[k] = uniform_load(a.az + z0 + k);
        const R azlo = az[0] < az[ZR - 1] ? az[0] : az[ZR - 1];
        const R azhi = az[0] < az[ZR - 1] ? az[ZR - 1] : az[0];
        // block corner coordinates
        const int xa = tx * 16, xb = min(xa + 15, a.Nx - 1);
        const int ya = ty * 16, yb = min(ya + 15, a.Ny - 1);
        const float cx0 = (float)a.ax[xa], cx1 = (float)a.ax[xb];
        const float cy0 = (float)a.ay[ya], cy1 = (float)a.ay[yb];

        // pass 1: angles whose footprint fits the LDS tile.  Wave 0 evaluates
        // a group's 64 footprints (one per lane) and sorts the staged angles
        // into inner ones (every tap row inside the detector: gathers without
        // range tests) and edge ones; each kind runs in a loop of its own, so
        // the accumulators never pass a branch merge (which costs a copy of
        // all of them per angle).  Angles missing the detector are dropped;
        // those whose box exceeds the tile are left to pass 2.
        bool need_hbm = !STAGED;
        if constexpr (STAGED) {
            auto boxes_of_group = [&](int g) {
                if (wave != 0) return;
                const int bb = g * 64 + lane, slot = g & 1;
                int4 q = make_int4(0, 0, 0, 0);
                if (bb < a.nang)
                    q = footprint(a.ang[2 * bb], a.ang[2 * bb + 1], cx0, cx1, cy0, cy1, azlo, azhi,
                                  a.cX, a.cZ, a.D, a.Nu, a.Nv, BOX_H);
                boxes[slot * 64 + lane] = q;
                const bool inner = q.z > 0 && (q.z & 0x10000);
                const bool edge = q.z > 0 && !(q.z & 0x10000);
                const uint64_t mi = __ballot(inner), me = __ballot(edge);
                const uint64_t mh = __ballot(q.z < 0);
                // rank among the lanes below of the same kind
                const int ri = __builtin_amdgcn_mbcnt_hi((unsigned)(mi >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)mi, 0));
                const int re = __builtin_amdgcn_mbcnt_hi((unsigned)(me >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)me, 0));
                if (inner) order[slot * 64 + ri] = (unsigned char)lane;
                if (edge) order[slot * 64 + 63 - re] = (unsigned char)lane;
                if (lane == 0) {
                    counts[2 * slot] = __popcll(mi);
                    counts[2 * slot + 1] = __popcll(me);
                    if (mh) hbm_flag = 1;
                }
            };
            auto uniform_box = [&](int slot, int i) {
                const int4 q = boxes[slot * 64 + i];
                return make_int4(__builtin_amdgcn_readfirstlane(q.x),
                                 __builtin_amdgcn_readfirstlane(q.y),
                                 __builtin_amdgcn_readfirstlane(q.z),
                                 __builtin_amdgcn_readfirstlane(q.w));
            };
            auto uniform_int = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
            // the tile starts finite: rows of a box past its last 16-row block
            // are never written
            for (int i = threadIdx.x; i < P * BOX_H; i += 256) boxmem[i] = 0.f;
            if (threadIdx.x == 0) hbm_flag = 0;
            __syncthreads();
            boxes_of_group(0);
            boxes_of_group(1);
            __syncthreads();
            BoxStage<S, BOX_H> stage;
            auto issue = [&](int b, int4 q) {
                stage.issue(a.stack + (size_t)b * a.angle_bytes, q.x, q.y, q.w, a.row_bytes, a.Nv);
            };
            // column weights and LDS offsets of one staged angle, then its gathers
            auto gather_angle = [&](auto fast_tag, int b, int4 bx) {
                constexpr bool FAST = decltype(fast_tag)::value;
                const R c = a.ang[2 * b], s = a.ang[2 * b + 1];
                // the reference's S, T, U and u with nvcc's contractions; the
                // packed-FP16 path, whose error is set by its FP16 weights,
                // takes U from the hardware reciprocal instead of the IEEE
                // quotient
                const R Sx = rfma(aX, c, aY * s);
                const R T = rfma(aY, c, -(aX * s));
                R U;
                if constexpr (PACKED)
                    U = a.D * __builtin_amdgcn_rcpf(a.D - Sx);
                else
                    U = a.D / (a.D - Sx);
                const R u = rfma(T, U, a.cX);
                const bool act = inside && u >= (R)0 && u <= a.umax;
                const R fu = rfract(u);
                const int iu = act ? (int)u : 0;            // u >= 0: trunc = floor
                const R W = act ? U * U * (a.scales[b] * wnorm) : (R)0;
                const R wu1 = W * fu, wu0 = W - wu1;
                half2_t wuh = {0, 0};
                if constexpr (PACKED) wuh = half2_t{(_Float16)wu0, (_Float16)wu1};
                // LDS byte offset of sample (row iv, column iu) is
                // iv*4*P + colb; inactive lanes sit at v = bv0, offset 0
                const int colb = (act ? (iu - bx.x) * 4 : 0) - bx.y * (4 * P);
                const R Uv = act ? U : (R)0;
                const R cz = act ? a.cZ : (R)bx.y;
                gather_run<FAST, ST::PAIRS, ST::MIX, ZR, LB, P>(acc, az, boxmem, Uv, cz, colb,
                                                                wu0, wu1, wuh, a.vmax, a.one);
            };
            using fast_t = std::integral_constant<bool, true>;
            using edge_t = std::integral_constant<bool, false>;
            const int ngroups = (a.nang + 63) >> 6;
            int pf = -1;   // angle whose box the stage registers are loading
            for (int g = 0; g < ngroups; ++g) {
                const int slot = g & 1, g0 = g * 64;
                if (g > 0) {
                    // the slot of group g-1 is free once every wave is here
                    __syncthreads();
                    boxes_of_group(g + 1);
                    __syncthreads();
                }
                // edge angles: loaded and gathered one at a time
                const int ne = uniform_int(counts[2 * slot + 1]);
                for (int j = 0; j < ne; ++j) {
                    const int i = uniform_int(order[slot * 64 + 63 - j]);
                    const int4 bx = uniform_box(slot, i);
                    issue(g0 + i, bx);
                    __syncthreads();   // every wave is done with the previous box
                    stage.commit(boxmem, bx.w);
                    __syncthreads();
                    gather_angle(edge_t{}, g0 + i, bx);
                }
                // inner angles: the next one's box is loaded while one is gathered
                const int ni = uniform_int(counts[2 * slot]);
                if (ni > 0) {
                    const int i = uniform_int(order[slot * 64]);
                    if (pf != g0 + i) issue(g0 + i, uniform_box(slot, i));
                }
                pf = -1;
                for (int j = 0; j < ni; ++j) {
                    const int i = uniform_int(order[slot * 64 + j]);
                    const int4 bx = uniform_box(slot, i);
                    __syncthreads();   // every wave is done with the previous box
                    stage.commit(boxmem, bx.w);
                    __syncthreads();
                    // the next inner angle: this group's, else the next group's
                    // first when that group has no edge angles (whose loads
                    // would take the stage registers)
                    int nb = -1;
                    int4 nq = make_int4(0, 0, 0, -1);   // none: an empty range, zeros
                    if (j + 1 < ni) {
                        const int k = uniform_int(order[slot * 64 + j + 1]);
                        nb = g0 + k;
                        nq = uniform_box(slot, k);
                    } else if (g + 1 < ngroups && uniform_int(counts[2 * (slot ^ 1) + 1]) == 0 &&
                               uniform_int(counts[2 * (slot ^ 1)]) > 0) {
                        const int k = uniform_int(order[(slot ^ 1) * 64]);
                        nb = g0 + 64 + k;
                        nq = uniform_box(slot ^ 1, k);
                    }
                    issue(nb >= 0 ? nb : 0, nq);
                    pf = nb;
                    gather_angle(fast_t{}, g0 + i, bx);
                }
            }
            need_hbm = uniform_int(hbm_flag) != 0;
        }

        // pass 2: angles gathered straight from HBM -- all of them for FP64,
        // and those whose box exceeds the LDS tile for the staged types
        if (need_hbm && __any(inside)) {
            const char *proj = a.stack;
            for (int b = 0; b < a.nang; ++b, proj += a.angle_bytes) {
                if constexpr (STAGED) {
                    const int4 bx = footprint(a.ang[2 * b], a.ang[2 * b + 1], cx0, cx1, cy0, cy1,
                                              azlo, azhi, a.cX, a.cZ, a.D, a.Nu, a.Nv, BOX_H);
                    if (bx.z >= 0) continue;
                }
                const R c = a.ang[2 * b], s = a.ang[2 * b + 1];
                const R Sx = rfma(aX, c, aY * s);
                const R T = rfma(aY, c, -(aX * s));
                const R U = a.D / (a.D - Sx);
                const R u = rfma(T, U, a.cX);
                const bool act = inside && u >= (R)0 && u <= a.umax;
                const R fu = rfract(u);
                const int iu = act ? (int)u : 0;
                const R W = act ? U * U * (a.scales[b] * wnorm) : (R)0;
                const R wu1 = W * fu, wu0 = W - wu1;
                const R Uv = act ? U : (R)0;
                const R cz = act ? a.cZ : (R)0;
                const char *colp = proj + (size_t)iu * ST::BYTES;
#pragma unroll
                for (int k = 0; k < ZR; ++k) {
                    const R v = rfma(az[k], Uv, cz);
                    const bool ok = act && v >= (R)0 && v <= a.vmax;
                    const int iv = ok ? (int)v : 0;
                    const R fv = rfract(v);
                    const char *p0 = colp + (size_t)iv * a.row_bytes;
                    // rows are zero-padded past Nu; row Nv-1 is only sampled
                    // with fv = 0, and its missing neighbour is replaced by itself
                    const int dn = iv + 1 < a.Nv ? (int)a.row_bytes : 0;
                    const R q00 = ST::get(p0), q01 = ST::get(p0 + ST::BYTES);
                    const R q10 = ST::get(p0 + dn), q11 = ST::get(p0 + dn + ST::BYTES);
                    const R c0 = rfma(fv, q10 - q00, q00);
                    const R c1 = rfma(fv, q11 - q01, q01);
                    acc[k] = rfma(ok ? wu1 : (R)0, c1, rfma(ok ? wu0 : (R)0, c0, acc[k]));
                }
            }
        }
    }

    if (!valid_xy) return;
    const size_t plane = (size_t)a.Nx * a.Ny;
    R *out = a.vol + (size_t)y * a.Nx + x;
#pragma unroll
    for (int k = 0; k < ZR; ++k) {
        const int z = z0 + k;
        if (z < a.Nz) {
            const R r = inside ? acc[k] * (a.out_scale * wback) : (R)0;
            if (a.accumulate)
                out[(size_t)z * plane] += r;
            else
                out[(size_t)z * plane] = r;
        }
    }
}

int bp_zrun(int dtype)
{
    switch (dtype) {
    case FDK_F32: return Store<st_f32>::ZR;
    case FDK_F16: return Store<st_f16>::ZR;
    case FDK_F16_PACKED: return Store<st_f16p>::ZR;
    case FDK_BF16: return Store<st_bf16>::ZR;
    case FDK_F64: return Store<st_f64>::ZR;
    default: return 0;
    }
}

int sample_bytes_of(int dtype)
{
    switch (dtype) {
    case FDK_F32: return 4;
    case FDK_F16: return 2;
    case FDK_F16_PACKED: return 2;
    case FDK_BF16: return 2;
    case FDK_F64: return 8;
    default: return 0;
    }
}

template <typename S>
static int launch_bp_t(const BPLaunch &L, hipStream_t st)
{
    using R = typename Store<S>::R;
    constexpr int ZR = Store<S>::ZR;
    if (L.pitch < stack_pitch(L.Nu))
        return fail(FDK_ERR_INVALID_ARGUMENT, "backproject: stack pitch below the padded width");
    BPArgs<R> a;
    a.stack = (const char *)L.stack;
    a.row_bytes = (uint32_t)((size_t)L.pitch * Store<S>::BYTES);
    a.angle_bytes = (size_t)a.row_bytes * L.Nv;
    if (a.angle_bytes > 0x7fffffffULL)
        return fail(FDK_ERR_UNSUPPORTED, "backproject: projection of 2 GiB or more");
    a.ang = (const R *)L.ang;
    a.scales = (const R *)L.scales;
    a.ax = (const R *)L.ax;
    a.ay = (const R *)L.ay;
    a.az = (const R *)L.az;
    a.vol = (R *)L.vol;
    a.Nx = L.Nx;
    a.Ny = L.Ny;
    a.Nz = L.Nz;
    a.Nu = L.Nu;
    a.Nv = L.Nv;
    a.nang = L.nang;
    a.cX = (R)L.cX;
    a.cZ = (R)L.cZ;
    a.D = (R)L.D;
    a.umax = (R)(L.Nu - 1);
    a.vmax = (R)(L.Nv - 1);
    a.out_scale = (R)L.out_scale;
    a.one = 1.0f;
    a.maskR2 = L.maskR2;
    a.accumulate = L.accumulate;
    a.tiles_x = (L.Nx + 15) / 16;
    a.tiles_xy = a.tiles_x * ((L.Ny + 15) / 16);
    const long long nzrun = (L.Nz + ZR - 1) / ZR;
    const long long nb = (long long)a.tiles_xy * nzrun;
    if (nb == 0) return FDK_OK;
    if (!grid_fits(nb, 256)) return fail(FDK_ERR_UNSUPPORTED, "backproject: volume too large");
    a.nblocks = (int)nb;
    hipLaunchKernelGGL((bp_kernel<S, ZR, (ZR < TKV_BP_LB ? ZR : TKV_BP_LB)>), dim3((unsigned)nb),
                       dim3(256), 0, st, a);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

int launch_backproject(int dtype, const BPLaunch &a, hipStream_t st)
{
    switch (dtype) {
    case FDK_F32: return launch_bp_t<st_f32>(a, st);
    case FDK_F16: return launch_bp_t<st_f16>(a, st);
    case FDK_F16_PACKED: return launch_bp_t<st_f16p>(a, st);
    case FDK_BF16: return launch_bp_t<st_bf16>(a, st);
    case FDK_F64: return launch_bp_t<st_f64>(a, st);
    default: return fail(FDK_ERR_INVALID_ARGUMENT, "backproject: unknown storage type");
    }
}

}  // namespace tkv
