// fdk_filter.hip -- cosine pre-weighting, LDS-resident Stockham FFT ramp filter,
// and the repack of filtered rows into the backprojector's stored stack (gfx950).
//
// Reference path: RB_comma3DKernel (FBPG/kernel.cu:179-201), fConvolution with
// cuFFT R2C/C2R per detector row (FBPG/kernel.cu:557-781), FiltrationKernel
// (FBPG/kernel.cu:40-53).  Here one workgroup transforms whole rows inside LDS:
// two real rows travel as the real and imaginary parts of one complex
// sequence (the spectrum is Hermitian-preserving, so the two filtered rows come
// back as the real and imaginary parts), zero-padded to p = 2*nextpow2(Nu),
// forward Stockham FFT (radix-16 passes and at most one radix-2/4/8 pass; Van
// Loan, Computational Frameworks for the FFT, 1992), bin-wise product with
// the spectrum, inverse FFT without 1/p (the backprojector applies it, as the
// reference does at FBPG/kernel.cu:575, :113), crop to Nu.  HBM traffic: one
// read and one write of each detector sample; everything else stays in LDS.
#include "fdk_internal.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

// waves per SIMD the radix-16 row filter is compiled for (4: 128 VGPRs
// without spills at p = 4096; 3 and 5 measured 41 and 48 ms against 34 ms at
// 1800 x 2048^2)
#ifndef TKV_FILTER_WAVES
#define TKV_FILTER_WAVES 4
#endif
// 1: FP32 rows with p >= 2048 use radix-8 passes over 512 threads (8 values
// per thread, 76 VGPRs, 6 waves per SIMD) -- one pass more per transform;
// measured 38 ms against the radix-16 plan's 34 ms, so off by default
#ifndef TKV_FILTER_R8
#define TKV_FILTER_R8 0
#endif
#ifndef TKV_FILTER_R8_WAVES
#define TKV_FILTER_R8_WAVES 6
#endif

namespace tkv {

template <typename R>
struct cpx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Cosine weight of one detector sample (kernel.cu:187, :197-198):
// D / sqrt(ce^2 + D^2 + cs^2), the sum contracted left to right into fmas as
// nvcc compiles the reference expression.
template <typename R>
__device__ __forceinline__ R rfma_w(R a, R b, R c);
template <>
__device__ __forceinline__ float rfma_w(float a, float b, float c) { return fmaf(a, b, c); }
template <>
__device__ __forceinline__ double rfma_w(double a, double b, double c) { return fma(a, b, c); }

template <typename R>
__device__ __forceinline__ R cos_weight(int u, int v, R cX, R cZ, R D)
{
    R ce = (R)u - cX;
    R cs = (R)v - cZ;
    R t = rfma_w(cs, cs, rfma_w(ce, ce, D * D));
    return D / sqrt(t);
}

// ---------------------------------------------------------------- cosine weight
// Coalesced: one wave per detector row (the reference walks a whole row per
// thread, kernel.cu:195-199), lanes on consecutive samples, four per lane and
// step with 16-byte accesses when the rows allow; four rows per workgroup.
// (One workgroup per 256 samples had made the pass bound by workgroup
// dispatch: 1.2 ms for 0.47 GB.)
template <typename R>
__global__ void __launch_bounds__(256) weight_kernel(R *__restrict__ sino, int Nu, int Nv,
                                                     int nang, R cX, R cZ, R D, int vec4)
{
    const long long rows = (long long)Nv * nang;
    const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    const int lane = threadIdx.x & 63;
    const int v = (int)(row / nang);
    R *r = sino + row * Nu;
    if (sizeof(R) == 4 && vec4) {
        // every row starts 16-byte aligned (checked at launch)
        float4 *r4 = (float4 *)r;
        for (int g = lane; 4 * g < Nu; g += 64) {
            float4 q = r4[g];
            const int u = 4 * g;
            q.x *= (float)cos_weight<R>(u, v, cX, cZ, D);
            q.y *= (float)cos_weight<R>(u + 1, v, cX, cZ, D);
            q.z *= (float)cos_weight<R>(u + 2, v, cX, cZ, D);
            q.w *= (float)cos_weight<R>(u + 3, v, cX, cZ, D);
            r4[g] = q;
        }
        return;
    }
    for (int u = lane; u < Nu; u += 64) r[u] *= cos_weight<R>(u, v, cX, cZ, D);
}

int launch_weight(int real_bytes, void *sino, int Nu, int Nv, int nang, float cX, float cZ,
                  float D, hipStream_t st)
{
    const long long rows = (long long)Nv * nang;
    if (rows == 0 || Nu == 0) return FDK_OK;
    if (!grid_fits((rows + 3) / 4, 256)) return fail(FDK_ERR_UNSUPPORTED, "weight: too many rows");
    const dim3 grid((unsigned)((rows + 3) / 4));
    const int vec4 = (Nu & 3) == 0 && ((uintptr_t)sino & 15) == 0;
    if (real_bytes == 8)
        hipLaunchKernelGGL(weight_kernel<double>, grid, dim3(256), 0, st, (double *)sino, Nu,
                           Nv, nang, (double)cX, (double)cZ, (double)D, 0);
    else
        hipLaunchKernelGGL(weight_kernel<float>, grid, dim3(256), 0, st, (float *)sino, Nu, Nv,
                           nang, cX, cZ, D, vec4);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

// ---------------------------------------------------------------- Stockham FFT
template <typename R>
__device__ __forceinline__ R group_max(R m, int width)
{
    for (int off = 1; off < width && off < 64; off <<= 1) {
        R o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    return m;
}

__device__ __forceinline__ unsigned float_bits_abs(float m) { return __float_as_uint(m); }
__device__ __forceinline__ unsigned float_bits_abs(double m) { return __float_as_uint((float)m); }

// ---------------------------------------------------------------- radix-16 row filter
// Mixed-radix Stockham plan: passes of radix RMAX (16, or 8), with the one
// smaller pass that log2(p) mod log2(RMAX) requires placed second, so the first
// and the last pass are both of radix RMAX whenever p >= RMAX^2.  A transform
// takes TPT = min(p/RMAX, BLOCK) threads; every thread holds E = P/TPT complex
// values through all passes.
template <int P_, int RMAX = 16, int BLOCK_ = 256>
struct Plan {
    static_assert(RMAX == 16 || RMAX == 8, "radix");
    static constexpr int P = P_;
    static constexpr int BLOCK = BLOCK_;
    static constexpr int LR = RMAX == 16 ? 4 : 3;
    static constexpr int LOG2 = __builtin_ctz(P);
    static constexpr int NR = LOG2 / LR;
    static constexpr int REM = 1 << (LOG2 % LR);
    static constexpr int NPASS = NR == 0 ? 1 : NR + (REM > 1 ? 1 : 0);
    static constexpr int TPT = P < RMAX ? 1 : (P / RMAX < BLOCK ? P / RMAX : BLOCK);
    static constexpr int NT = BLOCK / TPT;
    static constexpr int E = P / TPT;
    static constexpr int radix(int pass)
    {
        return NR == 0 ? REM : (pass == 0 ? RMAX : (REM > 1 && pass == 1 ? REM : RMAX));
    }
    static constexpr int RFIRST = radix(0);
    static constexpr int RLAST = radix(NPASS - 1);
};

// LDS index with one pad slot per 2^TKV_FILTER_PAD_SHIFT values (bank-conflict
// relief; 0: no padding), when the padded transforms still fit 64 KiB per
// workgroup
#ifndef TKV_FILTER_PAD_SHIFT
#define TKV_FILTER_PAD_SHIFT 4
#endif
template <typename R, typename PL>
struct Pad {
    static constexpr int P = PL::P;
    static constexpr int SH = TKV_FILTER_PAD_SHIFT;
    static constexpr bool ON =
        SH > 0 && sizeof(cpx<R>) * (size_t)PL::NT * (P + (P >> SH)) <= 65536;
    static constexpr int SIZE = ON ? P + (P >> SH) : P;
    static __device__ __forceinline__ int at(int n) { return ON ? n + (n >> SH) : n; }
};

// forward twiddle W_N^q = exp(-2 pi i q / N), N in {8, 16}: angles are
// multiples of pi/8; cos and sin of k*pi/8 for k = 0..3 (the rest by symmetry)
template <typename R, int N, int Q, bool INV>
__device__ __forceinline__ cpx<R> wconst()
{
    constexpr int k = (Q * (16 / N)) % 16;        // angle k*pi/8, k in [0,16)
    constexpr double C[4] = {1.0, 0.92387953251128675613, 0.70710678118654752440,
                             0.38268343236508977173};
    constexpr int q = k % 4, quad = k / 4;        // k*pi/8 = quad*pi/2 + q*pi/8
    constexpr double c0 = C[q], s0 = q == 0 ? 0.0 : C[4 - q];
    constexpr double c = quad == 0 ? c0 : quad == 1 ? -s0 : quad == 2 ? -c0 : s0;
    constexpr double s = quad == 0 ? s0 : quad == 1 ? c0 : quad == 2 ? -s0 : -c0;
    return {(R)c, (R)(INV ? s : -s)};
}

template <typename R, bool INV>
__device__ __forceinline__ void dft4v(cpx<R> &a0, cpx<R> &a1, cpx<R> &a2, cpx<R> &a3)
{
    cpx<R> v0 = {a0.x + a2.x, a0.y + a2.y}, v1 = {a0.x - a2.x, a0.y - a2.y};
    cpx<R> v2 = {a1.x + a3.x, a1.y + a3.y}, d = {a1.x - a3.x, a1.y - a3.y};
    cpx<R> v3 = INV ? cpx<R>{-d.y, d.x} : cpx<R>{d.y, -d.x};
    a0 = {v0.x + v2.x, v0.y + v2.y};
    a1 = {v1.x + v3.x, v1.y + v3.y};
    a2 = {v0.x - v2.x, v0.y - v2.y};
    a3 = {v1.x - v3.x, v1.y - v3.y};
}

// In-register DFT of length RAD on x[0..RAD), in place: X[k] = sum_n x[n] W^(nk)
// ends up in x[DFT::pos(k)].  RAD = 8, 16 are split as N = 4*N2 with
// n = N2*n1 + n2, k = k1 + 4*k2: length-4 DFTs over n1 on the elements
// {n2 + N2*n1} (in place, y(n2,k1) lands at n2 + N2*k1), twiddles W_N^(n2 k1),
// then length-N2 DFTs over n2 on the contiguous groups {N2*k1 + n2}, leaving
// X[k1 + 4*k2] at N2*k1 + k2.  No temporaries beyond the RAD values.
template <typename R, int RAD, bool INV>
struct DFT;
template <typename R, bool INV>
struct DFT<R, 2, INV> {
    static constexpr int pos(int k) { return k; }
    static __device__ __forceinline__ void run(cpx<R> *x)
    {
        cpx<R> t = x[0];
        x[0] = {t.x + x[1].x, t.y + x[1].y};
        x[1] = {t.x - x[1].x, t.y - x[1].y};
    }
};
template <typename R, bool INV>
struct DFT<R, 4, INV> {
    static constexpr int pos(int k) { return k; }
    static __device__ __forceinline__ void run(cpx<R> *x) { dft4v<R, INV>(x[0], x[1], x[2], x[3]); }
};
template <typename R, bool INV>
struct DFT<R, 8, INV> {
    static constexpr int pos(int k) { return 2 * (k % 4) + k / 4; }
    static __device__ __forceinline__ void run(cpx<R> *x)
    {
        dft4v<R, INV>(x[0], x[2], x[4], x[6]);
        dft4v<R, INV>(x[1], x[3], x[5], x[7]);
        x[3] = cmul(x[3], wconst<R, 8, 1, INV>());
        x[5] = cmul(x[5], wconst<R, 8, 2, INV>());
        x[7] = cmul(x[7], wconst<R, 8, 3, INV>());
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cpx<R> a = x[2 * k1], b = x[2 * k1 + 1];
            x[2 * k1] = {a.x + b.x, a.y + b.y};
            x[2 * k1 + 1] = {a.x - b.x, a.y - b.y};
        }
    }
};
template <typename R, bool INV>
struct DFT<R, 16, INV> {
    static constexpr int pos(int k) { return 4 * (k % 4) + k / 4; }
    static __device__ __forceinline__ void run(cpx<R> *x)
    {
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) dft4v<R, INV>(x[n2], x[n2 + 4], x[n2 + 8], x[n2 + 12]);
        // y(n2, k1) at n2 + 4*k1 gets W16^(n2*k1)
        x[5] = cmul(x[5], wconst<R, 16, 1, INV>());
        x[9] = cmul(x[9], wconst<R, 16, 2, INV>());
        x[13] = cmul(x[13], wconst<R, 16, 3, INV>());
        x[6] = cmul(x[6], wconst<R, 16, 2, INV>());
        x[10] = cmul(x[10], wconst<R, 16, 4, INV>());
        x[14] = cmul(x[14], wconst<R, 16, 6, INV>());
        x[7] = cmul(x[7], wconst<R, 16, 3, INV>());
        x[11] = cmul(x[11], wconst<R, 16, 6, INV>());
        x[15] = cmul(x[15], wconst<R, 16, 9, INV>());
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1)
            dft4v<R, INV>(x[4 * k1], x[4 * k1 + 1], x[4 * k1 + 2], x[4 * k1 + 3]);
    }
};

// one Stockham pass on the thread's registers: butterfly i = t + b*TPT holds
// its inputs (indices i + m*P/RAD) in reg[b*RAD + m]; outputs belong at
// ((i - k)*RAD + k) + m*pc with k = i mod pc
template <typename R, typename PL, int RAD, bool INV>
__device__ __forceinline__ void pass_compute(cpx<R> *reg, const cpx<R> *__restrict__ tw, int t,
                                             int pc)
{
    constexpr int P = PL::P;
    // the inverse passes read the forward passes' twiddles: an opaque copy of
    // the table pointer per pass keeps the compiler from holding every pass's
    // twiddles in registers from the forward transform to the inverse one
    const cpx<R> *twp = tw;
    asm volatile("" : "+s"(twp));
#pragma unroll
    for (int b = 0; b < PL::E / RAD; ++b) {
        const int i = t + b * PL::TPT;
        const int k = i & (pc - 1);
        cpx<R> *x = reg + b * RAD;
        if (pc > 1) {
            // twiddles W^(m k stride): the power-of-two multiples m = 1, 2, 4, 8
            // from the table (exactly rounded), the others as products of two
            // earlier ones (m = 2^j + r, r < 2^j: at most 3 roundings); a lane
            // then reads 4 table entries instead of RAD-1 scattered ones
            const int kk = k * (P / (RAD * pc));
            cpx<R> wp[RAD];
#pragma unroll
            for (int m = 1; m < RAD; ++m) {
                const int hi = 1 << (31 - __builtin_clz(m));   // largest power of 2 <= m
                wp[m] = hi == m ? twp[m * kk] : cmul(wp[hi], wp[m - hi]);
            }
#pragma unroll
            for (int m = 1; m < RAD; ++m) {
                cpx<R> w = wp[m];
                if (INV) w.y = -w.y;
                x[m] = cmul(x[m], w);
            }
        }
        DFT<R, RAD, INV>::run(x);
    }
}

template <typename R, typename PL, int RAD>
__device__ __forceinline__ void pass_gather(cpx<R> *reg, const cpx<R> *s, int t)
{
    constexpr int P = PL::P;
#pragma unroll
    for (int b = 0; b < PL::E / RAD; ++b) {
        const int i = t + b * PL::TPT;
#pragma unroll
        for (int m = 0; m < RAD; ++m) reg[b * RAD + m] = s[Pad<R, PL>::at(i + m * (P / RAD))];
    }
}

template <typename R, typename PL, int RAD>
__device__ __forceinline__ void pass_scatter(const cpx<R> *reg, cpx<R> *s, int t, int pc)
{
#pragma unroll
    for (int b = 0; b < PL::E / RAD; ++b) {
        const int i = t + b * PL::TPT;
        const int k = i & (pc - 1);
        const int j = (i - k) * RAD + k;
#pragma unroll
        for (int m = 0; m < RAD; ++m)
            s[Pad<R, PL>::at(j + m * pc)] = reg[b * RAD + DFT<R, RAD, false>::pos(m)];
    }
}

// all passes of one transform; on entry reg holds the first pass's inputs, on
// exit the outputs X[i + m*P/RLAST] of butterfly i = t + b*TPT in
// reg[b*RLAST + DFT::pos(m)]
template <typename R, typename PL, bool INV, int PASS>
struct Passes {
    static __device__ __forceinline__ void run(cpx<R> *reg, cpx<R> *s,
                                               const cpx<R> *__restrict__ tw, int t, int pc)
    {
        constexpr int RAD = PL::radix(PASS);
        if (PASS > 0) {
            __syncthreads();   // the previous pass's outputs are in LDS
            pass_gather<R, PL, RAD>(reg, s, t);
        }
        pass_compute<R, PL, RAD, INV>(reg, tw, t, pc);
        if (PASS + 1 < PL::NPASS) {
            __syncthreads();   // every thread has read its inputs
            pass_scatter<R, PL, RAD>(reg, s, t, pc);
            Passes<R, PL, INV, (PASS + 1 < PL::NPASS ? PASS + 1 : PASS)>::run(reg, s, tw, t,
                                                                           pc * RAD);
        }
    }
};

template <typename R, typename PL, int WAVES>
__global__ void __launch_bounds__(PL::BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES)))
    filter_kernel(const R *in, R *out, int Nu, int Nv, int nang, int weight, R cX, R cZ, R D,
                  const R *__restrict__ Hre, const R *__restrict__ Him, int conj,
                  const cpx<R> *__restrict__ tw, unsigned *__restrict__ maxabs)
{
    constexpr int P = PL::P;
    constexpr int E = PL::E;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int tid = threadIdx.x;
    const int f = tid / PL::TPT;          // transform within the workgroup
    const int t = tid - f * PL::TPT;      // thread within the transform
    cpx<R> *s = reinterpret_cast<cpx<R> *>(smem_raw) + f * Pad<R, PL>::SIZE;
    const long long rows = (long long)Nv * nang;
    const long long rA = 2 * ((long long)blockIdx.x * PL::NT + f);
    const long long rB = rA + 1;
    const bool hasA = rA < rows, hasB = rB < rows;
    const int vA = hasA ? (int)(rA / nang) : 0;
    const int vB = hasB ? (int)(rB / nang) : 0;
    const R *inA = in + (hasA ? rA : 0) * Nu;
    const R *inB = in + (hasB ? rB : 0) * Nu;

    cpx<R> reg[E];
    // first pass inputs straight from HBM: two rows as (re, im), optional
    // fused cosine weight (one evaluation when both rows share a detector
    // row), zero padding beyond Nu.  Nu <= P/2 (checked at launch), so the
    // inputs m >= RFIRST/2 of every butterfly (n >= P/2) are zero at compile
    // time and the first pass folds them away.
    const bool same_v = vA == vB;
#pragma unroll
    for (int b = 0; b < E / PL::RFIRST; ++b) {
        const int i = t + b * PL::TPT;
        // all of the butterfly's loads first (column clamped into the row; a
        // missing row is row 0), masked after: they are in flight together
        // while the weights are formed, not one pair per round trip
        constexpr int MH = (PL::RFIRST + 1) / 2;
        R la[MH], lc[MH];
#pragma unroll
        for (int m = 0; m < MH; ++m) {
            const int n = i + m * (P / PL::RFIRST);
            const int nc = n < Nu ? n : Nu - 1;
            la[m] = inA[nc];
            lc[m] = inB[nc];
        }
#pragma unroll
        for (int m = 0; m < PL::RFIRST; ++m) {
            const int n = i + m * (P / PL::RFIRST);
            R a = 0, c = 0;
            if (m < MH) {
                a = hasA && n < Nu ? la[m] : (R)0;
                c = hasB && n < Nu ? lc[m] : (R)0;
                if (weight) {   // finite for every n < P/2; zeros stay zero
                    const R wa = cos_weight<R>(n, vA, cX, cZ, D);
                    a *= wa;
                    if (same_v)
                        c *= wa;
                    else
                        c *= cos_weight<R>(n, vB, cX, cZ, D);
                }
            }
            reg[b * PL::RFIRST + m] = {a, c};
        }
    }
    Passes<R, PL, false, 0>::run(reg, s, tw, t, 1);

    // bin-wise product with the spectrum (FiltrationKernel, kernel.cu:40-53), in
    // registers: X[i + m*P/RLAST] sits at reg[b*RLAST + pos(m)]; the product is
    // stored back in natural order m, the inverse first pass's input order
    using DL = DFT<R, PL::RLAST, false>;
    cpx<R> nat[E];
#pragma unroll
    for (int b = 0; b < E / PL::RLAST; ++b) {
        const int i = t + b * PL::TPT;
#pragma unroll
        for (int m = 0; m < PL::RLAST; ++m) {
            const int n = i + m * (P / PL::RLAST);
            cpx<R> z = reg[b * PL::RLAST + DL::pos(m)];
            R hr = Hre[n];
            if (Him) {
                R hi = conj ? -Him[n] : Him[n];
                nat[b * PL::RLAST + m] = {z.x * hr - z.y * hi, z.x * hi + z.y * hr};
            } else {
                nat[b * PL::RLAST + m] = {z.x * hr, z.y * hr};
            }
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) reg[e] = nat[e];
    if (PL::RFIRST != PL::RLAST) {
        // hand the spectrum to the inverse transform's first pass through LDS
        __syncthreads();
#pragma unroll
        for (int b = 0; b < E / PL::RLAST; ++b) {
            const int i = t + b * PL::TPT;
#pragma unroll
            for (int m = 0; m < PL::RLAST; ++m)
                s[Pad<R, PL>::at(i + m * (P / PL::RLAST))] = reg[b * PL::RLAST + m];
        }
        __syncthreads();
        pass_gather<R, PL, PL::RFIRST>(reg, s, t);
    }
    Passes<R, PL, true, 0>::run(reg, s, tw, t, 1);

    // crop to Nu straight to HBM, per-projection max |value|
    using DI = DFT<R, PL::RLAST, true>;
    R mA = 0, mB = 0;
    R *outA = out + (hasA ? rA : 0) * Nu;
    R *outB = out + (hasB ? rB : 0) * Nu;
    // (outputs m >= RLAST/2, n >= P/2 >= Nu, are cropped: their arithmetic in
    // the last pass is dead and removed at compile time)
#pragma unroll
    for (int b = 0; b < E / PL::RLAST; ++b) {
        const int i = t + b * PL::TPT;
#pragma unroll
        for (int m = 0; m < (PL::RLAST + 1) / 2; ++m) {
            const int n = i + m * (P / PL::RLAST);
            if (n < Nu) {
                cpx<R> z = reg[b * PL::RLAST + DI::pos(m)];
                if (hasA) outA[n] = z.x;
                if (hasB) outB[n] = z.y;
                R ax = z.x < 0 ? -z.x : z.x, ay = z.y < 0 ? -z.y : z.y;
                mA = ax > mA ? ax : mA;
                mB = ay > mB ? ay : mB;
            }
        }
    }
    if (maxabs) {
        mA = group_max(mA, PL::TPT);
        mB = group_max(mB, PL::TPT);
        bool leader = (PL::TPT >= 64) ? ((tid & 63) == 0) : (t == 0);
        if (leader) {
            if (hasA) atomicMax(&maxabs[rA % nang], float_bits_abs(mA));
            if (hasB) atomicMax(&maxabs[rB % nang], float_bits_abs(mB));
        }
    }
}

template <typename R, typename PL, int WAVES>
static int launch_filter_plan(const FilterLaunch &a, hipStream_t st)
{
    constexpr int P = PL::P;
    long long rows = (long long)a.Nv * a.nang;
    long long pairs = (rows + 1) / 2;
    long long blocks = (pairs + PL::NT - 1) / PL::NT;
    if (blocks == 0) return FDK_OK;
    if (!grid_fits(blocks, PL::BLOCK)) return fail(FDK_ERR_UNSUPPORTED, "filter: too many rows");
    if (2 * a.Nu > P) return fail(FDK_ERR_INVALID_ARGUMENT, "filter: rows longer than p/2");
    size_t lds = PL::NPASS > 1 || PL::RFIRST != PL::RLAST
                     ? sizeof(cpx<R>) * (size_t)PL::NT * Pad<R, PL>::SIZE
                     : 0;
    auto kern = filter_kernel<R, PL, WAVES>;
    // beyond 64 KiB (one transform of p = 16384 in FP32, 8192 in FP64) the
    // workgroup takes up to the 160 KiB of LDS of a gfx950 CU
    if (lds > 163840) return fail(FDK_ERR_UNSUPPORTED, "filter: transform does not fit LDS");
    if (lds > 65536)
        TKV_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(PL::BLOCK), lds, st, (const R *)a.in,
                       (R *)a.out, a.Nu, a.Nv, a.nang, a.weight, (R)a.cX, (R)a.cZ, (R)a.D,
                       (const R *)a.Hre, (const R *)a.Him, a.conj, (const cpx<R> *)a.tw,
                       a.maxabs);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

// radix 16 over 256 threads, at most 16 values per thread: FP32 p = 8192 one
// transform per workgroup of 512 threads (over 256 threads, 32 values each
// spilled registers: 33.7 against 25.9 ms for 3072-sample rows x 1.1 M), the
// longest rows (FP32 p = 16384, FP64 p = 8192) one per workgroup of 1024 / 512
// threads (the transform alone fills 128 KiB of LDS); TKV_FILTER_R8: long FP32
// rows (p >= 2048) in radix-8 passes over 512 threads, 8 values per thread
template <typename R, int P>
static int launch_filter_p(const FilterLaunch &a, hipStream_t st)
{
    if constexpr (sizeof(R) == 4 && P > 8192)
        return launch_filter_plan<R, Plan<P, 16, 1024>, TKV_FILTER_WAVES>(a, st);
    else if constexpr (sizeof(R) == 4 && P == 8192)
        return launch_filter_plan<R, Plan<P, 16, 512>, TKV_FILTER_WAVES>(a, st);
    else if constexpr (sizeof(R) == 8 && P > 4096)
        return launch_filter_plan<R, Plan<P, 16, 512>, 2>(a, st);
    else if constexpr (TKV_FILTER_R8 && sizeof(R) == 4 && P >= 2048)
        return launch_filter_plan<R, Plan<P, 8, 512>, TKV_FILTER_R8_WAVES>(a, st);
    else
        // FP64 values take twice the registers: compiled for 3 waves
        return launch_filter_plan<R, Plan<P>, sizeof(R) == 4 ? TKV_FILTER_WAVES : 3>(a, st);
}

template <typename R>
static int launch_filter_t(const FilterLaunch &a, hipStream_t st)
{
    switch (a.p) {
    case 4: return launch_filter_p<R, 4>(a, st);
    case 8: return launch_filter_p<R, 8>(a, st);
    case 16: return launch_filter_p<R, 16>(a, st);
    case 32: return launch_filter_p<R, 32>(a, st);
    case 64: return launch_filter_p<R, 64>(a, st);
    case 128: return launch_filter_p<R, 128>(a, st);
    case 256: return launch_filter_p<R, 256>(a, st);
    case 512: return launch_filter_p<R, 512>(a, st);
    case 1024: return launch_filter_p<R, 1024>(a, st);
    case 2048: return launch_filter_p<R, 2048>(a, st);
    case 4096: return launch_filter_p<R, 4096>(a, st);
    case 8192: return launch_filter_p<R, 8192>(a, st);
    case 16384:
        if constexpr (sizeof(R) == 4) return launch_filter_p<R, 16384>(a, st);
        break;
    default: break;
    }
    return fail(FDK_ERR_UNSUPPORTED,
                "filter: padded length " + std::to_string(a.p) + " is not supported");
}

int launch_filter(int real_bytes, const FilterLaunch &a, hipStream_t st)
{
    return real_bytes == 8 ? launch_filter_t<double>(a, st) : launch_filter_t<float>(a, st);
}

// ---------------------------------------------------------------- tap gradient
// For the learnable filter y = crop(IFFT_unnorm(H . FFT(pad q))), H = DFT_p(h):
//   dL/dh[j] = p * sum_rows sum_m G[m] q[(m - j) mod p] = sum_k R[k] e^{-2 pi i k j/p},
//   R[k] = sum_rows Q[k] conj(G^[k])          (Q, G^: DFTs of the padded rows).
// Two rows A, B travel as one complex sequence (Zq = Q_A + i Q_B, likewise ZG);
// then Q_A conj(G_A) + Q_B conj(G_B) = (P[k] + conj(P[-k]))/2 with
// P[k] = Zq[k] conj(ZG[k]).  Each workgroup accumulates P over its row pairs in
// registers and writes one partial spectrum; tapgrad_reduce_kernel sums the
// partials in double, and the host symmetrises and applies the final length-p
// transform (an FFT in double).
// Per transform of the workgroup (the row filter's radix-16 Stockham plan PL),
// over row pairs pr = blockIdx.x*NT + f, stepping by gridDim.x*NT: the forward
// transforms of the q pair and of the G pair leave their bins in the same
// register slots, so P[k] accumulates slot by slot in registers; at the end the
// NT transforms' sums are added through LDS in natural bin order.
template <typename R, typename PL>
__device__ __forceinline__ void tapgrad_load_pair(cpx<R> *reg, const R *__restrict__ x, int Nu,
                                                  long long rA, bool hasA, bool hasB, int t)
{
    constexpr int P = PL::P;
    const R *xa = x + (hasA ? rA : 0) * Nu;
    const R *xb = x + (hasB ? rA + 1 : 0) * Nu;
#pragma unroll
    for (int b = 0; b < PL::E / PL::RFIRST; ++b) {
        const int i = t + b * PL::TPT;
        // Nu <= P/2: inputs m >= RFIRST/2 are zero at compile time.  The
        // loads are unconditional (column clamped into the row, a missing row
        // read as row 0) and masked after, so they are in flight together
        constexpr int MH = (PL::RFIRST + 1) / 2;
        R la[MH], lb[MH];
#pragma unroll
        for (int m = 0; m < MH; ++m) {
            const int n = i + m * (P / PL::RFIRST);
            const int nc = n < Nu ? n : Nu - 1;
            la[m] = xa[nc];
            lb[m] = xb[nc];
        }
#pragma unroll
        for (int m = 0; m < PL::RFIRST; ++m) {
            const int n = i + m * (P / PL::RFIRST);
            R va = 0, vb = 0;
            if (m < MH) {
                va = hasA && n < Nu ? la[m] : (R)0;
                vb = hasB && n < Nu ? lb[m] : (R)0;
            }
            reg[b * PL::RFIRST + m] = {va, vb};
        }
    }
}

template <typename R, typename PL>
__global__ void __launch_bounds__(PL::BLOCK)
    tapgrad_kernel(const R *__restrict__ q, const R *__restrict__ G, int Nu, long long rows,
                   const cpx<R> *__restrict__ tw, cpx<R> *__restrict__ partial)
{
    constexpr int P = PL::P;
    constexpr int E = PL::E;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int tid = threadIdx.x;
    const int f = tid / PL::TPT;
    const int t = tid - f * PL::TPT;
    cpx<R> *s = reinterpret_cast<cpx<R> *>(smem_raw) + f * Pad<R, PL>::SIZE;
    cpx<R> acc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = {0, 0};
    const long long pairs = (rows + 1) / 2;
    // the bound is the same for every thread of the workgroup (the passes'
    // barriers are workgroup-wide); a transform past the last pair adds zeros
    for (long long pr0 = (long long)blockIdx.x * PL::NT; pr0 < pairs;
         pr0 += (long long)gridDim.x * PL::NT) {
        const long long pr = pr0 + f;
        const long long rA = 2 * pr;
        const bool hasA = pr < pairs && rA < rows, hasB = pr < pairs && rA + 1 < rows;
        cpx<R> zq[E], reg[E];
        tapgrad_load_pair<R, PL>(reg, q, Nu, rA, hasA, hasB, t);
        Passes<R, PL, false, 0>::run(reg, s, tw, t, 1);
#pragma unroll
        for (int e = 0; e < E; ++e) zq[e] = reg[e];
        tapgrad_load_pair<R, PL>(reg, G, Nu, rA, hasA, hasB, t);
        Passes<R, PL, false, 0>::run(reg, s, tw, t, 1);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            acc[e].x += zq[e].x * reg[e].x + zq[e].y * reg[e].y;   // Zq * conj(ZG)
            acc[e].y += zq[e].y * reg[e].x - zq[e].x * reg[e].y;
        }
    }
    // bins in natural order per transform, then the NT transforms' sum
    using DL = DFT<R, PL::RLAST, false>;
    cpx<R> *red = reinterpret_cast<cpx<R> *>(smem_raw);
    __syncthreads();   // every transform is done with its LDS
#pragma unroll
    for (int b = 0; b < E / PL::RLAST; ++b) {
        const int i = t + b * PL::TPT;
#pragma unroll
        for (int m = 0; m < PL::RLAST; ++m)
            red[f * P + i + m * (P / PL::RLAST)] = acc[b * PL::RLAST + DL::pos(m)];
    }
    __syncthreads();
    for (int n = tid; n < P; n += PL::BLOCK) {
        cpx<R> sum = {0, 0};
        for (int ff = 0; ff < PL::NT; ++ff) {
            sum.x += red[ff * P + n].x;
            sum.y += red[ff * P + n].y;
        }
        partial[(size_t)blockIdx.x * P + n] = sum;
    }
}

// Sum of the workgroups' partial spectra, in double: out[k] = sum_b partial[b][k].
// A workgroup owns 64 consecutive bins (one per lane, so each block's row is
// read in 64-bin segments); its 16 waves take the blocks b = w, w + 16, ...
// and their 16 sums are added in a fixed tree, so the result does not depend
// on scheduling.
template <typename R>
__global__ void __launch_bounds__(1024)
    tapgrad_reduce_kernel(const cpx<R> *__restrict__ partial, int nblocks, int p,
                          double2 *__restrict__ out)
{
    __shared__ double2 part[16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    double sx = 0., sy = 0.;
    if (k < p)
        for (int b = w; b < nblocks; b += 16) {
            const cpx<R> v = partial[(size_t)b * p + k];
            sx += (double)v.x;
            sy += (double)v.y;
        }
    part[w][lane] = make_double2(sx, sy);
    __syncthreads();
#pragma unroll
    for (int h = 8; h > 0; h >>= 1) {
        if (w < h) {
            const double2 o = part[w + h][lane];
            part[w][lane] = make_double2(part[w][lane].x + o.x, part[w][lane].y + o.y);
        }
        __syncthreads();
    }
    if (w == 0 && k < p) out[k] = part[0][lane];
}

template <typename R, int P>
static int launch_tapgrad_p(const TapGradLaunch &a, hipStream_t st)
{
    using PL = Plan<P, 16, 256>;
    // the passes' transforms, and the final sum in natural order
    const size_t lds = sizeof(cpx<R>) * (size_t)PL::NT *
                       (Pad<R, PL>::SIZE > P ? (size_t)Pad<R, PL>::SIZE : (size_t)P);
    auto kern = tapgrad_kernel<R, PL>;
    if (lds > 163840) return fail(FDK_ERR_UNSUPPORTED, "tap gradient: transform does not fit LDS");
    if (lds > 65536)
        TKV_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)a.nblocks), dim3(PL::BLOCK), lds, st, (const R *)a.q,
                       (const R *)a.G, a.Nu, (long long)a.rows, (const cpx<R> *)a.tw,
                       (cpx<R> *)a.partial);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

template <typename R>
static int launch_tapgrad_t(const TapGradLaunch &a, hipStream_t st)
{
    switch (a.p) {
    case 4: return launch_tapgrad_p<R, 4>(a, st);
    case 8: return launch_tapgrad_p<R, 8>(a, st);
    case 16: return launch_tapgrad_p<R, 16>(a, st);
    case 32: return launch_tapgrad_p<R, 32>(a, st);
    case 64: return launch_tapgrad_p<R, 64>(a, st);
    case 128: return launch_tapgrad_p<R, 128>(a, st);
    case 256: return launch_tapgrad_p<R, 256>(a, st);
    case 512: return launch_tapgrad_p<R, 512>(a, st);
    case 1024: return launch_tapgrad_p<R, 1024>(a, st);
    case 2048: return launch_tapgrad_p<R, 2048>(a, st);
    case 4096:
        if (sizeof(R) == 4) return launch_tapgrad_p<R, 4096>(a, st);
        break;
    default: break;
    }
    return fail(FDK_ERR_UNSUPPORTED,
                "tap gradient: padded length " + std::to_string(a.p) + " is not supported");
}

// The final transform on the device: dh[j] = Re sum_k R[k] e^{-2 pi i k j / p}
// with R[k] = (P[k] + conj(P[p-k]))/2 the Hermitian part of the summed
// spectrum P; one thread per tap j, twiddles from the exact table
// tw[n] = e^{-2 pi i n / p} at n = k*j mod p.
__global__ void __launch_bounds__(256)
    tapgrad_dft_kernel(const double2 *__restrict__ P, const double2 *__restrict__ tw, int p,
                       double *__restrict__ dh)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= p) return;
    double s = 0.;
    unsigned n = 0;   // k*j mod p, p a power of two
    for (int k = 0; k < p; ++k, n = (n + (unsigned)j) & (unsigned)(p - 1)) {
        const int m = (p - k) & (p - 1);
        const double2 a = P[k], b = P[m];
        const double rr = 0.5 * (a.x + b.x), ri = 0.5 * (a.y - b.y);
        const double2 w = tw[n];   // (cos, -sin) of 2 pi k j / p
        s += rr * w.x - ri * w.y;
    }
    dh[j] = s;
}

int launch_tap_gradient_dft(const void *spectrum, const void *tw64, int p, double *dh,
                            hipStream_t st)
{
    hipLaunchKernelGGL(tapgrad_dft_kernel, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, st,
                       (const double2 *)spectrum, (const double2 *)tw64, p, dh);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

int launch_tap_gradient(int real_bytes, const TapGradLaunch &a, hipStream_t st)
{
    int rc = real_bytes == 8 ? launch_tapgrad_t<double>(a, st) : launch_tapgrad_t<float>(a, st);
    if (rc) return rc;
    const dim3 grid((unsigned)((a.p + 63) / 64));
    if (real_bytes == 8)
        hipLaunchKernelGGL(tapgrad_reduce_kernel<double>, grid, dim3(1024), 0, st,
                           (const cpx<double> *)a.partial, a.nblocks, a.p, (double2 *)a.spectrum);
    else
        hipLaunchKernelGGL(tapgrad_reduce_kernel<float>, grid, dim3(1024), 0, st,
                           (const cpx<float> *)a.partial, a.nblocks, a.p, (double2 *)a.spectrum);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

// ---------------------------------------------------------------- repack
// filtered [v][beta][u] (R)  ->  stored stack [beta][v][pitch] in the storage
// type, samples divided by scale_beta, zero for u >= Nu.  A workgroup writes
// four stack rows; a lane converts four consecutive samples per step.
template <typename Out>
struct Pack4;
template <>
struct Pack4<float> {
    using V = float4;
    static __device__ __forceinline__ V make(float a, float b, float c, float d)
    {
        return {a, b, c, d};
    }
};
template <>
struct Pack4<double> {
    using V = double4;
    static __device__ __forceinline__ V make(double a, double b, double c, double d)
    {
        return {a, b, c, d};
    }
};
struct f16_t;
struct bf16_t;
template <>
struct Pack4<f16_t> {
    using V = uint2;
    static __device__ __forceinline__ unsigned h(float x)
    {
        return __half_as_ushort(__float2half_rn(x));
    }
    static __device__ __forceinline__ V make(float a, float b, float c, float d)
    {
        return {h(a) | (h(b) << 16), h(c) | (h(d) << 16)};
    }
};
template <>
struct Pack4<bf16_t> {
    using V = uint2;
    static __device__ __forceinline__ unsigned h(float x)
    {
        return __bfloat16_as_ushort(__float2bfloat16(x));
    }
    static __device__ __forceinline__ V make(float a, float b, float c, float d)
    {
        return {h(a) | (h(b) << 16), h(c) | (h(d) << 16)};
    }
};

// FP16: the largest filtered magnitude of a projection maps to 2^13; values
// 2^27 below it are still normal FP16 numbers.
constexpr float F16_TARGET = 8192.0f;

template <typename R, typename Out, bool SCALED>
__global__ void __launch_bounds__(256)
    pack_kernel(const R *__restrict__ filt, int Nu, int Nv, int nang, int pitch,
                const unsigned *__restrict__ maxabs, typename Pack4<Out>::V *__restrict__ stack,
                R *__restrict__ scales, int vec4)
{
    const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= (long long)nang * Nv) return;
    const int beta = (int)(row / Nv), v = (int)(row - (long long)beta * Nv);
    R inv = 1, sc = 1;
    if (SCALED) {
        const float m = __uint_as_float(maxabs[beta]);
        sc = m > 0 ? (R)(m / F16_TARGET) : (R)1;
        inv = (R)1 / sc;
    }
    const int lane = threadIdx.x & 63;
    if (v == 0 && lane == 0) scales[beta] = sc;
    const R *src = filt + ((size_t)v * nang + beta) * Nu;
    typename Pack4<Out>::V *dst = stack + (size_t)row * (pitch / 4);
    if constexpr (sizeof(R) == 4) {
        if (vec4) {
            // rows start 16-byte aligned (checked at launch): one 16-byte load
            // per lane and group
            const float4 *src4 = (const float4 *)src;
            for (int g = lane; g < pitch / 4; g += 64) {
                float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
                if (4 * g < Nu) q = src4[g];
                dst[g] = Pack4<Out>::make(q.x * inv, q.y * inv, q.z * inv, q.w * inv);
            }
            return;
        }
    }
    for (int g = lane; g < pitch / 4; g += 64) {
        const int u = 4 * g;
        R q[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) q[i] = u + i < Nu ? src[u + i] * inv : (R)0;
        dst[g] = Pack4<Out>::make(q[0], q[1], q[2], q[3]);
    }
}

int launch_pack(int dtype_in, const void *filt, int Nu, int Nv, int nang, int pitch,
                const unsigned *maxabs, int store_dtype, void *stack, void *scales,
                hipStream_t st)
{
    if (nang == 0 || Nv == 0) return FDK_OK;
    if (pitch % 4 != 0) return fail(FDK_ERR_INVALID_ARGUMENT, "pack: pitch not a multiple of 4");
    const long long rows = (long long)nang * Nv;
    if (!grid_fits((rows + 3) / 4, 256)) return fail(FDK_ERR_UNSUPPORTED, "pack: stack too large");
    dim3 grid((unsigned)((rows + 3) / 4));
    const int vec4 = (Nu & 3) == 0 && ((uintptr_t)filt & 15) == 0;
    if (dtype_in == FDK_F64) {
        if (store_dtype != FDK_F64)
            return fail(FDK_ERR_INVALID_ARGUMENT, "pack: F64 input needs F64 storage");
        hipLaunchKernelGGL((pack_kernel<double, double, false>), grid, dim3(256), 0, st,
                           (const double *)filt, Nu, Nv, nang, pitch, maxabs, (double4 *)stack,
                           (double *)scales, vec4);
    } else {
        const float *f = (const float *)filt;
        switch (store_dtype) {
        case FDK_F32:
            hipLaunchKernelGGL((pack_kernel<float, float, false>), grid, dim3(256), 0, st, f, Nu,
                               Nv, nang, pitch, maxabs, (float4 *)stack, (float *)scales, vec4);
            break;
        case FDK_F16:
        case FDK_F16_PACKED:   // same stored stack; only the backprojector differs
            if (!maxabs) return fail(FDK_ERR_INVALID_ARGUMENT, "pack: FP16 needs maxabs");
            hipLaunchKernelGGL((pack_kernel<float, f16_t, true>), grid, dim3(256), 0, st, f, Nu,
                               Nv, nang, pitch, maxabs, (uint2 *)stack, (float *)scales, vec4);
            break;
        case FDK_BF16:
            hipLaunchKernelGGL((pack_kernel<float, bf16_t, false>), grid, dim3(256), 0, st, f,
                               Nu, Nv, nang, pitch, maxabs, (uint2 *)stack, (float *)scales, vec4);
            break;
        default:
            return fail(FDK_ERR_INVALID_ARGUMENT, "pack: F32 input needs F32/F16/BF16 storage");
        }
    }
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

// ---------------------------------------------------------------- accumulate
// CollectionKernel (kernel.cu:203-215): slab += partial, flat and coalesced.
__global__ void __launch_bounds__(256) accumulate_kernel(float *__restrict__ dst,
                                                         const float *__restrict__ src,
                                                         size_t n)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) dst[i] += src[i];
}

int launch_accumulate(float *dst, const float *src, size_t n, hipStream_t st)
{
    if (n == 0) return FDK_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(accumulate_kernel, dim3((unsigned)blocks), dim3(256), 0, st, dst, src,
                       n);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

// zero fill by a kernel (no memset node, so the sequence stays graph-capturable)
__global__ void __launch_bounds__(256) zero_u32_kernel(unsigned *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = 0u;
}

int launch_zero_u32(unsigned *dst, size_t n, hipStream_t st)
{
    if (n == 0) return FDK_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, dst, n);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

// ---------------------------------------------------------------- compare
// per-block partial sums in double, combined with double atomics:
// out[0] += sum (a-b)^2, out[1] += sum b^2, out[2] = max |a-b| (as ordered bits)
__global__ void __launch_bounds__(256) compare_kernel(const float *__restrict__ a,
                                                      const float *__restrict__ b, size_t n,
                                                      double *__restrict__ out)
{
    __shared__ double s0[256], s1[256], s2[256];
    double e = 0, r = 0, m = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        double d = (double)a[i] - (double)b[i];
        e += d * d;
        r += (double)b[i] * (double)b[i];
        double ad = d < 0 ? -d : d;
        m = ad > m ? ad : m;
    }
    s0[threadIdx.x] = e;
    s1[threadIdx.x] = r;
    s2[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            s0[threadIdx.x] += s0[threadIdx.x + off];
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] = s2[threadIdx.x] > s2[threadIdx.x + off] ? s2[threadIdx.x]
                                                                     : s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        unsafeAtomicAdd(&out[0], s0[0]);
        unsafeAtomicAdd(&out[1], s1[0]);
        atomicMax((unsigned long long *)&out[2], (unsigned long long)__double_as_longlong(s2[0]));
    }
}

__global__ void zero_f64x3(double *out)
{
    if (threadIdx.x < 3) out[threadIdx.x] = 0.0;
}

int launch_compare(const float *a, const float *b, size_t n, double *d_out3, hipStream_t st)
{
    hipLaunchKernelGGL(zero_f64x3, dim3(1), dim3(64), 0, st, d_out3);
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(compare_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a, b, n, d_out3);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

}  // namespace tkv
