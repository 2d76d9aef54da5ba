// fdk_phantom.hip -- synthetic cone-beam projections on the GPU (benchmark and
// test input; the reference reads measured data from disk instead).
//
// Exact line integrals through a set of ellipsoids for every detector pixel of
// the engine's virtual detector (isocentre plane, pitch tau): source at
// D*(cos b, sin b, 0), pixel (u,v) at (u-cX)*(-sin b, cos b, 0) + (0,0,v-cZ),
// the frame in which the backprojector's S, T, U relations hold.  FP64
// throughout.  Then optional Poisson counting noise (counter-based Philox4x32-10,
// Salmon et al. SC'11; PTRS transformed rejection, Hoermann 1993) and the log.
#include "fdk_internal.hpp"

namespace tkv {

constexpr int MAX_ELL = 128;

// LDS per ellipsoid: rows of R scaled by 1/semi-axis (9), the source position
// in the ellipsoid frame m (3), |m|^2 - 1, density -> 14 doubles.  The
// workgroup shares one projection (one source position), so m is per block.
__global__ void __launch_bounds__(256)
    phantom_kernel(const double *__restrict__ ells, int n, const double *__restrict__ ang,
                   float *__restrict__ sino, int Nu, int Nv, int nang, double cX, double cZ,
                   double D, double tau)
{
    __shared__ double E[MAX_ELL * 14];
    const size_t row = blockIdx.y + (size_t)gridDim.y * blockIdx.z;   // v * nang + b
    const bool row_ok = row < (size_t)Nv * nang;
    const int v = row_ok ? (int)(row / (size_t)nang) : 0;
    const int b = row_ok ? (int)(row - (size_t)v * nang) : 0;
    const double c = ang[2 * b], s = ang[2 * b + 1];
    const double src[3] = {D * c, D * s, 0.0};
    for (int i = threadIdx.x; i < n; i += 256) {
        const double *e = ells + 19 * i;
        double *o = E + 14 * i;
        double mm = 0.0;
        for (int k = 0; k < 3; ++k) {
            double inv = 1.0 / e[3 + k];
            double mk = 0.0;
            for (int j = 0; j < 3; ++j) {
                double r = e[6 + 3 * k + j] * inv;
                o[3 * k + j] = r;
                mk += r * (src[j] - e[j]);
            }
            o[9 + k] = mk;
            mm += mk * mk;
        }
        o[12] = mm - 1.0;
        o[13] = e[15];
    }
    __syncthreads();
    const int u = blockIdx.x * 256 + threadIdx.x;
    if (u >= Nu || !row_ok) return;
    const double du = (double)u - cX;
    const double d[3] = {-du * s - src[0], du * c - src[1], (double)v - cZ};
    const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double acc = 0.0;
    for (int i = 0; i < n; ++i) {
        const double *o = E + 14 * i;
        double w0 = o[0] * d[0] + o[1] * d[1] + o[2] * d[2];
        double w1 = o[3] * d[0] + o[4] * d[1] + o[5] * d[2];
        double w2 = o[6] * d[0] + o[7] * d[1] + o[8] * d[2];
        double A = w0 * w0 + w1 * w1 + w2 * w2;
        double Bq = o[9] * w0 + o[10] * w1 + o[11] * w2;
        double disc = Bq * Bq - A * o[12];
        if (disc > 0) acc += o[13] * 2.0 * sqrt(disc) / A;
    }
    sino[row * Nu + u] = (float)(acc * len * tau);
}

int launch_phantom(const PhantomLaunch &a, hipStream_t st)
{
    if (a.n > MAX_ELL) return fail(FDK_ERR_UNSUPPORTED, "phantom: at most 128 ellipsoids");
    size_t rows = (size_t)a.Nv * a.nang;
    if (rows == 0 || a.Nu == 0) return FDK_OK;
    unsigned gy = (unsigned)(rows < 65535 ? rows : 65535);
    unsigned gz = (unsigned)((rows + gy - 1) / gy);
    hipLaunchKernelGGL(phantom_kernel, dim3((a.Nu + 255) / 256, gy, gz), dim3(256), 0, st, a.ells,
                       a.n, a.ang, a.sino, a.Nu, a.Nv, a.nang, a.cX, a.cZ, a.D, a.tau);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

// ---------------------------------------------------------------- Poisson noise
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 0x1.0p-53;
}

// One Poisson(lam) draw for element idx; draw i uses counter (idx, i, 0).
__device__ double poisson_draw(double lam, uint64_t seed, uint64_t idx)
{
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t c[4];
    if (lam <= 0) return 0;
    if (lam < 10.0) {
        c[0] = (uint32_t)idx; c[1] = (uint32_t)(idx >> 32); c[2] = 0; c[3] = 0;
        philox4x32(c, k0, k1);
        double U = u53(c[0], c[1]);
        double pk = exp(-lam), F = pk;
        int k = 0;
        while (U > F && k < 1000) {
            ++k;
            pk *= lam / k;
            F += pk;
        }
        return k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (uint32_t it = 0; it < 64; ++it) {
        c[0] = (uint32_t)idx; c[1] = (uint32_t)(idx >> 32); c[2] = it; c[3] = 0;
        philox4x32(c, k0, k1);
        double U = u53(c[0], c[1]) - 0.5;
        double V = u53(c[2], c[3]);
        double us = 0.5 - fabs(U);
        double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0))
            return k;
    }
    return floor(lam + 0.5);
}

__global__ void __launch_bounds__(256)
    poisson_kernel(float *__restrict__ p, size_t n, double I0, uint64_t seed)
{
    // grid-stride over a capped grid: every sample keeps its own stream index
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256) {
        double lam = I0 * exp(-(double)p[i]);
        double cnt = poisson_draw(lam, seed, (uint64_t)i);
        if (cnt < 1.0) cnt = 1.0;
        p[i] = (float)(-log(cnt / I0));
    }
}

int launch_poisson_log(float *p, size_t n, double I0, uint64_t seed, hipStream_t st)
{
    if (n == 0) return FDK_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > (1u << 20)) blocks = 1u << 20;   // 2^28 threads, within a dispatch's range
    hipLaunchKernelGGL(poisson_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, I0, seed);
    TKV_HIP(hipGetLastError());
    return FDK_OK;
}

}  // namespace tkv
