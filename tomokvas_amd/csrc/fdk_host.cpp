// fdk_host.cpp -- host-side geometry and filter construction (FP64).
#include "fdk_internal.hpp"

#include <cmath>
#include <cstdio>

namespace tkv {

// ---------------------------------------------------------------- errors
static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? FDK_ERR_OUT_OF_MEMORY : FDK_ERR_HIP;
}

const char *last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------- geometry
// p = 2 * nextpow2(Nu) with a minimum of 2 before doubling (kernel.cu:570-573)
int padded_length(int Nu)
{
    int p = 2;
    while (p < Nu) p *= 2;
    return 2 * p;
}

// CalcSinCos (kernel.cu:430-477): sin/cos of k*delta_f + rot evaluated in double.
// The angle itself is formed in double from the float inputs (the reference
// forms k*delta_f in float; the two differ by < Nang*dbeta*2^-24 rad).
void angle_tables(const Geometry &g, std::vector<double> &s, std::vector<double> &c)
{
    s.resize(g.Nang);
    c.resize(g.Nang);
    for (int k = 0; k < g.Nang; ++k) {
        double a = (double)k * (double)g.dbeta + (double)g.rot;
        s[k] = std::sin(a);
        c[k] = std::cos(a);
    }
}

int slab_count_reference(int Nz, int B)
{
    float nr_Z = (float)(Nz - 1) / (float)B;
    return (int)std::ceil(nr_Z);
}

int slab_count(int Nz, int B) { return (Nz + B - 1) / B; }

// Voxel coordinates as the reference kernel computes them in float
// (kernel.cu:82-88): a = scale*(idx + d), d = -centre, sign flip by XOR of the
// sign bit (kernel.cu:501-506).  z: slab n covers global slices n*B+z with the
// per-slab offset of kernel.cu:821 and the centring correction of
// kernel.cu:398-400 (float/double mix of the reference).  Entries beyond the
// last slice continue the same formula (they are never stored).
void axis_tables_f32(const Geometry &g, int nz_pad, std::vector<float> &ax,
                     std::vector<float> &ay, std::vector<float> &az)
{
    const float dX = -g.cx, dY = -g.cy, dZ = -g.cz;
    ax.resize(g.Nx);
    ay.resize(g.Ny);
    az.resize(nz_pad);
    for (int x = 0; x < g.Nx; ++x) {
        float v = g.scale * ((float)x + dX);
        ax[x] = g.flipX ? -v : v;
    }
    for (int y = 0; y < g.Ny; ++y) {
        float v = g.scale * ((float)y + dY);
        ay[y] = g.flipY ? -v : v;
    }
    const int B = g.B;
    const float nr_Z = (float)(g.Nz - 1) / (float)B;
    const int nZref = (int)std::ceil(nr_Z);
    const float dcorZ = ((float)nZref - nr_Z) * (float)B / 2.0f;
    for (int gz = 0; gz < nz_pad; ++gz) {
        int n = gz / B, z = gz - n * B;
        float dZn = (float)((double)dZ + ((double)n - (double)(float)nZref / 2.0) * (double)B +
                            (double)dcorZ);
        float v = g.scale * ((float)z + dZn);
        az[gz] = g.flipZ ? -v : v;
    }
}

// Exact coordinates for the FP64 instantiation: a = scale*(idx - centre),
// z measured from (Nz-1)/2 + cz (the value the slab recentring yields).
void axis_tables_f64(const Geometry &g, int nz_pad, std::vector<double> &ax,
                     std::vector<double> &ay, std::vector<double> &az)
{
    ax.resize(g.Nx);
    ay.resize(g.Ny);
    az.resize(nz_pad);
    for (int x = 0; x < g.Nx; ++x) {
        double v = (double)g.scale * ((double)x - (double)g.cx);
        ax[x] = g.flipX ? -v : v;
    }
    for (int y = 0; y < g.Ny; ++y) {
        double v = (double)g.scale * ((double)y - (double)g.cy);
        ay[y] = g.flipY ? -v : v;
    }
    for (int z = 0; z < nz_pad; ++z) {
        double v = (double)g.scale * ((double)z - 0.5 * (double)(g.Nz - 1) - (double)g.cz);
        az[z] = g.flipZ ? -v : v;
    }
}

// ---------------------------------------------------------------- filter
// Discrete Ram-Lak kernel (kernel.cu:792-802; Kak & Slaney eq. 3.61), in double.
static double ramlak(int n, double tau)
{
    if (n == 0) return 1. / (4. * tau * tau);
    if (n % 2 == 0) return 0.;
    return -1. / ((double)n * (double)n * M_PI * M_PI * tau * tau);
}

// Shepp-Logan kernel h(n) = -2/(pi^2 tau^2 (4n^2-1)) (extension).
static double shepplogan(int n, double tau)
{
    return -2. / (M_PI * M_PI * tau * tau * (4. * (double)n * (double)n - 1.));
}

// DFT_p of real x for bins [0, kmax), twiddle index reduced mod p.
static void dft(const std::vector<double> &x, int p, int kmax, std::vector<double> &re,
                std::vector<double> &im)
{
    std::vector<double> c(p), s(p);
    for (int j = 0; j < p; ++j) {
        double a = 2.0 * M_PI * (double)j / (double)p;
        c[j] = std::cos(a);
        s[j] = std::sin(a);
    }
    std::vector<int> nz;
    for (int n = 0; n < p; ++n)
        if (x[n] != 0.) nz.push_back(n);
    re.assign(kmax, 0.);
    im.assign(kmax, 0.);
    for (int k = 0; k < kmax; ++k) {
        double sr = 0., si = 0.;
        for (int n : nz) {
            int j = (int)(((long long)k * n) % p);
            sr += x[n] * c[j];
            si -= x[n] * s[j];
        }
        re[k] = sr;
        im[k] = si;
    }
}

// FillFilter + R2C + AbsKernel (kernel.cu:585, 620-646, 783-790): taps
// t[i] = h(i - Nu/2)/2 for i < Nu, zero-padded to p; spectrum tau*|DFT(t)|,
// evenly extended to the full length p (Hermitian symmetry of the R2C output).
void build_spectrum(int kind, const Geometry &g, const std::vector<double> &user_taps,
                    std::vector<double> &Hre, std::vector<double> &Him)
{
    const int p = padded_length(g.Nu);
    Hre.assign(p, 0.);
    Him.assign(p, 0.);
    if (kind == FDK_FILTER_USER) {
        std::vector<double> re, im;
        dft(user_taps, p, p, re, im);
        Hre = re;
        Him = im;
        return;
    }
    const double tau = (double)g.tau;
    std::vector<double> t(p, 0.);
    for (int i = 0; i < g.Nu; ++i) {
        int n = i - g.Nu / 2;
        t[i] = (kind == FDK_FILTER_SHEPPLOGAN ? shepplogan(n, tau) : ramlak(n, tau)) / 2.;
    }
    std::vector<double> re, im;
    dft(t, p, p / 2 + 1, re, im);
    for (int k = 0; k <= p / 2; ++k) {
        double a = tau * std::sqrt(re[k] * re[k] + im[k] * im[k]);
        Hre[k] = a;
        if (k > 0 && k < p / 2) Hre[p - k] = a;
    }
}

void fft_pow2(std::vector<double> &re, std::vector<double> &im)
{
    const size_t n = re.size();
    // bit-reversal permutation
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len >> 1;
        for (size_t k = 0; k < half; ++k) {
            // e^{-2 pi i k / len}
            const double a = -2.0 * M_PI * (double)k / (double)len;
            const double wr = std::cos(a), wi = std::sin(a);
            for (size_t s0 = 0; s0 < n; s0 += len) {
                const size_t i0 = s0 + k, i1 = i0 + half;
                const double xr = re[i1] * wr - im[i1] * wi;
                const double xi = re[i1] * wi + im[i1] * wr;
                re[i1] = re[i0] - xr;
                im[i1] = im[i0] - xi;
                re[i0] += xr;
                im[i0] += xi;
            }
        }
    }
}

}  // namespace tkv
