/*
 * fdk_oracle.c -- CPU restatement of the reference FDK path (Lukas-kV/TomoKvas,
 * FBPG/kernel.cu) plus the synthetic phantom generators used to feed it.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tomokvas_amd's HIP kernels.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#define _GNU_SOURCE
#include "fdk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ============================ filter construction ============================ */

/* p = 2 * (smallest power of two >= Nu, at least 2)  -- kernel.cu:570-573 */
int ofdk_padded_length(int Nu)
{
    int p = 2;
    while (p < Nu) p *= 2;
    return p * 2;
}

/* RamLak() exactly as kernel.cu:792-802: double arithmetic (PI is a double
 * literal, kernel.cuh:23), result rounded to float by the return type. */
float ofdk_ramlak_ref(int n, float tau)
{
    if (n == 0) return (float)(1. / (4. * tau * tau));
    if (n % 2 == 0) return 0.f;
    return (float)(-1. / (n * n * M_PI * M_PI * tau * tau));
}

/* The same discrete Ram-Lak kernel (Kak & Slaney eq. 3.61) kept in double. */
double ofdk_ramlak(int n, double tau)
{
    if (n == 0) return 1. / (4. * tau * tau);
    if (n % 2 == 0) return 0.;
    return -1. / ((double)n * (double)n * M_PI * M_PI * tau * tau);
}

/* Shepp-Logan apodised ramp kernel h(n) = -2 / (pi^2 tau^2 (4 n^2 - 1))
 * (Shepp & Logan 1974); an extension -- the reference ships only Ram-Lak. */
double ofdk_shepplogan(int n, double tau)
{
    return -2. / (M_PI * M_PI * tau * tau * (4. * (double)n * (double)n - 1.));
}

/* FillFilter (kernel.cu:783-790): t[i] = h(i - Nu/2)/2 for i < Nu, zero up to p */
void ofdk_fill_taps(int kind, int Nu, double tau, int p, double *taps)
{
    for (int i = 0; i < p; ++i) taps[i] = 0.;
    for (int i = 0; i < Nu; ++i) {
        int n = i - Nu / 2;
        taps[i] = (kind == 1 ? ofdk_shepplogan(n, tau) : ofdk_ramlak(n, tau)) / 2.;
    }
}

/* Direct DFT in double with the twiddle index reduced mod p (exact angles). */
static void dft_real(const double *x, int p, int kmax, double *re, double *im)
{
    double *c = (double *)malloc(sizeof(double) * (size_t)p);
    double *s = (double *)malloc(sizeof(double) * (size_t)p);
    for (int j = 0; j < p; ++j) {
        double a = 2.0 * M_PI * (double)j / (double)p;
        c[j] = cos(a); s[j] = sin(a);
    }
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < kmax; ++k) {
        double sr = 0., si = 0.;
        for (int n = 0; n < p; ++n) {
            if (x[n] == 0.) continue;
            int j = (int)(((long long)k * n) % p);
            sr += x[n] * c[j];
            si -= x[n] * s[j];
        }
        re[k] = sr; im[k] = si;
    }
    free(c); free(s);
}

/* R2C transform of the taps + AbsKernel (kernel.cu:620-646):
 * Hf[k] = tau * |DFT_p(t)[k]|, k = 0..p/2 */
void ofdk_spectrum_abs(const double *taps, int p, double tau, double *Hf)
{
    int nb = p / 2 + 1;
    double *re = (double *)malloc(sizeof(double) * (size_t)nb);
    double *im = (double *)malloc(sizeof(double) * (size_t)nb);
    dft_real(taps, p, nb, re, im);
    for (int k = 0; k < nb; ++k) Hf[k] = tau * sqrt(re[k] * re[k] + im[k] * im[k]);
    free(re); free(im);
}

void ofdk_spectrum_complex(const double *h, int p, double *H_re, double *H_im)
{
    dft_real(h, p, p, H_re, H_im);
}

/* ============================== geometry tables ============================== */

/* CalcSinCos (kernel.cu:430-477): beta_k = k*delta_f + rot; sin/cos in double.
 * The reference forms k*delta_f in float (kernel.cu:438, int*float); here the
 * angle is formed in double from the float inputs (FP64 geometry). */
void ofdk_sincos(const ofdk_geom *g, double *sinb, double *cosb)
{
    for (int b = 0; b < g->Nang; ++b) {
        double a = (double)b * (double)g->dbeta + (double)g->rot;
        sinb[b] = sin(a);
        cosb[b] = cos(a);
    }
}

/* nr_of_Z_blocks = ceil((rec_Z - 1)/B)  (kernel.cu:398-399).  When (Nz-1) is a
 * multiple of B this drops the last slice (and gives 0 slabs for Nz = 1). */
int ofdk_slab_count_reference(int Nz, int B)
{
    float nr_Z = (float)(Nz - 1) / (float)B;
    return (int)ceilf(nr_Z);
}

int ofdk_slab_count(int Nz, int B)
{
    return (Nz + B - 1) / B;
}

/* Voxel coordinates in float as FBPConeKernel forms them (kernel.cu:82-88):
 * a = scale * (idx + d) with d = -centre, then the sign-bit XOR of the axis
 * flips (kernel.cu:501-506).  For z, the per-slab offset of
 * CalculateReconstruction (kernel.cu:821) with the centring correction of
 * Initialize (kernel.cu:398-400), evaluated in the reference's mixed
 * float/double order; slabs beyond the reference count use the same formula. */
void ofdk_axis_tables(const ofdk_geom *g, float *ax, float *ay, float *az)
{
    const float dX = -g->cx, dY = -g->cy, dZ = -g->cz;
    for (int x = 0; x < g->Nx; ++x) {
        float v = g->scale * ((float)x + dX);
        ax[x] = g->flipX ? -v : v;
    }
    for (int y = 0; y < g->Ny; ++y) {
        float v = g->scale * ((float)y + dY);
        ay[y] = g->flipY ? -v : v;
    }
    const int B = g->B;
    const float nr_Z = (float)(g->Nz - 1) / (float)B;
    const int nZref = (int)ceilf(nr_Z);
    const float dcorZ = ((float)nZref - nr_Z) * (float)B / 2.0f;
    const int nslab = ofdk_slab_count(g->Nz, B);
    for (int n = 0; n < nslab; ++n) {
        float dZn = (float)((double)dZ + ((double)n - (double)(float)nZref / 2.0) * (double)B
                            + (double)dcorZ);
        for (int z = 0; z < B; ++z) {
            int gz = n * B + z;
            if (gz >= g->Nz) break;
            float v = g->scale * ((float)z + dZn);
            az[gz] = g->flipZ ? -v : v;
        }
    }
}

/* CollectionKernel (kernel.cu:203-215): slab += partial */
void ofdk_accumulate_f32(size_t n, const float *partial, float *slab)
{
    for (size_t i = 0; i < n; ++i) slab[i] += partial[i];
}

/* ============================ precision instances ============================ */

#define REAL float
#define SUF(x) x##_f32
#define OFDK_F64 0
#include "fdk_oracle_real.h"
#undef REAL
#undef SUF
#undef OFDK_F64

#define REAL double
#define SUF(x) x##_f64
#define OFDK_F64 1
#include "fdk_oracle_real.h"
#undef REAL
#undef SUF
#undef OFDK_F64

/* Direct convolution restatement of the filter for the pin of the FFT path:
 * out[m] = tau * p * sum_n x[n] * t(m - n), t(j) = h(j)/2 on j in [-Nu/2, Nu/2-1]. */
void ofdk_filter_direct_f64(const ofdk_geom *g, int kind, double *sino)
{
    const int Nu = g->Nu, p = ofdk_padded_length(Nu);
    const double tau = (double)g->tau;
    double *t = (double *)malloc(sizeof(double) * (size_t)(2 * Nu + 1));
    for (int j = -Nu; j <= Nu; ++j) {
        double h = (j >= -Nu / 2 && j <= Nu / 2 - 1)
                       ? (kind == 1 ? ofdk_shepplogan(j, tau) : ofdk_ramlak(j, tau)) / 2.
                       : 0.;
        t[j + Nu] = h;
    }
    const size_t rows = (size_t)g->Nv * g->Nang;
    #pragma omp parallel
    {
        double *tmp = (double *)malloc(sizeof(double) * (size_t)Nu);
        #pragma omp for schedule(static)
        for (size_t r = 0; r < rows; ++r) {
            double *row = sino + r * Nu;
            for (int m = 0; m < Nu; ++m) {
                double s = 0.;
                for (int n = 0; n < Nu; ++n) s += row[n] * t[m - n + Nu];
                tmp[m] = tau * (double)p * s;
            }
            memcpy(row, tmp, sizeof(double) * (size_t)Nu);
        }
        free(tmp);
    }
    free(t);
}

/* d<G, y>/dh for y = crop(p * circconv(pad(q), h)):
 * dh[j] = sum_rows sum_{m<Nu} G[m] * p * qpad[(m - j) mod p]. */
void ofdk_tap_gradient_f64(const ofdk_geom *g, const double *q, const double *G,
                           double *dh)
{
    const int Nu = g->Nu, p = ofdk_padded_length(Nu);
    const size_t rows = (size_t)g->Nv * g->Nang;
    for (int j = 0; j < p; ++j) dh[j] = 0.;
    #pragma omp parallel
    {
        double *loc = (double *)calloc((size_t)p, sizeof(double));
        #pragma omp for schedule(static)
        for (size_t r = 0; r < rows; ++r) {
            const double *qr = q + r * Nu, *Gr = G + r * Nu;
            for (int j = 0; j < p; ++j) {
                double s = 0.;
                for (int m = 0; m < Nu; ++m) {
                    int k = ((m - j) % p + p) % p;
                    if (k < Nu) s += Gr[m] * qr[k];
                }
                loc[j] += s * (double)p;
            }
        }
        #pragma omp critical
        for (int j = 0; j < p; ++j) dh[j] += loc[j];
        free(loc);
    }
}

/* ============================== synthetic data =============================== */

static void rot_z(double deg, double *R)
{
    double a = deg * M_PI / 180.0, c = cos(a), s = sin(a);
    /* rows = rotated x and y axes */
    R[0] = c;  R[1] = s;  R[2] = 0;
    R[3] = -s; R[4] = c;  R[5] = 0;
    R[6] = 0;  R[7] = 0;  R[8] = 1;
}

/* 3D Shepp-Logan head phantom: the 10-ellipsoid set of the 3D extension of
 * Shepp & Logan (1974) described by Kak & Slaney (ch. 3), original densities.
 * Columns: rho, semi-axes (a,b,c), centre (x,y,z), rotation about z (deg). */
int ofdk_shepp_logan(double half_fov, ofdk_ellipsoid *out)
{
    static const double T[10][8] = {
        { 1.00, .6900, .9200, .810,  0.00,  0.0000,  0.00,   0},
        {-0.98, .6624, .8740, .780,  0.00, -0.0184,  0.00,   0},
        {-0.02, .1100, .3100, .220,  0.22,  0.0000,  0.00, -18},
        {-0.02, .1600, .4100, .280, -0.22,  0.0000,  0.00,  18},
        { 0.01, .2100, .2500, .410,  0.00,  0.3500, -0.15,   0},
        { 0.01, .0460, .0460, .050,  0.00,  0.1000,  0.25,   0},
        { 0.01, .0460, .0460, .050,  0.00, -0.1000,  0.25,   0},
        { 0.01, .0460, .0230, .050, -0.08, -0.6050,  0.00,   0},
        { 0.01, .0230, .0230, .020,  0.00, -0.6060,  0.00,   0},
        { 0.01, .0230, .0460, .020,  0.06, -0.6050,  0.00,   0},
    };
    for (int i = 0; i < 10; ++i) {
        out[i].rho = T[i][0];
        for (int k = 0; k < 3; ++k) {
            out[i].ax[k] = T[i][1 + k] * half_fov;
            out[i].c[k] = T[i][4 + k] * half_fov;
        }
        rot_z(T[i][7], out[i].R);
    }
    return 10;
}

/* splitmix64 (Steele, Lea & Flood 2014) -> uniform double in [0,1) */
static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double unif(uint64_t *s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

int ofdk_random_ellipsoids(uint64_t seed, int n, double fov, ofdk_ellipsoid *out)
{
    uint64_t st = seed;
    const double rad = 0.4 * fov;
    for (int i = 0; i < n; ++i) {
        double c[3];
        do {  /* rejection sampling in the cube: uniform in the ball */
            for (int k = 0; k < 3; ++k) c[k] = (2.0 * unif(&st) - 1.0) * rad;
        } while (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] > rad * rad);
        for (int k = 0; k < 3; ++k) {
            out[i].c[k] = c[k];
            out[i].ax[k] = (0.02 + 0.23 * unif(&st)) * fov;
        }
        out[i].rho = 0.1 + 0.9 * unif(&st);
        /* uniform random rotation: Shoemake (1992) unit quaternion */
        double u1 = unif(&st), u2 = unif(&st), u3 = unif(&st);
        double a = sqrt(1 - u1), b = sqrt(u1);
        double qw = a * sin(2 * M_PI * u2), qx = a * cos(2 * M_PI * u2);
        double qy = b * sin(2 * M_PI * u3), qz = b * cos(2 * M_PI * u3);
        double *R = out[i].R;
        R[0] = 1 - 2 * (qy * qy + qz * qz); R[1] = 2 * (qx * qy - qz * qw); R[2] = 2 * (qx * qz + qy * qw);
        R[3] = 2 * (qx * qy + qz * qw); R[4] = 1 - 2 * (qx * qx + qz * qz); R[5] = 2 * (qy * qz - qx * qw);
        R[6] = 2 * (qx * qz - qy * qw); R[7] = 2 * (qy * qz + qx * qw); R[8] = 1 - 2 * (qx * qx + qy * qy);
    }
    return n;
}

/* chord length of the line s + t*d (all t) through ellipsoid e, in units of |d| */
static double chord(const ofdk_ellipsoid *e, const double s[3], const double d[3])
{
    double m[3], w[3];
    for (int k = 0; k < 3; ++k) {
        double ps = 0, pd = 0;
        for (int j = 0; j < 3; ++j) {
            ps += e->R[3 * k + j] * (s[j] - e->c[j]);
            pd += e->R[3 * k + j] * d[j];
        }
        m[k] = ps / e->ax[k];
        w[k] = pd / e->ax[k];
    }
    double A = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    double Bq = m[0] * w[0] + m[1] * w[1] + m[2] * w[2];
    double C = m[0] * m[0] + m[1] * m[1] + m[2] * m[2] - 1.0;
    double disc = Bq * Bq - A * C;
    if (disc <= 0) return 0.;
    return 2.0 * sqrt(disc) / A;   /* t2 - t1 */
}

/* Source at D*(cos b, sin b, 0); virtual detector through the isocentre with
 * pixel (u,v) at (u-cX)*(-sin b, cos b, 0) + (0,0,v-cZ): the frame in which
 * FBPConeKernel's S = a.(cos,sin), T = a.(-sin,cos), U = D/(D-S) hold. */
void ofdk_project_ellipsoids(const ofdk_geom *g, const ofdk_ellipsoid *e, int n,
                             double *sino)
{
    double *sb = (double *)malloc(sizeof(double) * (size_t)g->Nang);
    double *cb = (double *)malloc(sizeof(double) * (size_t)g->Nang);
    ofdk_sincos(g, sb, cb);
    #pragma omp parallel for collapse(2) schedule(static)
    for (int v = 0; v < g->Nv; ++v)
        for (int b = 0; b < g->Nang; ++b) {
            double s[3] = {g->D * cb[b], g->D * sb[b], 0.};
            for (int u = 0; u < g->Nu; ++u) {
                double q[3] = {-(u - (double)g->cX) * sb[b], (u - (double)g->cX) * cb[b],
                               v - (double)g->cZ};
                double d[3] = {q[0] - s[0], q[1] - s[1], q[2] - s[2]};
                double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                double acc = 0.;
                for (int i = 0; i < n; ++i) acc += e[i].rho * chord(&e[i], s, d);
                sino[((size_t)v * g->Nang + b) * g->Nu + u] = acc * len * (double)g->tau;
            }
        }
    free(sb); free(cb);
}

void ofdk_voxelize_ellipsoids(const ofdk_geom *g, const ofdk_ellipsoid *e, int n,
                              double *vol)
{
    #pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < g->Nz; ++z)
        for (int y = 0; y < g->Ny; ++y)
            for (int x = 0; x < g->Nx; ++x) {
                double a[3] = {(double)g->scale * (x - (double)g->cx),
                               (double)g->scale * (y - (double)g->cy),
                               (double)g->scale * (z - 0.5 * (g->Nz - 1) - (double)g->cz)};
                if (g->flipX) a[0] = -a[0];
                if (g->flipY) a[1] = -a[1];
                if (g->flipZ) a[2] = -a[2];
                double acc = 0.;
                for (int i = 0; i < n; ++i) {
                    double r = 0.;
                    for (int k = 0; k < 3; ++k) {
                        double t = 0.;
                        for (int j = 0; j < 3; ++j) t += e[i].R[3 * k + j] * (a[j] - e[i].c[j]);
                        t /= e[i].ax[k];
                        r += t * t;
                    }
                    if (r <= 1.0) acc += e[i].rho;
                }
                vol[((size_t)z * g->Ny + y) * g->Nx + x] = acc;
            }
}

/* ---- counter-based noise ------------------------------------------------------- */
static inline uint32_t mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * b) >> 32);
}

void ofdk_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = mulhi32(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = mulhi32(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 0x1.0p-53;
}

/* Poisson(lam) for pixel idx: inversion for lam < 10, otherwise the PTRS
 * transformed rejection of Hoermann (1993).  Draw i of pixel idx uses the
 * Philox block with counter (idx_lo, idx_hi, i, 0) and key (seed_lo, seed_hi).
 * At most 64 draws; the (never observed) exhaustion case returns round(lam). */
double ofdk_poisson_draw(double lam, uint64_t seed, uint64_t idx)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0, 0}, r[4];
    if (lam <= 0) return 0;
    if (lam < 10.0) {
        ofdk_philox4x32(ctr, key, r);
        double U = u53(r[0], r[1]);
        double pk = exp(-lam), F = pk;
        int k = 0;
        while (U > F && k < 1000) { ++k; pk *= lam / k; F += pk; }
        return k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (uint32_t it = 0; it < 64; ++it) {
        ctr[2] = it;
        ofdk_philox4x32(ctr, key, r);
        double U = u53(r[0], r[1]) - 0.5;
        double V = u53(r[2], r[3]);
        double us = 0.5 - fabs(U);
        double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0))
            return k;
    }
    return floor(lam + 0.5);
}

void ofdk_poisson_log(size_t n, double I0, uint64_t seed, float *p)
{
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
        double lam = I0 * exp(-(double)p[i]);
        double cnt = ofdk_poisson_draw(lam, seed, (uint64_t)i);
        if (cnt < 1.0) cnt = 1.0;
        p[i] = (float)(-log(cnt / I0));
    }
}
