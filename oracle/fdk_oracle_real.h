/*
 * fdk_oracle_real.h -- precision-generic part of the CPU oracle.
 * TEST INFRASTRUCTURE ONLY (see fdk_oracle.h).
 *
 * Included twice by fdk_oracle.c with
 *   REAL = float,  SUF(x) = x##_f32, OFDK_F64 = 0   (FP32 mirror)
 *   REAL = double, SUF(x) = x##_f64, OFDK_F64 = 1   (FP64 oracle)
 * Compiled with -ffp-contract=off so every float expression is evaluated
 * exactly as written.  Where nvcc's default contraction (-fmad=true) turns a
 * product feeding an addition of the reference's device code into one fused
 * multiply-add, this restatement writes that fma explicitly (SUF(ffma)); all
 * other operations are single roundings in the reference's order.
 */

static inline REAL SUF(ffma)(REAL a, REAL b, REAL c)
{
#if OFDK_F64
    return fma(a, b, c);
#else
    return fmaf(a, b, c);
#endif
}

/* ---- voxel coordinates ------------------------------------------------------- */
/* FP32 mirror: the float values the reference kernel forms (kernel.cu:82-88,
 * slab recentring kernel.cu:398-400 and :821).  FP64 oracle: the same
 * coordinates in exact arithmetic, aZ = scale*(g - (Nz-1)/2 - cz). */
static void SUF(coords)(const ofdk_geom *g, REAL *ax, REAL *ay, REAL *az)
{
#if OFDK_F64
    for (int x = 0; x < g->Nx; ++x) {
        double v = (double)g->scale * ((double)x - (double)g->cx);
        ax[x] = g->flipX ? -v : v;
    }
    for (int y = 0; y < g->Ny; ++y) {
        double v = (double)g->scale * ((double)y - (double)g->cy);
        ay[y] = g->flipY ? -v : v;
    }
    for (int z = 0; z < g->Nz; ++z) {
        double v = (double)g->scale *
                   ((double)z - 0.5 * (double)(g->Nz - 1) - (double)g->cz);
        az[z] = g->flipZ ? -v : v;
    }
#else
    ofdk_axis_tables(g, ax, ay, az);
#endif
}

/* Cylindrical field-of-view mask of kernel.cu:100, sqrt(aX^2+aY^2) < Nu/2,
 * evaluated as aX^2 + aY^2 < (Nu/2)^2 in double (identical decision except
 * within one float ulp of the rim; every implementation here uses this form). */
static inline int SUF(inside)(const ofdk_geom *g, REAL ax, REAL ay)
{
    double r2 = (double)ax * (double)ax + (double)ay * (double)ay;
    double R = 0.5 * (double)g->Nu;
    return r2 < R * R;
}

/* ---- cosine pre-weighting, RB_comma3DKernel (kernel.cu:179-201) -------------- */
void SUF(ofdk_weight)(const ofdk_geom *g, REAL *sino)
{
    const REAL D = (REAL)g->D;
    #pragma omp parallel for schedule(static)
    for (int v = 0; v < g->Nv; ++v) {
        REAL cs = (REAL)v - (REAL)g->cZ;
        for (int b = 0; b < g->Nang; ++b) {
            REAL *row = sino + ((size_t)v * g->Nang + b) * g->Nu;
            for (int u = 0; u < g->Nu; ++u) {
                REAL ce = (REAL)u - (REAL)g->cX;
                /* ce*ce + D*D + cs*cs, contracted left to right */
                REAL t = SUF(ffma)(cs, cs, SUF(ffma)(ce, ce, D * D));
                row[u] *= D / (REAL)sqrt((double)t);
            }
        }
    }
}

/* ---- radix-2 FFT (Cooley-Tukey, iterative, bit-reversed input) ---------------- */
/* Textbook decimation-in-time transform; twiddles exp(-+2 pi i k/p) are formed in
 * double from the exact index and rounded once to REAL. */
typedef struct { int p; REAL *wr, *wi; int *rev; } SUF(fftplan);

static void SUF(fft_init)(SUF(fftplan) *pl, int p)
{
    pl->p = p;
    pl->wr = (REAL *)malloc(sizeof(REAL) * (size_t)(p / 2 + 1));
    pl->wi = (REAL *)malloc(sizeof(REAL) * (size_t)(p / 2 + 1));
    pl->rev = (int *)malloc(sizeof(int) * (size_t)p);
    for (int k = 0; k < p / 2; ++k) {
        double a = 2.0 * M_PI * (double)k / (double)p;
        pl->wr[k] = (REAL)cos(a);
        pl->wi[k] = (REAL)(-sin(a));
    }
    int lg = 0;
    while ((1 << lg) < p) ++lg;
    for (int i = 0; i < p; ++i) {
        int r = 0;
        for (int b = 0; b < lg; ++b)
            if (i & (1 << b)) r |= 1 << (lg - 1 - b);
        pl->rev[i] = r;
    }
}

static void SUF(fft_free)(SUF(fftplan) *pl)
{
    free(pl->wr); free(pl->wi); free(pl->rev);
}

/* in-place, sign = -1 forward, +1 inverse (unnormalised) */
static void SUF(fft_exec)(const SUF(fftplan) *pl, REAL *re, REAL *im, int sign)
{
    const int p = pl->p;
    for (int i = 0; i < p; ++i) {
        int j = pl->rev[i];
        if (j > i) {
            REAL t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= p; len <<= 1) {
        int half = len >> 1, step = p / len;
        for (int s = 0; s < p; s += len) {
            for (int k = 0; k < half; ++k) {
                REAL wr = pl->wr[k * step];
                REAL wi = sign < 0 ? pl->wi[k * step] : -pl->wi[k * step];
                int a = s + k, b = a + half;
                REAL tr = re[b] * wr - im[b] * wi;
                REAL ti = re[b] * wi + im[b] * wr;
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] = re[a] + tr; im[a] = im[a] + ti;
            }
        }
    }
}

/* ---- row filtering, fConvolution + FiltrationKernel (kernel.cu:557-781) ------ */
/* Every row of length Nu is zero-padded to p, transformed, multiplied bin-wise by
 * the spectrum, inverse transformed WITHOUT the 1/p normalisation and cropped back
 * to Nu samples in place (the 1/p is applied by the backprojector, kernel.cu:575,
 * :113).  Hf holds p/2+1 real bins (kernel.cu:579); the even extension
 * Hf[p-k] = Hf[k] is the real-to-complex transform's Hermitian symmetry. */
static void SUF(filter_generic)(const ofdk_geom *g, const double *Hre,
                                const double *Him, int half_real, REAL *sino)
{
    const int p = ofdk_padded_length(g->Nu);
    const size_t rows = (size_t)g->Nv * g->Nang;
    SUF(fftplan) pl;
    SUF(fft_init)(&pl, p);
    REAL *Hr = (REAL *)malloc(sizeof(REAL) * (size_t)p);
    REAL *Hi = (REAL *)malloc(sizeof(REAL) * (size_t)p);
    for (int k = 0; k < p; ++k) {
        if (half_real) {
            int kk = k <= p / 2 ? k : p - k;
            Hr[k] = (REAL)Hre[kk]; Hi[k] = 0;
        } else {
            Hr[k] = (REAL)Hre[k]; Hi[k] = (REAL)Him[k];
        }
    }
    #pragma omp parallel
    {
        REAL *re = (REAL *)malloc(sizeof(REAL) * (size_t)p);
        REAL *im = (REAL *)malloc(sizeof(REAL) * (size_t)p);
        #pragma omp for schedule(static)
        for (size_t r = 0; r < rows; ++r) {
            REAL *row = sino + r * g->Nu;
            for (int i = 0; i < p; ++i) { re[i] = i < g->Nu ? row[i] : 0; im[i] = 0; }
            SUF(fft_exec)(&pl, re, im, -1);
            for (int k = 0; k < p; ++k) {
                REAL a = re[k], b = im[k];
                re[k] = a * Hr[k] - b * Hi[k];
                im[k] = a * Hi[k] + b * Hr[k];
            }
            SUF(fft_exec)(&pl, re, im, +1);
            for (int i = 0; i < g->Nu; ++i) row[i] = re[i];
        }
        free(re); free(im);
    }
    free(Hr); free(Hi);
    SUF(fft_free)(&pl);
}

void SUF(ofdk_filter)(const ofdk_geom *g, const double *Hf, REAL *sino)
{
    SUF(filter_generic)(g, Hf, NULL, 1, sino);
}

void SUF(ofdk_filter_complex)(const ofdk_geom *g, const double *Hre,
                              const double *Him, REAL *sino)
{
    SUF(filter_generic)(g, Hre, Him, 0, sino);
}

/* ---- detector interpolation, fQ3D (kernel.cu:121-177) ------------------------- */
/* Value of the filtered projection `beta` at (p + cX, f + cZ); zero outside the
 * closed detector rectangle [0,Nu-1] x [0,Nv-1]; ceil/floor formulation of the
 * reference (interpolation from the upper sample towards the lower one).
 * row0 > 0: sino holds rows [row0, row0 + Nv) of a taller detector and g->cZ is
 * relative to row0; the row position is evaluated on the taller detector, as
 * its full reconstruction does (f + cZ + row0), then shifted down by row0.  The
 * shift is exact (the difference is a multiple of the position's ulp below
 * the position itself), so the positions are bit-identical to the full
 * detector's. */
static inline REAL SUF(fq3d)(const ofdk_geom *g, const REAL *sino, int beta,
                             REAL T, REAL U, REAL aZ, int row0)
{
    /* fQ3D(T*U, aZ*U, ...) then _p + cX, _f + cZ: contracted to fmas */
    REAL scP = SUF(ffma)(T, U, (REAL)g->cX);
    REAL scF = row0 ? SUF(ffma)(aZ, U, (REAL)g->cZ + (REAL)row0) - (REAL)row0
                    : SUF(ffma)(aZ, U, (REAL)g->cZ);
    if (scP < 0 || scP > (REAL)(g->Nu - 1)) return 0;
    if (scF < 0 || scF > (REAL)(g->Nv - 1)) return 0;
    int af = (int)ceil(scF), bf = (int)floor(scF);
    int a = (int)ceil(scP), b = (int)floor(scP);
    const REAL *rowa = sino + ((size_t)af * g->Nang + beta) * g->Nu;
    const REAL *rowb = sino + ((size_t)bf * g->Nang + beta) * g->Nu;
    if (af != bf) {
        REAL va, vb;
        if (a != b) {
            va = SUF(ffma)(rowa[b] - rowa[a], (REAL)a - scP, rowa[a]);
            vb = SUF(ffma)(rowb[b] - rowb[a], (REAL)a - scP, rowb[a]);
        } else {
            va = rowa[a];
            vb = rowb[b];
        }
        return SUF(ffma)(vb - va, (REAL)af - scF, va);
    }
    if (a != b) return SUF(ffma)(rowa[b] - rowa[a], (REAL)a - scP, rowa[a]);
    return rowa[a];
}

/* ---- backprojection, FBPConeKernel (kernel.cu:68-119) -------------------------- */
/* vol[z][y][x] over the whole volume (all slabs, global slice index z).
 * Per voxel, angles are summed in ascending order, then scaled by
 * delta_f * (1/p) (kernel.cu:113, scaling from kernel.cu:575). */
void SUF(ofdk_backproject)(const ofdk_geom *g, const REAL *sino, REAL *vol)
{
    const int p = ofdk_padded_length(g->Nu);
    double *sb = (double *)malloc(sizeof(double) * (size_t)g->Nang);
    double *cb = (double *)malloc(sizeof(double) * (size_t)g->Nang);
    ofdk_sincos(g, sb, cb);
    REAL *ds = (REAL *)malloc(sizeof(REAL) * (size_t)g->Nang);
    REAL *dc = (REAL *)malloc(sizeof(REAL) * (size_t)g->Nang);
    for (int b = 0; b < g->Nang; ++b) { ds[b] = (REAL)sb[b]; dc[b] = (REAL)cb[b]; }
    REAL *ax = (REAL *)malloc(sizeof(REAL) * (size_t)g->Nx);
    REAL *ay = (REAL *)malloc(sizeof(REAL) * (size_t)g->Ny);
    REAL *az = (REAL *)malloc(sizeof(REAL) * (size_t)g->Nz);
    SUF(coords)(g, ax, ay, az);
    const REAL D = (REAL)g->D;
    const REAL fscale = (REAL)g->dbeta * ((REAL)1 / (REAL)p);
    #pragma omp parallel
    {
        REAL *acc = (REAL *)malloc(sizeof(REAL) * (size_t)g->Nx);
        #pragma omp for schedule(dynamic, 1) collapse(2)
        for (int z = 0; z < g->Nz; ++z) {
            for (int y = 0; y < g->Ny; ++y) {
                const REAL aY = ay[y], aZ = az[z];
                for (int x = 0; x < g->Nx; ++x) acc[x] = 0;
                for (int b = 0; b < g->Nang; ++b) {
                    const REAL s = ds[b], c = dc[b];
                    for (int x = 0; x < g->Nx; ++x) {
                        const REAL aX = ax[x];
                        if (!SUF(inside)(g, aX, aY)) continue;
                        REAL S = SUF(ffma)(aX, c, aY * s);
                        REAL T = SUF(ffma)(aY, c, -(aX * s));
                        REAL U = D / (D - S);
                        acc[x] = SUF(ffma)(U * U, SUF(fq3d)(g, sino, b, T, U, aZ, 0), acc[x]);
                    }
                }
                REAL *out = vol + ((size_t)z * g->Ny + y) * g->Nx;
                for (int x = 0; x < g->Nx; ++x)
                    out[x] = SUF(inside)(g, ax[x], aY) ? acc[x] * fscale : 0;
            }
        }
        free(acc);
    }
    free(sb); free(cb); free(ds); free(dc); free(ax); free(ay); free(az);
}

/* Same arithmetic as ofdk_backproject for an explicit list of voxels
 * (xyz[3*i+0..2] = x, y, global z); used for voxel subsamples of volumes too
 * large to reconstruct on the CPU.  row0: sino holds detector rows
 * [row0, row0 + Nv) of a taller detector, g->cZ relative to row0 (see fq3d). */
void SUF(ofdk_backproject_points)(const ofdk_geom *g, const REAL *sino, int row0, int npts,
                                  const int32_t *xyz, REAL *out)
{
    const int p = ofdk_padded_length(g->Nu);
    double *sb = (double *)malloc(sizeof(double) * (size_t)g->Nang);
    double *cb = (double *)malloc(sizeof(double) * (size_t)g->Nang);
    ofdk_sincos(g, sb, cb);
    REAL *ax = (REAL *)malloc(sizeof(REAL) * (size_t)g->Nx);
    REAL *ay = (REAL *)malloc(sizeof(REAL) * (size_t)g->Ny);
    REAL *az = (REAL *)malloc(sizeof(REAL) * (size_t)g->Nz);
    SUF(coords)(g, ax, ay, az);
    const REAL D = (REAL)g->D;
    const REAL fscale = (REAL)g->dbeta * ((REAL)1 / (REAL)p);
    #pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < npts; ++i) {
        const REAL aX = ax[xyz[3 * i]], aY = ay[xyz[3 * i + 1]], aZ = az[xyz[3 * i + 2]];
        if (!SUF(inside)(g, aX, aY)) { out[i] = 0; continue; }
        REAL acc = 0;
        for (int b = 0; b < g->Nang; ++b) {
            const REAL s = (REAL)sb[b], c = (REAL)cb[b];
            REAL S = SUF(ffma)(aX, c, aY * s);
            REAL T = SUF(ffma)(aY, c, -(aX * s));
            REAL U = D / (D - S);
            acc = SUF(ffma)(U * U, SUF(fq3d)(g, sino, b, T, U, aZ, row0), acc);
        }
        out[i] = acc * fscale;
    }
    free(sb); free(cb); free(ax); free(ay); free(az);
}

/* ---- adjoint of the backprojector ---------------------------------------------- */
/* Exact transpose of ofdk_backproject: every voxel value, weighted by
 * U^2 * delta_f / p, is splatted onto the (up to) four detector samples that
 * fq3d interpolates, with the same bilinear weights.  Parallel over angles, so
 * each thread owns whole projections and the result is deterministic. */
void SUF(ofdk_forward_project)(const ofdk_geom *g, const REAL *vol, REAL *sino)
{
    const int p = ofdk_padded_length(g->Nu);
    double *sb = (double *)malloc(sizeof(double) * (size_t)g->Nang);
    double *cb = (double *)malloc(sizeof(double) * (size_t)g->Nang);
    ofdk_sincos(g, sb, cb);
    REAL *ax = (REAL *)malloc(sizeof(REAL) * (size_t)g->Nx);
    REAL *ay = (REAL *)malloc(sizeof(REAL) * (size_t)g->Ny);
    REAL *az = (REAL *)malloc(sizeof(REAL) * (size_t)g->Nz);
    SUF(coords)(g, ax, ay, az);
    const REAL D = (REAL)g->D;
    const REAL fscale = (REAL)g->dbeta * ((REAL)1 / (REAL)p);
    const size_t nsin = (size_t)g->Nv * g->Nang * g->Nu;
    for (size_t i = 0; i < nsin; ++i) sino[i] = 0;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < g->Nang; ++b) {
        const REAL s = (REAL)sb[b], c = (REAL)cb[b];
        for (int z = 0; z < g->Nz; ++z)
        for (int y = 0; y < g->Ny; ++y)
        for (int x = 0; x < g->Nx; ++x) {
            const REAL aX = ax[x], aY = ay[y], aZ = az[z];
            if (!SUF(inside)(g, aX, aY)) continue;
            REAL val = vol[((size_t)z * g->Ny + y) * g->Nx + x];
            if (val == 0) continue;
            REAL S = SUF(ffma)(aX, c, aY * s);
            REAL T = SUF(ffma)(aY, c, -(aX * s));
            REAL U = D / (D - S);
            REAL scP = SUF(ffma)(T, U, (REAL)g->cX);
            REAL scF = SUF(ffma)(aZ, U, (REAL)g->cZ);
            if (scP < 0 || scP > (REAL)(g->Nu - 1)) continue;
            if (scF < 0 || scF > (REAL)(g->Nv - 1)) continue;
            int af = (int)ceil(scF), bf = (int)floor(scF);
            int a = (int)ceil(scP), bb = (int)floor(scP);
            REAL gu = -scP + (REAL)a;   /* weight of column bb */
            REAL gf = -scF + (REAL)af;  /* weight of row bf    */
            REAL w = U * U * val * fscale;
            REAL *ra = sino + ((size_t)af * g->Nang + b) * g->Nu;
            REAL *rb = sino + ((size_t)bf * g->Nang + b) * g->Nu;
            if (af != bf) {
                if (a != bb) {
                    ra[a]  += w * (1 - gf) * (1 - gu);
                    ra[bb] += w * (1 - gf) * gu;
                    rb[a]  += w * gf * (1 - gu);
                    rb[bb] += w * gf * gu;
                } else {
                    ra[a] += w * (1 - gf);
                    rb[a] += w * gf;
                }
            } else {
                if (a != bb) {
                    ra[a]  += w * (1 - gu);
                    ra[bb] += w * gu;
                } else {
                    ra[a] += w;
                }
            }
        }
    }
    free(sb); free(cb); free(ax); free(ay); free(az);
}

/* ---- complete FDK as CalculateReconstruction runs it (kernel.cu:804-897) ------ */
void SUF(ofdk_fdk)(const ofdk_geom *g, int kind, REAL *sino, REAL *vol)
{
    const int p = ofdk_padded_length(g->Nu);
    double *taps = (double *)malloc(sizeof(double) * (size_t)p);
    double *Hf = (double *)malloc(sizeof(double) * (size_t)(p / 2 + 1));
    ofdk_fill_taps(kind, g->Nu, (double)g->tau, p, taps);
    ofdk_spectrum_abs(taps, p, (double)g->tau, Hf);
    SUF(ofdk_weight)(g, sino);
    SUF(ofdk_filter)(g, Hf, sino);
    SUF(ofdk_backproject)(g, sino, vol);
    free(taps); free(Hf);
}
