/*
 * fdk_oracle.h -- CPU restatement of the reference FDK algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in tomokvas_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  Nothing in the product links it.
 *
 * Two instantiations of every numeric routine:
 *   *_f32  "FP32 mirror": the reference's float arithmetic and operation order
 *          (kernel.cu), OpenMP-parallel; also the timed CPU baseline.
 *   *_f64  "FP64 oracle": the same algorithm evaluated in double precision.
 *
 * Reference citations are paths inside Lukas-kV/TomoKvas (FBPG/...).
 *
 * Conventions shared with the reference engine (FBPG/kernel.cuh:25-61):
 *   projections  sinogram-major [v][beta][u]  (Nv x Nang x Nu, FBPG/mucat_file.cpp:244-267)
 *   volume       [z][y][x]                      (Nz x Ny x Nx)
 *   lengths      virtual detector pixels at the isocentre; tau = pixel pitch
 *
 * How each function is pinned (checked against something other than itself)
 * is listed in the header comment of tests/test_oracle.py.
 */
#ifndef TOMOKVAS_FDK_ORACLE_H
#define TOMOKVAS_FDK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry exactly as the reference engine receives it: the values that
 * FilteredBackProjection::SetSinogramProperties / SetReconstructionProperties /
 * SetScaleXYZ / SetAxisDirection / SetBlockSize take (FBPG/kernel.cuh:89-101),
 * all float at that boundary. */
typedef struct {
    int   Nu, Nang, Nv;     /* sin_X, sin_Y, sin_Z                              */
    float cX, cZ;           /* detector centre column / row                     */
    float dbeta;            /* delta_f: angular step (rad)                      */
    float D;                /* source-isocentre distance in virtual px          */
    float tau;              /* sampling: virtual pixel pitch                    */
    int   Nx, Ny, Nz;       /* rec_X, rec_Y, rec_Z                              */
    float cx, cy, cz;       /* centre_X/Y/Z passed to SetReconstructionProperties */
    float rot;              /* rotation offset (rad)                            */
    float scale;            /* voxel size in virtual px (SetScaleXYZ)           */
    int   flipX, flipY, flipZ;  /* SetAxisDirection                             */
    int   B;                /* slices per slab (SetBlockSize)                   */
} ofdk_geom;

/* ---- filter construction (FBPG/kernel.cu:570-655, 783-802) ---------------- */
int    ofdk_padded_length(int Nu);                       /* p = 2*nextpow2(Nu)  */
float  ofdk_ramlak_ref(int n, float tau);                /* float, as kernel.cu:792 */
double ofdk_ramlak(int n, double tau);                   /* same formula in double */
double ofdk_shepplogan(int n, double tau);               /* apodised extension     */
/* taps t[i] = h(i - Nu/2)/2, i in [0,Nu), zero-padded to p (kernel.cu:783-790).
 * kind 0 = Ram-Lak, 1 = Shepp-Logan. */
void   ofdk_fill_taps(int kind, int Nu, double tau, int p, double *taps);
/* Hf[k] = tau*|DFT_p(t)[k]|, k in [0, p/2]  (kernel.cu:620-646), direct FP64 DFT */
void   ofdk_spectrum_abs(const double *taps, int p, double tau, double *Hf);
/* H[k] = DFT_p(h)[k] (complex, k in [0,p)) for user taps h of length p        */
void   ofdk_spectrum_complex(const double *h, int p, double *H_re, double *H_im);

/* ---- geometry tables (FBPG/kernel.cu:430-477, 82-88, 398-400, 815-821) ---- */
void   ofdk_sincos(const ofdk_geom *g, double *sinb, double *cosb);
/* reference slab count ceil((Nz-1)/B) (kernel.cu:398-399) */
int    ofdk_slab_count_reference(int Nz, int B);
/* slab count that covers every slice: ceil(Nz/B) */
int    ofdk_slab_count(int Nz, int B);
/* per-axis voxel coordinates in virtual px after the axis flips, float as the
 * reference computes them; aZ uses the slab recentring of kernel.cu:821 */
void   ofdk_axis_tables(const ofdk_geom *g, float *ax, float *ay, float *az);

/* ---- pipeline stages, FP32 mirror -------------------------------------------- */
void ofdk_weight_f32(const ofdk_geom *g, float *sino);
void ofdk_filter_f32(const ofdk_geom *g, const double *Hf, float *sino);
void ofdk_filter_complex_f32(const ofdk_geom *g, const double *Hre, const double *Him,
                             float *sino);
void ofdk_backproject_f32(const ofdk_geom *g, const float *sino, float *vol);
void ofdk_backproject_points_f32(const ofdk_geom *g, const float *sino, int row0, int npts,
                                 const int32_t *xyz, float *out);
void ofdk_forward_project_f32(const ofdk_geom *g, const float *vol, float *sino);
void ofdk_accumulate_f32(size_t n, const float *partial, float *slab);
/* complete FDK: weight -> filter(kind) -> backproject, sino is consumed */
void ofdk_fdk_f32(const ofdk_geom *g, int kind, float *sino, float *vol);

/* ---- pipeline stages, FP64 oracle -------------------------------------------- */
void ofdk_weight_f64(const ofdk_geom *g, double *sino);
void ofdk_filter_f64(const ofdk_geom *g, const double *Hf, double *sino);
void ofdk_filter_complex_f64(const ofdk_geom *g, const double *Hre, const double *Him,
                             double *sino);
void ofdk_backproject_f64(const ofdk_geom *g, const double *sino, double *vol);
void ofdk_backproject_points_f64(const ofdk_geom *g, const double *sino, int row0, int npts,
                                 const int32_t *xyz, double *out);
void ofdk_forward_project_f64(const ofdk_geom *g, const double *vol, double *sino);
void ofdk_fdk_f64(const ofdk_geom *g, int kind, double *sino, double *vol);
/* direct O(Nu^2) row convolution with g(n)=h(n)/2 on [-Nu/2, Nu/2-1], times tau*p
 * (equal to the FFT filter when the centred kernel has a positive real DFT) */
void ofdk_filter_direct_f64(const ofdk_geom *g, int kind, double *sino);
/* gradient of <G, filter_H(q)> w.r.t. the spatial taps h (length p): for every
 * row, d/dh[n] = sum_m G[m] * p * q[(m-n) mod p]   (real, length p)            */
void ofdk_tap_gradient_f64(const ofdk_geom *g, const double *q, const double *G,
                           double *dh);

/* ---- synthetic data ------------------------------------------------------------ */
typedef struct {
    double c[3];      /* centre, virtual px in the (flipped) voxel frame          */
    double ax[3];     /* semi-axes, virtual px                                    */
    double R[9];      /* rows = ellipsoid axes in world coordinates               */
    double rho;       /* density                                                  */
} ofdk_ellipsoid;

/* 3D Shepp-Logan, 10 ellipsoids; unit cube [-1,1] scaled by half_fov (px)     */
int  ofdk_shepp_logan(double half_fov, ofdk_ellipsoid *out);
/* n random ellipsoids from a seed (splitmix64): centres uniform in the sphere of
 * radius 0.4*fov, semi-axes U(0.02,0.25)*fov, density U(0.1,1), uniform random
 * rotation (Shoemake quaternion) */
int  ofdk_random_ellipsoids(uint64_t seed, int n, double fov, ofdk_ellipsoid *out);
/* exact line integrals (density * chord length * tau) for every detector pixel,
 * sinogram-major [v][beta][u], angles from ofdk_sincos */
void ofdk_project_ellipsoids(const ofdk_geom *g, const ofdk_ellipsoid *e, int n,
                             double *sino);
/* density at every voxel centre of the volume (point sampling) */
void ofdk_voxelize_ellipsoids(const ofdk_geom *g, const ofdk_ellipsoid *e, int n,
                              double *vol);
/* Philox4x32-10 (Salmon et al., SC'11) block for (counter, key) */
void ofdk_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Poisson(I0*exp(-p)) counts then -ln(max(n,1)/I0), in place, seeded */
void ofdk_poisson_log(size_t n, double I0, uint64_t seed, float *p);
/* one Poisson draw for pixel index idx (same streams as ofdk_poisson_log) */
double ofdk_poisson_draw(double lam, uint64_t seed, uint64_t idx);

#ifdef __cplusplus
}
#endif
#endif
