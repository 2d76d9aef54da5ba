// fdk_internal.hpp -- shared declarations of the MI355X FDK library (not installed).
//
// Device data layouts (all in HBM, one allocation each; 288 GB holds every
// configuration without tiling):
//   input stack   float [v][beta][u]          as the host hands it over
//                 (sinogram-major, FBPG/mucat_file.cpp:244-267)
//   stored stack  S     [beta][v][pitch]      filtered projections, angle-major,
//                 in the storage type S (FP32, FP16 / scale_beta, BF16, FP64);
//                 pitch = stack_pitch(Nu) >= Nu + 68, zero past column Nu-1,
//                 so the backprojector's staged boxes stay inside a row.
//   volume        float [z][y][x]             the full volume, slabs are z-ranges
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tomokvas/fdk.h"

namespace tkv {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define TKV_HIP(call)                                            \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return ::tkv::hip_fail(e_, #call); \
    } while (0)

// ---------------------------------------------------------------- geometry
// Float-valued parameters exactly as the reference engine's setters receive them
// (FBPG/kernel.cuh:89-101).
struct Geometry {
    int Nu = 0, Nang = 0, Nv = 0;
    float cX = 0, cZ = 0, dbeta = 0, D = 0, tau = 0;
    int Nx = 0, Ny = 0, Nz = 0;
    float cx = 0, cy = 0, cz = 0, rot = 0, scale = 1;
    int flipX = 0, flipY = 0, flipZ = 0;
    int B = 16;
};

int padded_length(int Nu);
// FP64 angle tables: angle_k = k*dbeta + rot (kernel.cu:430-441)
void angle_tables(const Geometry &g, std::vector<double> &s, std::vector<double> &c);
// float voxel coordinates as FBPConeKernel forms them (kernel.cu:82-88, 398-400,
// 821); az is extended linearly to nz_pad >= Nz entries
void axis_tables_f32(const Geometry &g, int nz_pad, std::vector<float> &ax,
                     std::vector<float> &ay, std::vector<float> &az);
void axis_tables_f64(const Geometry &g, int nz_pad, std::vector<double> &ax,
                     std::vector<double> &ay, std::vector<double> &az);
// reference slab count ceil((Nz-1)/B) and the covering count ceil(Nz/B)
int slab_count_reference(int Nz, int B);
int slab_count(int Nz, int B);

// filter spectrum in FP64 (kernel.cu:570-655, 783-802):
//   kind RAMLAK / SHEPPLOGAN: H[k] = tau*|DFT_p(t)[k]| real, k in [0,p)
//   kind USER: H[k] = DFT_p(h)[k] complex for the caller's p spatial taps
void build_spectrum(int kind, const Geometry &g, const std::vector<double> &user_taps,
                    std::vector<double> &Hre, std::vector<double> &Him);
// in-place forward DFT X[j] = sum_k x[k] e^{-2 pi i k j / n} in double, n a
// power of two (iterative radix-2; twiddles from the exact index)
void fft_pow2(std::vector<double> &re, std::vector<double> &im);

// ---------------------------------------------------------------- kernel launchers
// A dispatch holds its size in work-items in a 32-bit field per dimension, so a
// one-dimensional launch covers at most 2^32 - 1 threads; the launch call does
// not reject a larger grid, which then falls short of its range.  Launchers
// check their grids with this, or loop over a capped grid.
inline bool grid_fits(long long blocks, int threads)
{
    return blocks >= 0 && (unsigned long long)blocks * (unsigned long long)threads <= 0xffffffffull;
}

struct FilterLaunch {
    const void *in;         // [rows][Nu] rows = Nv * nang, Real
    void *out;              // may alias in
    int Nu, Nv, nang, p;
    int weight;             // fuse the cosine pre-weight into the load
    float cX, cZ, D;
    const void *Hre, *Him;  // device spectrum [p] (Real), Him may be null (real H)
    int conj;               // multiply by conj(H) (adjoint)
    const void *tw;         // device twiddles exp(-2 pi i j/p), [p] complex Real
    unsigned *maxabs;       // per local angle, may be null (float bits)
};
int launch_weight(int real_bytes, void *sino, int Nu, int Nv, int nang, float cX, float cZ,
                  float D, hipStream_t st);
int launch_filter(int real_bytes, const FilterLaunch &a, hipStream_t st);
int launch_pack(int dtype_in, const void *filt, int Nu, int Nv, int nang, int pitch,
                const unsigned *maxabs, int store_dtype, void *stack, void *scales,
                hipStream_t st);
int launch_accumulate(float *dst, const float *src, size_t n, hipStream_t st);
int launch_zero_u32(unsigned *dst, size_t n, hipStream_t st);
int launch_compare(const float *a, const float *b, size_t n, double *d_out3, hipStream_t st);

struct BPLaunch {
    const void *stack;      // [nang][Nv][pitch] in the storage type
    const void *scales;     // per local angle dequantisation factor (Real)
    const void *ang;        // per local angle (cos, sin) pairs (Real)
    const void *ax, *ay, *az;  // coordinate tables (Real), az padded to the z-run
    void *vol;              // [Nz][Ny][Nx] Real
    int Nu, Nv, nang, pitch;
    int Nx, Ny, Nz;
    double cX, cZ, D, out_scale, maskR2;
    int accumulate;
};
int launch_backproject(int store_dtype, const BPLaunch &a, hipStream_t st);
int bp_zrun(int store_dtype);
int sample_bytes_of(int store_dtype);   // bytes of one stored sample
int stack_pitch(int Nu);                // row pitch of the stored stack, samples
constexpr int STACK_TAIL_ROWS = 128;    // spare rows after the last projection (>= box rows)

struct FPLaunch {           // adjoint of the backprojector
    const void *vol;        // [Nz][Ny][Nx]
    void *sino;             // [Nv][nang][Nu], accumulated into (caller clears)
    const void *ang, *ax, *ay, *az;
    int Nu, Nv, nang;
    int Nx, Ny, Nz;
    double cX, cZ, D, out_scale, maskR2;
    // FP32 path: bound on the sum of the LDS contributions to one detector
    // sample, per unit of the workgroup's largest |voxel| (fixes the scale of
    // its fixed-point accumulation)
    double fx_bound;
};
int launch_forward_project(int real_bytes, const FPLaunch &a, hipStream_t st);

struct TapGradLaunch {
    const void *q, *G;      // [rows][Nu] weighted input rows and output gradients
    int Nu, rows, p;
    const void *tw;
    void *partial;          // [nblocks][p] complex Real
    int nblocks;
    void *spectrum;         // [p] complex double: the sum of the partials
};
int launch_tap_gradient(int real_bytes, const TapGradLaunch &a, hipStream_t st);
// dh[j] = Re sum_k R[k] tw[k*j mod p], R the Hermitian part of the spectrum
// (p complex doubles), tw64 the context's FP64 twiddles; dh on the device
int launch_tap_gradient_dft(const void *spectrum, const void *tw64, int p, double *dh,
                            hipStream_t st);

struct PhantomLaunch {
    const double *ells;     // [n][19]
    int n;
    const double *ang;      // (cos, sin) per local angle
    float *sino;            // [Nv][nang][Nu]
    int Nu, Nv, nang, ang0_stride;
    double cX, cZ, D, tau;
};
int launch_phantom(const PhantomLaunch &a, hipStream_t st);
int launch_poisson_log(float *p, size_t n, double I0, uint64_t seed, hipStream_t st);

// Host buffer of floats: page-locked when the HIP runtime grants it, so that
// copies between it and the device run at full link speed; ordinary memory
// otherwise (no GPU present).  Contents start at zero unless asked otherwise.
class HostBuffer {
  public:
    HostBuffer() = default;
    HostBuffer(const HostBuffer &) = delete;
    HostBuffer &operator=(const HostBuffer &) = delete;
    ~HostBuffer() { release(); }
    bool alloc(size_t n, bool zero = true)
    {
        release();
        if (hipHostMalloc((void **)&p_, n * sizeof(float), hipHostMallocDefault) == hipSuccess) {
            pinned_ = true;
        } else {
            (void)hipGetLastError();
            p_ = (float *)std::malloc(n * sizeof(float));
            pinned_ = false;
        }
        if (p_ && zero) std::memset(p_, 0, n * sizeof(float));
        n_ = p_ ? n : 0;
        return p_ != nullptr;
    }
    float *data() { return p_; }
    size_t size() const { return n_; }
    bool pinned() const { return pinned_; }

  private:
    void release()
    {
        if (p_ && pinned_) (void)hipHostFree(p_);
        else std::free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    float *p_ = nullptr;
    size_t n_ = 0;
    bool pinned_ = false;
};

}  // namespace tkv
