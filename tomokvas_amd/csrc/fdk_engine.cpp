// fdk_engine.cpp -- the FDK engine behind the C ABI of include/tomokvas/fdk.h.
//
// Replaces class FilteredBackProjection (FBPG/kernel.cuh:82-151,
// FBPG/kernel.cu:218-911).  Differences by design, all MI355X-first:
//  * the projection stack is ONE HBM allocation (the reference splits it into
//    at most 10 pitched chunks of AllocationLimit bytes, kernel.cu:342-371);
//  * the whole volume is reconstructed in one launch and then handed to the
//    slab callback slab by slab (the reference launches once per Z-slab,
//    kernel.cu:819-888), so slab size never limits the kernel's parallelism;
//  * geometry tables and the filter spectrum are built in FP64 on the host;
//  * the filtered stack is repacked angle-major into the storage type, FP32,
//    FP16 (per-projection scale) or BF16, with zero-padded rows;
//  * destroy frees only this context's memory (no device reset).
#include "fdk_internal.hpp"

#include <algorithm>
#include <numeric>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <mutex>

namespace tkv {
const char *last_error();
}

using tkv::fail;

// ---- the allocation-free tap gradient's device scratch ------------------------------
// One file-scope array per device, used only by fdk_tap_gradient_noalloc.  In
// host code the array's name is a host-side stand-in, not a device address:
// host code uses only the address hipGetSymbolAddress returns on each device
// (looked up once per device, with that device current, on its first call).
namespace {
constexpr size_t TAPGRAD_SCRATCH = 32u << 20;   // bytes
static __device__ __align__(256) unsigned char tapgrad_scratch[TAPGRAD_SCRATCH];

constexpr int MAX_DEVICES = 64;
std::mutex tapgrad_scratch_mu;
void *tapgrad_scratch_ptr[MAX_DEVICES] = {};

// device address of the scratch on device `dev`, which must be current
int tapgrad_scratch_of_device(int dev, void **out)
{
    if (dev < 0 || dev >= MAX_DEVICES)
        return fail(FDK_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(tapgrad_scratch_mu);
    if (!tapgrad_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(tapgrad_scratch));
        if (e != hipSuccess) return tkv::hip_fail(e, "hipGetSymbolAddress(tapgrad_scratch)");
        if (!ptr) return fail(FDK_ERR_HIP, "hipGetSymbolAddress(tapgrad_scratch): null address");
        tapgrad_scratch_ptr[dev] = ptr;
    }
    *out = tapgrad_scratch_ptr[dev];
    return FDK_OK;
}
}  // namespace

struct fdk_ctx {
    int dev = 0;
    tkv::Geometry g;
    int collecting = 0, show = 0;
    long long alloc_limit = 1024LL * 1024 * 1024;
    int filter_kind = FDK_FILTER_RAMLAK;
    std::vector<double> user_taps;
    int storage = FDK_F32;
    const float *host_sino = nullptr;
    bool have_sino = false, have_rec = false;

    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {};
    float stage_ms[4] = {0, 0, 0, 0};
    // slab pipeline: slabs leave on copy_stream while the next chunk of the
    // volume is backprojected on stream
    int volume_mode = FDK_VOLUME_RESIDENT;
    hipStream_t copy_stream = nullptr;
    hipStream_t prep_stream = nullptr;   // fdk_reconstruct_host: filter and pack of a chunk
    hipEvent_t ev_bp[2] = {}, ev_cp[2] = {}, ev_up[2] = {}, ev_f[2] = {};
    hipEvent_t ev_caller = nullptr;   // the caller's NULL-stream work (after_caller_work)
    int ring_slices = 0;     // streamed volume: slices held in HBM (circular)

    // tables (prepared)
    bool prepared = false;
    int p = 0, nz_pad = 0;
    float *d_ang32 = nullptr, *d_ax32 = nullptr, *d_ay32 = nullptr, *d_az32 = nullptr;
    double *d_ang64 = nullptr, *d_ax64 = nullptr, *d_ay64 = nullptr, *d_az64 = nullptr;
    float *d_H32re = nullptr, *d_H32im = nullptr, *d_tw32 = nullptr, *d_one32 = nullptr;
    double *d_H64re = nullptr, *d_H64im = nullptr, *d_tw64 = nullptr, *d_one64 = nullptr;
    bool complex_H = false;

    // engine buffers (initialized)
    bool initialized = false;
    float *d_sino = nullptr, *d_filt = nullptr, *d_vol = nullptr, *d_partial = nullptr;
    float *d_ring = nullptr;   // streamed volume: ring_slices slices, slice z at z % ring_slices
    void *d_stack = nullptr;
    float *d_scales = nullptr;
    unsigned *d_maxabs = nullptr;
    void *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    // bytes held by each engine buffer: fdk_initialize keeps a buffer whose
    // capacity still fits (freeing and allocating tens of GB on every angle
    // block has been seen to slow down from milliseconds to seconds per call)
    size_t cap_sino = 0, cap_filt = 0, cap_stack = 0, cap_scales = 0, cap_maxabs = 0;
    size_t cap_vol = 0, cap_ring = 0, cap_partial = 0;
};

namespace {

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

void free_tables(fdk_ctx *c)
{
    dfree(c->d_ang32); dfree(c->d_ax32); dfree(c->d_ay32); dfree(c->d_az32);
    dfree(c->d_ang64); dfree(c->d_ax64); dfree(c->d_ay64); dfree(c->d_az64);
    dfree(c->d_H32re); dfree(c->d_H32im); dfree(c->d_tw32); dfree(c->d_one32);
    dfree(c->d_H64re); dfree(c->d_H64im); dfree(c->d_tw64); dfree(c->d_one64);
    c->prepared = false;
}

void free_engine(fdk_ctx *c)
{
    dfree(c->d_sino); dfree(c->d_filt); dfree(c->d_vol); dfree(c->d_partial);
    dfree(c->d_ring);
    dfree(c->d_stack); dfree(c->d_scales); dfree(c->d_maxabs); dfree(c->d_scratch);
    c->scratch_bytes = 0;
    c->cap_sino = c->cap_filt = c->cap_stack = c->cap_scales = c->cap_maxabs = 0;
    c->cap_vol = c->cap_ring = c->cap_partial = 0;
    c->initialized = false;
}

template <typename T>
int upload(T *&d, const std::vector<T> &h)
{
    dfree(d);
    TKV_HIP(hipMalloc((void **)&d, sizeof(T) * std::max<size_t>(h.size(), 1)));
    if (!h.empty()) TKV_HIP(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return FDK_OK;
}

int dalloc(void **p, size_t bytes, const char *what);

// *p holds at least `need` bytes afterwards: kept when its capacity fits,
// else freed and allocated again
int ensure(void **p, size_t &cap, size_t need, const char *what)
{
    if (*p && cap >= need) return FDK_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    int rc = dalloc(p, need, what);
    if (!rc) cap = need;
    return rc;
}

int dalloc(void **p, size_t bytes, const char *what)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(e == hipErrorOutOfMemory ? FDK_ERR_OUT_OF_MEMORY : FDK_ERR_HIP,
                    std::string(what) + ": hipMalloc of " + std::to_string(bytes) +
                        " bytes failed: " + hipGetErrorString(e));
    }
    return FDK_OK;
}

// stage-level calls run on the caller's stream; NULL is the default stream
// (ordered with the caller's other work), never the engine's private stream
hipStream_t pick(fdk_ctx *, void *stream)
{
    return (hipStream_t)stream;
}

// workgroups of the tap-gradient kernel that give every one a transform of
// its own: row pairs / transforms per workgroup (256 threads; the radix-16
// plan gives a transform min(p/16, 256) threads, one below p = 16)
long long tap_blocks(int p, long long rows)
{
    const int tpt = p < 16 ? 1 : std::min(p / 16, 256), nt = 256 / tpt;
    return ((rows + 1) / 2 + nt - 1) / nt;
}

int check_geometry(const tkv::Geometry &g)
{
    if (g.Nu < 1 || g.Nv < 1 || g.Nang < 1)
        return fail(FDK_ERR_INVALID_ARGUMENT, "sinogram sizes must be positive");
    if (g.Nx < 1 || g.Ny < 1 || g.Nz < 1)
        return fail(FDK_ERR_INVALID_ARGUMENT, "reconstruction sizes must be positive");
    if (g.B < 1) return fail(FDK_ERR_INVALID_ARGUMENT, "block size must be positive");
    if (!(g.tau > 0)) return fail(FDK_ERR_INVALID_ARGUMENT, "sampling must be positive");
    if (!(g.D > 0.5f * g.Nu))
        return fail(FDK_ERR_INVALID_ARGUMENT,
                    "source distance must exceed Nu/2 (the field-of-view radius)");
    if (tkv::padded_length(g.Nu) > 16384)
        return fail(FDK_ERR_UNSUPPORTED, "detector rows longer than 8192 are not supported");
    if (!(g.scale > 0)) return fail(FDK_ERR_INVALID_ARGUMENT, "scale must be positive");
    return FDK_OK;
}

int set_device(fdk_ctx *c)
{
    TKV_HIP(hipSetDevice(c->dev));
    return FDK_OK;
}

int prepare(fdk_ctx *c)
{
    if (!c->have_sino || !c->have_rec)
        return fail(FDK_ERR_NOT_INITIALIZED,
                    "set sinogram and reconstruction properties before preparing");
    int rc = check_geometry(c->g);
    if (rc) return rc;
    if ((rc = set_device(c))) return rc;
    free_tables(c);
    const tkv::Geometry &g = c->g;
    c->p = tkv::padded_length(g.Nu);
    // z tables cover every z-run of every kernel (largest run 64) plus slack
    c->nz_pad = ((g.Nz + 63) / 64) * 64 + 64;

    std::vector<double> s, co;
    tkv::angle_tables(g, s, co);
    std::vector<double> a64(2 * g.Nang);
    std::vector<float> a32(2 * g.Nang);
    for (int k = 0; k < g.Nang; ++k) {
        a64[2 * k] = co[k];
        a64[2 * k + 1] = s[k];
        a32[2 * k] = (float)co[k];
        a32[2 * k + 1] = (float)s[k];
    }
    if ((rc = upload(c->d_ang32, a32)) || (rc = upload(c->d_ang64, a64))) return rc;

    std::vector<float> ax, ay, az;
    tkv::axis_tables_f32(g, c->nz_pad, ax, ay, az);
    if ((rc = upload(c->d_ax32, ax)) || (rc = upload(c->d_ay32, ay)) ||
        (rc = upload(c->d_az32, az)))
        return rc;
    std::vector<double> bx, by, bz;
    tkv::axis_tables_f64(g, c->nz_pad, bx, by, bz);
    if ((rc = upload(c->d_ax64, bx)) || (rc = upload(c->d_ay64, by)) ||
        (rc = upload(c->d_az64, bz)))
        return rc;

    if (c->filter_kind == FDK_FILTER_USER && (int)c->user_taps.size() != c->p)
        return fail(FDK_ERR_INVALID_ARGUMENT,
                    "user filter needs p = " + std::to_string(c->p) + " taps, got " +
                        std::to_string(c->user_taps.size()));
    std::vector<double> Hre, Him;
    tkv::build_spectrum(c->filter_kind, g, c->user_taps, Hre, Him);
    c->complex_H = c->filter_kind == FDK_FILTER_USER;
    std::vector<float> H32re(Hre.begin(), Hre.end()), H32im(Him.begin(), Him.end());
    if ((rc = upload(c->d_H32re, H32re)) || (rc = upload(c->d_H32im, H32im)) ||
        (rc = upload(c->d_H64re, Hre)) || (rc = upload(c->d_H64im, Him)))
        return rc;

    // twiddles exp(-2 pi i j / p), formed in double from the exact index
    std::vector<double> tw64(2 * c->p);
    std::vector<float> tw32(2 * c->p);
    for (int j = 0; j < c->p; ++j) {
        double ang = 2.0 * M_PI * (double)j / (double)c->p;
        tw64[2 * j] = std::cos(ang);
        tw64[2 * j + 1] = -std::sin(ang);
        tw32[2 * j] = (float)tw64[2 * j];
        tw32[2 * j + 1] = (float)tw64[2 * j + 1];
    }
    if ((rc = upload(c->d_tw32, tw32)) || (rc = upload(c->d_tw64, tw64))) return rc;
    std::vector<float> one32(g.Nang, 1.0f);
    std::vector<double> one64(g.Nang, 1.0);
    if ((rc = upload(c->d_one32, one32)) || (rc = upload(c->d_one64, one64))) return rc;
    c->prepared = true;
    return FDK_OK;
}

int need_prepared(fdk_ctx *c)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    if (!c->prepared) return fail(FDK_ERR_NOT_INITIALIZED, "call fdk_prepare or fdk_initialize first");
    return set_device(c);
}

int sample_bytes(int dtype) { return tkv::sample_bytes_of(dtype); }

double out_scale(const fdk_ctx *c, int dtype)
{
    // result *= delta_f * scaling, scaling = 1/p (kernel.cu:113, :575); float
    // product for the float paths, double for FP64
    if (dtype == FDK_F64) return (double)c->g.dbeta * (1.0 / (double)c->p);
    return (double)(c->g.dbeta * (1.0f / (float)c->p));
}

int stage_filter(fdk_ctx *c, int dtype, const void *in, void *out, int nang, int weight,
                 int adjoint, unsigned *maxabs, hipStream_t st)
{
    tkv::FilterLaunch L;
    L.in = in;
    L.out = out;
    L.Nu = c->g.Nu;
    L.Nv = c->g.Nv;
    L.nang = nang;
    L.p = c->p;
    L.weight = weight;
    L.cX = c->g.cX;
    L.cZ = c->g.cZ;
    L.D = c->g.D;
    const bool f64 = dtype == FDK_F64;
    L.Hre = f64 ? (const void *)c->d_H64re : (const void *)c->d_H32re;
    L.Him = c->complex_H ? (f64 ? (const void *)c->d_H64im : (const void *)c->d_H32im) : nullptr;
    L.conj = adjoint;
    L.tw = f64 ? (const void *)c->d_tw64 : (const void *)c->d_tw32;
    L.maxabs = maxabs;
    return tkv::launch_filter(f64 ? 8 : 4, L, st);
}

// Slices [z0, z0 + nz) of the volume into vol (nz < 0: every slice).
int stage_backproject(fdk_ctx *c, int store_dtype, const void *stack, const void *scales,
                      int ang0, int nang, void *vol, int accumulate, hipStream_t st,
                      int z0 = 0, int nz = -1)
{
    const bool f64 = store_dtype == FDK_F64;
    tkv::BPLaunch L;
    L.stack = stack;
    L.scales = scales ? scales : (f64 ? (const void *)c->d_one64 : (const void *)c->d_one32);
    L.ang = f64 ? (const void *)(c->d_ang64 + 2 * ang0) : (const void *)(c->d_ang32 + 2 * ang0);
    L.ax = f64 ? (const void *)c->d_ax64 : (const void *)c->d_ax32;
    L.ay = f64 ? (const void *)c->d_ay64 : (const void *)c->d_ay32;
    // the z tables run nz_pad >= Nz + 64 entries, so a z-run starting at any
    // slice stays inside them
    L.az = f64 ? (const void *)(c->d_az64 + z0) : (const void *)(c->d_az32 + z0);
    L.vol = vol;
    L.Nu = c->g.Nu;
    L.Nv = c->g.Nv;
    L.nang = nang;
    L.pitch = tkv::stack_pitch(c->g.Nu);
    L.Nx = c->g.Nx;
    L.Ny = c->g.Ny;
    L.Nz = nz < 0 ? c->g.Nz : nz;
    L.cX = c->g.cX;
    L.cZ = c->g.cZ;
    L.D = c->g.D;
    L.out_scale = out_scale(c, store_dtype);
    L.maskR2 = 0.25 * (double)c->g.Nu * (double)c->g.Nu;
    L.accumulate = accumulate;
    return tkv::launch_backproject(store_dtype, L, st);
}

// Largest backprojection chunk of the slab pipeline: 1/parts of the volume
// rounded up to a multiple of the backprojector's z-run, so that no chunk
// ends in a partly filled run.
int chunk_target(const tkv::Geometry &g, int zrun, int parts)
{
    const int t = std::max((g.Nz + parts - 1) / parts, zrun);
    return (t + zrun - 1) / zrun * zrun;
}

// Slices of the circular volume buffer of a streamed volume: two of the
// largest chunks of any storage type (z-run <= 32) plus one slab, or the
// whole volume when that is no larger.
int ring_slices_for(const tkv::Geometry &g)
{
    const long long w = 2LL * chunk_target(g, 32, 8) + g.B;
    return (int)std::min<long long>(w, g.Nz);
}

// Chunk boundaries z_0 = 0 < z_1 < ... = Nz: chunks of at most C slices; the
// last ones taper (half the rest, rounded up to the z-run) so that the slab
// transfers left after the last backprojection are short.  Each launch ends
// with a partly idle GPU while its last workgroups drain (about 2 ms at
// 1024^3), so a resident volume starts with half of it: C = Nz/2 gives
// chunks of 1/2, 1/4, ... down to one z-run.
std::vector<int> chunk_bounds(int Nz, int C, int zrun)
{
    std::vector<int> zb{0};
    int z = 0;
    while (z < Nz) {
        const int rest = Nz - z;
        int n = std::min(C, std::max(zrun, ((rest + 1) / 2 + zrun - 1) / zrun * zrun));
        if (n >= rest) n = rest;
        z += n;
        zb.push_back(z);
    }
    return zb;
}

int check_angles(fdk_ctx *c, int ang0, int nang)
{
    if (ang0 < 0 || nang < 0 || ang0 + nang > c->g.Nang)
        return fail(FDK_ERR_INVALID_ARGUMENT, "angle range outside [0, Nang)");
    return FDK_OK;
}

// Angle chunks of a whole-volume backprojection: the angle loop split into
// up to TKV_BP_ANGLE_CHUNKS launches of at least TKV_BP_CHUNK_MIN_ANGLES
// angles, the later ones adding onto the volume.  Within one long launch the
// workgroups drift apart in their progress through the angles, so the
// projections read at any moment spread over many angles; a launch boundary
// every ~200 angles keeps them together (measured at 1024^3 from 1800
// projections, FP16, on one GPU: 519 ms in one launch, 487 ms in 8, 484 in
// 12, 480 in 16, 487 in 24, 503 in 36 -- launch tails and the volume's
// read-modify-write grow with the count).
#ifndef TKV_BP_ANGLE_CHUNKS
#define TKV_BP_ANGLE_CHUNKS 16
#endif
#ifndef TKV_BP_CHUNK_MIN_ANGLES
#define TKV_BP_CHUNK_MIN_ANGLES 100
#endif
// the same inside each chunk of a streamed volume (1024^3 from 1800
// projections, FP16, one GPU: 610 ms with 1 launch per chunk, 591 with 4, 575
// with 16)
#ifndef TKV_STREAM_ANGLE_CHUNKS
#define TKV_STREAM_ANGLE_CHUNKS 16
#endif
std::vector<int> angle_bounds(int nang, int max_chunks = TKV_BP_ANGLE_CHUNKS)
{
    const int K = std::max(1, std::min(max_chunks, nang / TKV_BP_CHUNK_MIN_ANGLES));
    std::vector<int> ab(K + 1);
    for (int k = 0; k <= K; ++k) ab[k] = (int)((long long)nang * k / K);
    return ab;
}

// Angles [ab[k], ab[k+1]) of the stored stack onto vol, each chunk after the
// first adding onto it (the first also when accumulate is set).
int backproject_in_angle_chunks(fdk_ctx *c, const std::vector<int> &ab, void *vol,
                                int accumulate, hipStream_t st)
{
    const size_t stored = (size_t)sample_bytes(c->storage) * tkv::stack_pitch(c->g.Nu) * c->g.Nv;
    for (size_t k = 0; k + 1 < ab.size(); ++k) {
        const int a0 = ab[k], nc = ab[k + 1] - a0;
        int rc = stage_backproject(c, c->storage, (const char *)c->d_stack + (size_t)a0 * stored,
                                   c->d_scales + a0, a0, nc, vol, accumulate || k > 0, st);
        if (rc) return rc;
    }
    return FDK_OK;
}

// The slab hand-out of a resident volume: the angles in TKV_HANDOUT_ANGLE_CHUNKS
// chunks (those of a whole-volume pass, so both give the same volume), all but
// the last TKV_HANDOUT_PIPED_CHUNKS over the whole volume, those through the
// slab pipeline in z-chunks of at most 1/TKV_HANDOUT_ZPARTS of the volume (see
// backproject_and_hand_out).  The piped angles must last longer than the slab
// transfers (4.3 GB at 57 GB/s for 1024^3), and the last z-chunks' transfers
// are the ones left exposed.  Measured at 1024^3 from 1800 projections, FP16,
// page-locked slabs (FDK wall time; 559 ms without hand-out): 8 chunks with the
// last piped in halves 610-629 ms; 16 with the last 4 in 1/8 parts 581, in 1/4
// parts 589, in 1/16 parts 601; the last 6 in 1/8 parts 577; the last 8, 582.
#ifndef TKV_HANDOUT_ANGLE_CHUNKS
#define TKV_HANDOUT_ANGLE_CHUNKS TKV_BP_ANGLE_CHUNKS
#endif
#ifndef TKV_HANDOUT_PIPED_CHUNKS
#define TKV_HANDOUT_PIPED_CHUNKS 6
#endif
#ifndef TKV_HANDOUT_ZPARTS
#define TKV_HANDOUT_ZPARTS 8
#endif

// Largest chunk for the current storage type: 1/TKV_HANDOUT_ZPARTS of the
// volume when it is resident; an eighth when it is streamed, within
// 2 C + B <= ring_slices (see backproject_and_hand_out).
int pipeline_chunk(const fdk_ctx *c)
{
    const tkv::Geometry &g = c->g;
    const int zr = tkv::bp_zrun(c->storage);
    if (c->volume_mode != FDK_VOLUME_STREAMED) return chunk_target(g, zr, TKV_HANDOUT_ZPARTS);
    int C = chunk_target(g, zr, 8);
    if (c->ring_slices < g.Nz) C = std::min(C, (c->ring_slices - g.B) / 2 / zr * zr);
    return C;
}

// The volume in chunks: chunk k+1 is backprojected on the compute stream
// while the slabs that chunk k completed go to the host on the copy stream
// (with a collecting run each slab's partial sum comes up first and is added
// on the device) -- the slab loop of kernel.cu:838-884, overlapped with the
// backprojection.  Chunks are cut at multiples of the z-run, slabs at
// multiples of B; a slab leaves once the chunk holding its last slice is done.
//
// A streamed volume lives in a circular buffer of W = ring_slices slices
// (slice z at z % W; chunk launches and slab copies that cross the end are
// split).  With chunks of at most C slices and W >= 2 C + B, chunk k+1
// overwrites only slices below z_k - B, all in slabs that ended by z_k and so
// left in iteration k-1: the compute stream waits for that iteration's copies.
// Records ev[3] after the last backprojection and ev[4] after the last copy,
// so ev[3] -> ev[4] is the transfer time left exposed.
// Angles [ang0, ang0 + nang) of the stored stack (stack and scales point at
// angle ang0), added onto the volume when accumulate is set; a resident
// volume's chunks are backprojected in the angle launches `sub` (bounds
// relative to ang0, from 0 to nang).
int backproject_and_hand_out(fdk_ctx *c, fdk_get_block_fn get_block, void *user,
                             const void *stack, const float *scales, int ang0, int nang,
                             int accumulate, const std::vector<int> &sub_resident)
{
    const tkv::Geometry &g = c->g;
    hipStream_t st = c->stream, cs = c->copy_stream;
    const bool streamed = c->volume_mode == FDK_VOLUME_STREAMED;
    const size_t plane = (size_t)g.Nx * g.Ny;
    const int W = streamed ? c->ring_slices : g.Nz;
    float *ring = streamed ? c->d_ring : c->d_vol;
    const bool wraps = W < g.Nz;
    const int C = pipeline_chunk(c);
    const int zr = tkv::bp_zrun(c->storage);
    if (C < 1 || (wraps && 2 * C + g.B > W))
        return fail(FDK_ERR_INVALID_ARGUMENT, "slab pipeline: volume buffer too small");
    // a ring slot holds other slices before its chunk is backprojected: only a
    // resident volume can be added onto
    if (accumulate && streamed)
        return fail(FDK_ERR_INVALID_ARGUMENT, "slab pipeline: cannot accumulate a streamed volume");
    const std::vector<int> zb = chunk_bounds(g.Nz, C, zr);
    const int nchunk = (int)zb.size() - 1;
    const int nslab = tkv::slab_count(g.Nz, g.B);

    // a streamed volume's chunks also cut their angles (the first launch
    // writes the chunk's slot, the later ones add onto it)
    const std::vector<int> sub =
        streamed ? angle_bounds(nang, TKV_STREAM_ANGLE_CHUNKS) : sub_resident;
    if (sub.size() < 2 || sub.front() != 0 || sub.back() != nang)
        return fail(FDK_ERR_INVALID_ARGUMENT, "slab pipeline: angle launches do not cover the angles");
    const size_t stored = (size_t)sample_bytes(c->storage) * tkv::stack_pitch(g.Nu) * g.Nv;
    auto bp = [&](int k) -> int {
        if (wraps && k >= 2) TKV_HIP(hipStreamWaitEvent(st, c->ev_cp[(k - 2) & 1], 0));
        for (int z = zb[k]; z < zb[k + 1];) {
            const int pos = z % W, n = std::min(zb[k + 1] - z, W - pos);
            for (size_t j = 0; j + 1 < sub.size(); ++j) {
                const int s0 = sub[j], ns = sub[j + 1] - s0;
                int rc = stage_backproject(c, c->storage, (const char *)stack + (size_t)s0 * stored,
                                           scales + s0, ang0 + s0, ns, ring + (size_t)pos * plane,
                                           accumulate || j > 0, st, z, n);
                if (rc) return rc;
            }
            z += n;
        }
        TKV_HIP(hipEventRecord(c->ev_bp[k & 1], st));
        if (k == nchunk - 1) TKV_HIP(hipEventRecord(c->ev[3], st));
        return FDK_OK;
    };
    int rc = bp(0);
    if (rc) return rc;
    int n = 0;   // next slab to hand out
    for (int k = 0; k < nchunk; ++k) {
        if (k + 1 < nchunk && (rc = bp(k + 1))) return rc;
        TKV_HIP(hipStreamWaitEvent(cs, c->ev_bp[k & 1], 0));
        for (; n < nslab && std::min((n + 1) * g.B, g.Nz) <= zb[k + 1]; ++n) {
            const int z0 = n * g.B, nz = std::min(g.B, g.Nz - z0);
            // the previous slab has landed before its buffer goes to the writer
            TKV_HIP(hipStreamSynchronize(cs));
            float *block = get_block(user, n, c->collecting);
            if (!block)
                return fail(FDK_ERR_CALLBACK, "slab callback returned NULL for slab " +
                                                  std::to_string(n));
            if (c->collecting)
                TKV_HIP(hipMemcpyAsync(c->d_partial, block, nz * plane * sizeof(float),
                                       hipMemcpyHostToDevice, cs));
            for (int z = z0; z < z0 + nz;) {   // at most two pieces (ring end)
                const int pos = z % W, m = std::min(z0 + nz - z, W - pos);
                float *dev = ring + (size_t)pos * plane;
                const size_t off = (size_t)(z - z0) * plane;
                if (c->collecting &&
                    (rc = tkv::launch_accumulate(dev, c->d_partial + off, m * plane, cs)))
                    return rc;
                TKV_HIP(hipMemcpyAsync(block + off, dev, m * plane * sizeof(float),
                                       hipMemcpyDeviceToHost, cs));
                z += m;
            }
        }
        TKV_HIP(hipEventRecord(c->ev_cp[k & 1], cs));
    }
    TKV_HIP(hipEventRecord(c->ev[4], cs));
    return FDK_OK;
}

// The context's streams and events, all on device c->dev.  A failure leaves
// every handle either valid or null, so destroy_streams may always follow.
int create_streams(fdk_ctx *c)
{
    TKV_HIP(hipSetDevice(c->dev));
    TKV_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    TKV_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    TKV_HIP(hipStreamCreateWithFlags(&c->prep_stream, hipStreamNonBlocking));
    for (auto &e : c->ev) TKV_HIP(hipEventCreate(&e));
    TKV_HIP(hipEventCreateWithFlags(&c->ev_caller, hipEventDisableTiming));
    for (int i = 0; i < 2; ++i) {
        TKV_HIP(hipEventCreateWithFlags(&c->ev_bp[i], hipEventDisableTiming));
        TKV_HIP(hipEventCreateWithFlags(&c->ev_cp[i], hipEventDisableTiming));
        TKV_HIP(hipEventCreateWithFlags(&c->ev_up[i], hipEventDisableTiming));
        TKV_HIP(hipEventCreateWithFlags(&c->ev_f[i], hipEventDisableTiming));
    }
    return FDK_OK;
}

// The context's streams are non-blocking, so nothing orders them after work
// the caller queued on the NULL stream before a call -- filling the device
// projection stack through stage-level calls or torch, say.  Entry points that
// run on the context's streams first make each of them wait for an event
// recorded on the NULL stream.
int after_caller_work(fdk_ctx *c)
{
    TKV_HIP(hipEventRecord(c->ev_caller, nullptr));
    TKV_HIP(hipStreamWaitEvent(c->stream, c->ev_caller, 0));
    TKV_HIP(hipStreamWaitEvent(c->copy_stream, c->ev_caller, 0));
    TKV_HIP(hipStreamWaitEvent(c->prep_stream, c->ev_caller, 0));
    return FDK_OK;
}

// destroys every stream and event of c (the current device must be c->dev)
// and nulls the handles
void destroy_streams(fdk_ctx *c)
{
    auto drop_event = [](hipEvent_t &e) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    };
    auto drop_stream = [](hipStream_t &s) {
        if (s) (void)hipStreamDestroy(s);
        s = nullptr;
    };
    for (auto &e : c->ev) drop_event(e);
    drop_event(c->ev_caller);
    for (int i = 0; i < 2; ++i) {
        drop_event(c->ev_bp[i]);
        drop_event(c->ev_cp[i]);
        drop_event(c->ev_up[i]);
        drop_event(c->ev_f[i]);
    }
    drop_stream(c->stream);
    drop_stream(c->copy_stream);
    drop_stream(c->prep_stream);
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char *fdk_last_error_string(void) { return tkv::last_error(); }

int fdk_create(int gpu_id, fdk_ctx **out)
{
    if (!out) return fail(FDK_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    int n = 0;
    TKV_HIP(hipGetDeviceCount(&n));
    if (gpu_id < 0 || gpu_id >= n)
        return fail(FDK_ERR_INVALID_ARGUMENT, "GPU id " + std::to_string(gpu_id) +
                                                  " out of range (" + std::to_string(n) +
                                                  " devices)");
    fdk_ctx *c = new fdk_ctx();
    c->dev = gpu_id;
    const int rc = create_streams(c);
    if (rc) {
        fdk_destroy(c);
        return rc;
    }
    *out = c;
    return FDK_OK;
}

int fdk_destroy(fdk_ctx *c)
{
    if (!c) return FDK_OK;
    (void)hipSetDevice(c->dev);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    if (c->prep_stream) (void)hipStreamSynchronize(c->prep_stream);
    free_engine(c);
    free_tables(c);
    destroy_streams(c);
    delete c;
    return FDK_OK;
}

size_t fdk_get_free_mem_info(int gpu_id)
{
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return 0;
    if (hipSetDevice(gpu_id) != hipSuccess) return 0;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
    (void)hipSetDevice(cur);
    return fr;
}

int fdk_set_gpu_id(fdk_ctx *c, int gpu_id)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    if (c->prepared || c->initialized)
        return fail(FDK_ERR_INVALID_ARGUMENT, "GPU id must be set before initialisation");
    int n = 0;
    TKV_HIP(hipGetDeviceCount(&n));
    if (gpu_id < 0 || gpu_id >= n) return fail(FDK_ERR_INVALID_ARGUMENT, "GPU id out of range");
    // every stream and event belongs to the device it was created on: all of
    // them are made again on the new device (also for the same id, so that one
    // path serves every call; SetGPUid only records the id, kernel.cu:283-286,
    // and the device is selected when the engine runs)
    (void)hipSetDevice(c->dev);
    destroy_streams(c);
    c->dev = gpu_id;
    return create_streams(c);
}

int fdk_set_block_size(fdk_ctx *c, int slices)
{
    if (!c || slices < 1) return fail(FDK_ERR_INVALID_ARGUMENT, "block size must be positive");
    c->g.B = slices;
    c->prepared = false;
    c->initialized = false;   // sizes or tables change: fdk_initialize again
    return FDK_OK;
}

int fdk_set_allocation_limit(fdk_ctx *c, long long bytes)
{
    if (!c || bytes <= 0) return fail(FDK_ERR_INVALID_ARGUMENT, "allocation limit must be positive");
    c->alloc_limit = bytes;
    return FDK_OK;
}

int fdk_set_collecting(fdk_ctx *c, int collecting)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    if (collecting && !c->d_partial) c->initialized = false;   // needs its collect slab
    c->collecting = collecting ? 1 : 0;
    return FDK_OK;
}

int fdk_set_show_info(fdk_ctx *c, int show)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    c->show = show ? 1 : 0;
    return FDK_OK;
}

// SetAxisDirection (kernel.cu:499-507): the reference only ever sets a flip
// (calling it twice cannot clear one); here each call states all three.
int fdk_set_axis_direction(fdk_ctx *c, int fx, int fy, int fz)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    c->g.flipX = fx ? 1 : 0;
    c->g.flipY = fy ? 1 : 0;
    c->g.flipZ = fz ? 1 : 0;
    c->prepared = false;
    c->initialized = false;   // sizes or tables change: fdk_initialize again
    return FDK_OK;
}

int fdk_set_sinogram_properties(fdk_ctx *c, const float *sino, int Nu, float cX, int Nang,
                                int Nv, float cZ, float dbeta, float distance, float sampling)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    c->host_sino = sino;
    c->g.Nu = Nu;
    c->g.cX = cX;
    c->g.Nang = Nang;
    c->g.Nv = Nv;
    c->g.cZ = cZ;
    c->g.dbeta = dbeta;
    c->g.D = distance;
    c->g.tau = sampling;
    c->have_sino = true;
    c->prepared = false;
    c->initialized = false;   // sizes or tables change: fdk_initialize again
    return FDK_OK;
}

int fdk_set_reconstruction_properties(fdk_ctx *c, int Nx, float cx, int Ny, float cy, int Nz,
                                      float cz, float rot)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    c->g.Nx = Nx;
    c->g.cx = cx;
    c->g.Ny = Ny;
    c->g.cy = cy;
    c->g.Nz = Nz;
    c->g.cz = cz;
    c->g.rot = rot;
    c->have_rec = true;
    c->prepared = false;
    c->initialized = false;   // sizes or tables change: fdk_initialize again
    return FDK_OK;
}

int fdk_set_scale_xyz(fdk_ctx *c, float scale)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    c->g.scale = scale;
    c->prepared = false;
    c->initialized = false;   // sizes or tables change: fdk_initialize again
    return FDK_OK;
}

int fdk_set_filter(fdk_ctx *c, int kind, const double *taps, int ntaps)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    if (kind != FDK_FILTER_RAMLAK && kind != FDK_FILTER_SHEPPLOGAN && kind != FDK_FILTER_USER)
        return fail(FDK_ERR_INVALID_ARGUMENT, "unknown filter kind");
    if (kind == FDK_FILTER_USER && (!taps || ntaps < 1))
        return fail(FDK_ERR_INVALID_ARGUMENT, "user filter needs taps");
    c->filter_kind = kind;
    c->user_taps.assign(taps && kind == FDK_FILTER_USER ? taps : nullptr,
                        taps && kind == FDK_FILTER_USER ? taps + ntaps : nullptr);
    c->prepared = false;
    c->initialized = false;   // sizes or tables change: fdk_initialize again
    return FDK_OK;
}

int fdk_set_storage(fdk_ctx *c, int dtype)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    if (dtype != FDK_F32 && dtype != FDK_F16 && dtype != FDK_BF16 && dtype != FDK_F16_PACKED)
        return fail(FDK_ERR_INVALID_ARGUMENT, "storage must be F32, F16, BF16 or F16_PACKED");
    if (c->initialized && dtype != c->storage) {
        int rc = ensure(&c->d_stack, c->cap_stack, fdk_stack_bytes(c, dtype, c->g.Nang),
                        "stored stack");
        if (rc) return rc;
    }
    c->storage = dtype;
    return FDK_OK;
}

int fdk_prepare(fdk_ctx *c)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    return prepare(c);
}

size_t fdk_stack_bytes(fdk_ctx *c, int store_dtype, int nang)
{
    if (!c) return 0;
    // nang projections plus tkv::STACK_TAIL_ROWS spare rows: the backprojector's
    // box loads past the last detector row are range-checked to zero, and the
    // spare rows keep even those addresses inside the allocation
    const size_t row = (size_t)sample_bytes(store_dtype) * (size_t)tkv::stack_pitch(c->g.Nu);
    return row * ((size_t)c->g.Nv * (size_t)nang + tkv::STACK_TAIL_ROWS);
}

int fdk_initialize(fdk_ctx *c)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    int rc = prepare(c);
    if (rc) return rc;
    c->initialized = false;
    const tkv::Geometry &g = c->g;
    const size_t nsino = (size_t)g.Nu * g.Nv * g.Nang;
    const size_t nvol = (size_t)g.Nx * g.Ny * g.Nz;
    const size_t slab = (size_t)g.Nx * g.Ny * g.B;
    if ((rc = ensure((void **)&c->d_sino, c->cap_sino, nsino * sizeof(float), "projection stack")) ||
        (rc = ensure((void **)&c->d_filt, c->cap_filt, nsino * sizeof(float), "filtered stack")) ||
        (rc = ensure(&c->d_stack, c->cap_stack, fdk_stack_bytes(c, c->storage, g.Nang),
                     "stored stack")) ||
        (rc = ensure((void **)&c->d_scales, c->cap_scales, g.Nang * sizeof(float), "scales")) ||
        (rc = ensure((void **)&c->d_maxabs, c->cap_maxabs, g.Nang * sizeof(unsigned), "maxabs"))) {
        free_engine(c);
        return rc;
    }
    if (c->volume_mode == FDK_VOLUME_STREAMED) {
        dfree(c->d_vol);
        c->cap_vol = 0;
        c->ring_slices = ring_slices_for(g);
        rc = ensure((void **)&c->d_ring, c->cap_ring,
                    (size_t)c->ring_slices * g.Nx * g.Ny * sizeof(float), "volume ring");
    } else {
        dfree(c->d_ring);
        c->cap_ring = 0;
        rc = ensure((void **)&c->d_vol, c->cap_vol, nvol * sizeof(float), "volume");
    }
    if (!rc && c->collecting)
        rc = ensure((void **)&c->d_partial, c->cap_partial, slab * sizeof(float), "collect slab");
    if (rc) {
        free_engine(c);
        return rc;
    }
    if (c->host_sino) {
        hipError_t e = hipMemcpy(c->d_sino, c->host_sino, nsino * sizeof(float),
                                 hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            free_engine(c);
            return tkv::hip_fail(e, "projection upload");
        }
    }
    c->host_sino = nullptr;  // borrowed only until here (para_TC.cpp:150 frees it)
    c->initialized = true;
    if (c->show) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c->dev) == hipSuccess)
            std::cout << "GPU " << c->dev << ": " << prop.name << ", "
                      << prop.multiProcessorCount << " CUs, "
                      << prop.totalGlobalMem / (1024 * 1024) << " MiB\n";
    }
    return FDK_OK;
}

float *fdk_device_sinogram(fdk_ctx *c) { return c && c->initialized ? c->d_sino : nullptr; }
const float *fdk_device_volume(fdk_ctx *c) { return c && c->initialized ? c->d_vol : nullptr; }

int fdk_set_volume_mode(fdk_ctx *c, int mode)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    if (mode != FDK_VOLUME_RESIDENT && mode != FDK_VOLUME_STREAMED)
        return fail(FDK_ERR_INVALID_ARGUMENT, "volume mode must be FDK_VOLUME_RESIDENT or "
                                              "FDK_VOLUME_STREAMED");
    if (mode != c->volume_mode) c->initialized = false;   // other volume buffers
    c->volume_mode = mode;
    return FDK_OK;
}

int fdk_get_angle_chunks(fdk_ctx *c, int *chunks)
{
    if (!c || !chunks) return fail(FDK_ERR_INVALID_ARGUMENT, "null argument");
    *chunks = (int)angle_bounds(c->g.Nang).size() - 1;
    return FDK_OK;
}

int fdk_get_volume_chunk(fdk_ctx *c, int *slices)
{
    if (!c || !slices) return fail(FDK_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->initialized) return fail(FDK_ERR_NOT_INITIALIZED, "call fdk_initialize first");
    *slices = pipeline_chunk(c);
    return FDK_OK;
}

int fdk_calculate_reconstruction(fdk_ctx *c, fdk_get_block_fn get_block, void *user)
{
    if (!c) return fail(FDK_ERR_INVALID_ARGUMENT, "null context");
    if (!c->initialized) return fail(FDK_ERR_NOT_INITIALIZED, "call fdk_initialize first");
    int rc = set_device(c);
    if (rc) return rc;
    if ((rc = after_caller_work(c))) return rc;
    const tkv::Geometry &g = c->g;
    hipStream_t st = c->stream;
    if ((rc = tkv::launch_zero_u32(c->d_maxabs, g.Nang, st))) return rc;
    TKV_HIP(hipEventRecord(c->ev[0], st));
    // cosine weight fused into the row filter (kernel.cu:807-808)
    if ((rc = stage_filter(c, FDK_F32, c->d_sino, c->d_filt, g.Nang, 1, 0, c->d_maxabs, st)))
        return rc;
    TKV_HIP(hipEventRecord(c->ev[1], st));
    if ((rc = tkv::launch_pack(FDK_F32, c->d_filt, g.Nu, g.Nv, g.Nang, tkv::stack_pitch(g.Nu),
                               c->d_maxabs, c->storage, c->d_stack, c->d_scales, st)))
        return rc;
    TKV_HIP(hipEventRecord(c->ev[2], st));
    if (!get_block) {
        if (c->volume_mode == FDK_VOLUME_STREAMED)
            return fail(FDK_ERR_INVALID_ARGUMENT, "a streamed volume needs a slab callback");
        if ((rc = backproject_in_angle_chunks(c, angle_bounds(g.Nang), c->d_vol, 0, st)))
            return rc;
        TKV_HIP(hipEventRecord(c->ev[3], st));
        TKV_HIP(hipEventRecord(c->ev[4], st));
    } else {
        // resident: all angle chunks but the last TKV_HANDOUT_PIPED_CHUNKS
        // over the whole volume, those through the slab pipeline (slabs leave
        // as its z-chunks finish, so the piped angles must last long enough
        // to hide most of the slab transfers);
        // streamed: one chunk of every angle through the pipeline
        const bool streamed = c->volume_mode == FDK_VOLUME_STREAMED;
        const std::vector<int> ab = streamed ? std::vector<int>{0, g.Nang}
                                             : angle_bounds(g.Nang, TKV_HANDOUT_ANGLE_CHUNKS);
        const int K = (int)ab.size() - 1;
        const int Pc = streamed ? 1 : std::min(K, TKV_HANDOUT_PIPED_CHUNKS);
        const size_t stored =
            (size_t)sample_bytes(c->storage) * tkv::stack_pitch(g.Nu) * g.Nv;
        if (K > Pc && (rc = backproject_in_angle_chunks(
                           c, std::vector<int>(ab.begin(), ab.end() - Pc), c->d_vol, 0, st)))
            return rc;
        const int a0 = ab[K - Pc];
        std::vector<int> sub;
        for (int k = K - Pc; k <= K; ++k) sub.push_back(ab[k] - a0);
        if ((rc = backproject_and_hand_out(c, get_block, user,
                                           (const char *)c->d_stack + (size_t)a0 * stored,
                                           c->d_scales + a0, a0, g.Nang - a0, K > Pc, sub))) {
            (void)hipStreamSynchronize(c->copy_stream);
            return rc;
        }
    }
    TKV_HIP(hipStreamSynchronize(st));
    TKV_HIP(hipStreamSynchronize(c->copy_stream));
    for (int i = 0; i < 4; ++i)
        TKV_HIP(hipEventElapsedTime(&c->stage_ms[i], c->ev[i], c->ev[i + 1]));
    return FDK_OK;
}

// Initialize + CalculateReconstruction with the host stack borrowed for the
// call: the angles are cut into K chunks; chunk k+1 crosses the host link on
// the copy stream (a 2-D copy of its [v][a0, a1)[u] runs into a compact
// [v][nc][u] block) while chunk k is weighted, filtered, packed and
// backprojected -- the first chunk replacing the volume, the others added onto
// it -- on the compute stream; the last chunk's backprojection goes through
// the slab pipeline.  Filter and pack run on a third stream, so chunk k+1's
// overlap chunk k's backprojection.  A streamed volume cannot be accumulated
// over chunks, so it takes one chunk.  The order of host calls matters for pageable memory,
// whose copies return only when done: each chunk's kernels are queued before
// the next chunk's copy is issued.
int fdk_reconstruct_host(fdk_ctx *c, const float *host_sino, fdk_get_block_fn get_block,
                         void *user)
{
    if (!c || !host_sino) return fail(FDK_ERR_INVALID_ARGUMENT, "null argument");
    if (!get_block && c->volume_mode == FDK_VOLUME_STREAMED)
        return fail(FDK_ERR_INVALID_ARGUMENT, "a streamed volume needs a slab callback");
    c->host_sino = nullptr;   // allocate without an upload
    int rc = fdk_initialize(c);
    if (rc) return rc;
    if ((rc = set_device(c))) return rc;
    if ((rc = after_caller_work(c))) return rc;
    const tkv::Geometry &g = c->g;
    hipStream_t st = c->stream, cs = c->copy_stream;
    const size_t per_ang = (size_t)g.Nu * g.Nv;
    const int pitch = tkv::stack_pitch(g.Nu);
    const size_t stored = (size_t)sample_bytes(c->storage) * pitch * g.Nv;
    const std::vector<int> ab =
        c->volume_mode == FDK_VOLUME_STREAMED ? std::vector<int>{0, g.Nang} : angle_bounds(g.Nang);
    const int K = (int)ab.size() - 1;
    const size_t host_row = (size_t)g.Nang * g.Nu * sizeof(float);
    auto upload = [&](int k) -> int {
        const int a0 = ab[k], nc = ab[k + 1] - a0;
        const size_t w = (size_t)nc * g.Nu * sizeof(float);
        TKV_HIP(hipMemcpy2DAsync(c->d_sino + (size_t)a0 * per_ang, w, host_sino + (size_t)a0 * g.Nu,
                                 host_row, w, g.Nv, hipMemcpyHostToDevice, cs));
        TKV_HIP(hipEventRecord(c->ev_up[k & 1], cs));
        return FDK_OK;
    };
    hipStream_t ps = c->prep_stream;
    if ((rc = tkv::launch_zero_u32(c->d_maxabs, g.Nang, ps))) return rc;
    TKV_HIP(hipEventRecord(c->ev[0], cs));
    if ((rc = upload(0))) return rc;
    for (int k = 0; k < K; ++k) {
        const int a0 = ab[k], nc = ab[k + 1] - a0;
        float *in = c->d_sino + (size_t)a0 * per_ang, *filt = c->d_filt + (size_t)a0 * per_ang;
        void *stack = (char *)c->d_stack + (size_t)a0 * stored;
        float *scales = c->d_scales + a0;
        // the chunk's filter and pack on the prep stream once it has landed,
        // so they can run while the previous chunk is backprojected; its
        // backprojection on the compute stream once they are done
        TKV_HIP(hipStreamWaitEvent(ps, c->ev_up[k & 1], 0));
        if ((rc = stage_filter(c, FDK_F32, in, filt, nc, 1, 0, c->d_maxabs + a0, ps)) ||
            (rc = tkv::launch_pack(FDK_F32, filt, g.Nu, g.Nv, nc, pitch, c->d_maxabs + a0,
                                   c->storage, stack, scales, ps)))
            break;
        TKV_HIP(hipEventRecord(c->ev_f[k & 1], ps));
        TKV_HIP(hipStreamWaitEvent(st, c->ev_f[k & 1], 0));
        if (k + 1 < K || !get_block) {
            if ((rc = stage_backproject(c, c->storage, stack, scales, a0, nc, c->d_vol, k > 0, st)))
                break;
            if (k + 1 < K) {
                if ((rc = upload(k + 1))) break;
            } else {
                TKV_HIP(hipEventRecord(c->ev[3], st));
                TKV_HIP(hipEventRecord(c->ev[4], st));
            }
        } else if ((rc = backproject_and_hand_out(c, get_block, user, stack, scales, a0, nc,
                                                  K > 1, std::vector<int>{0, nc}))) {
            break;
        }
    }
    if (rc) {
        (void)hipStreamSynchronize(cs);
        (void)hipStreamSynchronize(ps);
        (void)hipStreamSynchronize(st);
        return rc;
    }
    TKV_HIP(hipStreamSynchronize(st));
    TKV_HIP(hipStreamSynchronize(cs));
    TKV_HIP(hipStreamSynchronize(ps));
    c->stage_ms[0] = c->stage_ms[1] = 0.f;
    TKV_HIP(hipEventElapsedTime(&c->stage_ms[2], c->ev[0], c->ev[3]));
    TKV_HIP(hipEventElapsedTime(&c->stage_ms[3], c->ev[3], c->ev[4]));
    return FDK_OK;
}

int fdk_get_stage_times(fdk_ctx *c, float *ms4)
{
    if (!c || !ms4) return fail(FDK_ERR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < 4; ++i) ms4[i] = c->stage_ms[i];
    return FDK_OK;
}

int fdk_get_geometry(fdk_ctx *c, int *ints8, float *floats12)
{
    if (!c || !ints8 || !floats12) return fail(FDK_ERR_INVALID_ARGUMENT, "null argument");
    const tkv::Geometry &g = c->g;
    int iv[8] = {g.Nu, g.Nang, g.Nv, g.Nx, g.Ny, g.Nz, g.B, c->p};
    float fv[12] = {g.cX, g.cZ, g.dbeta, g.D, g.tau, g.cx, g.cy, g.cz, g.rot, g.scale,
                    (float)(g.flipX | (g.flipY << 1) | (g.flipZ << 2)), (float)c->storage};
    std::memcpy(ints8, iv, sizeof iv);
    std::memcpy(floats12, fv, sizeof fv);
    return FDK_OK;
}

// ------------------------------------------------------------ stage level
int fdk_weight(fdk_ctx *c, int dtype, void *d_sino, int nang, void *stream)
{
    int rc = need_prepared(c);
    if (rc) return rc;
    if (dtype != FDK_F32 && dtype != FDK_F64)
        return fail(FDK_ERR_INVALID_ARGUMENT, "weight: F32 or F64 data");
    if (nang < 0) return fail(FDK_ERR_INVALID_ARGUMENT, "weight: negative angle count");
    return tkv::launch_weight(dtype == FDK_F64 ? 8 : 4, d_sino, c->g.Nu, c->g.Nv, nang, c->g.cX,
                              c->g.cZ, c->g.D, pick(c, stream));
}

int fdk_filter(fdk_ctx *c, int dtype, void *d_sino, int nang, int weight, int adjoint,
               uint32_t *d_maxabs, void *stream)
{
    int rc = need_prepared(c);
    if (rc) return rc;
    if (dtype != FDK_F32 && dtype != FDK_F64)
        return fail(FDK_ERR_INVALID_ARGUMENT, "filter: F32 or F64 data");
    if (nang < 0) return fail(FDK_ERR_INVALID_ARGUMENT, "filter: negative angle count");
    return stage_filter(c, dtype, d_sino, d_sino, nang, weight, adjoint, d_maxabs,
                        pick(c, stream));
}

int fdk_pack(fdk_ctx *c, int dtype_in, const void *d_filtered, int nang,
             const uint32_t *d_maxabs, int store_dtype, void *d_stack, void *d_scales,
             void *stream)
{
    int rc = need_prepared(c);
    if (rc) return rc;
    if (!d_filtered || !d_stack || !d_scales)
        return fail(FDK_ERR_INVALID_ARGUMENT, "pack: null buffer");
    if ((uintptr_t)d_stack & 15)
        return fail(FDK_ERR_INVALID_ARGUMENT, "pack: stored stack not 16-byte aligned");
    return tkv::launch_pack(dtype_in, d_filtered, c->g.Nu, c->g.Nv, nang, tkv::stack_pitch(c->g.Nu),
                            d_maxabs, store_dtype, d_stack, d_scales, pick(c, stream));
}

int fdk_backproject(fdk_ctx *c, int store_dtype, const void *d_stack, const void *d_scales,
                    int ang0, int nang, void *d_vol, int accumulate, void *stream)
{
    int rc = need_prepared(c);
    if (rc) return rc;
    if ((rc = check_angles(c, ang0, nang))) return rc;
    if (!d_stack || !d_vol) return fail(FDK_ERR_INVALID_ARGUMENT, "backproject: null buffer");
    if (!sample_bytes(store_dtype))
        return fail(FDK_ERR_INVALID_ARGUMENT, "backproject: unknown storage type");
    // in the angle chunks of a whole pass (launch boundaries keep the
    // workgroups' projections together in L2, see angle_bounds)
    hipStream_t st = pick(c, stream);
    const std::vector<int> ab = angle_bounds(nang);
    const size_t stored = (size_t)sample_bytes(store_dtype) * tkv::stack_pitch(c->g.Nu) * c->g.Nv;
    const size_t sb = store_dtype == FDK_F64 ? sizeof(double) : sizeof(float);
    for (size_t k = 0; k + 1 < ab.size(); ++k) {
        const int a = ab[k], n = ab[k + 1] - a;
        const void *sc = d_scales ? (const void *)((const char *)d_scales + a * sb) : nullptr;
        if ((rc = stage_backproject(c, store_dtype, (const char *)d_stack + a * stored, sc,
                                    ang0 + a, n, d_vol, accumulate || k > 0, st)))
            return rc;
    }
    return FDK_OK;
}

int fdk_forward_project(fdk_ctx *c, int dtype, const void *d_vol, int ang0, int nang,
                        void *d_sino, void *stream)
{
    int rc = need_prepared(c);
    if (rc) return rc;
    if ((rc = check_angles(c, ang0, nang))) return rc;
    if (dtype != FDK_F32 && dtype != FDK_F64)
        return fail(FDK_ERR_INVALID_ARGUMENT, "forward_project: F32 or F64 data");
    const bool f64 = dtype == FDK_F64;
    tkv::FPLaunch L;
    L.vol = d_vol;
    L.sino = d_sino;
    L.ang = f64 ? (const void *)(c->d_ang64 + 2 * ang0) : (const void *)(c->d_ang32 + 2 * ang0);
    L.ax = f64 ? (const void *)c->d_ax64 : (const void *)c->d_ax32;
    L.ay = f64 ? (const void *)c->d_ay64 : (const void *)c->d_ay32;
    L.az = f64 ? (const void *)c->d_az64 : (const void *)c->d_az32;
    L.Nu = c->g.Nu;
    L.Nv = c->g.Nv;
    L.nang = nang;
    L.Nx = c->g.Nx;
    L.Ny = c->g.Ny;
    L.Nz = c->g.Nz;
    L.cX = c->g.cX;
    L.cZ = c->g.cZ;
    L.D = c->g.D;
    L.out_scale = out_scale(c, dtype);
    L.maskR2 = 0.25 * (double)c->g.Nu * (double)c->g.Nu;
    {
        // per voxel-update the four splats weigh at most U^2 |x| together;
        // columns inside the cylinder (|S| < r) have U <= D/(D-r), and the
        // staged path never sees U >= 100 (its footprints need D-S > D/100).
        // A sample of a 16 x 16-column, 64-slice workgroup's box receives at
        // most one splat per column and slice:
        //  - slices: those whose row v lies in [row-1, row+1), at most
        //    1 + 2/spacing with the slice spacing scale * U >= scale*D/(D+r);
        //  - columns: those whose u lies in [col-1, col+1).  Along a line of
        //    the tile axis closer to the T direction (|cos| or |sin| >=
        //    1/sqrt 2), u = cX + D T/(D-S) is monotone with
        //    |du/d(step)| = pitch D |c'(D-S) + T s'| / (D-S)^2 >=
        //    pitch D (D - 2r) / (sqrt 2 (D+r)^2) (|S|, |T| < r), so each of
        //    the 16 lines holds at most 1 + 2/du of them (all 16 if D <= 2r).
        const double r = std::sqrt(L.maskR2), D = c->g.D;
        const double pitch = std::fabs((double)c->g.scale);
        const double umax = D > r ? std::min(100.0, D / (D - r)) : 100.0;
        const double umin = D / (D + r);
        const double spacing = pitch * umin;
        const double spr = std::min(64.0, 1.0 + std::floor(2.0 / std::max(spacing, 1e-9)));
        const double du = D > 2.0 * r
                              ? pitch * D * (D - 2.0 * r) / (std::sqrt(2.0) * (D + r) * (D + r))
                              : 0.0;
        const double per_line = du > 0.0 ? std::min(16.0, 1.0 + std::floor(2.0 / du)) : 16.0;
        const double ncols = 16.0 * per_line;
        L.fx_bound = ncols * spr * umax * umax;
    }
    return tkv::launch_forward_project(f64 ? 8 : 4, L, pick(c, stream));
}

int fdk_tap_gradient(fdk_ctx *c, int dtype, const void *d_q, const void *d_G, int nang,
                     double *dh, void *stream)
{
    int rc = need_prepared(c);
    if (rc) return rc;
    if (dtype != FDK_F32 && dtype != FDK_F64)
        return fail(FDK_ERR_INVALID_ARGUMENT, "tap_gradient: F32 or F64 data");
    if (!dh) return fail(FDK_ERR_INVALID_ARGUMENT, "tap_gradient: null output");
    const int p = c->p;
    const long long rows = (long long)c->g.Nv * nang;
    const int nblocks = (int)std::max<long long>(1, std::min<long long>(tap_blocks(p, rows), 2048));
    const size_t rb = dtype == FDK_F64 ? 8 : 4;
    // partial spectra of the workgroups, then their sum (p complex doubles)
    const size_t pbytes = (size_t)nblocks * p * 2 * rb;
    const size_t sbytes = (size_t)p * 2 * sizeof(double);
    const size_t bytes = pbytes + sbytes;
    if (bytes > c->scratch_bytes) {
        dfree(c->d_scratch);
        c->scratch_bytes = 0;
        if ((rc = dalloc(&c->d_scratch, bytes, "tap gradient partials"))) return rc;
        c->scratch_bytes = bytes;
    }
    hipStream_t st = pick(c, stream);
    tkv::TapGradLaunch L;
    L.q = d_q;
    L.G = d_G;
    L.Nu = c->g.Nu;
    L.rows = (int)std::min<long long>(rows, 0x7fffffff);
    L.p = p;
    L.tw = dtype == FDK_F64 ? (const void *)c->d_tw64 : (const void *)c->d_tw32;
    L.partial = c->d_scratch;
    L.nblocks = nblocks;
    L.spectrum = (char *)c->d_scratch + pbytes;
    if ((rc = tkv::launch_tap_gradient((int)rb, L, st))) return rc;
    std::vector<double> P(2 * (size_t)p);
    TKV_HIP(hipMemcpyAsync(P.data(), L.spectrum, sbytes, hipMemcpyDeviceToHost, st));
    TKV_HIP(hipStreamSynchronize(st));
    // the Hermitian part R[k] = (P[k] + conj(P[-k]))/2, then
    // dh[j] += Re sum_k R[k] e^{-2 pi i k j / p}
    std::vector<double> Rr(p), Ri(p);
    for (int k = 0; k < p; ++k) {
        const int m = (p - k) % p;
        Rr[k] = 0.5 * (P[2 * k] + P[2 * m]);
        Ri[k] = 0.5 * (P[2 * k + 1] - P[2 * m + 1]);
    }
    tkv::fft_pow2(Rr, Ri);
    for (int j = 0; j < p; ++j) dh[j] += Rr[j];
    return FDK_OK;
}

// Allocation-free variant of fdk_tap_gradient for callers that cannot call
// hipMalloc (code under HIP graph capture, tight training loops, harnesses that
// forbid it).  The workgroups' partial spectra, their sum and the taps'
// gradient live in the file-scope device array tapgrad_scratch (TAPGRAD_SCRATCH
// bytes, held by every GPU the library runs on until the process ends); the
// number of workgroups is capped so that their partials fit.  Every call on
// one GPU shares that GPU's array: calls that overlap on one GPU race.  Nothing
// in it has to start at zero (every launch overwrites what it reads).  The
// final transform runs on the GPU and its result is added into dh through a
// buffer on the stack, so the call makes no heap allocation either.
int fdk_tap_gradient_noalloc(fdk_ctx *c, int dtype, const void *d_q, const void *d_G, int nang,
                             double *dh, void *stream)
{
    int rc = need_prepared(c);   // selects c->dev
    if (rc) return rc;
    if (dtype != FDK_F32 && dtype != FDK_F64)
        return fail(FDK_ERR_INVALID_ARGUMENT, "tap_gradient_noalloc: F32 or F64 data");
    if (!dh || !d_q || !d_G)
        return fail(FDK_ERR_INVALID_ARGUMENT, "tap_gradient_noalloc: null buffer");
    const int p = c->p;
    const size_t rb = dtype == FDK_F64 ? 8 : 4;
    // spectrum (p complex doubles) and gradient (p doubles) at the end of the
    // scratch, the partial spectra before them
    const size_t tail = (size_t)p * 3 * sizeof(double);
    const size_t per_block = (size_t)p * 2 * rb;
    if (tail + per_block > TAPGRAD_SCRATCH)
        return fail(FDK_ERR_UNSUPPORTED,
                    "tap_gradient_noalloc: padded length " + std::to_string(p) +
                        " needs more than the " + std::to_string(TAPGRAD_SCRATCH >> 20) +
                        " MiB scratch");
    const long long rows = (long long)c->g.Nv * nang;
    const long long cap = (long long)((TAPGRAD_SCRATCH - tail) / per_block);
    const int nblocks = (int)std::max<long long>(
        1, std::min<long long>({tap_blocks(p, rows), cap, 2048LL}));
    char *scr = nullptr;
    if ((rc = tapgrad_scratch_of_device(c->dev, (void **)&scr))) return rc;
    hipStream_t st = pick(c, stream);
    tkv::TapGradLaunch L;
    L.q = d_q;
    L.G = d_G;
    L.Nu = c->g.Nu;
    L.rows = (int)std::min<long long>(rows, 0x7fffffff);
    L.p = p;
    L.tw = dtype == FDK_F64 ? (const void *)c->d_tw64 : (const void *)c->d_tw32;
    L.partial = scr;
    L.nblocks = nblocks;
    L.spectrum = scr + (TAPGRAD_SCRATCH - tail);
    double *d_dh = (double *)((char *)L.spectrum + (size_t)p * 2 * sizeof(double));
    if ((rc = tkv::launch_tap_gradient((int)rb, L, st))) return rc;
    if ((rc = tkv::launch_tap_gradient_dft(L.spectrum, c->d_tw64, p, d_dh, st))) return rc;
    double part[512];
    for (int j0 = 0; j0 < p; j0 += 512) {
        const int n = std::min(512, p - j0);
        TKV_HIP(hipMemcpyAsync(part, d_dh + j0, (size_t)n * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        TKV_HIP(hipStreamSynchronize(st));
        for (int j = 0; j < n; ++j) dh[j0 + j] += part[j];
    }
    return FDK_OK;
}

int fdk_accumulate(float *d_dst, const float *d_src, size_t n, void *stream)
{
    if (!d_dst || !d_src) return fail(FDK_ERR_INVALID_ARGUMENT, "accumulate: null buffer");
    return tkv::launch_accumulate(d_dst, d_src, n, (hipStream_t)stream);
}

int fdk_compare(const float *d_a, const float *d_b, size_t n, double *out3, void *stream)
{
    if (!d_a || !d_b || !out3) return fail(FDK_ERR_INVALID_ARGUMENT, "compare: null buffer");
    hipStream_t st = (hipStream_t)stream;
    double *d_out = nullptr;
    int rc = dalloc((void **)&d_out, 3 * sizeof(double), "compare");
    if (rc) return rc;
    rc = tkv::launch_compare(d_a, d_b, n, d_out, st);
    hipError_t e = rc ? hipSuccess : hipMemcpyAsync(out3, d_out, 3 * sizeof(double),
                                                    hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && !rc) e = hipStreamSynchronize(st);
    (void)hipFree(d_out);
    if (rc) return rc;
    if (e != hipSuccess) return tkv::hip_fail(e, "compare");
    return FDK_OK;
}

int fdk_phantom_project(fdk_ctx *c, const double *ells, int n, int ang0, int nang,
                        float *d_sino, void *stream)
{
    int rc = need_prepared(c);
    if (rc) return rc;
    if ((rc = check_angles(c, ang0, nang))) return rc;
    if (n < 0 || (n > 0 && !ells) || !d_sino)
        return fail(FDK_ERR_INVALID_ARGUMENT, "phantom: bad arguments");
    hipStream_t st = pick(c, stream);
    double *d_e = nullptr;
    if ((rc = dalloc((void **)&d_e, sizeof(double) * 19 * std::max(n, 1), "ellipsoids"))) return rc;
    hipError_t e = hipMemcpyAsync(d_e, ells, sizeof(double) * 19 * n, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        (void)hipFree(d_e);
        return tkv::hip_fail(e, "phantom upload");
    }
    tkv::PhantomLaunch L;
    L.ells = d_e;
    L.n = n;
    L.ang = c->d_ang64 + 2 * ang0;
    L.sino = d_sino;
    L.Nu = c->g.Nu;
    L.Nv = c->g.Nv;
    L.nang = nang;
    L.ang0_stride = 0;
    L.cX = c->g.cX;
    L.cZ = c->g.cZ;
    L.D = c->g.D;
    L.tau = c->g.tau;
    rc = tkv::launch_phantom(L, st);
    (void)hipStreamSynchronize(st);
    (void)hipFree(d_e);
    return rc;
}

int fdk_poisson_log(float *d_p, size_t n, double I0, uint64_t seed, void *stream)
{
    if (!d_p || !(I0 > 0)) return fail(FDK_ERR_INVALID_ARGUMENT, "poisson: bad arguments");
    return tkv::launch_poisson_log(d_p, n, I0, seed, (hipStream_t)stream);
}

}  // extern "C"
