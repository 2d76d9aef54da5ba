/*
 * tomokvas/fdk.h -- C ABI of the MI355X (gfx950) FDK cone-beam reconstructor.
 *
 * Drop-in boundary for the reference engine `class FilteredBackProjection`
 * (Lukas-kV/TomoKvas FBPG/kernel.cuh:82-151) as FBPG/para_TC.cpp:130-153 drives
 * it.  Every engine call below names the reference member it replaces.  Plain C
 * types only: host buffers are borrowed for the duration of a call, device
 * buffers of the engine are owned by the context, streams are passed as
 * `void*` (a hipStream_t; NULL = the default stream).
 *
 * Return codes: 0 = success (the reference returns cudaSuccess), otherwise one
 * of the FDK_ERR_* values; fdk_last_error_string() describes the last failure
 * on the calling thread (the reference prints to cerr).
 *
 * Thread model: one context per host thread; calls on one context are not
 * thread-safe (like the reference engine).  Calls are synchronous unless a
 * stream is passed to a stage-level call.
 */
#ifndef TOMOKVAS_FDK_H
#define TOMOKVAS_FDK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define FDK_OK                    0
#define FDK_ERR_INVALID_ARGUMENT  1
#define FDK_ERR_NOT_INITIALIZED   2
#define FDK_ERR_HIP               3   /* a HIP runtime call failed            */
#define FDK_ERR_OUT_OF_MEMORY     4
#define FDK_ERR_CALLBACK          5   /* the slab callback returned NULL      */
#define FDK_ERR_UNSUPPORTED       6

/* ---- data types of stage-level buffers --------------------------------------------- */
#define FDK_F32   0
#define FDK_F16   1   /* filtered projections stored as FP16 with a per-projection scale;
                         interpolation weights, lerps and accumulation in FP32 */
#define FDK_BF16  2
#define FDK_F64   3
#define FDK_F16_PACKED 4 /* storage: FP16 as FDK_F16, but interpolated in packed FP16
                            (FP16 fractions and column weights, v_dot2 into FP32 sums):
                            faster, outside the FP32-interpolation precision form; opt-in */

/* ---- ramp filters ------------------------------------------------------------------ */
#define FDK_FILTER_RAMLAK      0   /* kernel.cu:783-802 (the reference's only filter)  */
#define FDK_FILTER_SHEPPLOGAN  1   /* apodised extension                              */
#define FDK_FILTER_USER        2   /* p spatial taps h, H = DFT_p(h)                  */

/* where the reconstructed volume lives during fdk_calculate_reconstruction */
#define FDK_VOLUME_RESIDENT    0   /* whole volume in HBM, readable with fdk_device_volume */
#define FDK_VOLUME_STREAMED    1   /* two chunks in HBM; slabs leave through the callback  */

typedef struct fdk_ctx fdk_ctx;

/* Host slab sink, replacing `float* bin_siz::GetBlock(int n, bool collecting)`
 * (FBPG/mucat_file.h:155, called at FBPG/kernel.cu:844).  Must return a host
 * buffer of Nx*Ny*B floats for slab n; when `collecting` is nonzero it holds the
 * partial sum of earlier angle blocks for that slab, and the engine adds the new
 * contribution to it (FBPG/kernel.cu:849-876) before overwriting it. */
typedef float *(*fdk_get_block_fn)(void *user, int slab, int collecting);

/* ---- engine lifecycle ----------------------------------------------------------------- */
/* FilteredBackProjection() + SetGPUid (kernel.cu:219-253, kernel.cuh:96) */
int    fdk_create(int gpu_id, fdk_ctx **out);
/* ~FilteredBackProjection (kernel.cu:255-281); frees device memory, does NOT
 * reset the device (the reference calls cudaDeviceReset, kernel.cu:274) */
int    fdk_destroy(fdk_ctx *ctx);
/* static size_t GetFreeMemInfo(int gpu) (kernel.cuh:108, kernel.cu:313-322);
 * queries the requested device (the reference queries the current one) */
size_t fdk_get_free_mem_info(int gpu_id);
const char *fdk_last_error_string(void);

/* ---- setters (kernel.cuh:89-101) -------------------------------------------------------- */
int fdk_set_gpu_id(fdk_ctx *ctx, int gpu_id);                       /* SetGPUid        */
int fdk_set_block_size(fdk_ctx *ctx, int slices);                  /* SetBlockSize    */
int fdk_set_allocation_limit(fdk_ctx *ctx, long long bytes);       /* SetAllocationLimit:
                            recorded only; the stack is one HBM allocation             */
int fdk_set_collecting(fdk_ctx *ctx, int collecting);              /* SetCollecting   */
int fdk_set_show_info(fdk_ctx *ctx, int show);                     /* SetShowInfo     */
int fdk_set_axis_direction(fdk_ctx *ctx, int flip_x, int flip_y, int flip_z); /* SetAxisDirection */
/* SetSinogramProperties(float* sino, Nu, cX, Nang, Nv, cZ, delta_fi, distance,
 * sampling) (kernel.cuh:89, kernel.cu:479-497).  `sino` is the host stack in
 * sinogram-major [v][beta][u] order, borrowed until fdk_initialize returns. */
int fdk_set_sinogram_properties(fdk_ctx *ctx, const float *sino, int Nu, float cX,
                                int Nang, int Nv, float cZ, float dbeta,
                                float distance, float sampling);
/* SetReconstructionProperties(bin_siz*, Nx, cx, Ny, cy, Nz, cz, rot)
 * (kernel.cuh:92, kernel.cu:524-537); the bin_siz sink becomes the callback of
 * fdk_calculate_reconstruction */
int fdk_set_reconstruction_properties(fdk_ctx *ctx, int Nx, float cx, int Ny, float cy,
                                      int Nz, float cz, float rot);
int fdk_set_scale_xyz(fdk_ctx *ctx, float scale);                  /* SetScaleXYZ     */
/* extensions: ramp filter choice (default Ram-Lak, as the reference) and the
 * storage type of the filtered stack (default FDK_F32) */
int fdk_set_filter(fdk_ctx *ctx, int kind, const double *taps, int ntaps);
int fdk_set_storage(fdk_ctx *ctx, int dtype);
/* extension: volume residency (default FDK_VOLUME_RESIDENT), set before
 * fdk_initialize.  With a slab callback both modes backproject the volume in
 * chunks (1/2, 1/4, ... of it when resident; 1/8 when streamed, tapering at
 * the end) and move the slabs that chunk k completed to the host on a second
 * stream while chunk k+1 is computed; FDK_VOLUME_STREAMED keeps only about a
 * quarter of the volume plus one slab in HBM (a circular buffer), so the
 * volume may exceed HBM
 * (kernel.cu:819-888 computes one slab at a time without overlap). */
int fdk_set_volume_mode(fdk_ctx *ctx, int mode);
/* launches of one whole-volume backprojection: the angle loop is cut into up
 * to 16 chunks of >= 100 angles (workgroups of one long launch drift apart in
 * their progress through the angles; a launch boundary every ~200 angles keeps
 * the projections they read together) */
int fdk_get_angle_chunks(fdk_ctx *ctx, int *chunks);
/* largest backprojection chunk of the slab pipeline, slices (a multiple of
 * the backprojector's z-run for the current storage type) */
int fdk_get_volume_chunk(fdk_ctx *ctx, int *slices);

/* ---- engine run ---------------------------------------------------------------------- */
/* Initialize() (kernel.cu:324-428): allocate, upload the borrowed sinogram,
 * build the FP64 geometry tables and the FP64 filter spectrum. */
int fdk_initialize(fdk_ctx *ctx);
/* CalculateReconstruction() (kernel.cu:804-897): cosine weight -> ramp filter ->
 * backprojection, then for each slab n = 0..ceil(Nz/B)-1 the callback is asked
 * for its host buffer and the slab's slices are copied into it.  The reference
 * visits ceil((Nz-1)/B) slabs and so skips the last slice when (Nz-1) % B == 0;
 * this engine always covers every slice.  Ordered after the work queued on the
 * NULL stream before the call (e.g. filling fdk_device_sinogram); work on
 * other streams must be complete.  Returns when the volume is finished. */
int fdk_calculate_reconstruction(fdk_ctx *ctx, fdk_get_block_fn get_block, void *user);
/* extension: fdk_initialize + fdk_calculate_reconstruction in one call with
 * the host stack [v][beta][u] borrowed for its duration (set the sinogram
 * properties with a NULL pointer).  With a resident volume the angles are cut
 * into up to 16 chunks of >= 100 angles and chunk k+1 is uploaded while chunk k
 * is filtered and backprojected, so the host-link transfer hides behind the
 * computation (page-locked host memory transfers fastest).  Stage times:
 * [0], [1] zero, [2] first upload to last backprojection, [3] slab transfers
 * left after it.  fdk_device_sinogram's contents are unspecified afterwards.
 * Ordered after the NULL stream's earlier work, as fdk_calculate_reconstruction. */
int fdk_reconstruct_host(fdk_ctx *ctx, const float *host_sino, fdk_get_block_fn get_block,
                         void *user);
/* Device copy of the last reconstructed volume [Nz][Ny][Nx] (owned by ctx). */
const float *fdk_device_volume(fdk_ctx *ctx);
/* Device projection stack [v][beta][u] of an initialised context.  With a NULL
 * sinogram pointer fdk_initialize allocates it without an upload, so callers
 * whose projections are already in HBM fill it here (on the NULL stream, or
 * with their own streams synchronised before fdk_calculate_reconstruction). */
float *fdk_device_sinogram(fdk_ctx *ctx);
/* Stage times of the last fdk_calculate_reconstruction in ms (device events):
 * [0] cosine weight + ramp filter, [1] pack to the storage type,
 * [2] backprojection, [3] slab transfers + accumulation. */
int fdk_get_stage_times(fdk_ctx *ctx, float *ms4);
int fdk_get_geometry(fdk_ctx *ctx, int *ints8, float *floats12);

/* ---- stage-level entry points (device buffers owned by the caller) ----------------------
 * These are the kernel launches of kernel.cu:539-555 (cosine weight), :557-781
 * (row filter), :68-177 (backprojection) and :203-215 (accumulate), exposed for
 * tests, benchmarks and the differentiable form.  `ang0, nang` select the angles
 * [ang0, ang0+nang) of the context geometry; sinogram buffers are then
 * [Nv][nang][Nu].  Geometry and tables come from the context, prepared by
 * fdk_prepare (no sinogram upload needed). */
int fdk_prepare(fdk_ctx *ctx);
/* RB_comma3DKernel (kernel.cu:179-201): P *= D/sqrt((u-cX)^2 + D^2 + (v-cZ)^2), in place */
int fdk_weight(fdk_ctx *ctx, int dtype, void *d_sino, int nang, void *stream);
/* fConvolution (kernel.cu:557-781): zero-pad rows to p, FFT, multiply by the
 * context spectrum (conj if `adjoint`), inverse FFT unnormalised, crop; in place.
 * weight != 0 fuses the cosine weight into the load.  d_maxabs (optional, nang
 * uint32 = float bits) receives the per-projection max |filtered value|. */
int fdk_filter(fdk_ctx *ctx, int dtype, void *d_sino, int nang, int weight, int adjoint,
               uint32_t *d_maxabs, void *stream);
/* bytes of the stored stack for nang projections in storage `store_dtype`:
 * [nang][Nv][pitch] samples, pitch >= Nu + 68 (zero-padded rows) */
size_t fdk_stack_bytes(fdk_ctx *ctx, int store_dtype, int nang);
/* repack filtered [Nv][nang][Nu] (F32 or F64) into the backprojector's stored
 * stack (16-byte aligned); FP16 stores value/scale with scale = maxabs/8192 per
 * projection (d_maxabs required), F32/BF16/F64 store the values (scale 1) */
int fdk_pack(fdk_ctx *ctx, int dtype_in, const void *d_filtered, int nang,
             const uint32_t *d_maxabs, int store_dtype, void *d_stack, void *d_scales,
             void *stream);
/* FBPConeKernel (kernel.cu:68-177) over the whole volume for angles
 * [ang0, ang0+nang): vol (+)= delta_f/p * sum U^2 * bilinear(Q).  d_stack and
 * d_scales hold those nang projections (d_scales NULL: all 1).  vol is F64 for
 * F64 storage, F32 otherwise.  Runs in the angle chunks of a whole FDK pass. */
int fdk_backproject(fdk_ctx *ctx, int store_dtype, const void *d_stack,
                    const void *d_scales, int ang0, int nang, void *d_vol, int accumulate,
                    void *stream);
/* Adjoint of fdk_backproject: voxel-driven transposed-bilinear forward projector
 * with atomic detector splats; sino [Nv][nang][Nu] is accumulated into. */
int fdk_forward_project(fdk_ctx *ctx, int dtype, const void *d_vol, int ang0, int nang,
                        void *d_sino, void *stream);
/* Gradient of sum(G * filter_h(q)) w.r.t. the p spatial taps h for rows
 * q, G [Nv][nang][Nu]; accumulated into the host array dh[p] (double). */
int fdk_tap_gradient(fdk_ctx *ctx, int dtype, const void *d_q, const void *d_G, int nang,
                     double *dh, void *stream);
/* fdk_tap_gradient without any allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight training
 * loops, harnesses that forbid hipMalloc).  Its partial spectra live in a
 * 32 MiB device array per GPU that the library holds until the process ends
 * (fewer workgroups than fdk_tap_gradient when their partials would not fit);
 * calls that overlap on one GPU race for it.  Same arguments and result as
 * fdk_tap_gradient (up to floating-point summation order); FDK_ERR_UNSUPPORTED
 * when a padded row length's spectrum exceeds the array (not for p <= 8192). */
int fdk_tap_gradient_noalloc(fdk_ctx *ctx, int dtype, const void *d_q, const void *d_G,
                             int nang, double *dh, void *stream);
/* CollectionKernel (kernel.cu:203-215): dst += src */
int fdk_accumulate(float *d_dst, const float *d_src, size_t n, void *stream);
/* Quality metric on device data (no reference counterpart): out3 = { sum (a-b)^2,
 * sum b^2, max |a-b| } accumulated in double; synchronous. */
int fdk_compare(const float *d_a, const float *d_b, size_t n, double *out3, void *stream);

/* ---- synthetic projections (benchmark / test input; no reference counterpart) ------ */
/* exact ellipsoid line integrals for angles [ang0, ang0+nang) into float
 * [Nv][nang][Nu]; ells = n rows of 19 doubles (centre 3, semi-axes 3,
 * rotation rows 9, density, 3 unused), coordinates in virtual pixels */
int fdk_phantom_project(fdk_ctx *ctx, const double *ells, int n, int ang0, int nang,
                        float *d_sino, void *stream);
/* Poisson(I0*exp(-p)) counts, then -ln(max(n,1)/I0), in place (Philox4x32-10) */
int fdk_poisson_log(float *d_p, size_t n, double I0, uint64_t seed, void *stream);

/* ---- measurement aid (no reference counterpart) -------------------------------------- */
/* L2-resident streaming-read bandwidth of the current device, in GB/s: every wave
 * reads 1 KiB blocks of a zeroed buffer of footprint_bytes (a power of two in
 * [64 KiB, 64 MiB]; 2 MiB stays in each XCD's L2) at pseudo-random block indices;
 * best of three timed launches after one warm-up; synchronous on the null stream.
 * The backprojector's gather roofline is priced against this figure. */
int fdk_probe_l2_read(size_t footprint_bytes, double *gbs);

#ifdef __cplusplus
}
#endif
#endif
