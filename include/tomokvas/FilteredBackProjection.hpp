// tomokvas/FilteredBackProjection.hpp -- header-only C++ drop-in for the reference
// engine class, forwarding to the C ABI of tomokvas/fdk.h.
//
// Same member names, argument order and meaning as `class FilteredBackProjection`
// (Lukas-kV/TomoKvas FBPG/kernel.cuh:82-151), so host code written like
// FBPG/para_TC.cpp:130-153 compiles unchanged except that status codes are `int`
// (0 = success) instead of cudaError_t.  The output sink of
// SetReconstructionProperties is any class with `float* GetBlock(int, bool)`
// (the reference's bin_siz, FBPG/mucat_file.h:155, qualifies as is).
#ifndef TOMOKVAS_FILTERED_BACK_PROJECTION_HPP
#define TOMOKVAS_FILTERED_BACK_PROJECTION_HPP

#include <cstddef>
#include <iostream>

#include "tomokvas/fdk.h"

namespace tomokvas {

class FilteredBackProjection {
public:
    FilteredBackProjection() { status_ = fdk_create(0, &ctx_); }
    ~FilteredBackProjection() { fdk_destroy(ctx_); }
    FilteredBackProjection(const FilteredBackProjection &) = delete;
    FilteredBackProjection &operator=(const FilteredBackProjection &) = delete;

    void SetSinogramProperties(float *sinogram, const int &sinogram_X, const float &centre_X,
                               const int &sinogram_Y, const int &sinogram_Z,
                               const float &centre_Z, const float &delta_fi, float distance,
                               float sampling)
    {
        keep(fdk_set_sinogram_properties(ctx_, sinogram, sinogram_X, centre_X, sinogram_Y,
                                         sinogram_Z, centre_Z, delta_fi, distance, sampling));
    }

    template <class Sink>
    void SetReconstructionProperties(Sink *output_file, const int &reconstruction_X,
                                     const float &centre_X, const int &reconstruction_Y,
                                     const float &centre_Y, const int &reconstruction_Z,
                                     const float &centre_Z, const float &rot)
    {
        sink_ = output_file;
        get_block_ = &trampoline<Sink>;
        keep(fdk_set_reconstruction_properties(ctx_, reconstruction_X, centre_X,
                                               reconstruction_Y, centre_Y, reconstruction_Z,
                                               centre_Z, rot));
    }

    void SetAxisDirection(const bool &X = false, const bool &Y = false, const bool &Z = false)
    {
        keep(fdk_set_axis_direction(ctx_, X, Y, Z));
    }
    void SetGPUid(const int &ID = 0) { keep(fdk_set_gpu_id(ctx_, ID)); }
    void SetBlockSize(const int &sz = 16) { keep(fdk_set_block_size(ctx_, sz)); }
    void SetAllocationLimit(const long long &memory_limit = 1024LL * 1024 * 1024)
    {
        keep(fdk_set_allocation_limit(ctx_, memory_limit));
    }
    void SetCollecting(const bool &collecting = false) { keep(fdk_set_collecting(ctx_, collecting)); }
    void SetShowInfo(const bool &show = false) { keep(fdk_set_show_info(ctx_, show)); }
    void SetScaleXYZ(const float &scale) { keep(fdk_set_scale_xyz(ctx_, scale)); }

    // extensions
    void SetFilter(int kind, const double *taps = nullptr, int ntaps = 0)
    {
        keep(fdk_set_filter(ctx_, kind, taps, ntaps));
    }
    void SetStorage(int dtype) { keep(fdk_set_storage(ctx_, dtype)); }
    // FDK_VOLUME_RESIDENT (default) or FDK_VOLUME_STREAMED, before Initialize
    void SetVolumeMode(int mode) { keep(fdk_set_volume_mode(ctx_, mode)); }

    int Initialize()
    {
        if (status_ != FDK_OK) return report(status_);
        return report(fdk_initialize(ctx_));
    }
    // the reference computes the angle tables inside Initialize as well
    int CalcSinCos() { return report(fdk_prepare(ctx_)); }
    int CalculateReconstruction()
    {
        return report(fdk_calculate_reconstruction(ctx_, get_block_, sink_));
    }
    // Initialize + CalculateReconstruction for a caller that keeps the host
    // stack [v][beta][u] until the call returns (SetSinogramProperties with a
    // NULL pointer): the upload runs in angle chunks behind the computation.
    int ReconstructHost(const float *host_sino)
    {
        return report(fdk_reconstruct_host(ctx_, host_sino, get_block_, sink_));
    }
    void fShowInfo() const
    {
        std::cout << "GPU free memory: " << GetFreeMemInfo(0) / 1024 / 1024 << " MiB" << std::endl;
    }
    static size_t GetFreeMemInfo(int GPUId) { return fdk_get_free_mem_info(GPUId); }

    fdk_ctx *handle() const { return ctx_; }

private:
    template <class Sink>
    static float *trampoline(void *user, int slab, int collecting)
    {
        return static_cast<Sink *>(user)->GetBlock(slab, collecting != 0);
    }
    // setters return void in the reference; keep the first failure for Initialize
    void keep(int rc)
    {
        if (rc != FDK_OK && status_ == FDK_OK) status_ = rc;
        if (rc != FDK_OK) std::cerr << fdk_last_error_string() << std::endl;
    }
    static int report(int rc)
    {
        if (rc != FDK_OK) std::cerr << fdk_last_error_string() << std::endl;
        return rc;
    }

    fdk_ctx *ctx_ = nullptr;
    int status_ = FDK_OK;
    void *sink_ = nullptr;
    fdk_get_block_fn get_block_ = nullptr;
};

}  // namespace tomokvas

#endif
