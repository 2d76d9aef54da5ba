// tsan_stubs.cpp -- CPU stand-ins for the ThreadSanitizer build of the host
// I/O code (tests/test_tsan.py): the four HIP runtime calls that code makes
// (page-locked allocation falls back to ordinary memory, as it does when no
// GPU is present), and an engine that does no reconstruction but follows the
// slab contract of fdk_reconstruct_host (kernel.cu:844-876): every slab of
// every angle block is asked for once, in order, with `collecting` after the
// first block, and receives (partial sum +) the block's projection checksum.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>

#include "tomokvas/fdk.h"
#include "fdk_internal.hpp"

namespace tkv {
const char *last_error();   // fdk_host.cpp
}

extern "C" {

hipError_t hipHostMalloc(void **ptr, size_t size, unsigned int)
{
    *ptr = std::malloc(size ? size : 1);
    return *ptr ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void *ptr)
{
    std::free(ptr);
    return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "stub"; }

struct fdk_ctx {
    int Nu = 0, Nv = 0, Nang = 0, Nx = 0, Ny = 0, Nz = 0, B = 16;
    int collecting = 0;
};

int fdk_create(int, fdk_ctx **out)
{
    *out = new fdk_ctx();
    return FDK_OK;
}
int fdk_destroy(fdk_ctx *c)
{
    delete c;
    return FDK_OK;
}
size_t fdk_get_free_mem_info(int) { return (size_t)1 << 30; }
const char *fdk_last_error_string(void) { return tkv::last_error(); }
int fdk_set_block_size(fdk_ctx *c, int b)
{
    c->B = b;
    return FDK_OK;
}
int fdk_set_allocation_limit(fdk_ctx *, long long) { return FDK_OK; }
int fdk_set_collecting(fdk_ctx *c, int collecting)
{
    c->collecting = collecting;
    return FDK_OK;
}
int fdk_set_show_info(fdk_ctx *, int) { return FDK_OK; }
int fdk_set_axis_direction(fdk_ctx *, int, int, int) { return FDK_OK; }
int fdk_set_sinogram_properties(fdk_ctx *c, const float *, int Nu, float, int Nang, int Nv, float,
                                float, float, float)
{
    c->Nu = Nu;
    c->Nang = Nang;
    c->Nv = Nv;
    return FDK_OK;
}
int fdk_set_reconstruction_properties(fdk_ctx *c, int Nx, float, int Ny, float, int Nz, float,
                                      float)
{
    c->Nx = Nx;
    c->Ny = Ny;
    c->Nz = Nz;
    return FDK_OK;
}
int fdk_set_scale_xyz(fdk_ctx *, float) { return FDK_OK; }
int fdk_set_storage(fdk_ctx *, int) { return FDK_OK; }
int fdk_set_volume_mode(fdk_ctx *, int) { return FDK_OK; }
int fdk_set_filter(fdk_ctx *, int, const double *, int) { return FDK_OK; }

int fdk_reconstruct_host(fdk_ctx *c, const float *sino, fdk_get_block_fn get_block, void *user)
{
    double sum = 0.;
    const size_t n = (size_t)c->Nu * c->Nv * c->Nang;
    for (size_t i = 0; i < n; ++i) sum += sino[i];
    const float value = (float)sum;
    const size_t plane = (size_t)c->Nx * c->Ny;
    const int slabs = (c->Nz + c->B - 1) / c->B;
    for (int s = 0; s < slabs; ++s) {
        float *buf = get_block(user, s, c->collecting);
        if (!buf) return tkv::fail(FDK_ERR_CALLBACK, "slab buffer");
        const size_t cnt = (size_t)std::min(c->B, c->Nz - s * c->B) * plane;
        for (size_t i = 0; i < cnt; ++i) buf[i] = (c->collecting ? buf[i] : 0.f) + value;
    }
    return FDK_OK;
}

}  // extern "C"
