// fdk_tiff_internal.hpp -- shared by the TIFF export's device part
// (fdk_tiff.hip) and host part (fdk_tiff_host.cpp); not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "tomokvas/fdk.h"

namespace tkv {

// u = (unsigned short)(65535 * (x - min) / (max - min)) in FP64, in that
// order and without contraction (both callers' sources turn contraction off),
// so the device and the host write the same bytes; values outside [min, max]
// clamp to 0 / 65535, NaN gives 0 (the reference's conversion of such values
// to unsigned short is undefined)
__host__ __device__ inline uint16_t scale_u16(float x, double vmin, double vmax)
{
    if (vmax == vmin) return 0;
    const double v = 65535 * ((double)x - vmin) / (vmax - vmin);
    if (!(v >= 0)) return 0;
    if (v >= 65535) return 65535;
    return (uint16_t)(unsigned)v;
}

// <dir>/<first+k>.tif for slices k in [0, count) of img
int write_slices(const uint16_t *img, int x, int y, int count, const char *dir, int first);

// threads that write the files of one slab (files are independent; the
// page-cache copy of one thread runs at a few GB/s)
int writer_threads(int count);

// slices [0, count) over `threads` threads, slice k on thread k % threads;
// produce(t, k) may fill the slice first (a no-op when it already holds it).
// Returns the lowest-numbered thread's failure with its message.
template <typename Produce>
int write_slices_parallel(int count, int threads, std::string &err, Produce produce)
{
    std::vector<int> rcs(threads, FDK_OK);
    std::vector<std::string> msgs(threads);
    auto part = [&](int t) {
        for (int k = t; k < count; k += threads) {
            const int rc = produce(t, k);
            if (rc) {
                rcs[t] = rc;
                msgs[t] = fdk_last_error_string();
                return;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) {
        try {
            pool.emplace_back(part, t);
        } catch (const std::system_error &) {
            part(t);   // no thread available: this one writes its share
        }
    }
    part(0);
    for (auto &th : pool) th.join();
    for (int t = 0; t < threads; ++t)
        if (rcs[t]) {
            err = msgs[t];
            return rcs[t];
        }
    return FDK_OK;
}

}  // namespace tkv
