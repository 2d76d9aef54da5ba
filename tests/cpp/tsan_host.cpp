// tsan_host.cpp -- drives the threaded host code under ThreadSanitizer
// (tests/test_tsan.py): the reconstruction program's loop with its
// angle-block loader thread and the .bin writer's I/O thread
// (fdk_reconstruct_file over the stub engine of tsan_stubs.cpp), the writer
// alone with collecting slabs, and the TIFF export's writer threads.
//
//   tsan_host driver DIR/NAME.CRA BUDGET_BYTES
//   tsan_host bin DIR
//   tsan_host tiff DIR
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tomokvas/fdk.h"
#include "tomokvas/fdk_driver.h"
#include "tomokvas/fdk_io.h"
#include "tomokvas/fdk_tiff.h"

static int check(int rc, const char *what)
{
    if (rc) std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, fdk_last_error_string());
    return rc;
}

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    if (mode == "driver" && argc == 4) {
        fdk_setting set;
        fdk_setting_defaults(&set);
        set.wait_for_enter = 0;
        fdk_run_options opt{FDK_F32, FDK_FILTER_RAMLAK, std::atoll(argv[3]), 0, -1};
        fdk_run_stats st;
        if (check(fdk_reconstruct_file(argv[2], &set, &opt, &st), "fdk_reconstruct_file")) return 1;
        std::printf("scan blocks %d angle blocks %d\n", st.scan_blocks, st.angle_blocks);
        return 0;
    }
    if (mode == "bin") {
        // 3 scan blocks x 2 angle blocks of 5 slabs of 3 slices (13 slices):
        // slab s of angle block k adds k + 1 + s, so the file holds 3 + 2s
        const int x = 7, y = 5, z = 13, B = 3, slabs = (z + B - 1) / B;
        fdk_bin *bin = nullptr;
        if (check(fdk_bin_open(argv[2], "tsan", x, y, z, 3, B, &bin), "fdk_bin_open")) return 1;
        for (int sb = 0; sb < 3; ++sb) {
            if (check(fdk_bin_set_scan_block(bin, sb), "fdk_bin_set_scan_block")) return 1;
            for (int k = 0; k < 2; ++k)
                for (int s = 0; s < slabs; ++s) {
                    float *b = fdk_bin_get_block(bin, s, k > 0);
                    if (!b) return check(1, "fdk_bin_get_block");
                    const int cnt = std::min(B, z - s * B) * x * y;
                    for (int i = 0; i < cnt; ++i) b[i] = (k > 0 ? b[i] : 0.f) + (float)(k + 1 + s);
                }
            if (check(fdk_bin_flush(bin), "fdk_bin_flush")) return 1;
        }
        return check(fdk_bin_close(bin), "fdk_bin_close") ? 1 : 0;
    }
    if (mode == "tiff") {
        const int x = 33, y = 17, z = 40;
        std::vector<float> vol((size_t)x * y * z);
        for (size_t i = 0; i < vol.size(); ++i) vol[i] = (float)((i * 2654435761u) % 1000) / 999.f;
        return check(fdk_tiff_export_host(vol.data(), x, y, z, 0.0, 1.0, argv[2], 0),
                     "fdk_tiff_export_host") ? 1 : 0;
    }
    return 2;
}
