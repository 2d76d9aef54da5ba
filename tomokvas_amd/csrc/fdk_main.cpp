// tomokvas-fbp: the reference's command line program (FBPG/para_TC.cpp:5-48)
// over the C ABI.  Usage: tomokvas-fbp <dir>/<name>.CRA
//   reads ./setting.xml (written with the defaults when absent), writes
//   <dir>/<name>.bin and <dir>/<name>.siz.
// Options beyond the reference: --storage f32|f16|bf16, --shepp-logan,
// --angle-block-mib N (size of one angle block instead of the free-HBM rule),
// --streamed / --resident (volume residency instead of the size rule),
// --setting PATH, --quiet.
#include "tomokvas/fdk.h"
#include "tomokvas/fdk_driver.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

int main(int argc, char **argv)
{
    std::printf("MI355X cone beam reconstruction (FDK)\n");
    const char *cra = nullptr, *setting = "setting.xml";
    fdk_run_options opt{FDK_F32, FDK_FILTER_RAMLAK, 0, 1, -1};
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--storage" && i + 1 < argc) {
            const std::string t = argv[++i];
            if (t == "f32") opt.storage = FDK_F32;
            else if (t == "f16") opt.storage = FDK_F16;
            else if (t == "bf16") opt.storage = FDK_BF16;
            else if (t == "f16-packed") opt.storage = FDK_F16_PACKED;
            else {
                std::fprintf(stderr, "unknown storage type %s\n", t.c_str());
                return 2;
            }
        } else if (a == "--shepp-logan") {
            opt.filter = FDK_FILTER_SHEPPLOGAN;
        } else if (a == "--angle-block-mib" && i + 1 < argc) {
            opt.angle_block_bytes = std::atoll(argv[++i]) * 1024LL * 1024LL;
        } else if (a == "--setting" && i + 1 < argc) {
            setting = argv[++i];
        } else if (a == "--streamed") {
            opt.volume_mode = FDK_VOLUME_STREAMED;
        } else if (a == "--resident") {
            opt.volume_mode = FDK_VOLUME_RESIDENT;
        } else if (a == "--quiet") {
            opt.verbose = 0;
        } else if (!cra && a.size() && a[0] != '-') {
            cra = argv[i];
        } else {
            cra = nullptr;
            break;
        }
    }
    if (!cra) {
        const char *p = std::strrchr(argv[0], '/');
        std::fprintf(stderr,
                     "usage: %s [--storage f32|f16|bf16] [--shepp-logan] [--angle-block-mib N] "
                     "[--streamed|--resident] [--setting setting.xml] [--quiet] file.cra\n",
                     p ? p + 1 : argv[0]);
        return 255;   // main() returns -1
    }
    fdk_setting set;
    if (fdk_setting_load(setting, &set) != FDK_OK) fdk_setting_save(setting, &set);
    fdk_run_stats st{};
    const int rc = fdk_reconstruct_file(cra, &set, &opt, &st);
    if (rc != FDK_OK) {
        std::fprintf(stderr, "error %d: %s\n", rc, fdk_last_error_string());
        return 1;
    }
    if (opt.verbose)
        std::printf("read %.1f MiB, wrote %.1f MiB; engine %.3f s (upload included), waiting "
                    "for projections %.3f s, final writes %.3f s\n",
                    st.bytes_read / 1048576.0, st.bytes_written / 1048576.0, st.seconds_engine,
                    st.seconds_load_wait, st.seconds_flush);
    return 0;
}
