// The reconstruction program of FBPG/para_TC.cpp as a library call: settings
// file, .CRA projections in angle blocks, one engine context reused over every
// block, .bin/.siz output through the slab writer.
//
// Differences from the reference's main(), each deliberate:
//  * the angle-block budget is what the engine needs on this GPU (filtered and
//    stored stacks next to the projections, the whole volume resident), rounded
//    down to whole projections so that no projection is left out of a block;
//  * the next angle block is read from disk on a loader thread while the engine
//    works on the current one (two page-locked host buffers, so the upload runs
//    at link speed);
//  * each chunk's slabs are written while the next chunk is backprojected; a
//    volume above half the free HBM is streamed through a circular buffer of
//    about a quarter of it, so the volume may exceed HBM;
//  * one engine context serves every block (the reference creates and destroys
//    one per block; fdk_initialize reallocates what the new block needs);
//  * settings attributes that are absent keep their defaults;
//  * paths are POSIX ("/"), and the enter prompt is only given on a terminal.
#include "tomokvas/fdk_driver.h"
#include "tomokvas/fdk.h"
#include "tomokvas/fdk_io.h"
#include "fdk_internal.hpp"

#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <system_error>
#include <thread>

using tkv::fail;

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0)
{
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

// ------------------------------------------------------------ settings XML
// The attributes of the first <name .../> or <name ...> element inside text,
// or false when there is none.
bool find_element(const std::string &xml, const std::string &name, std::string *attrs)
{
    size_t pos = 0;
    while ((pos = xml.find("<" + name, pos)) != std::string::npos) {
        const size_t after = pos + 1 + name.size();
        if (after < xml.size() && (std::isspace((unsigned char)xml[after]) || xml[after] == '/' ||
                                   xml[after] == '>')) {
            const size_t end = xml.find('>', after);
            if (end == std::string::npos) return false;
            *attrs = xml.substr(after, end - after);
            return true;
        }
        pos = after;
    }
    return false;
}

bool attribute(const std::string &attrs, const char *key, std::string *value)
{
    const std::string k = key;
    size_t pos = 0;
    while ((pos = attrs.find(k, pos)) != std::string::npos) {
        const bool starts = pos == 0 || std::isspace((unsigned char)attrs[pos - 1]);
        size_t q = pos + k.size();
        while (q < attrs.size() && std::isspace((unsigned char)attrs[q])) ++q;
        if (starts && q < attrs.size() && attrs[q] == '=') {
            ++q;
            while (q < attrs.size() && std::isspace((unsigned char)attrs[q])) ++q;
            if (q < attrs.size() && (attrs[q] == '"' || attrs[q] == '\'')) {
                const size_t close = attrs.find(attrs[q], q + 1);
                if (close == std::string::npos) return false;
                *value = attrs.substr(q + 1, close - q - 1);
                return true;
            }
        }
        pos += k.size();
    }
    return false;
}

void int_attribute(const std::string &attrs, const char *key, int *v)
{
    std::string s;
    if (attribute(attrs, key, &s)) *v = std::atoi(s.c_str());
}

void double_attribute(const std::string &attrs, const char *key, double *v)
{
    std::string s;
    if (attribute(attrs, key, &s)) *v = std::atof(s.c_str());
}

void bool_attribute(const std::string &attrs, const char *key, int *v)
{
    std::string s;
    if (attribute(attrs, key, &s)) *v = std::atoi(s.c_str()) != 0;   // BoolAttribute
}

// ------------------------------------------------------------ angle blocks
struct Block {
    int scan = 0, angle = 0;
    int nang = 0;
    double rot = 0;
    tkv::HostBuffer data;            // page-locked: the engine's upload runs at link speed
    int rc = FDK_OK;
    std::string err;
};

// Select and read one angle block; touches the reader, so only one thread at
// a time runs it.
void load_block(fdk_cra *cra, size_t proj_floats, Block *b)
{
    b->rc = FDK_OK;
    fdk_cra_set_scan_block(cra, b->scan);
    fdk_cra_set_angle_block(cra, b->angle);
    int rc = fdk_cra_block_projections(cra, &b->nang);
    if (!rc) rc = fdk_cra_rotation(cra, &b->rot);
    if (!rc) {
        const size_t need = proj_floats * (size_t)b->nang;
        if (need > b->data.size() && !b->data.alloc(need, false))   // overwritten by the load
            rc = fail(FDK_ERR_OUT_OF_MEMORY, "host buffer for an angle block");
    }
    if (!rc) rc = fdk_cra_load(cra, b->data.data());
    if (rc) {
        b->rc = rc;
        b->err = fdk_last_error_string();
    }
}

struct CraHandle {
    fdk_cra *p = nullptr;
    ~CraHandle() { if (p) fdk_cra_close(p); }
};
struct BinHandle {
    fdk_bin *p = nullptr;
    ~BinHandle() { if (p) fdk_bin_close(p); }
};
struct CtxHandle {
    fdk_ctx *p = nullptr;
    ~CtxHandle() { if (p) fdk_destroy(p); }
};

}  // namespace

extern "C" int fdk_setting_defaults(fdk_setting *s)
{
    if (!s) return fail(FDK_ERR_INVALID_ARGUMENT, "setting: null argument");
    s->gpu_id = 0;
    s->mem_limit_mib = 5 * 1024;
    s->alloc_limit_mib = 1024;
    s->slices_per_block = 16;
    s->x_direction = s->y_direction = s->z_direction = 0;
    s->vertical_center = 0;
    s->wait_for_enter = 1;
    return FDK_OK;
}

extern "C" int fdk_setting_load(const char *xml_path, fdk_setting *s)
{
    if (!s) return fail(FDK_ERR_INVALID_ARGUMENT, "setting: null argument");
    fdk_setting_defaults(s);
    std::ifstream f(xml_path ? xml_path : "setting.xml");
    std::stringstream ss;
    if (f.is_open()) ss << f.rdbuf();
    const std::string xml = ss.str();
    // LoadSetting (para_TC.cpp:192-297): root <GPU_FBP>, then <GPU>,
    // <RECONSTRUCTION> and <GENERAL>
    const size_t root = xml.find("<GPU_FBP");
    if (root == std::string::npos)
        return fail(FDK_ERR_INVALID_ARGUMENT, "setting: no <GPU_FBP> root element");
    const size_t root_end = xml.find("</GPU_FBP>", root);
    const std::string body = xml.substr(root, root_end == std::string::npos ? std::string::npos
                                                                              : root_end - root);
    std::string a;
    if (find_element(body, "GPU", &a)) {
        int_attribute(a, "ID", &s->gpu_id);
        int_attribute(a, "MemLimit_MiB", &s->mem_limit_mib);
        int_attribute(a, "AllocLimit_MiB", &s->alloc_limit_mib);
    }
    if (find_element(body, "RECONSTRUCTION", &a)) {
        int_attribute(a, "SimultaneouslyReconstructedSlices", &s->slices_per_block);
        bool_attribute(a, "X_direction", &s->x_direction);
        bool_attribute(a, "Y_direction", &s->y_direction);
        bool_attribute(a, "Z_direction", &s->z_direction);
        double_attribute(a, "DetectorVerticalCenterOffset", &s->vertical_center);
    }
    if (find_element(body, "GENERAL", &a)) bool_attribute(a, "WaitForEnter", &s->wait_for_enter);
    return FDK_OK;
}

extern "C" int fdk_setting_save(const char *xml_path, const fdk_setting *s)
{
    if (!s) return fail(FDK_ERR_INVALID_ARGUMENT, "setting: null argument");
    const char *path = xml_path ? xml_path : "setting.xml";
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(FDK_ERR_INVALID_ARGUMENT, std::string("setting: cannot write ") + path);
    // SaveSetting (para_TC.cpp:259-302)
    std::fprintf(f,
                 "<GPU_FBP>\n"
                 "    <GPU ID=\"%d\" MemLimit_MiB=\"%d\" AllocLimit_MiB=\"%d\" />\n"
                 "    <RECONSTRUCTION SimultaneouslyReconstructedSlices=\"%d\" X_direction=\"%d\" "
                 "Y_direction=\"%d\" Z_direction=\"%d\" DetectorVerticalCenterOffset=\"%.17g\" />\n"
                 "    <GENERAL WaitForEnter=\"%d\" />\n"
                 "</GPU_FBP>\n",
                 s->gpu_id, s->mem_limit_mib, s->alloc_limit_mib, s->slices_per_block,
                 s->x_direction ? 1 : 0, s->y_direction ? 1 : 0, s->z_direction ? 1 : 0,
                 s->vertical_center, s->wait_for_enter ? 1 : 0);
    const bool ok = std::fclose(f) == 0;
    return ok ? FDK_OK : fail(FDK_ERR_INVALID_ARGUMENT, std::string("setting: cannot write ") + path);
}

extern "C" int fdk_reconstruct_file(const char *cra_path, const fdk_setting *set,
                                    const fdk_run_options *opt_in, fdk_run_stats *stats)
{
    const Clock::time_point t_start = Clock::now();
    if (!cra_path || !set) return fail(FDK_ERR_INVALID_ARGUMENT, "driver: null argument");
    fdk_run_options opt{FDK_F32, FDK_FILTER_RAMLAK, 0, 1, -1};
    if (opt_in) opt = *opt_in;
    fdk_run_stats st{};
    if (set->slices_per_block < 1)
        return fail(FDK_ERR_INVALID_ARGUMENT, "driver: SimultaneouslyReconstructedSlices must be >= 1");

    // "<dir>/<name>.CRA" -> dir, name (para_TC.cpp:22-35)
    const std::string fn = cra_path;
    const size_t sl = fn.rfind('/');
    const std::string dir = sl == std::string::npos ? "." : fn.substr(0, sl == 0 ? 1 : sl);
    const size_t base = sl == std::string::npos ? 0 : sl + 1;
    size_t dt = fn.rfind('.');
    if (dt == std::string::npos || dt < base) dt = fn.size();
    const std::string name = fn.substr(base, dt - base);

    CraHandle cra;
    int rc = fdk_cra_open(dir.c_str(), name.c_str(), &cra.p);
    if (rc) return rc;
    fdk_cra_meta m;
    fdk_cra_get_meta(cra.p, &m);
    const bool verbose = opt.verbose != 0;
    if (verbose) {
        std::printf("\nMetaData:\n   file name:             %s\n   source name:           %s\n"
                    "   sample spacing:        %g u\n   distance:              %g c\n"
                    "   detector x pixels:     %d\n   detector z pixels:     %d\n"
                    "   number of projections: %d\n   number of blocks:      %d\n"
                    "   x dimension:           %d\n   z dimension:           %d\n"
                    "   rotate angle:          %g\n   clockwise:             %d\n",
                    m.raw_file, m.source_name, m.sample_spacing, m.distance, m.detector_x,
                    m.detector_z, m.num_of_projections, m.num_of_blocks, m.x_dimension,
                    m.z_dimension, m.rotate_angle, m.clockwise);
        for (int i = 0; i < m.num_of_blocks && i < m.num_centres; ++i)
            std::printf("   block[%d] centre:       %g\n", i, m.centre_of_rotation[i]);
        std::printf("   reconstruction scale:  %g\n", m.scale);
    }
    if (m.num_centres < m.num_of_blocks)
        return fail(FDK_ERR_INVALID_ARGUMENT, "driver: the header gives fewer CENTRE lines than "
                                              "NUMBER OF BLOCKS");

    // sizes (para_TC.cpp:53-63): the reconstruction is scaled as unsigned * double
    const int sx = m.detector_x, sz = m.detector_z;
    const unsigned rx0 = m.x_dimension == 0 ? (unsigned)m.detector_x : (unsigned)m.x_dimension;
    const int rec_x = (int)(unsigned)(rx0 * m.scale);
    const int rec_y = rec_x;
    const int rec_z = (int)(unsigned)((unsigned)m.z_dimension * m.scale);
    if (rec_x < 1 || rec_z < 1)
        return fail(FDK_ERR_INVALID_ARGUMENT, "driver: empty reconstruction after scaling");
    const int B = set->slices_per_block;
    const size_t proj_floats = (size_t)sx * sz;
    const unsigned long long proj_bytes = proj_floats * sizeof(float);
    const unsigned long long vol_bytes = (unsigned long long)rec_x * rec_y * rec_z * sizeof(float);
    const unsigned long long slab_bytes = (unsigned long long)rec_x * rec_y * B * sizeof(float);

    // volume residency: whole in HBM when it takes at most half of the free
    // memory, else streamed through a circular buffer of at most
    // 2 (Nz/8 + 32) + B slices
    const size_t free_mem = fdk_get_free_mem_info(set->gpu_id);
    const bool streamed = opt.volume_mode == FDK_VOLUME_STREAMED ||
                          (opt.volume_mode < 0 && (double)vol_bytes > 0.45 * (double)free_mem);
    const unsigned long long ring_bytes =
        (unsigned long long)std::min(2 * ((rec_z + 7) / 8 + 32) + B, rec_z) * rec_x * rec_y *
        sizeof(float);
    const unsigned long long held_vol = streamed ? ring_bytes : vol_bytes;
    // angle-block budget: the engine keeps the projections and the filtered
    // stack in FP32 plus the stored stack (4, 2 or 2 bytes a sample), the
    // volume (or its ring) and one collect slab; 10 % of the free memory is
    // left spare (para_TC.cpp:65 leaves 10 % of free minus one slab for the
    // projections)
    long long budget = opt.angle_block_bytes;
    if (budget <= 0) {
        if (free_mem == 0)
            return fail(FDK_ERR_HIP, "driver: cannot query the free memory of GPU " +
                                         std::to_string(set->gpu_id) + " (no such device?)");
        const double stored = opt.storage == FDK_F32 ? 4.0 : 2.0;
        const double usable = 0.9 * (double)free_mem - (double)held_vol - (double)slab_bytes;
        budget = usable > 0 ? (long long)(usable * 4.0 / (8.0 + stored)) : 0;
    }
    budget = (long long)((unsigned long long)budget / proj_bytes * proj_bytes);
    if (budget < (long long)proj_bytes)
        return fail(FDK_ERR_OUT_OF_MEMORY,
                    "driver: not even one projection fits beside the volume in free HBM (" +
                        std::to_string(free_mem >> 20) + " MiB free, volume held " +
                        std::to_string(held_vol >> 20) + " MiB)");
    int angle_blocks = 0;
    if ((rc = fdk_cra_split(cra.p, budget, &angle_blocks))) return rc;
    const int scan_blocks = m.num_of_blocks;
    if (verbose)
        std::printf("\nGPU memory\n  available : %.1f MiB\n  angle block budget: %.1f MiB\n"
                    "\nreconstruction:\n  size x: %d\n  size y: %d\n  size z: %d\n"
                    "  size: %.1f MiB per scan block\n  memory for %d slices: %.1f MiB\n"
                    "\nscan blocks: %d\n\nsinograms:\n  size x: %d\n  size y: %d\n  size z: %d\n"
                    "  distance: %g px\n  vertical offset: %g px\n  sample spacing: %g um\n"
                    "  delta fi: %g rad\n  split to %d block(s)\n",
                    free_mem / 1048576.0, budget / 1048576.0, rec_x, rec_y, rec_z,
                    vol_bytes / 1048576.0, B, slab_bytes / 1048576.0, scan_blocks, sx,
                    m.num_of_projections, sz, m.distance_px, set->vertical_center,
                    m.sample_spacing, m.delta_fi, angle_blocks);

    BinHandle bin;
    if ((rc = fdk_bin_open(dir.c_str(), name.c_str(), rec_x, rec_y, rec_z, scan_blocks, B, &bin.p)))
        return rc;
    CtxHandle ctx;
    if ((rc = fdk_create(set->gpu_id, &ctx.p))) return rc;

    // every (scan block, angle block) in order; block i+1 loads while i runs
    const int total = scan_blocks * angle_blocks;
    Block blk[2];
    blk[0].scan = 0;
    blk[0].angle = 0;
    load_block(cra.p, proj_floats, &blk[0]);
    std::thread loader;
    struct Joiner {                    // every return waits for the loader
        std::thread &t;
        ~Joiner() { if (t.joinable()) t.join(); }
    } joiner{loader};
    auto join_loader = [&]() {
        if (loader.joinable()) loader.join();
    };
    for (int i = 0; i < total; ++i) {
        Block &cur = blk[i & 1];
        if (i > 0) {
            const Clock::time_point tw = Clock::now();
            join_loader();
            st.seconds_load_wait += seconds_since(tw);
        }
        if (cur.rc) return fail(cur.rc, cur.err);
        st.bytes_read += (double)proj_bytes * cur.nang;
        if (i + 1 < total) {
            Block &nxt = blk[(i + 1) & 1];
            nxt.scan = (i + 1) / angle_blocks;
            nxt.angle = (i + 1) % angle_blocks;
            try {
                loader = std::thread(load_block, cra.p, proj_floats, &nxt);
            } catch (const std::system_error &) {
                load_block(cra.p, proj_floats, &nxt);   // no thread available: load now
            }
        }
        const int scan = cur.scan, k = cur.angle;
        const float centre = (float)m.centre_of_rotation[scan];
        if (k == 0) {
            if ((rc = fdk_bin_set_scan_block(bin.p, scan))) return rc;
            if (verbose)
                std::printf("\n*****************\ncalculating scan block: %d / %d\n"
                            "  centre of rotation: %g px\n\n",
                            scan + 1, scan_blocks, centre);
        }
        const float rot = (float)cur.rot;
        if (verbose)
            std::printf("calculating sinogram block %d / %d\n  rotation correction: %g rad\n", k + 1,
                        angle_blocks, rot);
        // para_TC.cpp:134-155, with Initialize and CalculateReconstruction
        // fused so that the upload of the block runs behind its filtering and
        // backprojection
        const Clock::time_point te = Clock::now();
        fdk_ctx *c = ctx.p;
        if (!(rc = fdk_set_block_size(c, B)) &&
            !(rc = fdk_set_allocation_limit(c, (long long)set->alloc_limit_mib * 1024 * 1024)) &&
            !(rc = fdk_set_collecting(c, k > 0)) && !(rc = fdk_set_show_info(c, verbose && i == 0)) &&
            !(rc = fdk_set_axis_direction(c, set->x_direction, set->y_direction, set->z_direction)) &&
            !(rc = fdk_set_sinogram_properties(c, nullptr, sx, centre, cur.nang, sz,
                                               (float)(((double)sz - 1) / 2. + set->vertical_center),
                                               m.delta_fi, m.distance_px,
                                               (float)(m.sample_spacing * 0.0001))) &&
            !(rc = fdk_set_reconstruction_properties(c, rec_x, (float)(((double)rec_x - 1) / 2.),
                                                     rec_y, (float)(((double)rec_y - 1) / 2.), rec_z,
                                                     0.f, rot)) &&
            !(rc = fdk_set_scale_xyz(c, (float)(1. / m.scale))) &&
            !(rc = fdk_set_storage(c, opt.storage)) &&
            !(rc = fdk_set_volume_mode(c, streamed ? FDK_VOLUME_STREAMED : FDK_VOLUME_RESIDENT)) &&
            !(rc = fdk_set_filter(c, opt.filter, nullptr, 0)))
            rc = fdk_reconstruct_host(c, cur.data.data(), fdk_bin_get_block, bin.p);
        st.seconds_engine += seconds_since(te);
        if (rc) return rc;
        if (k == angle_blocks - 1) {
            const Clock::time_point tf = Clock::now();
            if ((rc = fdk_bin_flush(bin.p))) return rc;
            st.seconds_flush += seconds_since(tf);
            if (verbose) std::printf("\nscan block %d complete\n", scan + 1);
        }
    }
    join_loader();
    const Clock::time_point tf = Clock::now();
    fdk_bin *b = bin.p;
    bin.p = nullptr;
    if ((rc = fdk_bin_close(b))) return rc;
    st.seconds_flush += seconds_since(tf);
    st.scan_blocks = scan_blocks;
    st.angle_blocks = angle_blocks;
    st.bytes_written = (double)vol_bytes * scan_blocks;
    st.seconds_total = seconds_since(t_start);
    if (verbose) std::printf("\ntotal time: %.3f s\n", st.seconds_total);
    if (stats) *stats = st;
    if (verbose && set->wait_for_enter && isatty(STDIN_FILENO)) {
        std::printf("\npress enter\n");
        std::fflush(stdout);
        std::getchar();
    }
    return FDK_OK;
}
