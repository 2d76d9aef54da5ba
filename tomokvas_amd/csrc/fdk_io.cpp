// fdk_io.cpp -- the reference's projection container (.CRA + raw floats) and
// volume file (.bin + .siz) on POSIX, host code only.
//
// Reader: cra_con (FBPG/mucat_file.cpp:6-428).  Writer: bin_siz
// (FBPG/mucat_file.cpp:432-793).  The writer keeps the reference's three-buffer
// rotation -- the engine fills one slab buffer while the I/O thread writes the
// previous slab and prefetches the next -- with two changes: the I/O thread is
// joined before every hand-out (the reference's completion flag is set by the
// thread itself when it starts, so a fast caller could skip the wait), and a
// slab requested again right after it was handed out (a one-slab volume
// collecting over angle blocks) continues from the buffer just filled instead
// of a prefetch that read the file before that buffer was written.
#include "tomokvas/fdk_io.h"
#include "tomokvas/fdk.h"
#include "fdk_internal.hpp"

#include <hip/hip_runtime.h>
#include <fcntl.h>
#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

using tkv::fail;

namespace {

constexpr double kPi = 3.14159265358979323846;

std::string join_path(const char *dir, const std::string &name)
{
    std::string d = dir && *dir ? dir : ".";
    if (d.back() != '/') d += '/';
    return d + name;
}

void copy_field(char (&dst)[256], const std::string &s)
{
    std::snprintf(dst, sizeof dst, "%s", s.c_str());
}

}  // namespace

// ============================================================== .CRA reader
struct fdk_cra {
    std::string dir;
    fdk_cra_meta m{};
    int scan_block = 0;
    int nblocks = 1;         // angle blocks
    int per_block = 0;       // projections per full block
    int rest = 0;            // projections of a shorter last block (0: none)
    int active = 0;          // active angle block
};

extern "C" int fdk_cra_open(const char *dir, const char *name, fdk_cra **out)
{
    if (!name || !out) return fail(FDK_ERR_INVALID_ARGUMENT, "cra: null argument");
    *out = nullptr;
    const std::string fn = join_path(dir, std::string(name) + ".CRA");
    std::ifstream file(fn);
    if (!file.is_open()) return fail(FDK_ERR_INVALID_ARGUMENT, "cra: cannot open " + fn);

    auto *c = new fdk_cra;
    c->dir = dir && *dir ? dir : ".";
    fdk_cra_meta &m = c->m;
    // defaults of the reference constructor (mucat_file.cpp:26-39)
    m.num_of_blocks = 1;
    m.scale = 1.0;
    // header keys, matched at the start of a line (mucat_file.cpp:15-28, 81-132)
    static const char *keys[14] = {
        "FILE NAME:              ", "SOURCE NAME:            ", "SAMPLE SPACING:         ",
        "DISTANCE:               ", "DETECTOR X PIXELS:      ", "DETECTOR Z PIXELS:      ",
        "NUMBER OF PROJECTIONS:  ", "X DIMENSION:            ", "Z DIMENSION:            ",
        "ROTATE ANGLE:           ", "CENTRE:                 ", "NUMBER OF BLOCKS:       ",
        "CLOCKWISE:",               "SCALE:                  "};
    std::string line;
    while (std::getline(file, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();   // CRLF files
        for (int i = 0; i < 14; ++i) {
            const size_t kl = std::strlen(keys[i]);
            if (line.compare(0, kl, keys[i]) != 0) continue;
            const std::string rest = line.substr(kl);
            std::istringstream stm(rest);
            switch (i) {
            case 0: copy_field(m.raw_file, rest); break;
            case 1: copy_field(m.source_name, rest); break;
            case 2: stm >> m.sample_spacing; break;
            case 3: stm >> m.distance; break;
            case 4: stm >> m.detector_x; break;
            case 5: stm >> m.detector_z; break;
            case 6: stm >> m.num_of_projections; break;
            case 7: stm >> m.x_dimension; break;
            case 8: stm >> m.z_dimension; break;
            case 9: stm >> m.rotate_angle; break;
            case 10:
                if (m.num_centres < 64) stm >> m.centre_of_rotation[m.num_centres++];
                break;
            case 11: stm >> m.num_of_blocks; break;
            case 12: m.clockwise = 1; break;   // the key alone marks a clockwise scan
            case 13: stm >> m.scale; break;
            }
        }
    }
    if (m.sample_spacing <= 0 || m.distance <= 0 || m.detector_x < 1 || m.detector_z < 1 ||
        m.num_of_projections < 1 || m.num_of_blocks < 1) {
        delete c;
        return fail(FDK_ERR_INVALID_ARGUMENT,
                    "cra: header lacks spacing, distance, detector size, projections or blocks");
    }
    // derived quantities and clamps (mucat_file.cpp:135-160), with the
    // reference's int / float arithmetic: detector_x / 2 and x_dimension / 2
    // are integer divisions, the rest float
    m.distance_px = (float)(m.distance * 10000. / m.sample_spacing);
    m.delta_fi = (float)(2 * kPi / m.num_of_projections);
    const float distance = m.distance_px;
    const float phi = std::atan((float)(m.detector_x / 2) / distance);
    unsigned tmp = 2u * (unsigned)((int)(distance * std::sin(phi) + 1));
    if ((unsigned)m.x_dimension > tmp || m.x_dimension < 1) m.x_dimension = (int)tmp;
    const float zf = (distance - (float)(m.x_dimension / 2)) / distance * (float)m.detector_z;
    tmp = zf > 0.f ? (unsigned)zf : 0u;
    tmp &= ~1u;
    if ((unsigned)m.z_dimension > tmp) m.z_dimension = (int)tmp;
    if (m.z_dimension < 1) m.z_dimension = 1;
    // one angle block until split
    c->nblocks = 1;
    c->per_block = m.num_of_projections;
    *out = c;
    return FDK_OK;
}

extern "C" void fdk_cra_close(fdk_cra *cra) { delete cra; }

extern "C" int fdk_cra_get_meta(const fdk_cra *cra, fdk_cra_meta *out)
{
    if (!cra || !out) return fail(FDK_ERR_INVALID_ARGUMENT, "cra: null argument");
    *out = cra->m;
    return FDK_OK;
}

extern "C" int fdk_cra_set_scan_block(fdk_cra *cra, int block)
{
    if (!cra) return fail(FDK_ERR_INVALID_ARGUMENT, "cra: null argument");
    cra->scan_block = block >= 0 && block < cra->m.num_of_blocks ? block : 0;
    return cra->scan_block;
}

extern "C" int fdk_cra_split(fdk_cra *cra, long long max_bytes, int *nblocks)
{
    if (!cra || !nblocks) return fail(FDK_ERR_INVALID_ARGUMENT, "cra: null argument");
    const fdk_cra_meta &m = cra->m;
    const unsigned long long proj = (unsigned long long)m.detector_x * m.detector_z * sizeof(float);
    if (max_bytes < (long long)proj)
        return fail(FDK_ERR_INVALID_ARGUMENT, "cra: angle-block budget below one projection");
    // CalculateSinBlocks (mucat_file.cpp:354-375)
    const unsigned long long total = proj * (unsigned long long)m.num_of_projections;
    const unsigned long long budget = (unsigned long long)max_bytes;
    int n = (int)(total / budget);
    // a budget above the whole scan holds every projection (clamped so the
    // count fits an int)
    const int per = (int)std::min<unsigned long long>(budget / proj, m.num_of_projections);
    int rest = 0;
    if (total % budget > 0) {
        rest = m.num_of_projections - n * per;
        n += 1;
    }
    // when the budget is not a multiple of one projection the split above can
    // leave projections in no block (no remainder although n * per < N) or give
    // the last block more projections than the budget holds: refuse both
    const int covered = rest > 0 ? (n - 1) * per + rest : n * per;
    if (covered != m.num_of_projections || rest > per)
        return fail(FDK_ERR_UNSUPPORTED,
                    "cra: a budget of " + std::to_string(max_bytes) +
                        " bytes splits the angles unevenly (projections left out or a "
                        "last block over budget); use a multiple of one projection's bytes");
    cra->nblocks = n;
    cra->per_block = per;
    cra->rest = rest;
    cra->active = 0;
    *nblocks = n;
    return FDK_OK;
}

extern "C" int fdk_cra_set_angle_block(fdk_cra *cra, int block)
{
    if (!cra) return fail(FDK_ERR_INVALID_ARGUMENT, "cra: null argument");
    cra->active = block >= 0 && block < cra->nblocks ? block : 0;
    return cra->active;
}

static int projections_of(const fdk_cra *c)
{
    // GetNoProjection (mucat_file.cpp:399-409)
    int n = c->active < c->nblocks ? c->per_block : 0;
    if (c->active == c->nblocks - 1 && c->rest > 0) n = c->rest;
    return n;
}

extern "C" int fdk_cra_block_projections(const fdk_cra *cra, int *n)
{
    if (!cra || !n) return fail(FDK_ERR_INVALID_ARGUMENT, "cra: null argument");
    *n = projections_of(cra);
    return FDK_OK;
}

extern "C" int fdk_cra_rotation(const fdk_cra *cra, double *rad)
{
    if (!cra || !rad) return fail(FDK_ERR_INVALID_ARGUMENT, "cra: null argument");
    // GetRotRad (mucat_file.cpp:288-316).  Anticlockwise: block k starts k
    // full blocks in.  Clockwise: the loaded order is reversed, so block k's
    // first projection is its last one on file; the terms below place it at
    // rotate_angle - (k*per + n_k - 1) * delta_fi, modulo 2 pi.
    const double dfi = cra->m.delta_fi;
    const int k = cra->active, nb = cra->nblocks, per = cra->per_block, rest = cra->rest;
    double r = 0;
    if (cra->m.clockwise) {
        r += dfi;
        if (nb > 1) {
            const int x = rest > 0 ? nb - 2 : nb - 1;
            r += dfi * rest + dfi * x * per;
            if (k == nb - 1 && rest > 0)
                r -= (double)(k - 1) * per * dfi + rest * dfi;
            else
                r -= (double)k * per * dfi;
        }
    } else {
        r += (double)k * per * dfi;
    }
    *rad = r + cra->m.rotate_angle * kPi / 180.;
    return FDK_OK;
}

extern "C" int fdk_cra_load(const fdk_cra *cra, float *dst)
{
    if (!cra || !dst) return fail(FDK_ERR_INVALID_ARGUMENT, "cra: null argument");
    const fdk_cra_meta &m = cra->m;
    const std::string fn = join_path(cra->dir.c_str(), m.raw_file);
    const int fd = open(fn.c_str(), O_RDONLY);
    if (fd < 0) return fail(FDK_ERR_INVALID_ARGUMENT, "cra: cannot open raw file " + fn);
    const size_t proj = (size_t)m.detector_x * m.detector_z;
    const int n = projections_of(cra);
    const size_t sin = (size_t)m.detector_x * n;   // one sinogram row of the block
    // LoadData (mucat_file.cpp:205-263): scan blocks follow one another, the
    // angle blocks are consecutive runs of projections within a scan block
    const unsigned long long scan = (unsigned long long)proj * m.num_of_projections;
    const unsigned long long off =
        ((unsigned long long)cra->active * cra->per_block * proj + scan * cra->scan_block) *
        sizeof(float);
    // projections [j0, j1) per thread: positioned reads of whole projections,
    // each scattered row by row into the sinogram-major block; the reads and
    // the scatter are memory/disk bound, so several run at once
    const int nthreads =
        std::max(1, std::min({n, 16, (int)std::max(1u, std::thread::hardware_concurrency())}));
    std::vector<int> bad(nthreads, 0);
    auto work = [&](int t) {
        std::vector<float> p(proj);
        const int j0 = (int)((long long)n * t / nthreads), j1 = (int)((long long)n * (t + 1) / nthreads);
        for (int j = j0; j < j1 && !bad[t]; ++j) {
            const size_t want = proj * sizeof(float);
            size_t got = 0;
            while (got < want) {
                const ssize_t r = pread(fd, (char *)p.data() + got, want - got,
                                        (off_t)(off + (unsigned long long)j * want + got));
                if (r <= 0) break;
                got += (size_t)r;
            }
            if (got != want) {
                bad[t] = 1;
                break;
            }
            const int a = m.clockwise ? n - 1 - j : j;
            for (int s = 0; s < m.detector_z; ++s)
                std::memcpy(dst + sin * s + (size_t)a * m.detector_x,
                            p.data() + (size_t)s * m.detector_x, m.detector_x * sizeof(float));
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) {
        try {
            pool.emplace_back(work, t);
        } catch (const std::system_error &) {
            work(t);   // no thread available: this share now
        }
    }
    work(0);
    for (auto &th : pool) th.join();
    close(fd);
    for (int v : bad)
        if (v)
            return fail(FDK_ERR_INVALID_ARGUMENT,
                        "cra: raw file shorter than its header says (" + fn + ")");
    return FDK_OK;
}

// ============================================================== .bin writer
namespace {

struct IoJob {
    const float *write_src = nullptr;   // slab to write
    unsigned long long write_pos = 0, write_bytes = 0;
    float *read_dst = nullptr;          // slab to prefetch
    unsigned long long read_pos = 0, read_bytes = 0;
};

}  // namespace

struct fdk_bin {
    std::string bin_path, siz_path;
    int x = 0, y = 0, z = 0, scan_blocks = 1, spb = 16;
    size_t im_size = 0, block_floats = 0;
    int blocks = 0, rest = 0;
    unsigned long long scan_off = 0;
    FILE *f = nullptr;
    tkv::HostBuffer buf[3];
    int work = -1, work_block = -1;   // buffer handed out last, its slab
    int pre = -1, pre_block = -1;     // buffer holding a prefetched slab
    std::thread io;
    int io_rc = 0;
    std::string io_err;

    unsigned long long slab_bytes(int n) const
    {
        const size_t sl = (n == blocks - 1 && rest) ? (size_t)rest : (size_t)spb;
        return (unsigned long long)im_size * sl * sizeof(float);
    }
    unsigned long long slab_pos(int n) const
    {
        return scan_off + (unsigned long long)im_size * n * spb * sizeof(float);
    }
};

static void run_io(fdk_bin *b, IoJob job)
{
    // write first, then read: a prefetch of the slab just written sees it
    if (job.write_src && job.write_bytes) {
        if (fseeko(b->f, (off_t)job.write_pos, SEEK_SET) != 0 ||
            std::fwrite(job.write_src, 1, job.write_bytes, b->f) != job.write_bytes) {
            b->io_rc = FDK_ERR_CALLBACK;
            b->io_err = "bin: write failed";
            return;
        }
    }
    if (job.read_dst && job.read_bytes) {
        std::fflush(b->f);
        if (fseeko(b->f, (off_t)job.read_pos, SEEK_SET) != 0 ||
            std::fread(job.read_dst, 1, job.read_bytes, b->f) != job.read_bytes) {
            b->io_rc = FDK_ERR_CALLBACK;
            b->io_err = "bin: read failed";
        }
    }
}

static int join_io(fdk_bin *b)
{
    if (b->io.joinable()) b->io.join();
    if (b->io_rc) {
        const int rc = b->io_rc;
        b->io_rc = 0;
        return fail(rc, b->io_err);
    }
    return FDK_OK;
}

extern "C" int fdk_bin_open(const char *dir, const char *name, int x, int y, int z,
                            int scan_blocks, int slices_per_block, fdk_bin **out)
{
    if (!name || !out) return fail(FDK_ERR_INVALID_ARGUMENT, "bin: null argument");
    *out = nullptr;
    if (x < 1 || y < 1 || z < 1 || scan_blocks < 1 || slices_per_block < 1)
        return fail(FDK_ERR_INVALID_ARGUMENT, "bin: sizes must be positive");
    auto *b = new fdk_bin;
    b->bin_path = join_path(dir, std::string(name) + ".bin");
    b->siz_path = join_path(dir, std::string(name) + ".siz");
    b->x = x;
    b->y = y;
    b->z = z;
    b->scan_blocks = scan_blocks;
    b->spb = slices_per_block;
    b->im_size = (size_t)x * y;
    b->block_floats = b->im_size * slices_per_block;
    // SetSlicesPerBlock (mucat_file.cpp:494-505)
    b->blocks = (z + slices_per_block - 1) / slices_per_block;
    b->rest = z % slices_per_block;
    // CreateFile (mucat_file.cpp:535-556): the whole file at its final size
    b->f = std::fopen(b->bin_path.c_str(), "w+b");
    const off_t size = (off_t)((unsigned long long)scan_blocks * b->im_size * z * sizeof(float));
    if (!b->f || ftruncate(fileno(b->f), size) != 0) {
        const std::string why = std::strerror(errno);
        if (b->f) std::fclose(b->f);
        delete b;
        return fail(FDK_ERR_INVALID_ARGUMENT, "bin: cannot create the volume file: " + why);
    }
    for (auto &v : b->buf) {
        if (!v.alloc(b->block_floats)) {
            std::fclose(b->f);
            delete b;
            return fail(FDK_ERR_OUT_OF_MEMORY, "bin: slab buffers");
        }
    }
    *out = b;
    return FDK_OK;
}

extern "C" int fdk_bin_set_scan_block(fdk_bin *b, int scan_block)
{
    if (!b) return fail(FDK_ERR_INVALID_ARGUMENT, "bin: null argument");
    if (scan_block < 0 || scan_block >= b->scan_blocks)
        return fail(FDK_ERR_INVALID_ARGUMENT, "bin: scan block out of range");
    int rc = fdk_bin_flush(b);
    if (rc) return rc;
    b->scan_off = (unsigned long long)b->im_size * scan_block * b->z * sizeof(float);
    return FDK_OK;
}

extern "C" float *fdk_bin_get_block(void *handle, int n, int collecting)
{
    auto *b = (fdk_bin *)handle;
    if (!b) {
        fail(FDK_ERR_INVALID_ARGUMENT, "bin: null argument");
        return nullptr;
    }
    if (n < 0 || n >= b->blocks) {
        fail(FDK_ERR_INVALID_ARGUMENT, "bin: slab " + std::to_string(n) + " out of range");
        return nullptr;
    }
    if (join_io(b)) return nullptr;
    const int w = b->work;
    // the buffer for slab n: the prefetched one, a copy of the slab just
    // handed out (the same slab again), or a free one (read now if collecting)
    int nb = -1;
    if (collecting && b->pre >= 0 && b->pre_block == n) nb = b->pre;
    auto free_buf = [&](int not1, int not2) {
        for (int i = 0; i < 3; ++i)
            if (i != not1 && i != not2) return i;
        return -1;
    };
    if (nb < 0) {
        nb = free_buf(w, b->pre);
        if (collecting && w >= 0 && b->work_block == n) {
            std::memcpy(b->buf[nb].data(), b->buf[w].data(), b->slab_bytes(n));
        } else if (collecting) {
            if (fseeko(b->f, (off_t)b->slab_pos(n), SEEK_SET) != 0 ||
                std::fread(b->buf[nb].data(), 1, b->slab_bytes(n), b->f) != b->slab_bytes(n)) {
                fail(FDK_ERR_CALLBACK, "bin: read failed");
                return nullptr;
            }
        }
    }
    IoJob job;
    if (w >= 0) {
        job.write_src = b->buf[w].data();
        job.write_pos = b->slab_pos(b->work_block);
        job.write_bytes = b->slab_bytes(b->work_block);
    }
    // prefetch the next slab -- after the last one, slab 0 for the next angle
    // block (mucat_file.cpp:614-627) -- unless it is the slab being handed out
    const int next = n == b->blocks - 1 ? 0 : n + 1;
    b->pre = -1;
    b->pre_block = -1;
    if ((collecting || n == b->blocks - 1) && next != n) {
        const int pb = free_buf(nb, w);
        job.read_dst = b->buf[pb].data();
        job.read_pos = b->slab_pos(next);
        job.read_bytes = b->slab_bytes(next);
        b->pre = pb;
        b->pre_block = next;
    }
    b->work = nb;
    b->work_block = n;
    if (job.write_bytes || job.read_bytes) {
        try {
            b->io = std::thread(run_io, b, job);
        } catch (const std::system_error &) {
            run_io(b, job);   // no thread available: do it now
        }
    }
    return b->buf[nb].data();
}

extern "C" int fdk_bin_flush(fdk_bin *b)
{
    if (!b) return fail(FDK_ERR_INVALID_ARGUMENT, "bin: null argument");
    int rc = join_io(b);
    if (rc) return rc;
    if (b->work >= 0) {
        IoJob job;
        job.write_src = b->buf[b->work].data();
        job.write_pos = b->slab_pos(b->work_block);
        job.write_bytes = b->slab_bytes(b->work_block);
        run_io(b, job);
        b->work = -1;
        b->work_block = -1;
        if ((rc = join_io(b))) return rc;
    }
    b->pre = -1;
    b->pre_block = -1;
    std::fflush(b->f);
    return FDK_OK;
}

extern "C" int fdk_bin_close(fdk_bin *b)
{
    if (!b) return fail(FDK_ERR_INVALID_ARGUMENT, "bin: null argument");
    int rc = fdk_bin_flush(b);
    if (std::fclose(b->f) != 0 && !rc) rc = fail(FDK_ERR_CALLBACK, "bin: close failed");
    // SaveHeader (mucat_file.cpp:522-531)
    FILE *h = std::fopen(b->siz_path.c_str(), "w");
    if (!h) {
        if (!rc) rc = fail(FDK_ERR_CALLBACK, "bin: cannot write " + b->siz_path);
    } else {
        std::fprintf(h, "%d %d %d 0", b->x, b->y, b->z * b->scan_blocks);
        std::fclose(h);
    }
    delete b;
    return rc;
}
