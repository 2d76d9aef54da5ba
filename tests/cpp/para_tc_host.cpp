// para_tc_host.cpp -- a host program driving the engine exactly as
// FBPG/para_TC.cpp:116-153 does, through the drop-in C++ class of
// include/tomokvas/FilteredBackProjection.hpp: one engine per angle ("sin")
// block, SetGPUid after construction, SetCollecting(block) so every block after
// the first adds onto the slabs, the projection buffer freed right after
// Initialize, slabs handed to a bin_siz-like sink through GetBlock(int, bool).
//
//   para_tc_host IN.raw OUT.raw Nu Nv Nang Nx Ny Nz B D tau scale blocks
//
// IN.raw: float32 projections [v][beta][u] of all Nang angles (beta_k = k*dbeta,
// dbeta = 2 pi / Nang in float); OUT.raw: the float32 volume [z][y][x].
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "tomokvas/FilteredBackProjection.hpp"

// Holds the whole volume in memory.  GetBlock(n, collecting) hands out one
// slab buffer of Nx*Ny*B floats, pre-filled with slab n's partial sum when
// collecting (the reference's I/O thread reads it back from the .bin file,
// FBPG/mucat_file.cpp:576-642); a slab is stored when the next one is asked for
// and by FlushBuffer.
class MemoryBinSiz {
  public:
    MemoryBinSiz(int nx, int ny, int nz, int b)
        : nx_(nx), ny_(ny), nz_(nz), b_(b), vol_((size_t)nx * ny * nz, 0.f),
          buf_((size_t)nx * ny * b, 0.f)
    {
    }
    float *GetBlock(int n, bool collecting)
    {
        FlushBuffer();
        const size_t plane = (size_t)nx_ * ny_;
        std::fill(buf_.begin(), buf_.end(), 0.f);
        if (collecting)
            std::memcpy(buf_.data(), vol_.data() + (size_t)n * b_ * plane,
                        slices(n) * plane * sizeof(float));
        pending_ = n;
        ++calls_;
        if (collecting) ++collect_calls_;
        return buf_.data();
    }
    void FlushBuffer()
    {
        if (pending_ < 0) return;
        const size_t plane = (size_t)nx_ * ny_;
        std::memcpy(vol_.data() + (size_t)pending_ * b_ * plane, buf_.data(),
                    slices(pending_) * plane * sizeof(float));
        pending_ = -1;
    }
    const std::vector<float> &volume() const { return vol_; }
    int calls() const { return calls_; }
    int collect_calls() const { return collect_calls_; }

  private:
    size_t slices(int n) const { return (size_t)std::min(b_, nz_ - n * b_); }
    int nx_, ny_, nz_, b_;
    std::vector<float> vol_, buf_;
    int pending_ = -1, calls_ = 0, collect_calls_ = 0;
};

int main(int argc, char **argv)
{
    if (argc != 14) {
        std::cerr << "usage: para_tc_host IN.raw OUT.raw Nu Nv Nang Nx Ny Nz B D tau scale "
                     "blocks\n";
        return 2;
    }
    const int Nu = std::atoi(argv[3]), Nv = std::atoi(argv[4]), Nang = std::atoi(argv[5]);
    const int Nx = std::atoi(argv[6]), Ny = std::atoi(argv[7]), Nz = std::atoi(argv[8]);
    const int B = std::atoi(argv[9]);
    const float D = (float)std::atof(argv[10]), tau = (float)std::atof(argv[11]);
    const float scale = (float)std::atof(argv[12]);
    const int blocks = std::atoi(argv[13]);
    const float dbeta = (float)(2.0 * M_PI / Nang);
    const size_t n = (size_t)Nu * Nv * Nang;

    std::vector<float> all(n);
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(all.data(), sizeof(float), n, f) != n) {
        std::cerr << "cannot read " << argv[1] << "\n";
        return 2;
    }
    std::fclose(f);

    MemoryBinSiz output_file(Nx, Ny, Nz, B);
    const float centre = 0.5f * (float)(Nu - 1);
    for (int SinBlock = 0; SinBlock < blocks; ++SinBlock) {
        const int a0 = (int)((long long)Nang * SinBlock / blocks);
        const int a1 = (int)((long long)Nang * (SinBlock + 1) / blocks);
        const int sy = a1 - a0;
        // the block's projections in [v][beta][u], as cra_con::LoadData leaves them
        float *data = new float[(size_t)Nu * Nv * sy];
        for (int v = 0; v < Nv; ++v)
            std::memcpy(data + (size_t)v * sy * Nu, all.data() + ((size_t)v * Nang + a0) * Nu,
                        (size_t)sy * Nu * sizeof(float));
        const float rot = (float)((double)a0 * (double)dbeta);

        tomokvas::FilteredBackProjection *FBP = new tomokvas::FilteredBackProjection();
        FBP->SetGPUid(0);
        FBP->SetBlockSize(B);
        FBP->SetAllocationLimit((long long)1024 * 1024 * 1024);
        FBP->SetCollecting(SinBlock);
        FBP->SetShowInfo(false);
        FBP->SetAxisDirection(false, false, false);
        FBP->SetSinogramProperties(data, Nu, centre, sy, Nv, ((double)Nv - 1) / 2., dbeta, D,
                                   tau);
        FBP->SetReconstructionProperties(&output_file, Nx, ((double)Nx - 1) / 2., Ny,
                                         ((double)Ny - 1) / 2., Nz, 0, rot);
        FBP->SetScaleXYZ(scale);
        if (FBP->Initialize() != 0) {
            delete FBP;
            delete[] data;
            return 1;
        }
        delete[] data;   // the projections are on the GPU now
        const int rc = FBP->CalculateReconstruction();
        delete FBP;
        if (rc != 0) return 1;
    }
    output_file.FlushBuffer();

    f = std::fopen(argv[2], "wb");
    const std::vector<float> &vol = output_file.volume();
    if (!f || std::fwrite(vol.data(), sizeof(float), vol.size(), f) != vol.size()) {
        std::cerr << "cannot write " << argv[2] << "\n";
        return 2;
    }
    std::fclose(f);
    std::cout << "slabs handed out " << output_file.calls() << ", collecting "
              << output_file.collect_calls() << std::endl;
    return 0;
}
