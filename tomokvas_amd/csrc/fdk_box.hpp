// fdk_box.hpp -- the detector footprint of a block of voxels for one angle,
// shared by the backprojector (which stages the footprint from HBM into LDS)
// and its adjoint (which accumulates its splats there).
#pragma once

#include <hip/hip_runtime.h>

namespace tkv {

// box: at most BOX_W - 1 columns (+1 neighbour) and box_rows<ZR>() - 1 rows
// (+1) for a block of 16 x 16 columns x ZR slices; the row budget covers a
// magnification of about 1.4 at two detector samples per voxel
constexpr int BOX_W = 64;
template <int ZR>
constexpr int box_rows() { return ZR > 32 ? 128 : ZR > 24 ? 96 : ZR > 16 ? 80 : 64; }
constexpr int BOX_ALIGN = 4;   // box columns start on a multiple of 4 samples

// ---- footprint box of a block for one angle ----------------------------------------
// Returns (bu0, bv0, bw | inner << 16, bh): the detector samples
// [bu0, bu0+bw) x [bv0, bv0+bh)
// that the block's bilinear taps can touch, widened by 2 samples on each side
// for the rounding of the per-column values (the columns evaluate the same
// float expressions as the corners); bw == 0: the block misses the detector;
// bw < 0: go through HBM directly (box beyond the LDS tile, or a block corner
// near or behind the source plane, where U is not monotone over the block).
__device__ inline int4 footprint(float c, float s, float x0, float x1, float y0, float y1,
                                 float azlo, float azhi, float cX, float cZ, float D, int Nu,
                                 int Nv, int box_h)
{
    float umin = 3.0e38f, umax = -3.0e38f, Umin = 3.0e38f, Umax = -3.0e38f;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float aX = (k & 1) ? x1 : x0, aY = (k & 2) ? y1 : y0;
        const float S = fmaf(aX, c, aY * s);
        const float T = fmaf(aY, c, -(aX * s));
        const float den = D - S;
        ok = ok && den > 0.01f * D;
        const float U = D / den;
        const float u = fmaf(T, U, cX);
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
        Umin = fminf(Umin, U);
        Umax = fmaxf(Umax, U);
    }
    if (!ok) return make_int4(0, 0, -1, 0);
    const float p0 = azlo * Umin, p1 = azlo * Umax, p2 = azhi * Umin, p3 = azhi * Umax;
    const float vlo = fminf(fminf(p0, p1), fminf(p2, p3)) + cZ;
    const float vhi = fmaxf(fmaxf(p0, p1), fmaxf(p2, p3)) + cZ;
    // clamp before converting; anything beyond the detector by 4 samples is empty
    const float fu0 = floorf(fminf(fmaxf(umin, -8.0f), (float)Nu + 8.0f));
    const float fu1 = floorf(fminf(fmaxf(umax, -8.0f), (float)Nu + 8.0f));
    const float fv0 = floorf(fminf(fmaxf(vlo, -8.0f), (float)Nv + 8.0f));
    const float fv1 = floorf(fminf(fmaxf(vhi, -8.0f), (float)Nv + 8.0f));
    const int bu0 = max((int)fu0 - 2, 0) & ~(BOX_ALIGN - 1);
    const int bu1 = min((int)fu1 + 3, Nu - 1);
    const int bv0 = max((int)fv0 - 2, 0);
    const int bv1 = min((int)fv1 + 3, Nv - 1);
    const int bw = bu1 - bu0 + 1, bh = bv1 - bv0 + 1;
    if (bw <= 0 || bh <= 0) return make_int4(0, 0, 0, 0);
    if (bw >= BOX_W || bh >= box_h) return make_int4(0, 0, -1, 0);
    // bit 16 of the width: every sample row of the block lies inside
    // [0.02, Nv-1.02] (the per-column rows differ from the corner bounds by
    // rounding only), so the gathers need no range test
    const bool inner = vlo >= 0.02f && vhi <= (float)(Nv - 1) - 0.02f;
    return make_int4(bu0, bv0, bw | (inner ? 0x10000 : 0), bh);
}

}  // namespace tkv
