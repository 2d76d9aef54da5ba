/* tomokvas/fdk_io.h -- the reference's file formats on POSIX (host code, no GPU).
 *
 * .CRA projection container: a text header (key lines) and a raw float file of
 * projections [beta][v][u]; the reader splits the angles into blocks that fit a
 * byte budget and loads one block at a time sinogram-major [v][beta][u]
 * (replaces cra_con, FBPG/mucat_file.h:37-95, FBPG/mucat_file.cpp:6-428).
 *
 * .bin / .siz volume: a raw float volume [scan block][z][y][x] and a text size
 * header "x y z 0"; the writer hands the engine one host buffer per slab and
 * writes the previous slab / prefetches the next one on an I/O thread, three
 * buffers in rotation (replaces bin_siz, FBPG/mucat_file.h:129-193,
 * FBPG/mucat_file.cpp:432-793).  fdk_bin_get_block has the signature of
 * fdk_get_block_fn and is passed to fdk_calculate_reconstruction directly.
 *
 * Return codes are the FDK_* values of tomokvas/fdk.h; fdk_last_error_string()
 * describes the last failure. */
#ifndef TOMOKVAS_FDK_IO_H
#define TOMOKVAS_FDK_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdk_cra fdk_cra;
typedef struct fdk_bin fdk_bin;

/* the parsed header (cra_header, FBPG/mucat_file.h:15-35) and the quantities
 * GetMetaData derives from it (FBPG/mucat_file.cpp:135-160) */
typedef struct {
    char raw_file[256];            /* "FILE NAME:", relative to the .CRA's folder */
    char source_name[256];
    double sample_spacing;         /* um */
    double distance;               /* cm */
    int detector_x, detector_z;    /* px */
    int num_of_projections;
    int num_of_blocks;             /* scan blocks stored one after another */
    int x_dimension, z_dimension;  /* reconstruction size after the clamps */
    double rotate_angle;           /* degrees */
    int clockwise;
    int num_centres;               /* "CENTRE:" lines read (at most 64) */
    double centre_of_rotation[64];
    double scale;
    float distance_px;             /* distance * 1e4 / sample_spacing */
    float delta_fi;                /* 2 pi / num_of_projections */
} fdk_cra_meta;

/* GetMetaData (mucat_file.cpp:60-160): read <dir>/<name>.CRA */
int fdk_cra_open(const char *dir, const char *name, fdk_cra **out);
void fdk_cra_close(fdk_cra *cra);
int fdk_cra_get_meta(const fdk_cra *cra, fdk_cra_meta *out);
/* SetActiveBlock (mucat_file.cpp:339-347): the scan block to load; out of
 * range selects block 0, as the reference does; returns the block set */
int fdk_cra_set_scan_block(fdk_cra *cra, int block);
/* SetSinBlockMaxSize + CalculateSinBlocks (mucat_file.cpp:349-375): split the
 * angles into blocks of at most max_bytes; *nblocks receives the count.  A
 * budget below one projection, or one whose split would leave projections
 * unassigned, is rejected. */
int fdk_cra_split(fdk_cra *cra, long long max_bytes, int *nblocks);
/* SetSinActiveBlock (mucat_file.cpp:389-397); out of range selects block 0 */
int fdk_cra_set_angle_block(fdk_cra *cra, int block);
/* GetNoProjection (mucat_file.cpp:399-409): projections in the active block */
int fdk_cra_block_projections(const fdk_cra *cra, int *n);
/* GetRotRad (mucat_file.cpp:288-316): angle of the active block's first loaded
 * projection, rotate_angle included, radians */
int fdk_cra_rotation(const fdk_cra *cra, double *rad);
/* LoadData (mucat_file.cpp:189-263): the active block's projections into
 * dst [detector_z][n][detector_x] (n = fdk_cra_block_projections), angle
 * order reversed for clockwise scans */
int fdk_cra_load(const fdk_cra *cra, float *dst);

/* bin_siz constructor, SetScanBlocksCount, SetImageSize (creates the .bin at
 * its full size) and SetSlicesPerBlock (FBPG/para_TC.cpp:102-105) in one call:
 * <dir>/<name>.bin of scan_blocks * z * y * x floats, slabs of
 * slices_per_block slices */
int fdk_bin_open(const char *dir, const char *name, int x, int y, int z, int scan_blocks,
                 int slices_per_block, fdk_bin **out);
/* SetScanBlockOffset (mucat_file.cpp:507-511) */
int fdk_bin_set_scan_block(fdk_bin *bin, int scan_block);
/* GetBlock (mucat_file.cpp:576-642) as an fdk_get_block_fn: the host buffer
 * for slab `slab` of the active scan block.  The slab handed out before it is
 * written to the file on the I/O thread; with collecting the buffer holds the
 * slab's current file contents (the partial sum of earlier angle blocks),
 * prefetched while the previous slab was computed.  NULL on failure. */
float *fdk_bin_get_block(void *bin, int slab, int collecting);
/* FlushBuffer (mucat_file.cpp:644-666): write the slab still held */
int fdk_bin_flush(fdk_bin *bin);
/* destructor (mucat_file.cpp:485-492): flush, write the .siz header, close */
int fdk_bin_close(fdk_bin *bin);

#ifdef __cplusplus
}
#endif
#endif
