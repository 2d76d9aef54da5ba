/* tomokvas/fdk_tiff.h -- TIFF export of reconstructed slices.
 *
 * Replaces class image_file (FBPG/image_file.h:19-46, FBPG/image_file.cpp):
 * every slice is scaled with the minimum and maximum of the whole volume to
 * 16 bits, u = (unsigned short)(65535 * (x - min) / (max - min)) in FP64
 * (ConvertToTiff, image_file.cpp:78-128), and written as <dir>/%08d.tif, a
 * baseline TIFF: 16-bit grey (min-is-black), one row per strip, uncompressed
 * (the tags TiffWrite sets through libtiff, image_file.cpp:130-230), written
 * here without libtiff.  Only the 16-bit type (BW16) is provided: the
 * reference's 8-bit and RGB types never convert their pixels.
 *
 * fdk_tiff_export_device works on a volume in HBM (after a resident
 * reconstruction, fdk_device_volume): min/max and the 16-bit conversion run on
 * the GPU, only 2 bytes per voxel cross to the host.  fdk_tiff_export_host
 * does the same from host memory.  Both write the files on up to 8 threads;
 * on a failure the first failing file's error is returned (files of other
 * slices may have been written). */
#ifndef TOMOKVAS_FDK_TIFF_H
#define TOMOKVAS_FDK_TIFF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* TiffWrite for BW16: one image of width x height samples */
int fdk_tiff_write_u16(const char *path, const uint16_t *img, int width, int height);
/* min and max of n floats in device memory (NaN ignored) -> minmax[2] */
int fdk_volume_minmax(const float *d_vol, size_t n, double *minmax, void *stream);
/* the 16-bit scaling of n floats on the device; max == min gives zeros (the
 * reference leaves its pixel buffer unwritten then) */
int fdk_volume_to_u16(const float *d_vol, size_t n, double vmin, double vmax, uint16_t *d_out,
                      void *stream);
/* AppendImages + ConvertToTiff for a volume [z][y][x] in device memory: files
 * <dir>/%08d.tif numbered from first_index; minmax (may be NULL) receives the
 * scaling range */
int fdk_tiff_export_device(const float *d_vol, int x, int y, int z, const char *dir,
                           int first_index, double *minmax);
/* the same from host memory with a given range (image_file's running min/max) */
int fdk_tiff_export_host(const float *vol, int x, int y, int z, double vmin, double vmax,
                         const char *dir, int first_index);

#ifdef __cplusplus
}
#endif
#endif
