/* tomokvas/fdk_rsq.h -- Scanco RSQ raw projections: header reading on the host,
 * preprocessing on the GPU.
 *
 * An RSQ file holds, after a header, sinograms of 16-bit counts: each is one
 * dark row, one flat row and sin_y projection rows of sin_x samples.  The
 * preprocessing turns them into line integrals -- flat/dark correction and the
 * beam-hardening line A*log(.)+B (FlatFieldCorrection, FBPG/rsq_load.cpp:
 * 299-325) -- and either cuts the half detector at the rotation axis
 * (CutHalfSin, :393-410) or rebins the half-fan scan into a full-fan one
 * (Half2FullFan, :327-391).  Replaces class rsq (FBPG/rsq_load.h:63-135); the
 * fan->parallel rebinning (Half2Para/Fan2Para) is not provided.
 *
 * fdk_rsq_process runs on the GPU over raw sinograms already in device memory
 * and writes [n][sin_y_cpl][sin_x_cpl] in FP64 (the reference's output type) or
 * FP32 -- for detector rows, exactly the engine's projection stack layout
 * [v][beta][u] (fdk_device_sinogram). */
#ifndef TOMOKVAS_FDK_RSQ_H
#define TOMOKVAS_FDK_RSQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdk_rsq fdk_rsq;

/* the header fields the reference prints (ima_data_type, rsq_load.h:13-49)
 * and the beam-hardening line "Corr at X0 X1 X2:  Y0 Y1 Y2" (byte 3432) with
 * A = (Y2-Y0)/(X2-X0), B = Y0 - A*X0 (LoadHeader, rsq_load.cpp:489-552) */
typedef struct {
    char check[17];
    int data_type, nr_of_bytes, nr_of_blocks, patient_index, scanner_id;
    int dimx_p, dimy_p, dimz_p, dimx_um, dimy_um, dimz_um;
    int mu_scaling, nr_of_samples, nr_of_projections, scandist_um, scanner_type;
    int reference_line_um, energy, intensity, data_offset;
    int bh_x[3];
    double bh_y[3];
    double bh_a, bh_b;
    long long file_size;
} fdk_rsq_header;

/* the reference's setters (rsq_load.h:78-92) with its constructor defaults
 * (rsq_load.cpp:6-37): 1024 columns, 2000 projections, data at byte 3584,
 * axis 21 columns in, distance 125810/24 px, delta_fi = 2 pi / sin_y */
typedef struct {
    int sin_x, sin_y;
    long long header_offset;
    int axis_move;
    double distance, delta_fi;
    int flip, invert, complete, minimal;
} fdk_rsq_params;

int fdk_rsq_default_params(fdk_rsq_params *p);
int fdk_rsq_open(const char *path, fdk_rsq **out);
void fdk_rsq_close(fdk_rsq *r);
int fdk_rsq_get_header(const fdk_rsq *r, fdk_rsq_header *h);
/* GetNum / GetSize (CalculateNumOfSinograms, CalculateSizeOfSinogram,
 * rsq_load.cpp:76-105): sinograms in the file and the processed size */
int fdk_rsq_sizes(const fdk_rsq *r, const fdk_rsq_params *p, int *num_sinograms,
                  int *sin_x_cpl, int *sin_y_cpl);
/* LoadImages(first, last) (rsq_load.cpp:183-231): raw sinograms first..last
 * (inclusive) into dst [last-first+1][sin_y+2][sin_x] */
int fdk_rsq_load(const fdk_rsq *r, const fdk_rsq_params *p, int first, int last, uint16_t *dst);
/* ProccessAndCopy (rsq_load.cpp:270-297) on the GPU: d_raw as fdk_rsq_load
 * writes it, d_out [nsin][sin_y_cpl][sin_x_cpl] of out_dtype (FDK_F32 or
 * FDK_F64), on `stream` (NULL: the default stream).  Without completing it
 * only queues its kernels; with completing (the full-fan rebinning) it
 * returns when they are done and frees what it allocated for them: the
 * per-column atan table and an FP64 line-integral scratch of at most 64 MiB
 * (or one sinogram, sin_y x (sin_x - axis_move) values, when that is larger).
 * At most 65535 projections per sinogram.  Angles that the reference would
 * read outside the scan (delta_fi other than 2 pi / sin_y) wrap around it. */
int fdk_rsq_process(const fdk_rsq_params *p, double bh_a, double bh_b, const uint16_t *d_raw,
                    int nsin, int out_dtype, void *d_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
