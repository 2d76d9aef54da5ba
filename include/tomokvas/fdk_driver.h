/* tomokvas/fdk_driver.h -- the reference's reconstruction program as a library
 * call (host code): settings file, .CRA input, engine runs per angle block,
 * .bin/.siz output.
 *
 * fdk_reconstruct_file replaces main() of FBPG/para_TC.cpp:5-175: read the .CRA
 * header, size the angle blocks to the free HBM, and for every scan block and
 * every angle block load the projections, run the engine (collecting from the
 * second angle block on) and stream the slabs into the volume file.  The next
 * angle block is read from disk on a loader thread while the engine runs.
 * fdk_setting_load / fdk_setting_save replace LoadSetting / SaveSetting
 * (FBPG/para_TC.cpp:192-302) and the settings struct replaces `setting`
 * (FBPG/para_TC.h:20-32).
 *
 * Return codes are the FDK_* values of tomokvas/fdk.h; fdk_last_error_string()
 * describes the last failure. */
#ifndef TOMOKVAS_FDK_DRIVER_H
#define TOMOKVAS_FDK_DRIVER_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int gpu_id;               /* <GPU ID>                                          */
    int mem_limit_mib;        /* <GPU MemLimit_MiB>: read and saved; the angle blocks
                                 are sized to the free HBM, as the reference does  */
    int alloc_limit_mib;      /* <GPU AllocLimit_MiB> -> fdk_set_allocation_limit  */
    int slices_per_block;     /* <RECONSTRUCTION SimultaneouslyReconstructedSlices> */
    int x_direction, y_direction, z_direction;   /* <RECONSTRUCTION X/Y/Z_direction> */
    double vertical_center;   /* <RECONSTRUCTION DetectorVerticalCenterOffset>, px   */
    int wait_for_enter;       /* <GENERAL WaitForEnter>                            */
} fdk_setting;

/* the reference's defaults (para_TC.cpp:287-296): GPU 0, 5 GiB, 1 GiB, 16
 * slices, no axis flips, no vertical offset, wait for enter */
int fdk_setting_defaults(fdk_setting *s);
/* LoadSetting: parse <GPU_FBP> of an XML file.  FDK_OK when the root element
 * was found; otherwise *s holds the defaults and FDK_ERR_INVALID_ARGUMENT is
 * returned (the reference then writes the defaults with SaveSetting).  An
 * element that is absent gives that element's defaults, as the reference does;
 * an attribute that is absent keeps its default. */
int fdk_setting_load(const char *xml_path, fdk_setting *s);
/* SaveSetting: write the settings in the reference's layout */
int fdk_setting_save(const char *xml_path, const fdk_setting *s);

typedef struct {
    int storage;                  /* FDK_F32 (the reference), FDK_F16 or FDK_BF16   */
    int filter;                   /* FDK_FILTER_RAMLAK (the reference) or _SHEPPLOGAN */
    long long angle_block_bytes;  /* 0: sized to the free HBM; else the budget of one
                                     angle block, rounded down to whole projections */
    int verbose;                  /* print the reference's progress report          */
    int volume_mode;              /* -1: streamed when the volume exceeds 45 % of the
                                     free HBM, else resident; or FDK_VOLUME_RESIDENT /
                                     FDK_VOLUME_STREAMED                            */
} fdk_run_options;

typedef struct {
    int scan_blocks, angle_blocks;
    double seconds_total;
    double seconds_load_wait;     /* main thread waiting on the projection loader   */
    double seconds_engine;        /* upload + reconstruct (fdk_reconstruct_host), slab
                                     hand-over included */
    double seconds_flush;         /* final slab writes and .siz                      */
    double bytes_read, bytes_written;
} fdk_run_stats;

/* main() of para_TC.cpp: reconstruct <cra_path> (".../name.CRA") into
 * ".../name.bin" + ".siz".  opt and stats may be NULL (defaults: FDK_F32,
 * Ram-Lak, budget from free HBM, verbose, residency by size). */
int fdk_reconstruct_file(const char *cra_path, const fdk_setting *set,
                         const fdk_run_options *opt, fdk_run_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
