/*
 * rt_report.h — the report files and config.txt parsing of the C++ renderer
 * layer (include/rt_renderer.hpp, bvhraytracer_amd/csrc/renderer.cpp) as C
 * entry points.  Bindings (the Python package) write the reference's files
 * through these, so each format has one implementation.  All functions
 * return RT_OK or RT_ERR_ARG with rt_last_error() set (a file that cannot be
 * opened or written is RT_ERR_ARG too).  No GPU call.
 */
#ifndef RT_REPORT_H
#define RT_REPORT_H

#include <stddef.h>
#include <stdint.h>

#include "rt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RayTracer::readConfig (rayTracer.cpp:756-877, configUtil.h:6-58): the
 * settings after reading `path` over the reference's defaults, as
 * "key=value" lines (integers; booleans true/false; scenarios and
 * cacheSizes comma-separated; one "warning=<text>" line per message the
 * reference prints to std::cerr).  The text and its terminating zero go to
 * text[0..cap); *len receives the text length.  A text that does not fit
 * returns RT_ERR_ARG with *len set, so the caller can retry.  Replaces the
 * reference's in-process parse of config.txt (RayTracer::readConfig). */
int rt_report_read_config(const char* path, char* text, size_t cap, size_t* len);

/* _Perf.txt of CameraFast::renderImages (cameraFast.h:147-161). */
int rt_report_write_perf(const char* path, const char* name, uint32_t branching, uint32_t leaf_size,
                         uint32_t gang, uint32_t node_memory, uint32_t leaf_memory, float total_s,
                         float ray_sum_s, float triangle_sum_s, float average_depth, uint64_t triangles);

/* _Info.txt of the instrumented renderer (cameraData.h:241-329) from the
 * rt_trace_histograms counters of the primary and the secondary (ambient)
 * rays; bvh_depth visit-per-depth rows.  Counts must be nonzero (the
 * reference divides by them). */
int rt_report_write_info(const char* path, const char* name, uint32_t branching, uint32_t leaf_size,
                         uint32_t bvh_depth, const rt_trace_hist* primary, const rt_trace_hist* secondary);

/* The four workgroup analysis files of CameraData::workGroupAnalysisPerImage
 * (cameraData.h:660-1109) under <path>/WorkGroupSize_<wg>_Version_<0|1>/.
 * tiles = (width / wg) * (height / wg).  The rt_trace_steps records of tile
 * i are primary[o_i, o_i + primary_len[i]) with o_i the sum of the lengths
 * before i (secondary: every ambient sample's records back to back, likewise
 * with secondary_len).  The visit arrays hold tiles * wg * wg values per ray
 * in tile order (secondary summed over the samples).  paths receives the
 * four file paths, one per line, zero-terminated (cap bytes). */
int rt_report_write_workgroup_analysis(const char* path, const char* name, const char* problem, int32_t camera_id,
                                       uint32_t width, uint32_t height, uint32_t wg, const rt_step_stat* primary,
                                       const uint32_t* primary_len, const rt_step_stat* secondary,
                                       const uint32_t* secondary_len, const uint32_t* primary_nodes,
                                       const uint32_t* primary_leaves, const uint32_t* secondary_nodes,
                                       const uint32_t* secondary_leaves, int32_t wide_alternative, char* paths,
                                       size_t cap);

/* Pixels in CameraFast tile order (the order of rt_camera_rays) to
 * row-major, `channels` bytes per pixel (cameraFast.h:833-849; a partial
 * last tile row holds height % wg rows). */
int rt_report_tile_order_to_rows(const uint8_t* tiles, uint32_t width, uint32_t height, uint32_t wg,
                                 uint32_t channels, uint8_t* rows);

/* 8-bit RGBA PNG of width x height pixels, rows top to bottom (the
 * _Perf.png of cameraFast.h:829-852, encodeTwoSteps of util.h:52-61). */
int rt_report_write_png(const char* path, const uint8_t* rgba, uint32_t width, uint32_t height);

/* Gray value per pixel of the ambient-occlusion image (cameraFast.h:
 * 404-439): a hit pixel is (uint8)((1 - occluded / samples) * 255) in
 * float32, a miss 0, and every pixel 0 without samples. */
int rt_report_ambient_gray(const uint8_t* hit, const uint32_t* occluded, uint32_t n, uint32_t samples,
                           uint8_t* gray);

/* ambientDistance of RayTracer::renderImage (rayTracer.cpp:607): cbrt of
 * the volume of the triangles' bounding box / 10, in float32. */
int rt_report_ambient_distance(const float* tris, uint32_t n_tris, float* out);

#ifdef __cplusplus
}
#endif

#endif /* RT_REPORT_H */
