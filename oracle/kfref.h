/*
 * kfref — CPU restatement of the KinectFusion per-frame path of
 * isikmustafa/kinectfusion-cuda, used ONLY as test infrastructure:
 *   - the checker in tests/, __graft_entry__.smoke() and
 *   - the CPU baseline leg of bench.py.
 * Nothing in the product library links, calls or loads this code.
 *
 * The same source (kfref.c) is compiled twice:
 *   kfo_*   : FP32, IEEE-754 basic operations in the reference's (glm's) operation
 *             order, no FMA contraction (-ffp-contract=off).  This is the oracle
 *             the HIP kernels are compared against (bit-exact where the north star
 *             asks for it, tolerance elsewhere).
 *   kfo64_* : the identical algorithm in FP64 — the "FP64 oracle" that pose, TSDF
 *             and raycast errors are stated against.
 *
 * Conventions (reference file:line in parentheses):
 *   - matrices are column-major float[9] / float[16] as glm stores them,
 *     m[col*rows + row] (test/example_test.cpp:42-51);
 *   - float4 maps carry validity in .w: -1 valid, -2 invalid
 *     (project/src/device_helper.cuh:9-10);
 *   - depth validity: d > 0.01 && isfinite(d) (device_helper.cuh:11-17);
 *   - the volume is n^3 voxels, x fastest, then y, then z, with separate planes
 *     F (tsdf) and W (weight) (voxel_grid.h:8-47, voxel_grid.cpp:43-80 store them
 *     interleaved; kfo_interleave converts).
 *
 * Per-function citations and "parity pinned by" notes are in kfref.c.
 */
#ifndef KFREF_H
#define KFREF_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z, w; } kfo_float4;

/* ---- scalar helpers ---- */
float kfo_u16_to_half_float(uint16_t code);
float kfo_expf(float x);
int   kfo_threads(void);
void  kfo_set_threads(int n);
int   kfo_normals_too_different(const float n[3], const float tn[3], const float R[9], float cos_thresh);
int   kfo_normals_too_different_acos(const float n[3], const float tn[3], const float R[9], float angle_thresh);
void  kfo_correspondence(const float s[3], const float R_prev[9], const float t_prev[3], const float K[9], int uv[2], float fuv[2]);
float kfo_cos_threshold(float angle);
void  kfo_intrinsics(float fx, float fy, float cx, float cy, int level, float K[9], float K_inv[9]);

/* ---- measurement (project/src/measurement.cu) ---- */
void kfo_convert_to_depth_meters(const uint16_t* raw, float* out, int w, int h, float scale, int unmirror);
void kfo_apply_bilateral_filter(const float* in, float* out, int w, int h);
void kfo_down_sample(const float* in, int w, int h, float* out);
void kfo_create_vertex_map(const float* depth, int w, int h, const float K_inv[9], kfo_float4* out);
void kfo_create_normal_map(const kfo_float4* vmap, int w, int h, kfo_float4* out);
void kfo_compute_normal(const kfo_float4* vmap, int w, int h, int i, int j, kfo_float4* out);

/* ---- TSDF (project/src/tsdf.cu) ---- */
void    kfo_initialize_grid(float* F, float* W, int n, float f0, float w0);
int64_t kfo_fuse(float* F, float* W, int n, float width, float mue,
                 const float* depth, int w, int h, const float inv_pose[16], const float K[9]);

/* precision form: snorm16 F (q / 32767), uint8 observation counts (cap 128) */
void    kfo_initialize_grid_q(int16_t* F, uint8_t* W, int n, float f0, float w0);
int64_t kfo_fuse_q(int16_t* F, uint8_t* W, int n, float width, float mue,
                   const float* depth, int w, int h, const float inv_pose[16], const float K[9]);

/* ---- raycast (project/src/raycast.cu) ---- */
void kfo_raycast(const float* F, int n, float width, float mue, const float pose[16], const float K_inv[9],
                 int w, int h, kfo_float4* vout, kfo_float4* nout);
float kfo_sample_tsdf(const float* F, int n, float u, float v, float s);
void kfo_raycast_q(const int16_t* F, int n, float width, float mue, const float pose[16], const float K_inv[9],
                   int w, int h, kfo_float4* vout, kfo_float4* nout);
/* measurement only: the distinct voxels the reference march reads (fmt 0 float, 1 snorm16) */
int64_t kfo_raycast_footprint(const void* F, int fmt, int n, float width, float mue, const float pose[16],
                              const float K_inv[9], int w, int h, uint8_t* mark);
int64_t kfo_count_marked(const uint8_t* mark, int64_t count);

/* ---- ICP (project/src/icp.cu, icp.cuh, icp.cpp, linear_least_squares.cpp) ---- */
void kfo_construct_icp_residuals(const kfo_float4* vmap, const kfo_float4* tvmap, const kfo_float4* tnmap,
                                 int w, int h, const float R_prev[9], const float t_prev[3],
                                 const float R_cur[9], const float t_cur[3], const float K[9],
                                 float dist_thresh, float cos_thresh, float* A, float* b);
int64_t kfo_normal_equations(const float* A, const float* b, int64_t n, double ata[36], double atb[6]);
int  kfo_cholesky_solve6(const double ata[36], const double atb[6], double x[6]);
int  kfo_llt_solve6_f32(const float ata[36], const float atb[6], float x[6]);
void kfo_update_pose(float R[9], float t[3], const float x[6]);
void kfo_pose_error(const float R1[9], const float t1[3], const float R2[9], const float t2[3], float out[2]);
void kfo_rigid_inverse(const float R[9], const float t[3], float inv_pose[16]);
int  kfo_compute_pose(const kfo_float4* const vpyr[], const kfo_float4* const tvpyr[], const kfo_float4* const tnpyr[],
                      int w0, int h0, int levels, const int iters[], const float Ks[][9],
                      float dist_thresh, float angle_thresh, float stop_angle, float stop_dist,
                      const float R_in[9], const float t_in[3], float R_out[9], float t_out[3],
                      double* sums_log, int* n_iters);

/* ---- interchange (project/src/voxel_grid.cpp:43-80) ---- */
void kfo_interleave(const float* F, const float* W, int64_t count, float* fw);

/* ---- the per-frame loop (project/src/kinect_fusion.cpp:44-183) ---- */
typedef struct kfo_pipeline kfo_pipeline;
kfo_pipeline* kfo_pipeline_create(int w, int h, int n, float width, float mue,
                                  float fx, float fy, float cx, float cy, float depth_scale,
                                  const int iters[3], float dist_thresh, float angle_thresh,
                                  float stop_angle, float stop_dist, const float init_pose[16]);
void   kfo_pipeline_destroy(kfo_pipeline* p);
int    kfo_pipeline_frame(kfo_pipeline* p, const uint16_t* raw);
void   kfo_pipeline_pose(const kfo_pipeline* p, float pose[16]);
const float* kfo_pipeline_tsdf(const kfo_pipeline* p);
const float* kfo_pipeline_weight(const kfo_pipeline* p);
const kfo_float4* kfo_pipeline_map(const kfo_pipeline* p, int which, int level);
const float* kfo_pipeline_depth(const kfo_pipeline* p, int which, int level);
double kfo_pipeline_stage_ms(const kfo_pipeline* p, int stage);
int64_t kfo_pipeline_touched(const kfo_pipeline* p);
int    kfo_pipeline_use_snorm16(kfo_pipeline* p);
/* count-weight update on unquantised planes (the precision form's rule without
 * the snorm16 rounding); the FP64 build's kfo64_* twin is its error reference */
int    kfo_pipeline_use_count_weights(kfo_pipeline* p);
int64_t kfo_fuse_counts(float* F, float* W, int n, float width, float mue,
                        const float* depth, int w, int h, const float inv_pose[16], const float K[9]);
const int16_t* kfo_pipeline_tsdf_q(const kfo_pipeline* p);
const uint8_t* kfo_pipeline_weight_q(const kfo_pipeline* p);

#ifdef __cplusplus
}
#endif
#endif
