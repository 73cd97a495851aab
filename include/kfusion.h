/*
 * kfusion.h — C ABI of the MI355X-native KinectFusion per-frame path.
 *
 * This is the drop-in boundary for the reference's per-frame path
 * (isikmustafa/kinectfusion-cuda).  Every entry point takes plain pointers and
 * sizes, returns an int status (KF_OK = 0) and never throws; the C++ class
 * surface in include/kfusion/kfusion.hpp re-throws std::runtime_error exactly
 * where the reference throws.  Citations are reference file:line.
 *
 * Conventions
 *  - Device buffers are raw device pointers (allocated with kf_malloc or by any
 *    other HIP allocator in the same process).
 *  - Depth maps: float[h][w] metres.  Raw depth: uint16[h][w] codes.
 *  - Vertex / normal maps: float4[h][w] (x, y, z, validity) with validity -1
 *    valid and -2 invalid (project/src/device_helper.cuh:9-10).
 *  - Matrices: glm column-major, m[col*rows + row] (test/example_test.cpp:42-51).
 *  - Every stage takes a stream (NULL = the default stream) and an optional
 *    float* elapsed_ms.  With elapsed_ms == NULL the call only enqueues work;
 *    with elapsed_ms != NULL it records events around the work, waits for it and
 *    stores the elapsed milliseconds — the reference's "every kernel wrapper
 *    returns its time in ms" contract (project/src/cuda_event.cpp:14-30,
 *    e.g. measurement.cu:139-158).
 */
#ifndef KFUSION_H
#define KFUSION_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status --- */
enum kf_status {
    KF_OK = 0,
    KF_ERR_DIMS = 1,      /* map sizes disagree (measurement.cu:144-147, raycast.cu:155-158, grid_map_pyramid.h:63-72) */
    KF_ERR_NONFINITE = 2, /* ICP update not finite (icp.cpp:67-73) */
    KF_ERR_NOT_PD = 3,    /* normal equations not positive definite (Eigen LLT NumericalIssue, ignored by the reference) */
    KF_ERR_OOM = 4,
    KF_ERR_HIP = 5,       /* HIP runtime error (the reference prints and continues, cuda_utils.h:7-23) */
    KF_ERR_BUDGET = 6,    /* HBM budget exhausted (serving admission) */
    KF_ERR_ARG = 7,       /* invalid argument */
    KF_ERR_IO = 8         /* file I/O (voxel_grid.cpp:43-80 save/load) */
};
const char* kf_status_string(int status);
/* Detail text of the last error on the calling thread ("" if none). */
const char* kf_last_error(void);

typedef void* kf_stream_t; /* a hipStream_t; NULL = default stream */

/* ------------------------------------------------- device and memory plumbing */
int kf_device_init(int device);
int kf_device_synchronize(void);
int kf_malloc(void** dptr, size_t bytes);
int kf_free(void* dptr);
int kf_memcpy_h2d(void* dst, const void* src, size_t bytes, kf_stream_t stream);
int kf_memcpy_d2h(void* dst, const void* src, size_t bytes, kf_stream_t stream);
int kf_memcpy_d2d(void* dst, const void* src, size_t bytes, kf_stream_t stream);
int kf_stream_create(kf_stream_t* stream);
int kf_stream_destroy(kf_stream_t stream);
int kf_stream_synchronize(kf_stream_t stream);
/* Events for timing on a given stream (cuda_event.h:5-17). */
typedef void* kf_event_t;
int kf_event_create(kf_event_t* ev);
int kf_event_destroy(kf_event_t ev);
int kf_event_record(kf_event_t ev, kf_stream_t stream);
int kf_event_elapsed_ms(kf_event_t start, kf_event_t end, float* ms);

/* ---------------------------------------------------------------- sensor --- */
/* Sensor (project/src/sensor.h:7-53), passed by value to the kernels as in the
 * reference.  Intrinsics per level: f_L = f / 2^L, c_L = (c_0 + 1/2) / 2^L - 1/2
 * (sensor.h:14-16 hard-codes those values for 640x480). */
typedef struct kf_sensor {
    float pose[16];        /* camera -> world */
    float inv_pose[16];    /* world -> camera */
    float intr[3][9];      /* K_L */
    float inv_intr[3][9];  /* K_L^-1 (glm::inverse, sensor.h:17-19) */
} kf_sensor;
/* Sensor(float fov) generalised to (fx, fy, cx, cy) (sensor.h:10-20); pose = I. */
int kf_sensor_init(kf_sensor* s, float fx, float fy, float cx, float cy);
/* Sensor::setPose (sensor.h:30-34); the inverse is [R^T | -R^T t] of the rigid pose. */
int kf_sensor_set_pose(kf_sensor* s, const float pose[16]);

/* ----------------------------------------------------------- measurement --- */
/* kernel::convertToDepthMeters (measurement.cuh:12, measurement.cu:139-158). */
int kf_convert_to_depth_meters(const uint16_t* raw, float* depth_m, int w, int h, float scale, int unmirror,
                               kf_stream_t stream, float* elapsed_ms);
/* kernel::applyBilateralFilter (measurement.cuh:13, measurement.cu:160-179). */
int kf_apply_bilateral_filter(const float* depth_m, float* filtered, int w, int h, kf_stream_t stream,
                              float* elapsed_ms);
/* kernel::downSample (measurement.cuh:14, measurement.cu:181-200); the output must be half the input. */
int kf_down_sample(const float* in, int w_in, int h_in, float* out, int w_out, int h_out, kf_stream_t stream,
                   float* elapsed_ms);
/* kernel::createVertexMap (measurement.cuh:15, measurement.cu:202-221). */
int kf_create_vertex_map(const float* depth, float* vertex_map, int w, int h, const float inv_k[9],
                         kf_stream_t stream, float* elapsed_ms);
/* kernel::createNormalMap (measurement.cuh:16, measurement.cu:223-242). */
int kf_create_normal_map(const float* vertex_map, float* normal_map, int w, int h, kf_stream_t stream,
                         float* elapsed_ms);
/* The whole measurement step of one frame in one launch (kinect_fusion.cpp:116-135
 * plus the normal maps ICP uses): raw codes -> unfiltered metres (for fusion),
 * filtered depth pyramid, vertex pyramid and normal pyramid, levels 0..levels-1.
 * Results are identical to the individual stages above. */
int kf_preprocess(const uint16_t* raw, int w, int h, float scale, int unmirror, int levels,
                  const float inv_k[][9], float* depth_m, float* const depth_pyr[], float* const vertex_pyr[],
                  float* const normal_pyr[], kf_stream_t stream, float* elapsed_ms);

/* ----------------------------------------------------------- voxel grid --- */
/* VoxelGrid / VoxelGridStruct / Voxel (voxel_grid.h:8-61).  The volume is an
 * opaque handle; its HBM layout is 8^3 bricks of separate f and w planes. */
typedef struct kf_volume kf_volume;
enum kf_voxel_format {
    KF_VOXEL_F32 = 0,        /* f: float, w: float (the reference's float2 Voxel) */
    KF_VOXEL_SNORM16_U8 = 1  /* f: snorm16, w: uint8 observation count (cap 128) */
};
/* n must be a multiple of 8; resolution = width / (n - 1) (voxel_grid.h:43);
 * mue_m is the truncation distance in metres (the reference passes
 * mue = mue_scale * resolution, voxel_grid.h:44; see kf_volume_create_scaled). */
int kf_volume_create(int n, float width_m, float mue_m, int format, kf_volume** out);
int kf_volume_create_scaled(int n, float width_m, float mue_scale, int format, kf_volume** out);
int kf_volume_destroy(kf_volume* v);
int kf_volume_info(const kf_volume* v, int* n, float* width_m, float* resolution, float* mue_m, int* format);
/* Device bytes held by the volume (both planes). */
int kf_volume_bytes(const kf_volume* v, size_t* bytes);
/* Total number of voxels updated by every fuse into this volume so far. */
int kf_volume_touched_total(const kf_volume* v, int64_t* count);
/* kernel::initializeGrid (tsdf.cuh:15, tsdf.cu:79-90); value (f, w).  For the
 * snorm16 format f below -1 clamps to -1, f above 1 is stored as the
 * unobserved code (read back as 2, the reference's initial value), and w is
 * rounded to an integer count. */
int kf_initialize_grid(kf_volume* v, float f, float w, kf_stream_t stream, float* elapsed_ms);
/* kernel::fuse (tsdf.cuh:16, tsdf.cu:92-103): projective TSDF update from an
 * unfiltered metric depth map seen by the sensor.  The reference hard-codes a
 * 640x480 frustum; here the frustum is the depth map's w x h. */
int kf_fuse(const float* depth_m, int w, int h, kf_volume* v, const kf_sensor* sensor, kf_stream_t stream,
            float* elapsed_ms);
/* kernel::raycast (raycast.cuh:11, raycast.cu:149-169): world-space vertex and
 * normal maps of level `level` seen from sensor.pose; pixels whose ray misses
 * the volume or the [0.4, 6] m range are left untouched (raycast.cu:61-73).
 * The volume keeps per-launch scratch (handed-over rays, tile order): issue a
 * volume's raycasts and fusions in order on one stream. */
int kf_raycast(const kf_volume* v, const kf_sensor* sensor, float* vertex_map, float* normal_map, int w, int h,
               int level, kf_stream_t stream, float* elapsed_ms);
/* VoxelGrid::saveVoxelGrid / loadVoxelGrid (voxel_grid.cpp:43-80): raw
 * n^3 x (f, w) float pairs, x fastest, no header.  upload/download move the same
 * layout between host memory and the volume. */
int kf_volume_download(const kf_volume* v, float* fw_host);
int kf_volume_upload(kf_volume* v, const float* fw_host);
int kf_volume_save(const kf_volume* v, const char* path);
int kf_volume_load(kf_volume* v, const char* path);

/* ------------------------------------------------------------------ ICP --- */
/* kernel::constructIcpResiduals (icp.cuh:34-37, icp.cu:86-106): one row
 * A_i (6 floats, row-major N x 6) and b_i per pixel, zero rows when rejected. */
int kf_construct_icp_residuals(const float* vertex_map, const float* target_vertex_map,
                               const float* target_normal_map, int w, int h, const float prev_rot[9],
                               const float prev_transl[3], const float cur_rot[9], const float cur_transl[3],
                               const float intrinsics[9], float distance_thresh, float angle_thresh,
                               float* mat_a, float* vec_b, kf_stream_t stream, float* elapsed_ms);
/* LinearLeastSquares::solve (linear_least_squares.h:19, .cpp:23-44): solves the
 * normal equations of A x = b for A (n x 6 row-major) and b on the device:
 * a 27-term reduction (21 of A^T A, 6 of A^T b) and an FP64 Cholesky.
 * result: 6 floats on the host.  KF_ERR_NOT_PD when A^T A is not positive definite. */
int kf_solve_normal_equations(const float* mat_a, const float* vec_b, unsigned int n_equations, float result[6],
                              kf_stream_t stream, float* elapsed_ms);
/* kf_solve_normal_equations without allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops).  It only
 * enqueues on stream: result_dev (6 floats, device) receives the solution and status_dev (one int,
 * device, optional) KF_OK or KF_ERR_NOT_PD / KF_ERR_NONFINITE, exactly as kf_solve_normal_equations
 * computes them.  Its working memory is a fixed device array of about 1 MiB per GPU, held until
 * the process ends and shared by every call on that GPU: calls that overlap on one GPU race.
 * n_equations <= 2^22 (KF_ERR_ARG above). */
int kf_solve_normal_equations_noalloc(const float* mat_a, const float* vec_b, unsigned int n_equations,
                                      float* result_dev, int* status_dev, kf_stream_t stream);
/* One ICP iteration's association fused with its reduction: the 27 sums of the
 * rows kf_construct_icp_residuals would produce, plus the number of accepted
 * rows, without materialising A and b.  sums_host: 21 upper-triangle entries of
 * A^T A in row order (p <= q), then A^T b; as doubles.  w x h <= 2^24
 * (KF_ERR_DIMS above). */
int kf_icp_reduce(const float* vertex_map, const float* target_vertex_map, const float* target_normal_map,
                  int w, int h, const float prev_rot[9], const float prev_transl[3], const float cur_rot[9],
                  const float cur_transl[3], const float intrinsics[9], float distance_thresh, float angle_thresh,
                  int compensated, double sums_host[27], int64_t* inliers, kf_stream_t stream, float* elapsed_ms);

/* IcpConfig (icp.h:11-20). */
typedef struct kf_icp_config {
    int iters_per_layer[3];              /* index 0 = finest level */
    int levels;                          /* 1..3 */
    int width, height;                   /* level-0 map size; width x height <= 2^24 */
    float distance_thresh;               /* metres */
    float angle_thresh;                  /* radians */
    float iteration_stop_thresh_angle;   /* poseError chord (general_helper.cpp:11-21); 0 disables */
    float iteration_stop_thresh_distance;/* metres; 0 disables */
    int compensated;                     /* 0: FP32 tree sums; 1: Kahan-compensated FP32 sums */
} kf_icp_config;
typedef struct kf_icp kf_icp;
/* ICP::ICP / ~ICP (icp.cpp:12-23). */
int kf_icp_create(const kf_icp_config* config, kf_icp** out);
int kf_icp_destroy(kf_icp* icp);
/* ICP::computePose (icp.h:35-36, icp.cpp:25-58): coarse-to-fine pose of the
 * current vertex pyramid against the predicted (raycast) pyramids, starting at
 * sensor->pose.  The whole loop runs on the device; the call waits once at the
 * end.  pose_out: camera -> world (16 floats).  KF_ERR_NONFINITE / KF_ERR_NOT_PD
 * leave pose_out at the last good estimate. */
int kf_icp_compute_pose(kf_icp* icp, const float* const vertex_pyr[], const float* const target_vertex_pyr[],
                        const float* const target_normal_pyr[], const kf_sensor* sensor, float pose_out[16],
                        kf_stream_t stream, float* elapsed_ms);
/* ICP::getExecutionTimes (icp.h:38): {association+reduction ms, solve ms} of the last computePose. */
int kf_icp_execution_times(const kf_icp* icp, float times[2]);
/* Iterations executed by the last computePose (after early stops). */
int kf_icp_iterations(const kf_icp* icp, int* iterations);

/* ------------------------------------------------------------- per frame --- */
/* KinectFusionConfig (kinect_fusion.h:17-36) restricted to the per-frame path:
 * one volume serves tracking and fusion. */
typedef struct kf_pipeline_config {
    int frame_width, frame_height;
    float depth_scale;          /* metres per raw code */
    float fx, fy, cx, cy;       /* level-0 intrinsics */
    int unmirror;               /* Kinect mirror flag of convertToDepthMeters */
    int volume_n;
    float volume_size;          /* cube side in metres, centred at the origin */
    float mue;                  /* truncation distance in metres */
    int voxel_format;           /* enum kf_voxel_format */
    float initial_pose[16];     /* camera -> world of the first frame */
    kf_icp_config icp;
    /* optional visualisation volume (KinectFusion's m_voxel_grid_visualization,
     * kinect_fusion.cpp:19,179-201): fused every frame like the tracking volume;
     * vis_volume_n = 0 disables it (renders then use the tracking volume) */
    int vis_volume_n;
    float vis_volume_size, vis_mue;
} kf_pipeline_config;
typedef struct kf_pipeline kf_pipeline;
enum kf_stage { KF_STAGE_PREPROCESS = 0, KF_STAGE_RAYCAST = 1, KF_STAGE_ICP = 2, KF_STAGE_FUSE = 3, KF_STAGE_COUNT = 4 };
int kf_pipeline_create(const kf_pipeline_config* config, kf_stream_t stream, kf_pipeline** out);
int kf_pipeline_destroy(kf_pipeline* p);
/* KinectFusion::startTracking's loop body (kinect_fusion.cpp:48-72): the first
 * frame is fused at the initial pose (initializeTracking, :75-94); later frames
 * run preprocess -> raycast L0..L2 at the previous pose -> ICP -> fuse at the new
 * pose.  raw: device pointer to one uint16 frame.  Enqueues only; frames whose
 * ICP fails keep the previous pose and are not fused (reported by kf_pipeline_sync). */
int kf_pipeline_frame(kf_pipeline* p, const uint16_t* raw);
/* Waits for the enqueued frames; status: KF_OK or the first ICP failure since the last sync. */
int kf_pipeline_sync(kf_pipeline* p, int* status);
int kf_pipeline_pose(kf_pipeline* p, float pose[16]);
int kf_pipeline_volume(kf_pipeline* p, kf_volume** v);
/* which: 0 unfiltered metres (level 0), 1 filtered depth, 2 vertex, 3 normal,
 *        4 predicted vertex, 5 predicted normal. */
int kf_pipeline_map(kf_pipeline* p, int which, int level, void** dev_ptr, int* w, int* h);
/* Per-stage event timing (kinect_fusion.cpp:216-224 prints these averages). */
int kf_pipeline_set_timing(kf_pipeline* p, int enabled);
int kf_pipeline_stage_times(kf_pipeline* p, float total_ms[KF_STAGE_COUNT], int64_t* frames);
/* visualizeCurrentModel (kinect_fusion.cpp:185-201) without a window: raycasts level 0 of
 * the visualisation volume (or the tracking volume) from pose (NULL = the current sensor pose)
 * into maps of its own (not the tracking maps), then writes shading (shading = 1,
 * display.cu:85-105) or the normal image (shading = 0, display.cu:64-83) as w x h RGBA8 pixels
 * to rgba (device).  Waits for the frame work enqueued before it. */
int kf_pipeline_render(kf_pipeline* p, const float pose[16], int shading, uint32_t* rgba);
/* Average duration of the fuse's brick-list kernel (the voxel sweep) over the frames timed since
 * timing was enabled (ms), from events recorded around that kernel alone. */
int kf_pipeline_fuse_kernel_ms(kf_pipeline* p, float* avg_ms, int64_t* launches);

/* --------------------------------------------------------- visualisation --- */
/* shadingToWindowContent (display.cu:85-105): per pixel min(|n . ray|, 1) in r, g, b with the
 * level-0 ray of the sensor; RGBA8 (r in the low byte, alpha 255), bytes trunc(v * 255)
 * clamped to [0, 255], NaN -> 0. */
int kf_render_shading(const float* normal_map, int w, int h, const kf_sensor* sensor, uint32_t* rgba,
                      kf_stream_t stream, float* elapsed_ms);
/* normalMapToWindowContent (display.cu:64-83): normalize(inverse_rotation * n) as (r, g, b). */
int kf_render_normals(const float* normal_map, int w, int h, const float inverse_rotation[9], uint32_t* rgba,
                      kf_stream_t stream, float* elapsed_ms);

/* -------------------------------------------------------- TUM sequences --- */
/* RgbdDataset (rgbd_dataset.cpp:6-97) on the host: depth.txt and groundtruth.txt of a TUM
 * RGB-D sequence; each depth frame is paired with the ground-truth pose nearest in time by the
 * reference's forward search.  Depth images are 16-bit grayscale PNGs (TUM: 5000 codes per
 * metre, so depth_scale = 1/5000). */
typedef struct kf_tum kf_tum;
int kf_tum_open(const char* dir, kf_tum** out);
int kf_tum_close(kf_tum* s);
int kf_tum_size(const kf_tum* s, int* frames, int* width, int* height);
/* Frame i: depth (host, width*height codes; NULL to skip) and the ground-truth camera -> world
 * pose (glm::mat4_cast of the quaternion with the translation, rgbd_dataset.cpp:67-69) of frame
 * i + gt_offset.  gt_offset = 1 reproduces the reference, which reads the pose after advancing
 * to the next image (kinect_fusion.cpp:165-166); 0 pairs each frame with its own pose. */
int kf_tum_frame(kf_tum* s, int i, int gt_offset, uint16_t* depth, float pose[16]);
/* poseError (general_helper.cpp:11-21): |R1 v - R2 v| for v = normalize(1, 1, 1), and |t1 - t2|. */
int kf_pose_error(const float pose1[16], const float pose2[16], float* angle, float* distance);

/* ---------------------------------------------------------------- batch ---
 * Many independent KinectFusion sessions (each its own volume, resolution,
 * intrinsics and pose) advanced together: one tick runs one frame of every
 * listed session with one launch per stage across all of them (preprocess,
 * raycast L0..L2, each ICP iteration, fuse).  The reference runs one session
 * per process (kinect_fusion.cpp:44-73); a session here computes exactly what
 * kf_pipeline computes for the same frames (same kernels, same summation
 * order), whatever else shares its ticks.  No reference counterpart: this
 * is the multi-session form of kf_pipeline_frame. */
typedef struct kf_session_config {
    int frame_width, frame_height;  /* divisible by 4 (3-level pyramid); width x height <= 2^24 */
    float depth_scale;
    float fx, fy, cx, cy;
    int volume_n;                   /* multiple of 8 */
    float volume_size;
    float mue;
    int voxel_format;
    float initial_pose[16];
} kf_session_config;
/* Launch shapes (workgroups per persistent launch); every shape gives the same
 * results.  Zero fields take the defaults. */
typedef struct kf_batch_shape {
    unsigned preprocess_blocks, raycast_blocks, icp_blocks, cull_blocks, fuse_blocks;
} kf_batch_shape;
typedef struct kf_batch_config {
    int max_sessions;
    kf_icp_config icp;        /* iterations, thresholds and summation shared by every session (width/height unused) */
    int use_graph;            /* 1: each tick is one hipGraph launch (captured once per table slot) */
    int autotune;             /* 1: the first ticks time candidate launch shapes per stage and keep the fastest */
    kf_batch_shape shape;
} kf_batch_config;
typedef struct kf_batch kf_batch;
int kf_batch_create(const kf_batch_config* config, kf_stream_t stream, kf_batch** out);
int kf_batch_destroy(kf_batch* b);
/* HBM a session holds while admitted, in bytes: volume, maps and ICP state, plus the most
 * it can add to the batch's shared fuse and deferred-ray lists (which grow 2x at a time). */
int kf_session_bytes(const kf_session_config* config, size_t* bytes);
int kf_batch_add_session(kf_batch* b, const kf_session_config* config, int* session);
int kf_batch_remove_session(kf_batch* b, int session);
/* Enqueues one tick: frame raw[k] (device uint16, the session's size) for
 * session sessions[k], k < n, each session at most once.  A session's first
 * frame is fused at its initial pose; later frames are tracked then fused. */
int kf_batch_tick(kf_batch* b, int n, const int* sessions, const uint16_t* const* raw);
/* Waits for every enqueued tick. */
int kf_batch_sync(kf_batch* b);
int kf_batch_pose(kf_batch* b, int session, float pose[16]);
/* First ICP failure of the session since the last call (KF_OK if none); clears it. */
int kf_batch_session_status(kf_batch* b, int session, int* status);
int kf_batch_volume(kf_batch* b, int session, kf_volume** v);
/* which as kf_pipeline_map: 0 metres, 1 filtered L0, 2 vertex, 4 predicted vertex, 5 predicted normal. */
int kf_batch_map(kf_batch* b, int session, int which, int level, void** dev_ptr, int* w, int* h);
int kf_batch_shape_get(kf_batch* b, kf_batch_shape* shape, int* tuned);
/* GPU time of the ticks since the last call (events around each tick). */
int kf_batch_tick_times(kf_batch* b, double* total_ms, int64_t* ticks, float* max_ms);
int kf_batch_touched_total(kf_batch* b, unsigned long long* voxels);
/* Per-kernel timing of the ticks enqueued while enabled: such ticks run as direct launches with
 * events between the kernels (results are the same as from the tick's graph).  Enabling or
 * disabling waits for the enqueued ticks and clears the totals.  total_ms: preprocess, raycast
 * first pass, raycast finishing pass, ICP, fuse brick cull, fuse sweep, whole tick (ms summed
 * over the timed ticks); ticks: their number. */
int kf_batch_set_stage_timing(kf_batch* b, int enabled);
int kf_batch_stage_times(kf_batch* b, double total_ms[7], int64_t* ticks);
/* Session checkpoint (VoxelGrid::saveVoxelGrid/loadVoxelGrid, voxel_grid.cpp:43-80, plus the
 * pose): path holds the volume in the reference's raw n^3 x (f, w) float format (any voxel
 * format converts through it) and path + ".pose" the frame count and the camera -> world pose
 * as text.  Restoring into a session of the same volume size resumes it: its next frame is
 * tracked from the saved pose.  Both wait for the enqueued ticks. */
int kf_batch_session_save(kf_batch* b, int session, const char* path);
int kf_batch_session_restore(kf_batch* b, int session, const char* path);

/* -------------------------------------------------------------- serving ---
 * A request loop over a kf_batch: sessions are admitted against an HBM budget
 * (queued FIFO until slots and memory free up), every admitted volume stays
 * resident, frames are queued per session as they arrive and each tick takes
 * the oldest waiting frame of every session (continuous batching, at most two
 * ticks in flight; launch_fill and max_wait_ms decide when a tick starts).  Per-frame latency = completion seen by the loop - arrival.
 * No reference counterpart (the reference serves one camera per process). */
typedef struct kf_server_config {
    kf_batch_config batch;       /* max_sessions = resident session slots */
    size_t hbm_budget_bytes;     /* admission budget: each resident session is charged kf_session_bytes
                                    plus its two device frame buffers (2 x 2 bytes per pixel) */
    int queue_depth;             /* frames buffered per session (pinned host memory); 0 = 8 */
    float launch_fill;           /* a tick is launched once frames of at least this fraction of the resident
                                    sessions wait; 0 = 0.5 */
    float max_wait_ms;           /* ...or, with no tick running, once the oldest waiting frame has waited this
                                    long; 0 = 2 ms, < 0 = never wait (launch whenever the GPU is idle) */
} kf_server_config;
typedef struct kf_server_stats {
    int64_t frames, ticks, admitted, queued_admissions, failed_frames;
    double wall_s, frames_per_s, mean_frames_per_tick;
    double offered_frames_per_s;  /* submitted frames / (last arrival - first arrival) */
    double latency_p50_ms, latency_p99_ms, latency_max_ms, latency_mean_ms;
    size_t peak_resident_bytes;
} kf_server_stats;
typedef struct kf_server kf_server;
int kf_server_create(const kf_server_config* config, kf_server** out);
int kf_server_destroy(kf_server* s);
/* Seconds on the loop's monotonic clock (arrival stamps use it). */
double kf_server_now(void);
/* Requests a session; *ticket names it from now on.  Admitted at once when a
 * slot and the budget allow, else when earlier sessions close. */
int kf_server_open(kf_server* s, const kf_session_config* config, int* ticket);
/* The session ends after its queued frames; its slot and memory are freed. */
int kf_server_close(kf_server* s, int ticket);
/* Copies one host frame (uint16, the session's size) into the session's queue.
 * KF_ERR_BUDGET when the queue is full. */
int kf_server_submit(kf_server* s, int ticket, const uint16_t* frame, double arrival_s);
/* One turn of the loop: retires finished ticks, admits waiting sessions,
 * launches a tick when frames wait and fewer than two ticks are in flight.
 * *launched (optional) = 1 when a tick was launched. */
int kf_server_poll(kf_server* s, int* launched);
/* Runs the loop until every queued frame is done and every closed session released. */
int kf_server_drain(kf_server* s);
int kf_server_stats_get(kf_server* s, kf_server_stats* out);
/* Latencies (ms) of every finished frame, in completion order. */
int kf_server_latencies(kf_server* s, double* out, size_t cap, size_t* count);
/* 1 admitted, 0 waiting for admission, -1 closed and released. */
int kf_server_state(kf_server* s, int ticket, int* state);
int kf_server_pose(kf_server* s, int ticket, float pose[16]);
int kf_server_volume(kf_server* s, int ticket, kf_volume** v);
/* kf_batch_session_save / _restore for a resident session (restore before its first frame). */
int kf_server_checkpoint(kf_server* s, int ticket, const char* path);
int kf_server_restore(kf_server* s, int ticket, const char* path);

/* A Poisson load: session k opens at opens_s[k] (seconds after the start) and
 * sends frames_per_session[k] frames with exponential inter-arrival times of
 * mean 1/rate_hz, each scaled by a uniform factor in [1 - jitter, 1 + jitter]
 * (seeded).  Frame f of session k is bank[k][p(f)] where p ping-pongs over the
 * bank_len[k] frames (0, 1, ..., n-1, n-2, ..., 1, 0, 1, ...) so the camera
 * path stays continuous.  The session closes after its last frame. */
typedef struct kf_poisson_load {
    int sessions;
    const kf_session_config* configs;
    const uint16_t* const* banks;
    const int* bank_len;
    const int* frames_per_session;
    const double* opens_s;
    double rate_hz, jitter;
    uint64_t seed;
    int keep_open;               /* 1: sessions stay resident after their last frame (close them yourself) */
} kf_poisson_load;
/* Drives the loop with the load in real time until every frame is done.
 * tickets (optional, sessions entries) receive the sessions' tickets. */
int kf_server_run_load(kf_server* s, const kf_poisson_load* load, int* tickets, kf_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* KFUSION_H */
