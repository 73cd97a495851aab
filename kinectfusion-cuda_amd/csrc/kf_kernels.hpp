// Launch interfaces shared between the library's translation units.  Each
// launcher enqueues kernels on `s` and returns a kf_status.
#pragma once

#include "kf_internal.hpp"

namespace kf {

// ---- measurement (measurement.hip)
int launch_convert(const uint16_t* raw, float* out, int w, int h, float scale, int unmirror, hipStream_t s);
int launch_bilateral(const float* in, float* out, int w, int h, hipStream_t s);
int launch_down_sample(const float* in, int w_in, int h_in, float* out, hipStream_t s);
int launch_vertex_map(const float* depth, float4* out, int w, int h, const float* inv_k, hipStream_t s);
int launch_normal_map(const float4* v, float4* out, int w, int h, hipStream_t s);
int launch_preprocess(const uint16_t* raw, int w, int h, float scale, int unmirror, int levels,
                      const float (*inv_k)[9], float* depth_m, float* const* dpyr, float4* const* vpyr,
                      float4* const* npyr, hipStream_t s);

// ---- TSDF (tsdf.hip)
// inv_pose: by-value world->camera matrix; dyn_state (may be null) overrides it
// with PoseState::inv_pose and skips the launch's work when the frame failed.
int launch_fuse(const float* depth, int w, int h, kf_volume* v, const float* K, const float* inv_pose,
                const PoseState* dyn_state, hipStream_t s);
int launch_initialize(kf_volume* v, float f, float wt, hipStream_t s);

// ---- raycast (raycast.hip)
struct RaycastVolume {
    const void* f;
    const float* brick_min;  // per-brick lower bound of F (null: no skipping)
    const int* super_key;    // per-super-brick lower bound of F, order_key form
    int n, nb, ns, format;
    float width, mue;
};
struct RaycastLevel {
    float Kinv[9];
    float4* vout;
    float4* nout;
    int w, h, tiles_x, tile0;  // tile0: first global tile of this level
};
// One (session, level) of a batched raycast; lives in device memory.
struct RaycastJob {
    RaycastVolume vol;
    const float* pose;   // device: camera-to-world pose of the session (PoseState::pose)
    const int* active;   // device flag, null = always active
    RaycastLevel L;
};
RaycastVolume raycast_volume(const kf_volume* v);
int launch_raycast_batch(const RaycastJob* jobs_dev, int njobs, int tiles, hipStream_t s);
int launch_raycast(const kf_volume* v, const float* pose, const PoseState* dyn_state, const float* inv_k,
                   float4* vout, float4* nout, int w, int h, hipStream_t s);
// All pyramid levels in one launch (level l is w0 >> l by h0 >> l).
int launch_raycast_levels(const kf_volume* v, const float* pose, const PoseState* dyn_state, int levels,
                          const float (*inv_k)[9], float4* const* vout, float4* const* nout, int w0, int h0,
                          hipStream_t s);

// ---- ICP (icp.hip)
constexpr int kPartialStride = 64;  // floats per block partial: 27 sums, 27 compensations, count
struct IcpLevel {
    const float4* V;
    const float4* TV;
    const float4* TN;
    int w, h;
    float K[9];
};
int icp_blocks_for(int w, int h);
// Enqueues the whole coarse-to-fine loop on `s` with the state on the device,
// ending with the commit of the estimate into state->pose / inv_pose.
int launch_icp_loop(kf_icp* icp, const IcpLevel* levels, int n_levels, hipStream_t s);
// begin: starts the frame's ICP from state->pose (or from pose16 if given);
// launch_icp_loop then enqueues one launch per iteration and the final commit.
int launch_icp_begin(kf_icp* icp, const float* pose16, hipStream_t s);

}  // namespace kf
