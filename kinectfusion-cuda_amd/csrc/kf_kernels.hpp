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

// ---- batch ticks (batch.hip): one table per tick in device memory, a header
// followed by the job arrays of every stage.  Counts are per tick; the fuse
// brick list, its counter and the touched slots belong to the batch.
struct BatchHeader {
    int n_pre, pre_tiles, pyr_tiles;  // preprocess jobs, L0 16x16 tiles, L2 16x16 tiles
    int n_ray, ray_tiles;             // raycast jobs (session x level), 8x8 tiles
    int n_ray_f32, ray_tiles_f32;     // the FP32-volume jobs come first: their count and tiles
    int n_icp, icp_blocks[3];         // ICP sessions, blocks per level
    int n_fuse;                       // fuse jobs
    unsigned fuse_bricks;             // bricks over all fuse jobs
    unsigned* fuse_count;
    unsigned long long* fuse_list;    // job << 32 | brick of the job's volume
    unsigned long long* touched;
    unsigned* ray_next;  // 16 tile counters (8 XCD bands x 2 voxel formats) and 2 deferred-ray
                         // counts (per format), cleared by the first kernel
    struct DeferredRay* ray_list;  // 2 x ray_list_cap deferred rays (F32 jobs, then snorm16 jobs)
    unsigned ray_list_cap;
};

// Preprocess of one session: raw codes -> metres (fuse input), bilateral L0,
// pyramid and vertex maps at 3 levels.
struct PreJob {
    const uint16_t* raw;
    float* depth_m;
    float* d0;         // bilateral-filtered level 0
    float* dmax;       // tile-max maps of depth_m (FuseArgs::dmax layout)
    float4* v[3];      // camera-space vertex maps
    float* d12[2];     // filtered depth of levels 1 and 2 (optional)
    int w, h;
    float scale;
    float Kinv[3][9];
    int tile0, tiles_x;    // level-0 16x16 tiles of this job
    int ptile0, ptiles_x;  // level-2 16x16 tiles of this job
};
int launch_preprocess_batch(const BatchHeader* hd, const PreJob* jobs, unsigned blocks, hipStream_t s);
// The same two passes for one session (w, h divisible by 4; no mirror).
int launch_measure(const PreJob& J, hipStream_t s);

// ---- TSDF (tsdf.hip)
constexpr unsigned kFuseBlocks = 2048;  // fuse workgroups: 256 CUs x 8 resident workgroups of 4 waves
struct FuseArgs {
    const float* depth;
    int w, h;
    float K[9];
    float inv_pose[16];
    const PoseState* dyn;  // when set: inv_pose comes from dyn->inv_pose, skip if dyn->frame_skip_fuse
    int n, nb;
    float half, res, mue;
    float inv_mue;  // 1 / mue correctly rounded (the exact reciprocal the quotient diff / mue is formed from)
    const float* zk;
    float4* zt;     // (m2 z + m3) per row of the inverse pose for each k slice, written by the fuse's cull
    float band;     // near-integer band of the projection (2^-21 max(w, h))
    void* f;
    void* wt;
    unsigned long long* touched;
    int* brick_key;
    int* super_key;
    int ns;
    int cull;     // 1 when K is a plain pinhole matrix (culling planes valid)
    // lengths of the culling planes' normals |(fx, cx)|, |(fx, w - cx)|, |(fy, cy)|,
    // |(fy, h - cy)|, rounded up (they scale the bounding-sphere radius)
    float plane_norm[4];
    int format;   // kf_voxel_format
    // occlusion test: max valid depth per 16x16 tile (tx16 x ty16) followed by
    // the max per 64x64 tile (tx64 x ty64); null = frustum test only
    const float* dmax;
    int tx16, ty16, tx64, ty64;
};
// Floats of the tile-max maps of a w x h depth map.
inline size_t depth_max_floats(int w, int h)
{
    return (size_t)((w + 15) / 16) * ((h + 15) / 16) + (size_t)((w + 63) / 64) * ((h + 63) / 64);
}
struct FuseJob {
    FuseArgs a;       // a.dyn set: the session's committed pose
    unsigned brick0;  // first global brick index of this job
};
FuseArgs fuse_args(const float* depth, int w, int h, const kf_volume* v, const float* K, const float* inv_pose,
                   const PoseState* dyn_state);
// mid (optional): an event recorded between the brick cull and the sweep.
int launch_fuse_batch(const BatchHeader* hd, const FuseJob* jobs, unsigned cull_blocks, unsigned list_blocks,
                      hipStream_t s, hipEvent_t mid = nullptr);
// inv_pose: by-value world->camera matrix; dyn_state (may be null) overrides it
// with PoseState::inv_pose and skips the launch's work when the frame failed.
// list_events (optional): two events recorded around the brick-list kernel alone.
// dmax_ready: the volume's tile-max maps (volume_depth_max) already hold this
// depth map's (the measurement pass wrote them).
int launch_fuse(const float* depth, int w, int h, kf_volume* v, const float* K, const float* inv_pose,
                const PoseState* dyn_state, hipStream_t s, hipEvent_t* list_events, bool dmax_ready);
// The volume's tile-max map buffer sized for w x h (grown if needed); null on failure.
float* volume_depth_max(kf_volume* v, int w, int h, hipStream_t s);
int launch_initialize(kf_volume* v, float f, float wt, hipStream_t s);
// kf_volume_create without the host wait: initialisation is enqueued on `s`.
int volume_create(int n, float width_m, float mue_m, int format, hipStream_t s, kf_volume** out);

// ---- raycast (raycast.hip)
struct RaycastVolume {
    const void* f;
    const int* brick_key;    // per-brick halo lower bound of F, order_key form (null: no skipping)
    const int* super_key;    // per-super-brick lower bound of F, order_key form
    int n, nb, ns, format;
    float width, mue;
    float inv_width;  // 1 / width when width is a power of two (then x * inv_width == x / width), else 0
    int p2;           // width and n are powers of two (the march's texel coordinates are then exact)
};
struct RaycastLevel {
    float Kinv[9];
    float4* vout;
    float4* nout;
    int w, h, tiles_x, tile0;  // tile0: first global tile of this level
};
// A ray the per-lane march handed to the finishing pass: its march state
// before the sample at t (that sample already counted in guard).
struct DeferredRay {
    int job;    // pyramid level (single launch) or raycast job (batch)
    int px;     // j * w + i
    int guard;  // samples counted, the one at t included
    int flags;  // bit 0: the last sample was skipped (prev comes from prev_t), bit 1: small steps
    float t, prev, prev_t, pad;
};
// One (session, level) of a batched raycast; lives in device memory.
struct RaycastJob {
    RaycastVolume vol;
    const float* pose;   // device: camera-to-world pose of the session (PoseState::pose)
    RaycastLevel L;
};
RaycastVolume raycast_volume(const kf_volume* v);
// mid (optional): an event recorded between the first pass and the finishing pass.
int launch_raycast_batch(const BatchHeader* hd, const RaycastJob* jobs, unsigned blocks, hipStream_t s,
                         hipEvent_t mid = nullptr);
int launch_raycast(const kf_volume* v, const float* pose, const PoseState* dyn_state, const float* inv_k,
                   float4* vout, float4* nout, int w, int h, hipStream_t s);
// All pyramid levels in one launch (level l is w0 >> l by h0 >> l).
int launch_raycast_levels(const kf_volume* v, const float* pose, const PoseState* dyn_state, int levels,
                          const float (*inv_k)[9], float4* const* vout, float4* const* nout, int w0, int h0,
                          hipStream_t s);

// ---- ICP (icp.hip)
constexpr int kPartialStride = 64;  // floats per block partial: 27 sums, 27 compensations, count
struct IcpArgs {
    const float4* V;
    const float4* TV;
    const float4* TN;
    int w, h;
    float K[9];
    float dist_thresh, cos_thresh;
    PoseState* st;  // pose source (R, t, Rp, tp) and early-stop flags
    int level;
    float* partials;
    int stride;     // partial-row count of this launch (or of this job)
    int pix;        // pixels per lane (icp_pix_for of the map's pixel count)
};
// State carried from one ICP launch to the next (double-buffered by launch parity).
struct IcpCarry {
    float R[9];
    float t[3];
    int status;
    int level_done[4];
    int iterations;
    int pending;         // the writing launch left partial rows that still need their solve
    int pending_level;
    int pending_blocks;  // partial-row count of that launch
    int pad[2];
};
// One session at one level of a batched ICP launch.
struct IcpJob {
    IcpArgs a;         // a.stride = the session's blocks at this level; a.partials unused
    IcpCarry* carry;   // 2 entries, by launch parity
    float* partials;   // 2 x kPartialStride x max_blocks, by launch parity
    int max_blocks;
    int block0;        // first block of this job in the launch
};
struct IcpLevel {
    const float4* V;
    const float4* TV;
    const float4* TN;
    int w, h;
    float K[9];
};
int icp_pix_for(int64_t n);
int64_t icp_max_map_pixels();         // largest w x h an ICP map may have
int icp_level_blocks(int w, int h);  // workgroups of one iteration launch over a w x h map
int icp_blocks_for(int w, int h);    // the most of icp_level_blocks over the three levels of a w x h frame
// Enqueues the whole coarse-to-fine loop on `s` with the state on the device,
// ending with the commit of the estimate into state->pose / inv_pose.
int launch_icp_loop(kf_icp* icp, const IcpLevel* levels, int n_levels, hipStream_t s);
// begin: starts the frame's ICP from state->pose (or from pose16 if given);
// launch_icp_loop then enqueues one launch per iteration and the final commit.
int launch_icp_begin(kf_icp* icp, const float* pose16, hipStream_t s);
// Batched: begin, then steps k = 0.. over the levels coarse to fine, then commit.
int launch_icp_batch(const BatchHeader* hd, const IcpJob* jobs, const int* iters, int levels, int max_sessions,
                     int compensated, float stop_angle, float stop_dist, unsigned step_blocks, hipStream_t s);

// ---- visualisation (render.hip)
int launch_render(const float4* nmap, int w, int h, const float* mat9, int shading, uint32_t* out, hipStream_t s);
void shading_matrix(const float* pose16, const float* kinv9, float* out9);

}  // namespace kf
