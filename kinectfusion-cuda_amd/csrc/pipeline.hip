// Per-frame orchestration of the KinectFusion path (project/src/kinect_fusion.cpp:44-183)
// with one volume for tracking and fusion.  Every stage of a frame is enqueued
// on one stream and reads the pose from the device-resident PoseState, so a
// frame needs no host synchronisation (the reference synchronises after every
// kernel and copies the normal equations to the host every ICP iteration).
#include <string.h>

#include <vector>

#include "kf_device.hpp"
#include "kf_kernels.hpp"

struct kf_pipeline {
    kf_pipeline_config cfg{};
    hipStream_t s = nullptr;
    bool own_stream = false;
    kf_volume* vol = nullptr;
    kf_volume* vis = nullptr;                     // optional visualisation volume
    float4* rv = nullptr;                         // render maps (level 0), allocated on first render
    float4* rn = nullptr;
    kf_icp* icp = nullptr;
    float* depth_m = nullptr;
    float* dpyr[3] = {nullptr, nullptr, nullptr};
    float4* vpyr[3] = {nullptr, nullptr, nullptr};
    float4* npyr[3] = {nullptr, nullptr, nullptr};
    float4* pv[3] = {nullptr, nullptr, nullptr};
    float4* pn[3] = {nullptr, nullptr, nullptr};
    float K[3][9], Ki[3][9];
    int levels = 3;
    int64_t frame = 0;
    bool normals_fresh = false;  // npyr holds the normals of the current vertex pyramid
    // timing: 7 events per frame (start, after preprocess, after raycast, after ICP, after fuse,
    // and the pair around the fuse's brick-list kernel)
    bool timing = false;
    std::vector<hipEvent_t> pending;
    float stage_ms[KF_STAGE_COUNT] = {0, 0, 0, 0};
    int64_t timed_frames = 0;
    double fuse_ms = 0.0;
    int64_t fuse_launches = 0;
};

namespace {

int alloc(void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    return e == hipSuccess ? KF_OK : kf::fail_hip(e, "kf_pipeline_create allocation");
}

// Folds the recorded events of finished frames into the stage totals.
int drain_events(kf_pipeline* p)
{
    for (size_t f = 0; f + 7 <= p->pending.size(); f += 7) {
        hipEvent_t* e = &p->pending[f];
        KF_HIP(hipEventSynchronize(e[4]));
        float ms[4], list_ms = 0.0f;
        for (int q = 0; q < 4; ++q) KF_HIP(hipEventElapsedTime(&ms[q], e[q], e[q + 1]));
        KF_HIP(hipEventElapsedTime(&list_ms, e[5], e[6]));
        for (int q = 0; q < 4; ++q) p->stage_ms[q] += ms[q];
        p->fuse_ms += list_ms;
        p->fuse_launches += 1;
        p->timed_frames += 1;
    }
    for (hipEvent_t e : p->pending) (void)hipEventDestroy(e);
    p->pending.clear();
    return KF_OK;
}

}  // namespace

extern "C" {

int kf_pipeline_destroy(kf_pipeline* p)
{
    if (!p) return KF_OK;
    if (p->s) (void)hipStreamSynchronize(p->s);
    for (hipEvent_t e : p->pending) (void)hipEventDestroy(e);
    kf_volume_destroy(p->vol);
    kf_volume_destroy(p->vis);
    (void)hipFree(p->rv);
    (void)hipFree(p->rn);
    kf_icp_destroy(p->icp);
    (void)hipFree(p->depth_m);
    for (int l = 0; l < 3; ++l) {
        (void)hipFree(p->dpyr[l]);
        (void)hipFree(p->vpyr[l]);
        (void)hipFree(p->npyr[l]);
        (void)hipFree(p->pv[l]);
        (void)hipFree(p->pn[l]);
    }
    if (p->own_stream && p->s) (void)hipStreamDestroy(p->s);
    delete p;
    return KF_OK;
}

int kf_pipeline_create(const kf_pipeline_config* cfg, kf_stream_t stream, kf_pipeline** out)
{
    if (!cfg || !out) return kf::fail(KF_ERR_ARG, "kf_pipeline_create: null argument");
    *out = nullptr;
    const int w = cfg->frame_width, h = cfg->frame_height, levels = cfg->icp.levels;
    if (w <= 0 || h <= 0 || levels < 1 || levels > 3) return kf::fail(KF_ERR_ARG, "kf_pipeline_create: bad sizes");
    if (cfg->icp.width != w || cfg->icp.height != h)
        return kf::fail(KF_ERR_DIMS, "kf_pipeline_create: ICP size differs from the frame size");
    const int e = 1 << (levels - 1);
    if (w % e || h % e)  // grid_map_pyramid.h:63-72
        return kf::fail(KF_ERR_DIMS, "Can't construct pyramid. Dimensions must be divisible by 2^(m_n_layers - 1).");
    kf_pipeline* p = new kf_pipeline();
    p->cfg = *cfg;
    p->levels = levels;
    int st = KF_OK;
    if (stream) {
        p->s = kf::stream_of(stream);
    } else {
        hipError_t he = hipStreamCreateWithFlags(&p->s, hipStreamNonBlocking);
        if (he != hipSuccess) st = kf::fail_hip(he, "kf_pipeline_create stream");
        p->own_stream = true;
    }
    if (st == KF_OK) st = kf_volume_create(cfg->volume_n, cfg->volume_size, cfg->mue, cfg->voxel_format, &p->vol);
    if (st == KF_OK && cfg->vis_volume_n > 0)
        st = kf_volume_create(cfg->vis_volume_n, cfg->vis_volume_size, cfg->vis_mue, cfg->voxel_format, &p->vis);
    if (st == KF_OK) st = kf_icp_create(&cfg->icp, &p->icp);
    kf_sensor sensor;
    if (st == KF_OK) st = kf_sensor_init(&sensor, cfg->fx, cfg->fy, cfg->cx, cfg->cy);
    if (st == KF_OK) st = kf_sensor_set_pose(&sensor, cfg->initial_pose);
    if (st == KF_OK) {
        memcpy(p->K, sensor.intr, sizeof(p->K));
        memcpy(p->Ki, sensor.inv_intr, sizeof(p->Ki));
        st = alloc((void**)&p->depth_m, sizeof(float) * (size_t)w * h);
    }
    for (int l = 0; l < levels && st == KF_OK; ++l) {
        size_t px = (size_t)(w >> l) * (h >> l);
        if ((st = alloc((void**)&p->dpyr[l], sizeof(float) * px))) break;
        if ((st = alloc((void**)&p->vpyr[l], sizeof(float4) * px))) break;
        if ((st = alloc((void**)&p->npyr[l], sizeof(float4) * px))) break;
        if ((st = alloc((void**)&p->pv[l], sizeof(float4) * px))) break;
        if ((st = alloc((void**)&p->pn[l], sizeof(float4) * px))) break;
        // predicted maps start invalid (the reference's arrays start uninitialised)
        std::vector<float4> inval(px, make_float4(0.0f, 0.0f, 0.0f, kf::kInvalid));
        hipError_t he = hipMemcpy(p->pv[l], inval.data(), sizeof(float4) * px, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(p->pn[l], inval.data(), sizeof(float4) * px, hipMemcpyHostToDevice);
        if (he != hipSuccess) st = kf::fail_hip(he, "kf_pipeline_create init maps");
    }
    if (st == KF_OK) {
        kf::PoseState hs;
        memset(&hs, 0, sizeof(hs));
        memcpy(hs.pose, sensor.pose, sizeof(hs.pose));
        memcpy(hs.inv_pose, sensor.inv_pose, sizeof(hs.inv_pose));
        hipError_t he = hipMemcpy(p->icp->state, &hs, sizeof(hs), hipMemcpyHostToDevice);
        if (he != hipSuccess) st = kf::fail_hip(he, "kf_pipeline_create state");
    }
    if (st != KF_OK) {
        kf_pipeline_destroy(p);
        return st;
    }
    *out = p;
    return KF_OK;
}

int kf_pipeline_frame(kf_pipeline* p, const uint16_t* raw)
{
    if (!p || !raw) return kf::fail(KF_ERR_ARG, "kf_pipeline_frame: null argument");
    const int w = p->cfg.frame_width, h = p->cfg.frame_height;
    hipStream_t s = p->s;
    hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (p->timing) {
        // timing-only events: no system-scope cache write-back when they complete
        // (their timestamps are read after a stream synchronisation)
        for (auto& e : ev) KF_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        KF_HIP(hipEventRecord(ev[0], s));
    }
    int st;
    kf::PoseState* state = p->icp->state;
    // the fused two-pass measurement (metres, bilateral, pyramid, vertex maps and
    // the fuse's tile-max maps) when the frame allows it; else the stage kernels
    const bool fused = !p->cfg.unmirror && p->levels == 3 && w % 4 == 0 && h % 4 == 0;
    if (fused) {
        kf::PreJob J;
        memset(&J, 0, sizeof(J));
        J.raw = raw;
        J.depth_m = p->depth_m;
        J.d0 = p->dpyr[0];
        J.dmax = kf::volume_depth_max(p->vol, w, h, s);
        if (!J.dmax) return kf::fail(KF_ERR_OOM, "kf_pipeline_frame: depth tile-max maps");
        for (int l = 0; l < 3; ++l) J.v[l] = p->vpyr[l];
        J.d12[0] = p->dpyr[1];
        J.d12[1] = p->dpyr[2];
        J.w = w;
        J.h = h;
        J.scale = p->cfg.depth_scale;
        memcpy(J.Kinv, p->Ki, sizeof(J.Kinv));
        J.tiles_x = (w + 15) / 16;
        J.ptiles_x = (w / 4 + 15) / 16;
        if ((st = kf::launch_measure(J, s))) return st;
        p->normals_fresh = false;
    }
    if (p->frame == 0) {
        // initializeTracking (kinect_fusion.cpp:75-94): convert and fuse at the initial pose
        if (!fused && (st = kf::launch_convert(raw, p->depth_m, w, h, p->cfg.depth_scale, p->cfg.unmirror, s)))
            return st;
        if (p->timing) {
            KF_HIP(hipEventRecord(ev[1], s));
            KF_HIP(hipEventRecord(ev[2], s));
            KF_HIP(hipEventRecord(ev[3], s));
        }
    } else {
        // depthFrameToVertexPyramid (kinect_fusion.cpp:116-135); ICP computes the live
        // normals itself, so the normal pyramid is only built when a caller asks for it
        if (!fused) {
            if ((st = kf::launch_preprocess(raw, w, h, p->cfg.depth_scale, p->cfg.unmirror, p->levels, p->Ki,
                                            p->depth_m, p->dpyr, p->vpyr, nullptr, s)))
                return st;
            p->normals_fresh = false;
        }
        if (p->timing) KF_HIP(hipEventRecord(ev[1], s));
        // raycastTsdf (kinect_fusion.cpp:137-144) at the previous pose
        if ((st = kf::launch_raycast_levels(p->vol, nullptr, state, p->levels, p->Ki, p->pv, p->pn, w, h, s)))
            return st;
        if (p->timing) KF_HIP(hipEventRecord(ev[2], s));
        // computePose (kinect_fusion.cpp:146-161)
        kf::IcpLevel lv[3];
        for (int l = 0; l < p->levels; ++l) {
            lv[l] = kf::IcpLevel{p->vpyr[l], p->pv[l], p->pn[l], w >> l, h >> l, {0}};
            memcpy(lv[l].K, p->K[l], sizeof(lv[l].K));
        }
        if ((st = kf::launch_icp_begin(p->icp, nullptr, s))) return st;
        if ((st = kf::launch_icp_loop(p->icp, lv, p->levels, s))) return st;
        if (p->timing) KF_HIP(hipEventRecord(ev[3], s));
    }
    // fuseCurrentDepthToTSDF (kinect_fusion.cpp:179-183): unfiltered metres at the new pose
    if ((st = kf::launch_fuse(p->depth_m, w, h, p->vol, p->K[0], nullptr, state, s, p->timing ? ev + 5 : nullptr,
                              fused)))
        return st;
    if (p->timing) KF_HIP(hipEventRecord(ev[4], s));
    // the visualisation volume takes the same depth at the same pose (kinect_fusion.cpp:179-183)
    if (p->vis && (st = kf::launch_fuse(p->depth_m, w, h, p->vis, p->K[0], nullptr, state, s, nullptr, false)))
        return st;
    if (p->timing)
        for (auto e : ev) p->pending.push_back(e);
    p->frame += 1;
    return KF_OK;
}

int kf_pipeline_sync(kf_pipeline* p, int* status)
{
    if (!p) return kf::fail(KF_ERR_ARG, "kf_pipeline_sync: null pipeline");
    KF_HIP(hipStreamSynchronize(p->s));
    int st = drain_events(p);
    if (st) return st;
    int sticky = 0;
    KF_HIP(hipMemcpy(&sticky, &p->icp->state->sticky_status, sizeof(int), hipMemcpyDeviceToHost));
    if (sticky) {
        int zero = 0;
        KF_HIP(hipMemcpy(&p->icp->state->sticky_status, &zero, sizeof(int), hipMemcpyHostToDevice));
    }
    if (status) *status = sticky;
    return KF_OK;
}

int kf_pipeline_pose(kf_pipeline* p, float pose[16])
{
    if (!p || !pose) return kf::fail(KF_ERR_ARG, "kf_pipeline_pose: null argument");
    KF_HIP(hipStreamSynchronize(p->s));
    KF_HIP(hipMemcpy(pose, p->icp->state->pose, sizeof(float) * 16, hipMemcpyDeviceToHost));
    return KF_OK;
}

int kf_pipeline_volume(kf_pipeline* p, kf_volume** v)
{
    if (!p || !v) return kf::fail(KF_ERR_ARG, "kf_pipeline_volume: null argument");
    *v = p->vol;
    return KF_OK;
}

int kf_pipeline_map(kf_pipeline* p, int which, int level, void** dev_ptr, int* w, int* h)
{
    if (!p || !dev_ptr || level < 0 || level >= p->levels) return kf::fail(KF_ERR_ARG, "kf_pipeline_map: bad argument");
    void* ptr = nullptr;
    switch (which) {
        case 0: ptr = p->depth_m; level = 0; break;
        case 1: ptr = p->dpyr[level]; break;
        case 2: ptr = p->vpyr[level]; break;
        case 3:
            // computed on request (createNormalMap, measurement.cu:223-242) from the current vertex maps
            if (!p->normals_fresh) {
                for (int l = 0; l < p->levels; ++l) {
                    int st = kf::launch_normal_map(p->vpyr[l], p->npyr[l], p->cfg.frame_width >> l,
                                                   p->cfg.frame_height >> l, p->s);
                    if (st) return st;
                }
                p->normals_fresh = true;
            }
            ptr = p->npyr[level];
            break;
        case 4: ptr = p->pv[level]; break;
        case 5: ptr = p->pn[level]; break;
        default: return kf::fail(KF_ERR_ARG, "kf_pipeline_map: unknown map");
    }
    *dev_ptr = ptr;
    if (w) *w = p->cfg.frame_width >> level;
    if (h) *h = p->cfg.frame_height >> level;
    return KF_OK;
}

int kf_pipeline_render(kf_pipeline* p, const float pose[16], int shading, uint32_t* rgba)
{
    if (!p || !rgba) return kf::fail(KF_ERR_ARG, "kf_pipeline_render: null argument");
    const int w = p->cfg.frame_width, h = p->cfg.frame_height;
    const size_t px = (size_t)w * h;
    if (!p->rv) {
        KF_HIP(hipMalloc(&p->rv, sizeof(float4) * px));
        KF_HIP(hipMalloc(&p->rn, sizeof(float4) * px));
        std::vector<float4> inval(px, make_float4(0.0f, 0.0f, 0.0f, kf::kInvalid));
        KF_HIP(hipMemcpy(p->rv, inval.data(), sizeof(float4) * px, hipMemcpyHostToDevice));
        KF_HIP(hipMemcpy(p->rn, inval.data(), sizeof(float4) * px, hipMemcpyHostToDevice));
    }
    float P[16];
    if (pose) {
        memcpy(P, pose, sizeof(P));
    } else {
        KF_HIP(hipStreamSynchronize(p->s));
        KF_HIP(hipMemcpy(P, p->icp->state->pose, sizeof(P), hipMemcpyDeviceToHost));
    }
    const kf_volume* v = p->vis ? p->vis : p->vol;
    int st = kf::launch_raycast(v, P, nullptr, p->Ki[0], p->rv, p->rn, w, h, p->s);
    if (st) return st;
    float m[9];
    if (shading) {
        kf::shading_matrix(P, p->Ki[0], m);
    } else {
        // glm::mat3(inverse pose) of a rigid pose: R^T
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) m[c * 3 + r] = P[r * 4 + c];
    }
    if ((st = kf::launch_render(p->rn, w, h, m, shading, rgba, p->s))) return st;
    KF_HIP(hipStreamSynchronize(p->s));
    return KF_OK;
}

int kf_pipeline_set_timing(kf_pipeline* p, int enabled)
{
    if (!p) return kf::fail(KF_ERR_ARG, "kf_pipeline_set_timing: null pipeline");
    KF_HIP(hipStreamSynchronize(p->s));
    int st = drain_events(p);
    p->timing = enabled != 0;
    for (float& v : p->stage_ms) v = 0.0f;
    p->timed_frames = 0;
    p->fuse_ms = 0.0;
    p->fuse_launches = 0;
    return st;
}

int kf_pipeline_stage_times(kf_pipeline* p, float total_ms[KF_STAGE_COUNT], int64_t* frames)
{
    if (!p || !total_ms) return kf::fail(KF_ERR_ARG, "kf_pipeline_stage_times: null argument");
    int st = drain_events(p);
    for (int q = 0; q < KF_STAGE_COUNT; ++q) total_ms[q] = p->stage_ms[q];
    if (frames) *frames = p->timed_frames;
    return st;
}

int kf_pipeline_fuse_kernel_ms(kf_pipeline* p, float* avg_ms, int64_t* launches)
{
    if (!p || !avg_ms) return kf::fail(KF_ERR_ARG, "kf_pipeline_fuse_kernel_ms: null argument");
    int st = drain_events(p);
    *avg_ms = p->fuse_launches ? (float)(p->fuse_ms / (double)p->fuse_launches) : 0.0f;
    if (launches) *launches = p->fuse_launches;
    return st;
}

}  // extern "C"
