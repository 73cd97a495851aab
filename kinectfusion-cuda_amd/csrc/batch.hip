// Batch ticks: many independent sessions (each with its own volume, frame
// size, intrinsics and pose) advanced one frame per tick, with one persistent
// launch per stage across every session of the tick.
//
// A tick is described by a table in device memory: a BatchHeader followed by
// the job arrays of every stage (PreJob, RaycastJob per session and level,
// IcpJob per level and session, FuseJob).  The host fills the table in pinned
// memory; the tick's first operation copies it to the device, so one captured
// hipGraph (per table slot) serves every tick whatever the set of sessions.
// Two table slots alternate, so the host can fill tick t+1 while tick t runs.
//
// Launch shapes (workgroups per persistent launch) never change a result; the
// optional autotuner times candidate shapes per stage on the first ticks
// (direct launches with events between the stages) and keeps the fastest.
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include <string>

#include <algorithm>
#include <vector>

#include "kf_device.hpp"
#include "kf_kernels.hpp"

namespace {

constexpr int kCandidates = 4;
constexpr int kTuneReps = 3;
constexpr unsigned kMaxFuseBlocks = 32768;  // touched slots

struct Session {
    bool used = false;
    kf_session_config cfg{};
    kf_volume* vol = nullptr;
    kf::PoseState* st = nullptr;
    kf::IcpCarry* carry = nullptr;
    float* partials = nullptr;
    int max_blocks = 0;
    float* depth_m = nullptr;
    float* d0 = nullptr;
    float* dmax = nullptr;
    float4* v[3] = {nullptr, nullptr, nullptr};
    float4* pv[3] = {nullptr, nullptr, nullptr};
    float4* pn[3] = {nullptr, nullptr, nullptr};
    float K[3][9], Ki[3][9];
    int64_t frames = 0;
    size_t bytes = 0;
};

struct Slot {
    void* host = nullptr;  // pinned copy of the table
    void* dev = nullptr;
    hipGraphExec_t exec = nullptr;
    // tick start, after each stage (preprocess, raycast, ICP, fuse); with per-kernel
    // timing also after the raycast's first pass (5) and after the fuse's brick cull (6)
    hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool fine = false;     // the pending tick recorded events 5 and 6
    bool pending = false;  // a tick using this slot may still run
    bool staged = false;   // stage events were recorded (direct launch)
    bool tuning = false;   // the pending tick is a tuning sample
    int cand = 0;
    bool has_stage[4] = {false, false, false, false};
};

const unsigned kCand[4][kCandidates] = {
    {2048, 4096, 8192, 16384},   // preprocess
    {2048, 4096, 8192, 16384},   // raycast
    {256, 512, 1024, 2048},      // ICP steps
    {4096, 8192, 16384, 32768},  // fuse list
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

__global__ void fill_invalid_kernel(float4* m, int64_t n)
{
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) m[q] = make_float4(0.0f, 0.0f, 0.0f, kf::kInvalid);
}

struct Pose32 {
    float m[32];  // pose, inverse pose
};
// A new session's PoseState (pose and its inverse, everything else zero) and
// zeroed ICP carries.
__global__ void pose_state_init_kernel(kf::PoseState* st, kf::IcpCarry* carry, Pose32 p)
{
    if (threadIdx.x != 0) return;
    kf::PoseState z;
    memset(&z, 0, sizeof(z));
    for (int q = 0; q < 16; ++q) {
        z.pose[q] = p.m[q];
        z.inv_pose[q] = p.m[16 + q];
    }
    *st = z;
    kf::IcpCarry c;
    memset(&c, 0, sizeof(c));
    carry[0] = c;
    carry[1] = c;
}

}  // namespace

struct kf_batch {
    kf_batch_config cfg{};
    hipStream_t s = nullptr;
    bool own_stream = false;
    std::vector<Session> sess;
    size_t off_pre = 0, off_ray = 0, off_icp = 0, off_fuse = 0, table_bytes = 0;
    Slot slot[2];
    int next = 0;
    unsigned* fuse_count = nullptr;
    unsigned* ray_next = nullptr;  // 16 raycast tile counters + 2 deferred-ray counts
    kf::DeferredRay* ray_list = nullptr;  // 2 x ray_cap deferred rays (per voxel format)
    size_t ray_cap = 0;
    unsigned long long* fuse_list = nullptr;
    size_t list_cap = 0;
    unsigned long long* touched = nullptr;
    kf_batch_shape shape{};
    // autotuner: best stage time per candidate (ms), samples taken
    bool tuning = false;
    float tune_best[4][kCandidates];
    int tune_count[4][kCandidates];
    int tune_tick = 0;
    // per-kernel timing (kf_batch_set_stage_timing)
    bool stage_timing = false;
    double stage_ms[7] = {0, 0, 0, 0, 0, 0, 0};
    int64_t stage_ticks = 0;
    // tick timing
    double tick_ms_total = 0.0;
    int64_t ticks_timed = 0;
    float tick_ms_max = 0.0f;
};

namespace {

int alloc(void** p, size_t bytes, size_t* acc)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return kf::fail_hip(e, "batch allocation");
    if (acc) *acc += bytes;
    return KF_OK;
}

size_t volume_meta_bytes(int n)
{
    const size_t nb = (size_t)n / 8, ns = (nb + 7) / 8;
    return (sizeof(float) + sizeof(float4)) * n + sizeof(unsigned long long) * kf::kFuseBlocks +
           nb * nb * nb * (sizeof(int) + sizeof(int)) + ns * ns * ns * sizeof(int) + 2 * sizeof(unsigned);
}

size_t maps_bytes(int w, int h)
{
    size_t px = (size_t)w * h, b = 2 * sizeof(float) * px + sizeof(float) * kf::depth_max_floats(w, h);
    for (int l = 0; l < 3; ++l) b += 3 * sizeof(float4) * (size_t)(w >> l) * (h >> l);
    return b + 2 * sizeof(kf::IcpCarry) + 2 * sizeof(float) * kf::kPartialStride * (size_t)kf::icp_blocks_for(w, h) +
           sizeof(kf::PoseState);
}

}  // namespace

size_t kf::batch_shared_bytes(const kf_session_config* c)
{
    const size_t nb = (size_t)c->volume_n / 8, w = (size_t)c->frame_width, h = (size_t)c->frame_height;
    const size_t rays = w * h + (w / 2) * (h / 2) + (w / 4) * (h / 4);
    return 2 * (sizeof(unsigned long long) * nb * nb * nb + 2 * sizeof(kf::DeferredRay) * rays);
}

namespace {

int check_session_config(const kf_session_config* c)
{
    if (c->frame_width <= 0 || c->frame_height <= 0 || c->frame_width % 4 || c->frame_height % 4)
        return kf::fail(KF_ERR_DIMS, "Can't construct pyramid. Dimensions must be divisible by 2^(m_n_layers - 1).");
    if ((int64_t)c->frame_width * c->frame_height > kf::icp_max_map_pixels())
        return kf::fail(KF_ERR_DIMS, "session: frame above 2^24 pixels");
    if (c->volume_n < 8 || c->volume_n % 8) return kf::fail(KF_ERR_ARG, "session: volume_n must be a multiple of 8");
    if (c->voxel_format != KF_VOXEL_F32 && c->voxel_format != KF_VOXEL_SNORM16_U8)
        return kf::fail(KF_ERR_ARG, "session: unknown voxel format");
    return KF_OK;
}

void free_session(Session& q)
{
    kf_volume_destroy(q.vol);
    (void)hipFree(q.st);
    (void)hipFree(q.carry);
    (void)hipFree(q.partials);
    (void)hipFree(q.depth_m);
    (void)hipFree(q.d0);
    (void)hipFree(q.dmax);
    for (int l = 0; l < 3; ++l) {
        (void)hipFree(q.v[l]);
        (void)hipFree(q.pv[l]);
        (void)hipFree(q.pn[l]);
    }
    q = Session();
}

int drop_graphs(kf_batch* b)
{
    for (Slot& sl : b->slot)
        if (sl.exec) {
            (void)hipGraphExecDestroy(sl.exec);
            sl.exec = nullptr;
        }
    return KF_OK;
}

// Total bricks of the resident sessions: the fuse list never holds more.
// Deferred raycast rays: room for every ray of the resident sessions, per
// voxel format, up to kMaxDeferred (beyond it a lane marches its ray itself).
constexpr size_t kMaxDeferred = size_t(1) << 21;

int ensure_ray_list(kf_batch* b)
{
    size_t need = 0;
    for (const Session& q : b->sess)
        if (q.used) {
            const size_t w = (size_t)q.cfg.frame_width, h = (size_t)q.cfg.frame_height;
            need += w * h + (w / 2) * (h / 2) + (w / 4) * (h / 4);
        }
    need = std::min(need, kMaxDeferred);
    if (need <= b->ray_cap) return KF_OK;
    need = std::min(std::max(need, 2 * b->ray_cap), kMaxDeferred);
    KF_HIP(hipStreamSynchronize(b->s));
    (void)hipFree(b->ray_list);
    b->ray_list = nullptr;
    b->ray_cap = 0;
    int st = alloc((void**)&b->ray_list, 2 * sizeof(kf::DeferredRay) * need, nullptr);
    if (st == KF_OK) b->ray_cap = need;
    return st;
}

int ensure_list(kf_batch* b)
{
    int rs = ensure_ray_list(b);
    if (rs) return rs;
    size_t need = 0;
    for (const Session& q : b->sess)
        if (q.used) need += (size_t)(q.cfg.volume_n / 8) * (q.cfg.volume_n / 8) * (q.cfg.volume_n / 8);
    if (need <= b->list_cap) return KF_OK;
    need = std::max(need, 2 * b->list_cap);  // grow geometrically: each growth waits for the ticks in flight
    KF_HIP(hipStreamSynchronize(b->s));
    (void)hipFree(b->fuse_list);
    b->fuse_list = nullptr;
    b->list_cap = 0;
    int st = alloc((void**)&b->fuse_list, sizeof(unsigned long long) * need, nullptr);
    if (st == KF_OK) b->list_cap = need;
    return st;
}

// The stages of one tick on the batch's stream.  ev (optional): 5 events
// recorded at the tick's start and after each stage.
int enqueue_tick(kf_batch* b, const Slot& sl, const kf_batch_shape& sh, hipEvent_t* ev, bool fine = false)
{
    const char* base = static_cast<const char*>(sl.dev);
    const auto* hd = reinterpret_cast<const kf::BatchHeader*>(base);
    const auto* pre = reinterpret_cast<const kf::PreJob*>(base + b->off_pre);
    const auto* ray = reinterpret_cast<const kf::RaycastJob*>(base + b->off_ray);
    const auto* icp = reinterpret_cast<const kf::IcpJob*>(base + b->off_icp);
    const auto* fuse = reinterpret_cast<const kf::FuseJob*>(base + b->off_fuse);
    hipStream_t s = b->s;
    KF_HIP(hipMemcpyAsync(sl.dev, sl.host, b->table_bytes, hipMemcpyHostToDevice, s));
    if (ev) KF_HIP(hipEventRecord(ev[0], s));
    int st;
    if ((st = kf::launch_preprocess_batch(hd, pre, sh.preprocess_blocks, s))) return st;
    if (ev) KF_HIP(hipEventRecord(ev[1], s));
    if ((st = kf::launch_raycast_batch(hd, ray, sh.raycast_blocks, s, fine ? ev[5] : nullptr))) return st;
    if (ev) KF_HIP(hipEventRecord(ev[2], s));
    const kf_icp_config& ic = b->cfg.icp;
    if ((st = kf::launch_icp_batch(hd, icp, ic.iters_per_layer, ic.levels, b->cfg.max_sessions, ic.compensated,
                                   ic.iteration_stop_thresh_angle, ic.iteration_stop_thresh_distance, sh.icp_blocks,
                                   s)))
        return st;
    if (ev) KF_HIP(hipEventRecord(ev[3], s));
    if ((st = kf::launch_fuse_batch(hd, fuse, sh.cull_blocks, sh.fuse_blocks, s, fine ? ev[6] : nullptr))) return st;
    if (ev) KF_HIP(hipEventRecord(ev[4], s));
    return KF_OK;
}

int capture(kf_batch* b, Slot& sl)
{
    hipGraph_t g = nullptr;
    KF_HIP(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
    int st = enqueue_tick(b, sl, b->shape, nullptr);
    hipError_t e = hipStreamEndCapture(b->s, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return kf::fail_hip(e, "batch tick capture");
    e = hipGraphInstantiate(&sl.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        sl.exec = nullptr;
        return kf::fail_hip(e, "batch tick instantiate");
    }
    return KF_OK;
}

// Folds a finished tick of `sl` into the timing and the tuner.
int retire(kf_batch* b, Slot& sl)
{
    if (!sl.pending) return KF_OK;
    KF_HIP(hipEventSynchronize(sl.ev[4]));
    float ms = 0.0f;
    KF_HIP(hipEventElapsedTime(&ms, sl.ev[0], sl.ev[4]));
    b->tick_ms_total += ms;
    b->ticks_timed += 1;
    b->tick_ms_max = std::max(b->tick_ms_max, ms);
    if (sl.tuning) {
        const int stage_ev[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}};
        for (int stg = 0; stg < 4; ++stg) {
            if (!sl.has_stage[stg]) continue;
            float t = 0.0f;
            KF_HIP(hipEventElapsedTime(&t, sl.ev[stage_ev[stg][0]], sl.ev[stage_ev[stg][1]]));
            float& best = b->tune_best[stg][sl.cand];
            if (b->tune_count[stg][sl.cand] == 0 || t < best) best = t;
            b->tune_count[stg][sl.cand] += 1;
        }
    }
    if (sl.fine) {
        const int pairs[7][2] = {{0, 1}, {1, 5}, {5, 2}, {2, 3}, {3, 6}, {6, 4}, {0, 4}};
        for (int k = 0; k < 7; ++k) {
            float t = 0.0f;
            KF_HIP(hipEventElapsedTime(&t, sl.ev[pairs[k][0]], sl.ev[pairs[k][1]]));
            b->stage_ms[k] += t;
        }
        b->stage_ticks += 1;
    }
    sl.pending = false;
    sl.tuning = false;
    sl.fine = false;
    return KF_OK;
}

// Ends tuning when every stage that ran has kTuneReps samples per candidate.
void finish_tuning(kf_batch* b)
{
    bool done = true;
    bool any = false;
    for (int stg = 0; stg < 4; ++stg) {
        int have = 0, ran = 0;
        for (int c = 0; c < kCandidates; ++c) {
            have += b->tune_count[stg][c] >= kTuneReps ? 1 : 0;
            ran += b->tune_count[stg][c];
        }
        if (ran) any = true;
        if (ran && have < kCandidates) done = false;
    }
    // stages that never had work keep their shapes; give up after many ticks
    if (!(any && done) && b->tune_tick < 8 * kCandidates * kTuneReps) return;
    unsigned* fields[4] = {&b->shape.preprocess_blocks, &b->shape.raycast_blocks, &b->shape.icp_blocks,
                           &b->shape.fuse_blocks};
    for (int stg = 0; stg < 4; ++stg) {
        int best = -1;
        for (int c = 0; c < kCandidates; ++c)
            if (b->tune_count[stg][c] > 0 && (best < 0 || b->tune_best[stg][c] < b->tune_best[stg][best])) best = c;
        if (best >= 0) *fields[stg] = kCand[stg][best];
    }
    b->tuning = false;
    drop_graphs(b);
}

void default_shape(kf_batch_shape& s, const kf_batch_shape& given)
{
    // the shapes the tuner picks for 64 sessions of 640x480 -> 512^3
    s.preprocess_blocks = given.preprocess_blocks ? given.preprocess_blocks : 8192;
    s.raycast_blocks = given.raycast_blocks ? given.raycast_blocks : 8192;
    s.icp_blocks = given.icp_blocks ? given.icp_blocks : 1024;
    // the brick cull (not tuned): per 64 x 512^3 tick 0.18 ms at 4096 workgroups, 0.16 at 8192,
    // 0.15 at 16384
    s.cull_blocks = given.cull_blocks ? given.cull_blocks : 16384;
    s.fuse_blocks = given.fuse_blocks ? given.fuse_blocks : 16384;
}

bool valid_shape(const kf_batch_shape& s)
{
    const unsigned v[5] = {s.preprocess_blocks, s.raycast_blocks, s.icp_blocks, s.cull_blocks, s.fuse_blocks};
    for (unsigned x : v)
        if (x == 0 || x % 8 || x > 65536) return false;
    return s.fuse_blocks <= kMaxFuseBlocks;
}

Session* session_of(kf_batch* b, int id)
{
    if (!b || id < 0 || id >= (int)b->sess.size() || !b->sess[id].used) return nullptr;
    return &b->sess[id];
}

}  // namespace

extern "C" {

int kf_session_bytes(const kf_session_config* c, size_t* bytes)
{
    if (!c || !bytes) return kf::fail(KF_ERR_ARG, "kf_session_bytes: null argument");
    int st = check_session_config(c);
    if (st) return st;
    const size_t n = (size_t)c->volume_n, vox = n * n * n;
    const size_t planes = vox * (c->voxel_format == KF_VOXEL_F32 ? 8 : 3);
    *bytes = planes + volume_meta_bytes(c->volume_n) + maps_bytes(c->frame_width, c->frame_height) +
             kf::batch_shared_bytes(c);
    return KF_OK;
}

int kf_batch_destroy(kf_batch* b)
{
    if (!b) return KF_OK;
    if (b->s) (void)hipStreamSynchronize(b->s);
    drop_graphs(b);
    for (Slot& sl : b->slot) {
        if (sl.host) (void)hipHostFree(sl.host);
        (void)hipFree(sl.dev);
        for (hipEvent_t e : sl.ev)
            if (e) (void)hipEventDestroy(e);
    }
    for (Session& q : b->sess)
        if (q.used) free_session(q);
    (void)hipFree(b->fuse_count);
    (void)hipFree(b->ray_next);
    (void)hipFree(b->ray_list);
    (void)hipFree(b->fuse_list);
    (void)hipFree(b->touched);
    if (b->own_stream && b->s) (void)hipStreamDestroy(b->s);
    delete b;
    return KF_OK;
}

int kf_batch_create(const kf_batch_config* cfg, kf_stream_t stream, kf_batch** out)
{
    if (!cfg || !out) return kf::fail(KF_ERR_ARG, "kf_batch_create: null argument");
    *out = nullptr;
    if (cfg->max_sessions < 1 || cfg->max_sessions > 4096)
        return kf::fail(KF_ERR_ARG, "kf_batch_create: max_sessions must be 1..4096");
    if (cfg->icp.levels < 1 || cfg->icp.levels > 3) return kf::fail(KF_ERR_ARG, "kf_batch_create: ICP levels 1..3");
    for (int l = 0; l < cfg->icp.levels; ++l)
        if (cfg->icp.iters_per_layer[l] < 0) return kf::fail(KF_ERR_ARG, "kf_batch_create: negative iteration count");
    kf_batch* b = new kf_batch();
    b->cfg = *cfg;
    default_shape(b->shape, cfg->shape);
    if (!valid_shape(b->shape)) {
        delete b;
        return kf::fail(KF_ERR_ARG, "kf_batch_create: launch shapes must be multiples of 8, fuse_blocks <= 8192");
    }
    b->tuning = cfg->autotune != 0;
    memset(b->tune_best, 0, sizeof(b->tune_best));
    memset(b->tune_count, 0, sizeof(b->tune_count));
    b->sess.resize((size_t)cfg->max_sessions);
    const size_t m = (size_t)cfg->max_sessions;
    b->off_pre = align256(sizeof(kf::BatchHeader));
    b->off_ray = b->off_pre + align256(sizeof(kf::PreJob) * m);
    b->off_icp = b->off_ray + align256(sizeof(kf::RaycastJob) * 3 * m);
    b->off_fuse = b->off_icp + align256(sizeof(kf::IcpJob) * 3 * m);
    b->table_bytes = b->off_fuse + align256(sizeof(kf::FuseJob) * m);
    int st = KF_OK;
    if (stream) {
        b->s = kf::stream_of(stream);
    } else {
        hipError_t e = hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking);
        if (e != hipSuccess) st = kf::fail_hip(e, "kf_batch_create stream");
        b->own_stream = true;
    }
    for (Slot& sl : b->slot) {
        if (st) break;
        hipError_t e = hipHostMalloc(&sl.host, b->table_bytes, hipHostMallocDefault);
        if (e != hipSuccess) st = kf::fail_hip(e, "kf_batch_create pinned table");
        if (!st) st = alloc(&sl.dev, b->table_bytes, nullptr);
        if (!st) memset(sl.host, 0, b->table_bytes);
        for (hipEvent_t& ev : sl.ev)
            if (!st && hipEventCreate(&ev) != hipSuccess) st = kf::fail(KF_ERR_HIP, "kf_batch_create events");
    }
    if (!st) st = alloc((void**)&b->fuse_count, sizeof(unsigned), nullptr);
    if (!st) st = alloc((void**)&b->ray_next, 18 * sizeof(unsigned), nullptr);
    if (!st) st = alloc((void**)&b->touched, sizeof(unsigned long long) * kMaxFuseBlocks, nullptr);
    if (!st) {
        hipError_t e = hipMemset(b->touched, 0, sizeof(unsigned long long) * kMaxFuseBlocks);
        if (e == hipSuccess) e = hipMemset(b->fuse_count, 0, sizeof(unsigned));
        if (e != hipSuccess) st = kf::fail_hip(e, "kf_batch_create clear");
    }
    if (st) {
        kf_batch_destroy(b);
        return st;
    }
    *out = b;
    return KF_OK;
}

int kf_batch_add_session(kf_batch* b, const kf_session_config* c, int* id)
{
    if (!b || !c || !id) return kf::fail(KF_ERR_ARG, "kf_batch_add_session: null argument");
    int st = check_session_config(c);
    if (st) return st;
    int slot = -1;
    for (int q = 0; q < (int)b->sess.size(); ++q)
        if (!b->sess[q].used) {
            slot = q;
            break;
        }
    if (slot < 0) return kf::fail(KF_ERR_BUDGET, "kf_batch_add_session: every session slot is in use");
    Session q;
    q.cfg = *c;
    const int w = c->frame_width, h = c->frame_height;
    kf_sensor sensor;
    // everything is initialised on the batch's stream, so admitting a session
    // does not wait for the ticks in flight
    st = kf::volume_create(c->volume_n, c->volume_size, c->mue, c->voxel_format, b->s, &q.vol);
    if (!st) st = kf_sensor_init(&sensor, c->fx, c->fy, c->cx, c->cy);
    if (!st) st = kf_sensor_set_pose(&sensor, c->initial_pose);
    if (!st) {
        memcpy(q.K, sensor.intr, sizeof(q.K));
        memcpy(q.Ki, sensor.inv_intr, sizeof(q.Ki));
        q.max_blocks = kf::icp_blocks_for(w, h);
        q.bytes = q.vol->bytes + volume_meta_bytes(c->volume_n);
        st = alloc((void**)&q.st, sizeof(kf::PoseState), &q.bytes);
    }
    if (!st) st = alloc((void**)&q.carry, 2 * sizeof(kf::IcpCarry), &q.bytes);
    if (!st) st = alloc((void**)&q.partials, 2 * sizeof(float) * kf::kPartialStride * (size_t)q.max_blocks, &q.bytes);
    if (!st) st = alloc((void**)&q.depth_m, sizeof(float) * (size_t)w * h, &q.bytes);
    if (!st) st = alloc((void**)&q.d0, sizeof(float) * (size_t)w * h, &q.bytes);
    if (!st) st = alloc((void**)&q.dmax, sizeof(float) * kf::depth_max_floats(w, h), &q.bytes);
    for (int l = 0; l < 3 && !st; ++l) {
        const size_t px = (size_t)(w >> l) * (h >> l);
        if ((st = alloc((void**)&q.v[l], sizeof(float4) * px, &q.bytes))) break;
        if ((st = alloc((void**)&q.pv[l], sizeof(float4) * px, &q.bytes))) break;
        if ((st = alloc((void**)&q.pn[l], sizeof(float4) * px, &q.bytes))) break;
        // predicted maps start invalid, as kf_pipeline's
        const unsigned blocks = (unsigned)((px + 255) / 256);
        fill_invalid_kernel<<<blocks, 256, 0, b->s>>>(q.pv[l], (int64_t)px);
        fill_invalid_kernel<<<blocks, 256, 0, b->s>>>(q.pn[l], (int64_t)px);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) st = kf::fail_hip(e, "kf_batch_add_session init maps");
    }
    if (!st) {
        Pose32 p;
        memcpy(p.m, sensor.pose, sizeof(sensor.pose));
        memcpy(p.m + 16, sensor.inv_pose, sizeof(sensor.inv_pose));
        pose_state_init_kernel<<<1, 64, 0, b->s>>>(q.st, q.carry, p);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) st = kf::fail_hip(e, "kf_batch_add_session state");
    }
    if (st) {
        free_session(q);
        return st;
    }
    q.used = true;
    b->sess[slot] = q;
    st = ensure_list(b);
    if (st) {
        free_session(b->sess[slot]);
        return st;
    }
    *id = slot;
    return KF_OK;
}

int kf_batch_remove_session(kf_batch* b, int id)
{
    Session* q = session_of(b, id);
    if (!q) return kf::fail(KF_ERR_ARG, "kf_batch_remove_session: unknown session");
    KF_HIP(hipStreamSynchronize(b->s));
    for (Slot& sl : b->slot) {
        int st = retire(b, sl);
        if (st) return st;
    }
    free_session(*q);
    return KF_OK;
}

int kf_batch_tick(kf_batch* b, int n, const int* ids, const uint16_t* const* raw)
{
    if (!b || n < 0 || (n > 0 && (!ids || !raw))) return kf::fail(KF_ERR_ARG, "kf_batch_tick: bad argument");
    if (n > b->cfg.max_sessions) return kf::fail(KF_ERR_ARG, "kf_batch_tick: more frames than sessions");
    std::vector<char> seen(b->sess.size(), 0);
    for (int k = 0; k < n; ++k) {
        if (!session_of(b, ids[k]) || !raw[k]) return kf::fail(KF_ERR_ARG, "kf_batch_tick: unknown session or null frame");
        if (seen[ids[k]]++) return kf::fail(KF_ERR_ARG, "kf_batch_tick: a session appears twice in one tick");
    }
    Slot& sl = b->slot[b->next];
    int st = retire(b, sl);  // the table of this slot is free once its last tick is done
    if (st) return st;
    if (b->tuning) finish_tuning(b);

    // ---- fill the table
    const size_t m = (size_t)b->cfg.max_sessions;
    char* base = static_cast<char*>(sl.host);
    auto* hd = reinterpret_cast<kf::BatchHeader*>(base);
    auto* pre = reinterpret_cast<kf::PreJob*>(base + b->off_pre);
    auto* ray = reinterpret_cast<kf::RaycastJob*>(base + b->off_ray);
    auto* icp = reinterpret_cast<kf::IcpJob*>(base + b->off_icp);
    auto* fuse = reinterpret_cast<kf::FuseJob*>(base + b->off_fuse);
    memset(hd, 0, sizeof(*hd));
    hd->fuse_count = b->fuse_count;
    hd->ray_next = b->ray_next;
    hd->ray_list = b->ray_list;
    hd->ray_list_cap = (unsigned)b->ray_cap;
    hd->fuse_list = b->fuse_list;
    hd->touched = b->touched;
    const kf_icp_config& ic = b->cfg.icp;
    const float cos_th = (float)cos((double)ic.angle_thresh);
    int ray_tiles = 0, pre_tiles = 0, pyr_tiles = 0;
    int icp_blocks[3] = {0, 0, 0};
    unsigned bricks = 0;
    int n_ray = 0, n_icp = 0;
    for (int k = 0; k < n; ++k) {
        Session& q = b->sess[ids[k]];
        const int w = q.cfg.frame_width, h = q.cfg.frame_height;
        kf::PreJob& P = pre[k];
        memset(&P, 0, sizeof(P));
        P.raw = raw[k];
        P.depth_m = q.depth_m;
        P.d0 = q.d0;
        P.dmax = q.dmax;
        for (int l = 0; l < 3; ++l) P.v[l] = q.v[l];
        P.w = w;
        P.h = h;
        P.scale = q.cfg.depth_scale;
        memcpy(P.Kinv, q.Ki, sizeof(P.Kinv));
        P.tile0 = pre_tiles;
        P.tiles_x = (w + 15) / 16;
        pre_tiles += P.tiles_x * ((h + 15) / 16);
        P.ptile0 = pyr_tiles;
        P.ptiles_x = (w / 4 + 15) / 16;
        pyr_tiles += P.ptiles_x * ((h / 4 + 15) / 16);
        if (q.frames > 0) {
            // tracking: raycast L0..L2 at the previous pose (jobs added below, FP32
            // volumes first), then ICP
            for (int l = 0; l < ic.levels; ++l) {
                kf::IcpJob& J = icp[(size_t)l * m + n_icp];
                memset(&J, 0, sizeof(J));
                J.a.V = q.v[l];
                J.a.TV = q.pv[l];
                J.a.TN = q.pn[l];
                J.a.w = w >> l;
                J.a.h = h >> l;
                memcpy(J.a.K, q.K[l], sizeof(J.a.K));
                J.a.dist_thresh = ic.distance_thresh;
                J.a.cos_thresh = cos_th;
                J.a.st = q.st;
                J.a.level = l;
                J.a.pix = kf::icp_pix_for((int64_t)(w >> l) * (h >> l));
                J.a.stride = kf::icp_level_blocks(w >> l, h >> l);
                J.carry = q.carry;
                J.partials = q.partials;
                J.max_blocks = q.max_blocks;
                J.block0 = icp_blocks[l];
                icp_blocks[l] += J.a.stride;
            }
            ++n_icp;
        }
        kf::FuseJob& F = fuse[k];
        memset(&F, 0, sizeof(F));
        F.a = kf::fuse_args(q.depth_m, w, h, q.vol, q.K[0], nullptr, q.st);
        F.a.dmax = q.dmax;
        F.brick0 = bricks;
        const unsigned nb = (unsigned)(q.cfg.volume_n / 8);
        bricks += nb * nb * nb;
    }
    // raycast jobs of the tracking sessions, FP32 volumes first (one launch per format)
    for (int pass = 0; pass < 2; ++pass) {
        const int fmt = pass == 0 ? KF_VOXEL_F32 : KF_VOXEL_SNORM16_U8;
        if (pass == 1) {
            hd->n_ray_f32 = n_ray;
            hd->ray_tiles_f32 = ray_tiles;
        }
        for (int k = 0; k < n; ++k) {
            Session& q = b->sess[ids[k]];
            if (q.frames == 0 || q.cfg.voxel_format != fmt) continue;
            const int w = q.cfg.frame_width, h = q.cfg.frame_height;
            for (int l = 0; l < ic.levels; ++l) {
                kf::RaycastJob& R = ray[n_ray++];
                memset(&R, 0, sizeof(R));
                R.vol = kf::raycast_volume(q.vol);
                R.pose = reinterpret_cast<const float*>(reinterpret_cast<const char*>(q.st) +
                                                        offsetof(kf::PoseState, pose));
                memcpy(R.L.Kinv, q.Ki[l], sizeof(R.L.Kinv));
                R.L.vout = q.pv[l];
                R.L.nout = q.pn[l];
                R.L.w = w >> l;
                R.L.h = h >> l;
                R.L.tiles_x = (R.L.w + 7) / 8;
                R.L.tile0 = ray_tiles;
                ray_tiles += R.L.tiles_x * ((R.L.h + 7) / 8);
            }
        }
    }
    hd->n_pre = n;
    hd->pre_tiles = pre_tiles;
    hd->pyr_tiles = pyr_tiles;
    hd->n_ray = n_ray;
    hd->ray_tiles = ray_tiles;
    hd->n_icp = n_icp;
    for (int l = 0; l < 3; ++l) hd->icp_blocks[l] = icp_blocks[l];
    hd->n_fuse = n;
    hd->fuse_bricks = bricks;

    // ---- launch
    if (b->tuning) {
        const int cand = b->tune_tick % kCandidates;
        kf_batch_shape sh = b->shape;
        sh.preprocess_blocks = kCand[0][cand];
        sh.raycast_blocks = kCand[1][cand];
        sh.icp_blocks = kCand[2][cand];
        sh.fuse_blocks = kCand[3][cand];
        st = enqueue_tick(b, sl, sh, sl.ev);
        sl.tuning = true;
        sl.cand = cand;
        sl.has_stage[0] = n > 0;
        sl.has_stage[1] = n_ray > 0;
        sl.has_stage[2] = n_icp > 0;
        sl.has_stage[3] = n > 0;
        b->tune_tick += 1;
    } else if (b->stage_timing) {
        st = enqueue_tick(b, sl, b->shape, sl.ev, true);
        sl.fine = true;
    } else if (b->cfg.use_graph) {
        if (!sl.exec) st = capture(b, sl);
        if (!st) {
            KF_HIP(hipEventRecord(sl.ev[0], b->s));
            hipError_t e = hipGraphLaunch(sl.exec, b->s);
            if (e != hipSuccess) st = kf::fail_hip(e, "batch tick graph launch");
            if (!st) KF_HIP(hipEventRecord(sl.ev[4], b->s));
        }
    } else {
        st = enqueue_tick(b, sl, b->shape, sl.ev);
    }
    if (st) return st;
    sl.pending = true;
    b->next ^= 1;
    for (int k = 0; k < n; ++k) b->sess[ids[k]].frames += 1;
    return KF_OK;
}

int kf_batch_set_stage_timing(kf_batch* b, int enabled)
{
    if (!b) return kf::fail(KF_ERR_ARG, "kf_batch_set_stage_timing: null batch");
    int st = kf_batch_sync(b);
    if (st) return st;
    b->stage_timing = enabled != 0;
    for (double& v : b->stage_ms) v = 0.0;
    b->stage_ticks = 0;
    return KF_OK;
}

int kf_batch_stage_times(kf_batch* b, double total_ms[7], int64_t* ticks)
{
    if (!b || !total_ms) return kf::fail(KF_ERR_ARG, "kf_batch_stage_times: null argument");
    int st = kf_batch_sync(b);
    if (st) return st;
    for (int k = 0; k < 7; ++k) total_ms[k] = b->stage_ms[k];
    if (ticks) *ticks = b->stage_ticks;
    return KF_OK;
}

int kf_batch_sync(kf_batch* b)
{
    if (!b) return kf::fail(KF_ERR_ARG, "kf_batch_sync: null batch");
    KF_HIP(hipStreamSynchronize(b->s));
    for (Slot& sl : b->slot) {
        int st = retire(b, sl);
        if (st) return st;
    }
    return KF_OK;
}

int kf_batch_pose(kf_batch* b, int id, float pose[16])
{
    Session* q = session_of(b, id);
    if (!q || !pose) return kf::fail(KF_ERR_ARG, "kf_batch_pose: bad argument");
    KF_HIP(hipStreamSynchronize(b->s));
    KF_HIP(hipMemcpy(pose, q->st->pose, sizeof(float) * 16, hipMemcpyDeviceToHost));
    return KF_OK;
}

int kf_batch_session_status(kf_batch* b, int id, int* status)
{
    Session* q = session_of(b, id);
    if (!q || !status) return kf::fail(KF_ERR_ARG, "kf_batch_session_status: bad argument");
    KF_HIP(hipStreamSynchronize(b->s));
    int sticky = 0;
    KF_HIP(hipMemcpy(&sticky, &q->st->sticky_status, sizeof(int), hipMemcpyDeviceToHost));
    if (sticky) {
        int zero = 0;
        KF_HIP(hipMemcpy(&q->st->sticky_status, &zero, sizeof(int), hipMemcpyHostToDevice));
    }
    *status = sticky;
    return KF_OK;
}

int kf_batch_volume(kf_batch* b, int id, kf_volume** v)
{
    Session* q = session_of(b, id);
    if (!q || !v) return kf::fail(KF_ERR_ARG, "kf_batch_volume: bad argument");
    *v = q->vol;
    return KF_OK;
}

int kf_batch_map(kf_batch* b, int id, int which, int level, void** dev_ptr, int* w, int* h)
{
    Session* q = session_of(b, id);
    if (!q || !dev_ptr || level < 0 || level > 2) return kf::fail(KF_ERR_ARG, "kf_batch_map: bad argument");
    void* ptr = nullptr;
    switch (which) {
        case 0: ptr = q->depth_m; level = 0; break;
        case 1: ptr = q->d0; level = 0; break;
        case 2: ptr = q->v[level]; break;
        case 4: ptr = q->pv[level]; break;
        case 5: ptr = q->pn[level]; break;
        default: return kf::fail(KF_ERR_ARG, "kf_batch_map: unknown map");
    }
    *dev_ptr = ptr;
    if (w) *w = q->cfg.frame_width >> level;
    if (h) *h = q->cfg.frame_height >> level;
    return KF_OK;
}

int kf_batch_shape_get(kf_batch* b, kf_batch_shape* shape, int* tuned)
{
    if (!b || !shape) return kf::fail(KF_ERR_ARG, "kf_batch_shape_get: null argument");
    *shape = b->shape;
    if (tuned) *tuned = (b->cfg.autotune && !b->tuning) ? 1 : 0;
    return KF_OK;
}

int kf_batch_tick_times(kf_batch* b, double* total_ms, int64_t* ticks, float* max_ms)
{
    if (!b) return kf::fail(KF_ERR_ARG, "kf_batch_tick_times: null batch");
    int st = kf_batch_sync(b);
    if (st) return st;
    if (total_ms) *total_ms = b->tick_ms_total;
    if (ticks) *ticks = b->ticks_timed;
    if (max_ms) *max_ms = b->tick_ms_max;
    b->tick_ms_total = 0.0;
    b->ticks_timed = 0;
    b->tick_ms_max = 0.0f;
    return KF_OK;
}

int kf_batch_session_save(kf_batch* b, int id, const char* path)
{
    Session* q = session_of(b, id);
    if (!q || !path) return kf::fail(KF_ERR_ARG, "kf_batch_session_save: bad argument");
    int st = kf_batch_sync(b);
    if (st) return st;
    float pose[32];  // pose, inverse pose (as the last fusion used it)
    KF_HIP(hipMemcpy(pose, q->st->pose, sizeof(float) * 16, hipMemcpyDeviceToHost));
    KF_HIP(hipMemcpy(pose + 16, q->st->inv_pose, sizeof(float) * 16, hipMemcpyDeviceToHost));
    if ((st = kf_volume_save(q->vol, path))) return st;
    const std::string pp = std::string(path) + ".pose";
    FILE* fp = fopen(pp.c_str(), "w");
    if (!fp) return kf::fail(KF_ERR_IO, "kf_batch_session_save: cannot open " + pp);
    // 9 significant digits round-trip a float exactly
    fprintf(fp, "frames %lld\npose", (long long)q->frames);
    for (int k = 0; k < 16; ++k) fprintf(fp, " %.9g", pose[k]);
    fprintf(fp, "\ninv_pose");
    for (int k = 16; k < 32; ++k) fprintf(fp, " %.9g", pose[k]);
    fprintf(fp, "\n");
    const bool ok = fclose(fp) == 0;
    return ok ? KF_OK : kf::fail(KF_ERR_IO, "kf_batch_session_save: write failed");
}

int kf_batch_session_restore(kf_batch* b, int id, const char* path)
{
    Session* q = session_of(b, id);
    if (!q || !path) return kf::fail(KF_ERR_ARG, "kf_batch_session_restore: bad argument");
    int st = kf_batch_sync(b);
    if (st) return st;
    const std::string pp = std::string(path) + ".pose";
    FILE* fp = fopen(pp.c_str(), "r");
    if (!fp) return kf::fail(KF_ERR_IO, "kf_batch_session_restore: cannot open " + pp);
    long long frames = 0;
    float pose[16], inv[16];
    int got = fscanf(fp, " frames %lld pose", &frames);
    for (int k = 0; k < 16 && got == 1; ++k) got = fscanf(fp, " %f", &pose[k]);
    bool have_inv = got == 1 && fscanf(fp, " inv_pose") == 0;
    for (int k = 0; k < 16 && have_inv; ++k) have_inv = fscanf(fp, " %f", &inv[k]) == 1;
    fclose(fp);
    if (got != 1 || frames < 0) return kf::fail(KF_ERR_IO, "kf_batch_session_restore: malformed " + pp);
    if ((st = kf_volume_load(q->vol, path))) return st;
    Pose32 p;
    memcpy(p.m, pose, sizeof(pose));
    if (have_inv) {
        memcpy(p.m + 16, inv, sizeof(inv));
    } else {
        kf_sensor sensor;  // Sensor::setPose's inverse (sensor.h:30-34)
        if ((st = kf_sensor_init(&sensor, q->cfg.fx, q->cfg.fy, q->cfg.cx, q->cfg.cy))) return st;
        if ((st = kf_sensor_set_pose(&sensor, pose))) return st;
        memcpy(p.m + 16, sensor.inv_pose, sizeof(sensor.inv_pose));
    }
    pose_state_init_kernel<<<1, 64, 0, b->s>>>(q->st, q->carry, p);
    KF_LAUNCHED("pose_state_init_kernel");
    KF_HIP(hipStreamSynchronize(b->s));
    q->frames = frames;
    return KF_OK;
}

int kf_batch_touched_total(kf_batch* b, unsigned long long* voxels)
{
    if (!b || !voxels) return kf::fail(KF_ERR_ARG, "kf_batch_touched_total: null argument");
    int st = kf_batch_sync(b);
    if (st) return st;
    std::vector<unsigned long long> h(kMaxFuseBlocks);
    KF_HIP(hipMemcpy(h.data(), b->touched, sizeof(unsigned long long) * kMaxFuseBlocks, hipMemcpyDeviceToHost));
    unsigned long long s = 0;
    for (unsigned long long x : h) s += x;
    *voxels = s;
    return KF_OK;
}

}  // extern "C"
