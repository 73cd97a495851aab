// Host-side internals of libkfusion: error plumbing, timing helper and the
// layouts of the opaque handles declared in include/kfusion.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "kfusion.h"

namespace kf {

// Records a detail message for kf_last_error() and returns `status`.
int fail(int status, const std::string& msg);
// Device bytes a session can add to a batch's shared lists (its bricks in the
// fuse list, its rays in the deferred-ray list of each voxel format), counted
// twice because both lists grow geometrically (batch.hip).
size_t batch_shared_bytes(const kf_session_config* c);
int fail_hip(hipError_t e, const char* what);
void clear_error();

#define KF_HIP(call)                                             \
    do {                                                         \
        hipError_t kf_e_ = (call);                               \
        if (kf_e_ != hipSuccess) return ::kf::fail_hip(kf_e_, #call); \
    } while (0)

// Launch-error check right after a kernel launch.
#define KF_LAUNCHED(name)                                              \
    do {                                                               \
        hipError_t kf_e_ = hipGetLastError();                          \
        if (kf_e_ != hipSuccess) return ::kf::fail_hip(kf_e_, name);   \
    } while (0)

inline hipStream_t stream_of(kf_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// The reference's CudaEvent pair around one wrapper call (cuda_event.cpp:14-30):
// active only when the caller asked for elapsed milliseconds.
class CallTimer {
public:
    CallTimer(float* elapsed_ms, hipStream_t s);
    ~CallTimer();
    int start();
    int finish();  // records the end event, waits, stores the elapsed time

private:
    float* out_;
    hipStream_t s_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

// Device-resident ICP / pose state shared by the ICP loop, the raycast and the
// fuse of a frame, so a whole frame runs without host round-trips.
struct alignas(256) PoseState {
    float R[9];          // current estimate (camera -> world rotation)
    float t[3];
    float Rp[9];         // pose at the start of the ICP (previous frame)
    float tp[3];
    float pose[16];      // committed sensor pose (camera -> world)
    float inv_pose[16];  // its inverse (world -> camera)
    int status;          // first ICP failure of the current frame (kf_status), 0 if none
    int sticky_status;   // first failure since the host last read it
    int level_done[4];   // early stop reached at that level
    int iterations;      // iterations executed in the current frame
    int frame_skip_fuse; // 1 when the current frame's ICP failed
    double last_sums[27];
    long long last_inliers;
    float last_x[6];
};

}  // namespace kf

struct kf_volume {
    int n = 0, nb = 0, format = KF_VOXEL_F32;
    float width = 0, res = 0, mue = 0;
    void* f = nullptr;        // float* or int16_t* plane
    void* w = nullptr;        // float* or uint8_t* plane
    float* zk = nullptr;      // z coordinate of each k slice (repeated addition, tsdf.cu:27-31)
    float4* zt = nullptr;     // per fuse: the inverse pose's z terms of each k slice (FuseArgs::zt)
    unsigned long long* touched = nullptr;  // per-workgroup counts of fused voxels (summed on read)
    int* brick_key = nullptr;               // per-brick halo lower bound of F, order_key form (raycast skipping)
    int* super_key = nullptr;               // per-super-brick lower bound of F (order_key)
    int ns = 0;                             // super-bricks per axis (ceil(nb / 8))
    int* brick_list = nullptr;              // visible bricks of the current fuse
    unsigned int* list_count = nullptr;     // list lengths, double-buffered by launch parity
    float* dmax = nullptr;                  // max valid depth per 16x16 and per 64x64 pixel tile
    size_t dmax_cap = 0;                    // floats allocated for dmax
    int64_t fuse_launches = 0;
    size_t bytes = 0;
    // raycast scratch (grown on demand, also by raycasts through a const volume):
    // deferred-ray list and its two launch-parity counters
    mutable void* ray_list = nullptr;
    mutable unsigned* ray_count = nullptr;
    mutable size_t ray_cap = 0;
    mutable int64_t ray_launches = 0;
    // first-pass dispatch order (tiles that took longest last time first):
    // 2 launch parities x order_cap slots, and its counters and wave times
    mutable int* tile_order = nullptr;
    mutable void* order_state = nullptr;
    mutable size_t order_cap = 0;
    mutable int64_t order_key = -1;  // (levels, width, height) the stored order was made for
};

struct kf_icp {
    kf_icp_config cfg{};
    float cos_thresh = 1.0f;
    kf::PoseState* state = nullptr;  // device
    void* carry = nullptr;           // device, 2 x IcpCarry (double-buffered by launch parity)
    float* partials = nullptr;       // device, 2 x [kPartialStride][max_blocks]
    int max_blocks = 0;
    float times[2] = {0, 0};
    int last_iterations = 0;
};
