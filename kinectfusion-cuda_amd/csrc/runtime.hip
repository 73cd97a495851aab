// Device/memory plumbing, error reporting, call timing and the Sensor value
// type of the C ABI (include/kfusion.h).
#include <math.h>
#include <string.h>

#include <string>

#include "kf_internal.hpp"

namespace kf {

static thread_local std::string g_last_error;

int fail(int status, const std::string& msg)
{
    g_last_error = msg;
    return status;
}
int fail_hip(hipError_t e, const char* what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? KF_ERR_OOM : KF_ERR_HIP;
}
void clear_error() { g_last_error.clear(); }

CallTimer::CallTimer(float* elapsed_ms, hipStream_t s) : out_(elapsed_ms), s_(s) {}
CallTimer::~CallTimer()
{
    if (a_) (void)hipEventDestroy(a_);
    if (b_) (void)hipEventDestroy(b_);
}
int CallTimer::start()
{
    if (!out_) return KF_OK;
    KF_HIP(hipEventCreate(&a_));
    KF_HIP(hipEventCreate(&b_));
    KF_HIP(hipEventRecord(a_, s_));
    return KF_OK;
}
int CallTimer::finish()
{
    if (!out_) return KF_OK;
    KF_HIP(hipEventRecord(b_, s_));
    KF_HIP(hipEventSynchronize(b_));
    float ms = 0.0f;
    KF_HIP(hipEventElapsedTime(&ms, a_, b_));
    *out_ = ms;
    return KF_OK;
}

}  // namespace kf

extern "C" {

const char* kf_status_string(int status)
{
    switch (status) {
        case KF_OK: return "ok";
        case KF_ERR_DIMS: return "dimension mismatch";
        case KF_ERR_NONFINITE: return "an element in the result vector is not finite or NAN";
        case KF_ERR_NOT_PD: return "normal equations are not positive definite";
        case KF_ERR_OOM: return "out of device memory";
        case KF_ERR_HIP: return "HIP runtime error";
        case KF_ERR_BUDGET: return "HBM budget exhausted";
        case KF_ERR_ARG: return "invalid argument";
        case KF_ERR_IO: return "file I/O error";
        default: return "unknown status";
    }
}

const char* kf_last_error(void) { return kf::g_last_error.c_str(); }

int kf_device_init(int device)
{
    kf::clear_error();
    KF_HIP(hipSetDevice(device));
    KF_HIP(hipFree(nullptr));  // creates the context
    return KF_OK;
}
int kf_device_synchronize(void)
{
    KF_HIP(hipDeviceSynchronize());
    return KF_OK;
}
int kf_malloc(void** dptr, size_t bytes)
{
    if (!dptr) return kf::fail(KF_ERR_ARG, "kf_malloc: null out pointer");
    *dptr = nullptr;
    if (bytes == 0) return KF_OK;
    KF_HIP(hipMalloc(dptr, bytes));
    return KF_OK;
}
int kf_free(void* dptr)
{
    if (dptr) KF_HIP(hipFree(dptr));
    return KF_OK;
}
int kf_memcpy_h2d(void* dst, const void* src, size_t bytes, kf_stream_t stream)
{
    if (bytes == 0) return KF_OK;
    KF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, kf::stream_of(stream)));
    KF_HIP(hipStreamSynchronize(kf::stream_of(stream)));
    return KF_OK;
}
int kf_memcpy_d2h(void* dst, const void* src, size_t bytes, kf_stream_t stream)
{
    if (bytes == 0) return KF_OK;
    KF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, kf::stream_of(stream)));
    KF_HIP(hipStreamSynchronize(kf::stream_of(stream)));
    return KF_OK;
}
int kf_memcpy_d2d(void* dst, const void* src, size_t bytes, kf_stream_t stream)
{
    if (bytes == 0) return KF_OK;
    KF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, kf::stream_of(stream)));
    return KF_OK;
}
int kf_stream_create(kf_stream_t* stream)
{
    hipStream_t s = nullptr;
    KF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return KF_OK;
}
int kf_stream_destroy(kf_stream_t stream)
{
    if (stream) KF_HIP(hipStreamDestroy(kf::stream_of(stream)));
    return KF_OK;
}
int kf_stream_synchronize(kf_stream_t stream)
{
    KF_HIP(hipStreamSynchronize(kf::stream_of(stream)));
    return KF_OK;
}
int kf_event_create(kf_event_t* ev)
{
    hipEvent_t e = nullptr;
    KF_HIP(hipEventCreate(&e));
    *ev = e;
    return KF_OK;
}
int kf_event_destroy(kf_event_t ev)
{
    if (ev) KF_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)));
    return KF_OK;
}
int kf_event_record(kf_event_t ev, kf_stream_t stream)
{
    KF_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), kf::stream_of(stream)));
    return KF_OK;
}
int kf_event_elapsed_ms(kf_event_t start, kf_event_t end, float* ms)
{
    KF_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(end)));
    KF_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(end)));
    return KF_OK;
}

// ---------------------------------------------------------------- sensor
// glm::inverse of the 3x3 intrinsics in its cofactor order (sensor.h:17-19).
static void intrinsics_level(float fx, float fy, float cx, float cy, int level, float* K, float* Ki)
{
    float s = 1.0f;
    for (int l = 0; l < level; ++l) s = s * 0.5f;
    float fxl = fx * s, fyl = fy * s;
    float cxl = (cx + 0.5f) * s - 0.5f, cyl = (cy + 0.5f) * s - 0.5f;
    float m[9] = {fxl, 0, 0, 0, fyl, 0, cxl, cyl, 1};
    memcpy(K, m, sizeof(m));
#define M(c, r) m[(c) * 3 + (r)]
    float det = (+M(0, 0) * (M(1, 1) * M(2, 2) - M(2, 1) * M(1, 2))
                 - M(1, 0) * (M(0, 1) * M(2, 2) - M(2, 1) * M(0, 2)))
                + M(2, 0) * (M(0, 1) * M(1, 2) - M(1, 1) * M(0, 2));
    float ood = 1.0f / det;
    Ki[0 * 3 + 0] = +(M(1, 1) * M(2, 2) - M(2, 1) * M(1, 2)) * ood;
    Ki[1 * 3 + 0] = -(M(1, 0) * M(2, 2) - M(2, 0) * M(1, 2)) * ood;
    Ki[2 * 3 + 0] = +(M(1, 0) * M(2, 1) - M(2, 0) * M(1, 1)) * ood;
    Ki[0 * 3 + 1] = -(M(0, 1) * M(2, 2) - M(2, 1) * M(0, 2)) * ood;
    Ki[1 * 3 + 1] = +(M(0, 0) * M(2, 2) - M(2, 0) * M(0, 2)) * ood;
    Ki[2 * 3 + 1] = -(M(0, 0) * M(2, 1) - M(2, 0) * M(0, 1)) * ood;
    Ki[0 * 3 + 2] = +(M(0, 1) * M(1, 2) - M(1, 1) * M(0, 2)) * ood;
    Ki[1 * 3 + 2] = -(M(0, 0) * M(1, 2) - M(1, 0) * M(0, 2)) * ood;
    Ki[2 * 3 + 2] = +(M(0, 0) * M(1, 1) - M(1, 0) * M(0, 1)) * ood;
#undef M
}

int kf_sensor_init(kf_sensor* s, float fx, float fy, float cx, float cy)
{
    if (!s) return kf::fail(KF_ERR_ARG, "kf_sensor_init: null sensor");
    memset(s, 0, sizeof(*s));
    for (int i = 0; i < 4; ++i) s->pose[i * 5] = s->inv_pose[i * 5] = 1.0f;
    for (int l = 0; l < 3; ++l) intrinsics_level(fx, fy, cx, cy, l, s->intr[l], s->inv_intr[l]);
    return KF_OK;
}

int kf_sensor_set_pose(kf_sensor* s, const float pose[16])
{
    if (!s || !pose) return kf::fail(KF_ERR_ARG, "kf_sensor_set_pose: null argument");
    memcpy(s->pose, pose, sizeof(s->pose));
    // [R^T | -R^T t], in double and rounded (same arithmetic as the device commit)
    memset(s->inv_pose, 0, sizeof(s->inv_pose));
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) s->inv_pose[c * 4 + r] = pose[r * 4 + c];
    for (int r = 0; r < 3; ++r) {
        double acc = 0;
        for (int q = 0; q < 3; ++q) acc += (double)pose[r * 4 + q] * (double)pose[12 + q];
        s->inv_pose[12 + r] = (float)(-acc);
    }
    s->inv_pose[15] = 1.0f;
    return KF_OK;
}

}  // extern "C"
