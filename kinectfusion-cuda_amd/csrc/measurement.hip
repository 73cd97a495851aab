// Measurement step of the per-frame path (project/src/measurement.cu): raw
// depth codes -> metres, bilateral filter, depth pyramid, vertex and normal maps.
//
// Maps are plain row-major arrays in HBM (float depth, float4 vertex/normal with
// validity in .w).  Every kernel is one thread per output pixel over 16x16
// tiles (4 waves of 64), with bounds checks, so any frame size works (the
// reference launches w/8 x h/8 blocks and skips the remainder).
#include "kf_device.hpp"
#include "kf_kernels.hpp"

namespace kf {

namespace {

constexpr int kTile = 16;

inline dim3 grid2d(int w, int h) { return dim3((w + kTile - 1) / kTile, (h + kTile - 1) / kTile); }

// measurement.cu:14-31
__global__ __launch_bounds__(256) void convert_kernel(const uint16_t* __restrict__ raw, float* __restrict__ out,
                                                      int w, int h, float scale, int unmirror)
{
    int i = blockIdx.x * kTile + threadIdx.x, j = blockIdx.y * kTile + threadIdx.y;
    if (i >= w || j >= h) return;
    float d = half_rne(raw[(int64_t)j * w + i]);
    int io = unmirror ? (w - 1 - i) : i;
    out[(int64_t)j * w + io] = d * scale;
}

// measurement.cu:33-66: 7x7, x outer / y inner, clamp-to-edge, sigma_s = 4, sigma_r = 0.25.
// The 22x22 input window of the tile is staged in LDS once.
__global__ __launch_bounds__(256) void bilateral_kernel(const float* __restrict__ in, float* __restrict__ out, int w,
                                                        int h)
{
    __shared__ float tile[kTile + 6][kTile + 6 + 1];
    const int i0 = blockIdx.x * kTile - 3, j0 = blockIdx.y * kTile - 3;
    const int t = threadIdx.y * kTile + threadIdx.x;
    for (int q = t; q < (kTile + 6) * (kTile + 6); q += kTile * kTile) {
        int ly = q / (kTile + 6), lx = q % (kTile + 6);
        int gi = clampi(i0 + lx, 0, w - 1), gj = clampi(j0 + ly, 0, h - 1);
        tile[ly][lx] = in[(int64_t)gj * w + gi];
    }
    __syncthreads();
    int i = blockIdx.x * kTile + threadIdx.x, j = blockIdx.y * kTile + threadIdx.y;
    if (i >= w || j >= h) return;
    const float inv_s = 1.0f / (4.0f * 4.0f);
    const float inv_r = 1.0f / (0.25f * 0.25f);
    // The LDS window was filled with clamp-to-edge reads, so tap (x, y) is
    // simply the window element at the thread's offset.
    float center = tile[threadIdx.y + 3][threadIdx.x + 3];
    float norm = 0.0f, acc = 0.0f;
#pragma unroll
    for (int x = -3; x <= 3; ++x) {
#pragma unroll
        for (int y = -3; y <= 3; ++y) {
            float cur = tile[threadIdx.y + 3 + y][threadIdx.x + 3 + x];
            float sd = (float)(x * x + y * y);
            float id = center - cur;
            id = id * id;
            float f = kf_expf(-sd * inv_s - id * inv_r);
            norm = norm + f;
            acc = acc + f * cur;
        }
    }
    out[(int64_t)j * w + i] = acc / norm;
}

// measurement.cu:68-102
__global__ __launch_bounds__(256) void down_sample_kernel(const float* __restrict__ in, int w, float* __restrict__ out,
                                                          int wo, int ho)
{
    int i = blockIdx.x * kTile + threadIdx.x, j = blockIdx.y * kTile + threadIdx.y;
    if (i >= wo || j >= ho) return;
    const float tsd = 3.0f * 0.25f;
    float f1 = in[(int64_t)(2 * j) * w + 2 * i], f2 = in[(int64_t)(2 * j) * w + 2 * i + 1];
    float f3 = in[(int64_t)(2 * j + 1) * w + 2 * i], f4 = in[(int64_t)(2 * j + 1) * w + 2 * i + 1];
    float acc = f1;
    int cnt = 1;
    if (fabsf(f1 - f2) <= tsd) { acc = acc + f2; ++cnt; }
    if (fabsf(f1 - f3) <= tsd) { acc = acc + f3; ++cnt; }
    if (fabsf(f1 - f4) <= tsd) { acc = acc + f4; ++cnt; }
    out[(int64_t)j * wo + i] = acc / (float)cnt;
}

struct Mat3 {
    float m[9];
};

// measurement.cu:104-126
__global__ __launch_bounds__(256) void vertex_kernel(const float* __restrict__ depth, float4* __restrict__ out, int w,
                                                     int h, Mat3 ki)
{
    int i = blockIdx.x * kTile + threadIdx.x, j = blockIdx.y * kTile + threadIdx.y;
    if (i >= w || j >= h) return;
    float d = depth[(int64_t)j * w + i];
    float4 o = make_float4(0.0f, 0.0f, 0.0f, kInvalid);
    if (is_depth_valid(d)) {
        v3 p = m3v(ki.m, mk3((float)i + 0.5f, (float)j + 0.5f, 1.0f));
        p = mul3s(p, d);
        o = make_float4(p.x, p.y, p.z, kValid);
    }
    out[(int64_t)j * w + i] = o;
}

// measurement.cu:128-135 with device_helper.cuh:81-110
__global__ __launch_bounds__(256) void normal_kernel(const float4* __restrict__ v, float4* __restrict__ out, int w,
                                                     int h)
{
    int i = blockIdx.x * kTile + threadIdx.x, j = blockIdx.y * kTile + threadIdx.y;
    if (i >= w || j >= h) return;
    out[(int64_t)j * w + i] = compute_normal(v, w, h, i, j);
}

}  // namespace

int launch_convert(const uint16_t* raw, float* out, int w, int h, float scale, int unmirror, hipStream_t s)
{
    convert_kernel<<<grid2d(w, h), dim3(kTile, kTile), 0, s>>>(raw, out, w, h, scale, unmirror);
    KF_LAUNCHED("convert_kernel");
    return KF_OK;
}
int launch_bilateral(const float* in, float* out, int w, int h, hipStream_t s)
{
    bilateral_kernel<<<grid2d(w, h), dim3(kTile, kTile), 0, s>>>(in, out, w, h);
    KF_LAUNCHED("bilateral_kernel");
    return KF_OK;
}
int launch_down_sample(const float* in, int w_in, int h_in, float* out, hipStream_t s)
{
    int wo = w_in / 2, ho = h_in / 2;
    down_sample_kernel<<<grid2d(wo, ho), dim3(kTile, kTile), 0, s>>>(in, w_in, out, wo, ho);
    KF_LAUNCHED("down_sample_kernel");
    return KF_OK;
}
int launch_vertex_map(const float* depth, float4* out, int w, int h, const float* inv_k, hipStream_t s)
{
    Mat3 ki;
    for (int q = 0; q < 9; ++q) ki.m[q] = inv_k[q];
    vertex_kernel<<<grid2d(w, h), dim3(kTile, kTile), 0, s>>>(depth, out, w, h, ki);
    KF_LAUNCHED("vertex_kernel");
    return KF_OK;
}
int launch_normal_map(const float4* v, float4* out, int w, int h, hipStream_t s)
{
    normal_kernel<<<grid2d(w, h), dim3(kTile, kTile), 0, s>>>(v, out, w, h);
    KF_LAUNCHED("normal_kernel");
    return KF_OK;
}

int launch_preprocess(const uint16_t* raw, int w, int h, float scale, int unmirror, int levels,
                      const float (*inv_k)[9], float* depth_m, float* const* dpyr, float4* const* vpyr,
                      float4* const* npyr, hipStream_t s)
{
    int st;
    if ((st = launch_convert(raw, depth_m, w, h, scale, unmirror, s))) return st;
    if ((st = launch_bilateral(depth_m, dpyr[0], w, h, s))) return st;
    if ((st = launch_vertex_map(dpyr[0], vpyr[0], w, h, inv_k[0], s))) return st;
    for (int l = 1; l < levels; ++l) {
        if ((st = launch_down_sample(dpyr[l - 1], w >> (l - 1), h >> (l - 1), dpyr[l], s))) return st;
        if ((st = launch_vertex_map(dpyr[l], vpyr[l], w >> l, h >> l, inv_k[l], s))) return st;
    }
    if (npyr)
        for (int l = 0; l < levels; ++l)
            if ((st = launch_normal_map(vpyr[l], npyr[l], w >> l, h >> l, s))) return st;
    return KF_OK;
}

}  // namespace kf

// ------------------------------------------------------------------ C ABI
extern "C" {

int kf_convert_to_depth_meters(const uint16_t* raw, float* depth_m, int w, int h, float scale, int unmirror,
                               kf_stream_t stream, float* elapsed_ms)
{
    if (!raw || !depth_m || w <= 0 || h <= 0) return kf::fail(KF_ERR_ARG, "convertToDepthMeters: bad argument");
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    int st;
    if ((st = t.start()) || (st = kf::launch_convert(raw, depth_m, w, h, scale, unmirror, s))) return st;
    return t.finish();
}

int kf_apply_bilateral_filter(const float* depth_m, float* filtered, int w, int h, kf_stream_t stream,
                              float* elapsed_ms)
{
    if (!depth_m || !filtered || w <= 0 || h <= 0) return kf::fail(KF_ERR_ARG, "applyBilateralFilter: bad argument");
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    int st;
    if ((st = t.start()) || (st = kf::launch_bilateral(depth_m, filtered, w, h, s))) return st;
    return t.finish();
}

int kf_down_sample(const float* in, int w_in, int h_in, float* out, int w_out, int h_out, kf_stream_t stream,
                   float* elapsed_ms)
{
    if (!in || !out || w_out <= 0 || h_out <= 0) return kf::fail(KF_ERR_ARG, "downSample: bad argument");
    // measurement.cu:186-189
    if (w_in / w_out != 2 || h_in / h_out != 2 || w_in != 2 * w_out || h_in != 2 * h_out)
        return kf::fail(KF_ERR_DIMS, "downSample: output has to be half size of the input surface object!");
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    int st;
    if ((st = t.start()) || (st = kf::launch_down_sample(in, w_in, h_in, out, s))) return st;
    return t.finish();
}

int kf_create_vertex_map(const float* depth, float* vertex_map, int w, int h, const float inv_k[9],
                         kf_stream_t stream, float* elapsed_ms)
{
    if (!depth || !vertex_map || !inv_k || w <= 0 || h <= 0) return kf::fail(KF_ERR_ARG, "createVertexMap: bad argument");
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    int st;
    if ((st = t.start()) ||
        (st = kf::launch_vertex_map(depth, reinterpret_cast<float4*>(vertex_map), w, h, inv_k, s)))
        return st;
    return t.finish();
}

int kf_create_normal_map(const float* vertex_map, float* normal_map, int w, int h, kf_stream_t stream,
                         float* elapsed_ms)
{
    if (!vertex_map || !normal_map || w <= 0 || h <= 0) return kf::fail(KF_ERR_ARG, "createNormalMap: bad argument");
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    int st;
    if ((st = t.start()) || (st = kf::launch_normal_map(reinterpret_cast<const float4*>(vertex_map),
                                                        reinterpret_cast<float4*>(normal_map), w, h, s)))
        return st;
    return t.finish();
}

int kf_preprocess(const uint16_t* raw, int w, int h, float scale, int unmirror, int levels, const float inv_k[][9],
                  float* depth_m, float* const depth_pyr[], float* const vertex_pyr[], float* const normal_pyr[],
                  kf_stream_t stream, float* elapsed_ms)
{
    if (!raw || !depth_m || !depth_pyr || !vertex_pyr || !inv_k || levels < 1 || levels > 3 || w <= 0 || h <= 0)
        return kf::fail(KF_ERR_ARG, "kf_preprocess: bad argument");
    // grid_map_pyramid.h:63-72
    int e = 1 << (levels - 1);
    if (w % e || h % e)
        return kf::fail(KF_ERR_DIMS, "Can't construct pyramid. Dimensions must be divisible by 2^(m_n_layers - 1).");
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    float4* vp[3] = {nullptr, nullptr, nullptr};
    float4* np[3] = {nullptr, nullptr, nullptr};
    for (int l = 0; l < levels; ++l) {
        vp[l] = reinterpret_cast<float4*>(vertex_pyr[l]);
        if (normal_pyr) np[l] = reinterpret_cast<float4*>(normal_pyr[l]);
    }
    int st;
    if ((st = t.start()) || (st = kf::launch_preprocess(raw, w, h, scale, unmirror, levels, inv_k, depth_m, depth_pyr,
                                                        vp, normal_pyr ? np : nullptr, s)))
        return st;
    return t.finish();
}

}  // extern "C"
