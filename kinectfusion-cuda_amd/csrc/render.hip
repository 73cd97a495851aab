// Headless visualisation of raycast maps (project/src/display.cu:64-105): the
// shading image and the normal image the reference draws into its window, as
// RGBA8 pixels (r in the low byte, alpha 255: the reference's packed
// ((255 << 8 | b) << 8 | g) << 8 | r).  A channel value v in [0, 1] becomes
// the byte trunc(v * 255); values below 0 (and NaN) become 0, above 1 become
// 255 (the reference's float -> unsigned char conversion, defined here for
// every input).
#include "kf_device.hpp"
#include "kf_kernels.hpp"

namespace kf {

namespace {

struct Mat3Arg {
    float m[9];
};

__device__ __forceinline__ uint32_t to_byte(float v)
{
    const float s = v * 255.0f;
    if (!(s > 0.0f)) return 0u;
    if (s >= 255.0f) return 255u;
    return (uint32_t)s;
}

__device__ __forceinline__ uint32_t pack(float r, float g, float b)
{
    return (255u << 24) | (to_byte(b) << 16) | (to_byte(g) << 8) | to_byte(r);
}

// shadingToWindowContentKernel (display.cu:85-105): radiance = min(|n . d|, 1)
// with d = normalize((R K0^-1) (i + 1/2, j + 1/2, 1)).
__global__ __launch_bounds__(256) void shade_kernel(const float4* __restrict__ nmap, int w, int h, Mat3Arg rk,
                                                    uint32_t* __restrict__ out)
{
    const int i = blockIdx.x * 16 + threadIdx.x, j = blockIdx.y * 16 + threadIdx.y;
    if (i >= w || j >= h) return;
    const float4 nv = nmap[(int64_t)j * w + i];
    const v3 d = normalize3(m3v(rk.m, mk3((float)i + 0.5f, (float)j + 0.5f, 1.0f)));
    float rad = fabsf(dot3(mk3(nv.x, nv.y, nv.z), d));
    rad = rad < 1.0f ? rad : 1.0f;
    out[(int64_t)j * w + i] = pack(rad, rad, rad);
}

// normalMapToWindowContentKernel (display.cu:64-83): n' = normalize(R^-1 n) as (r, g, b).
__global__ __launch_bounds__(256) void normal_image_kernel(const float4* __restrict__ nmap, int w, int h,
                                                           Mat3Arg rinv, uint32_t* __restrict__ out)
{
    const int i = blockIdx.x * 16 + threadIdx.x, j = blockIdx.y * 16 + threadIdx.y;
    if (i >= w || j >= h) return;
    const float4 nv = nmap[(int64_t)j * w + i];
    const v3 n = normalize3(m3v(rinv.m, mk3(nv.x, nv.y, nv.z)));
    out[(int64_t)j * w + i] = pack(n.x, n.y, n.z);
}

}  // namespace

int launch_render(const float4* nmap, int w, int h, const float* mat9, int shading, uint32_t* out, hipStream_t s)
{
    Mat3Arg m;
    for (int q = 0; q < 9; ++q) m.m[q] = mat9[q];
    const dim3 grid((unsigned)((w + 15) / 16), (unsigned)((h + 15) / 16)), block(16, 16);
    if (shading)
        shade_kernel<<<grid, block, 0, s>>>(nmap, w, h, m, out);
    else
        normal_image_kernel<<<grid, block, 0, s>>>(nmap, w, h, m, out);
    KF_LAUNCHED("render kernel");
    return KF_OK;
}

// glm::mat3(pose) * inv_intr (the order display.cu:95 evaluates), column-major.
void shading_matrix(const float* pose16, const float* kinv9, float* out9)
{
    const float R[9] = {pose16[0], pose16[1], pose16[2], pose16[4], pose16[5], pose16[6], pose16[8], pose16[9], pose16[10]};
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            out9[c * 3 + r] = (R[0 * 3 + r] * kinv9[c * 3 + 0] + R[1 * 3 + r] * kinv9[c * 3 + 1]) + R[2 * 3 + r] * kinv9[c * 3 + 2];
}

}  // namespace kf

extern "C" {

int kf_render_shading(const float* normal_map, int w, int h, const kf_sensor* sensor, uint32_t* rgba,
                      kf_stream_t stream, float* elapsed_ms)
{
    if (!normal_map || !sensor || !rgba || w <= 0 || h <= 0)
        return kf::fail(KF_ERR_ARG, "shadingToWindowContent: bad argument");
    float rk[9];
    kf::shading_matrix(sensor->pose, sensor->inv_intr[0], rk);
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    int st;
    if ((st = t.start()) || (st = kf::launch_render(reinterpret_cast<const float4*>(normal_map), w, h, rk, 1, rgba, s)))
        return st;
    return t.finish();
}

int kf_render_normals(const float* normal_map, int w, int h, const float inverse_rotation[9], uint32_t* rgba,
                      kf_stream_t stream, float* elapsed_ms)
{
    if (!normal_map || !inverse_rotation || !rgba || w <= 0 || h <= 0)
        return kf::fail(KF_ERR_ARG, "normalMapToWindowContent: bad argument");
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    int st;
    if ((st = t.start()) ||
        (st = kf::launch_render(reinterpret_cast<const float4*>(normal_map), w, h, inverse_rotation, 0, rgba, s)))
        return st;
    return t.finish();
}

}  // extern "C"
