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
// The 22x22 input window of the tile is staged in LDS once, filled with
// clamp-to-edge reads, so tap (x, y) is simply the window element at the
// thread's offset.
typedef float BilateralTile[kTile + 6][kTile + 6 + 1];

// The weight's argument is -sd / sigma_s^2 - id / sigma_r^2 with
// 1 / sigma_s^2 = 1/16 and 1 / sigma_r^2 = 16: both products are exact (powers
// of two; 16 id overflows exactly when the product does), so it is one
// rounding of c - 16 id with c = -(x^2 + y^2) / 16 exact, i.e. one fused
// multiply-add on a per-tap constant.
__device__ __forceinline__ float bilateral_at(const BilateralTile& tile, int tx, int ty)
{
    float center = tile[ty + 3][tx + 3];
    float norm = 0.0f, acc = 0.0f;
    // the outer loop stays rolled: unrolled, all 49 taps' tile reads were held
    // at once (106 VGPRs, 4 waves per SIMD); rolled, 27 VGPRs and 8 waves
#pragma unroll 1
    for (int x = -3; x <= 3; ++x) {
        const float cx = (float)(x * x) * -0.0625f;
#pragma unroll
        for (int y = -3; y <= 3; ++y) {
            float cur = tile[ty + 3 + y][tx + 3 + x];
            float id = center - cur;
            id = id * id;
            float f = kf_expf_nonpos(fmaf(id, -16.0f, cx + (float)(y * y) * -0.0625f));
            norm = norm + f;
            acc = acc + f * cur;
        }
    }
    return acc / norm;
}

__global__ __launch_bounds__(256) void bilateral_kernel(const float* __restrict__ in, float* __restrict__ out, int w,
                                                        int h)
{
    __shared__ BilateralTile tile;
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
    out[(int64_t)j * w + i] = bilateral_at(tile, threadIdx.x, threadIdx.y);
}

// measurement.cu:68-102: the taps within 3 sigma_r of the top-left tap, averaged.
__device__ __forceinline__ float down4(float f1, float f2, float f3, float f4)
{
    const float tsd = 3.0f * 0.25f;
    float acc = f1;
    int cnt = 1;
    if (fabsf(f1 - f2) <= tsd) { acc = acc + f2; ++cnt; }
    if (fabsf(f1 - f3) <= tsd) { acc = acc + f3; ++cnt; }
    if (fabsf(f1 - f4) <= tsd) { acc = acc + f4; ++cnt; }
    return acc / (float)cnt;
}

__global__ __launch_bounds__(256) void down_sample_kernel(const float* __restrict__ in, int w, float* __restrict__ out,
                                                          int wo, int ho)
{
    int i = blockIdx.x * kTile + threadIdx.x, j = blockIdx.y * kTile + threadIdx.y;
    if (i >= wo || j >= ho) return;
    float f1 = in[(int64_t)(2 * j) * w + 2 * i], f2 = in[(int64_t)(2 * j) * w + 2 * i + 1];
    float f3 = in[(int64_t)(2 * j + 1) * w + 2 * i], f4 = in[(int64_t)(2 * j + 1) * w + 2 * i + 1];
    out[(int64_t)j * wo + i] = down4(f1, f2, f3, f4);
}

// measurement.cu:104-126
__device__ __forceinline__ float4 vertex_at(float d, int i, int j, const float* ki)
{
    float4 o = make_float4(0.0f, 0.0f, 0.0f, kInvalid);
    if (is_depth_valid(d)) {
        v3 p = m3v(ki, mk3((float)i + 0.5f, (float)j + 0.5f, 1.0f));
        p = mul3s(p, d);
        o = make_float4(p.x, p.y, p.z, kValid);
    }
    return o;
}

struct Mat3 {
    float m[9];
};

__global__ __launch_bounds__(256) void vertex_kernel(const float* __restrict__ depth, float4* __restrict__ out, int w,
                                                     int h, Mat3 ki)
{
    int i = blockIdx.x * kTile + threadIdx.x, j = blockIdx.y * kTile + threadIdx.y;
    if (i >= w || j >= h) return;
    out[(int64_t)j * w + i] = vertex_at(depth[(int64_t)j * w + i], i, j, ki.m);
}

// ---- the measurement step in two launches (one session or many).
// Pass 1, per level-0 16x16 tile lt of job J: raw codes -> metres in LDS (the
// fuse input is written back, with the tile's max valid depth), bilateral
// filter, level-0 vertex map.
__device__ __forceinline__ void measure_tile(const PreJob& J, int lt, BilateralTile& tile, float* wave_max)
{
    const int t = threadIdx.y * kTile + threadIdx.x;
    const int i0 = (lt % J.tiles_x) * kTile - 3, j0 = (lt / J.tiles_x) * kTile - 3;
    __syncthreads();  // the previous tile's reads of the window are over
    for (int q = t; q < (kTile + 6) * (kTile + 6); q += kTile * kTile) {
        int ly = q / (kTile + 6), lx = q % (kTile + 6);
        int gi = clampi(i0 + lx, 0, J.w - 1), gj = clampi(j0 + ly, 0, J.h - 1);
        tile[ly][lx] = half_rne(J.raw[(int64_t)gj * J.w + gi]) * J.scale;
    }
    __syncthreads();
    const int i = i0 + 3 + threadIdx.x, j = j0 + 3 + threadIdx.y;
    float dv = 0.0f;  // this pixel's contribution to the tile's max valid depth
    if (i < J.w && j < J.h) {
        const int64_t q = (int64_t)j * J.w + i;
        const float dm = tile[threadIdx.y + 3][threadIdx.x + 3];
        J.depth_m[q] = dm;
        dv = is_depth_valid(dm) ? dm : 0.0f;
        const float d = bilateral_at(tile, threadIdx.x, threadIdx.y);
        J.d0[q] = d;
        J.v[0][q] = vertex_at(d, i, j, J.Kinv[0]);
    }
    dv = kf::wave_max(dv);  // (wave_max below is the LDS array)
    if ((t & 63) == 0) wave_max[t >> 6] = dv;
    __syncthreads();
    if (t == 0) J.dmax[lt] = fmaxf(fmaxf(wave_max[0], wave_max[1]), fmaxf(wave_max[2], wave_max[3]));
}

// Pass 2, one thread per level-2 pixel of level-2 tile lt: its 4x4 level-0
// block gives 4 level-1 depths (downSample) and those the level-2 depth; vertex
// maps of both levels (and the depths, when J.d12 is set).  A workgroup covers
// a 64x64 level-0 region and also writes its max valid depth.
__device__ __forceinline__ void pyramid_tile(const PreJob& J, int lt)
{
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const int tx16 = J.tiles_x, ty16 = (J.h + 15) / 16;
        const int rx = lt % J.ptiles_x, ry = lt / J.ptiles_x;
        float m = 0.0f;
        for (int qy = ry * 4; qy < min(ry * 4 + 4, ty16); ++qy)
            for (int qx = rx * 4; qx < min(rx * 4 + 4, tx16); ++qx) m = fmaxf(m, J.dmax[qy * tx16 + qx]);
        J.dmax[tx16 * ty16 + lt] = m;
    }
    const int i2 = (lt % J.ptiles_x) * kTile + threadIdx.x, j2 = (lt / J.ptiles_x) * kTile + threadIdx.y;
    const int w1 = J.w >> 1, w2 = J.w >> 2;
    if (i2 >= w2 || j2 >= (J.h >> 2)) return;
    float l1[2][2];
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const int i1 = 2 * i2 + dx, j1 = 2 * j2 + dy;
            const float* r0 = J.d0 + (int64_t)(2 * j1) * J.w + 2 * i1;
            const float* r1 = r0 + J.w;
            const float d = down4(r0[0], r0[1], r1[0], r1[1]);
            l1[dy][dx] = d;
            J.v[1][(int64_t)j1 * w1 + i1] = vertex_at(d, i1, j1, J.Kinv[1]);
            if (J.d12[0]) J.d12[0][(int64_t)j1 * w1 + i1] = d;
        }
    const float d2 = down4(l1[0][0], l1[0][1], l1[1][0], l1[1][1]);
    J.v[2][(int64_t)j2 * w2 + i2] = vertex_at(d2, i2, j2, J.Kinv[2]);
    if (J.d12[1]) J.d12[1][(int64_t)j2 * w2 + i2] = d2;
}

// One session: one workgroup per tile, the job by value.
__global__ __launch_bounds__(256) void measure_kernel(PreJob J)
{
    __shared__ BilateralTile tile;
    __shared__ float wave_max[4];
    measure_tile(J, blockIdx.x, tile, wave_max);
}
__global__ __launch_bounds__(256) void pyramid_kernel(PreJob J) { pyramid_tile(J, blockIdx.x); }

// Many sessions per launch (batch ticks), persistent over the tick's tiles.
// The first kernel also clears the tick's fuse-list counter and raycast tile
// counters, which later passes of this tick use.
__global__ __launch_bounds__(256) void measure_batch_kernel(const BatchHeader* hd, const PreJob* jobs)
{
    __shared__ BilateralTile tile;
    __shared__ float wave_max[4];
    if (blockIdx.x == 0 && threadIdx.y == 0) {
        if (threadIdx.x == 0) *hd->fuse_count = 0u;
        hd->ray_next[threadIdx.x] = 0u;  // kTile == 16 tile counters
        if (threadIdx.x < 2) hd->ray_next[16 + threadIdx.x] = 0u;  // deferred-ray counts
    }
    const int tiles = hd->pre_tiles, njobs = hd->n_pre;
    for (int tl = blockIdx.x; tl < tiles; tl += gridDim.x) {
        const PreJob& J = jobs[job_of(jobs, njobs, tl, [](const PreJob& x) { return x.tile0; })];
        measure_tile(J, tl - J.tile0, tile, wave_max);
    }
}

__global__ __launch_bounds__(256) void pyramid_batch_kernel(const BatchHeader* hd, const PreJob* jobs)
{
    const int tiles = hd->pyr_tiles, njobs = hd->n_pre;
    for (int tl = blockIdx.x; tl < tiles; tl += gridDim.x) {
        const PreJob& J = jobs[job_of(jobs, njobs, tl, [](const PreJob& x) { return x.ptile0; })];
        pyramid_tile(J, tl - J.ptile0);
    }
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

int launch_measure(const PreJob& J, hipStream_t s)
{
    const int tiles = J.tiles_x * ((J.h + 15) / 16), ptiles = J.ptiles_x * ((J.h / 4 + 15) / 16);
    measure_kernel<<<(unsigned)tiles, dim3(kTile, kTile), 0, s>>>(J);
    KF_LAUNCHED("measure_kernel");
    pyramid_kernel<<<(unsigned)ptiles, dim3(kTile, kTile), 0, s>>>(J);
    KF_LAUNCHED("pyramid_kernel");
    return KF_OK;
}

int launch_preprocess_batch(const BatchHeader* hd, const PreJob* jobs, unsigned blocks, hipStream_t s)
{
    measure_batch_kernel<<<blocks, dim3(kTile, kTile), 0, s>>>(hd, jobs);
    KF_LAUNCHED("measure_batch_kernel");
    pyramid_batch_kernel<<<blocks, dim3(kTile, kTile), 0, s>>>(hd, jobs);
    KF_LAUNCHED("pyramid_batch_kernel");
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
