// TSDF volume: allocation, initialisation, projective fusion, and the raw
// interchange format (project/src/tsdf.cu, voxel_grid.h, voxel_grid.cpp).
//
// HBM layout: two planes (f, w) of n^3 voxels in 8^3 bricks (kf_device.hpp).
//   KF_VOXEL_F32        : f float, w float          (8 B / voxel)
//   KF_VOXEL_SNORM16_U8 : f snorm16, w uint8 count  (3 B / voxel)
// One workgroup of 256 lanes owns one brick and fuses its 512 voxels (two per
// lane, 4 z-slices apart), so every plane access of a wave is a contiguous
// 256 B (f32) run.  Bricks whose bounding sphere lies outside the view frustum
// are skipped before any voxel is visited; the test is conservative, so the
// result is identical to visiting every voxel.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "kf_device.hpp"
#include "kf_kernels.hpp"

namespace kf {

namespace {

struct FuseArgs {
    const float* depth;
    int w, h;
    float K[9];
    float inv_pose[16];
    const PoseState* dyn;  // when set: inv_pose comes from dyn->inv_pose, skip if dyn->frame_skip_fuse
    int n, nb;
    float half, res, mue;
    const float* zk;
    void* f;
    void* wt;
    unsigned long long* touched;
    int cull;  // 1 when K is a plain pinhole matrix (culling planes valid)
};

__device__ __forceinline__ bool brick_visible(const FuseArgs& a, const float* ip, int bi, int bj, int bk)
{
    // Bounding sphere of the brick's voxel centres (span 7 * res per axis).
    float cx = a.half + ((float)(bi * 8) + 3.5f) * a.res;
    float cy = a.half + ((float)(bj * 8) + 3.5f) * a.res;
    float cz = a.half + ((float)(bk * 8) + 3.5f) * a.res;
    float r = 3.5f * 1.7320508f * a.res * 1.01f + 1e-4f;
    float ex = (ip[0] * cx + ip[4] * cy) + (ip[8] * cz + ip[12]);
    float ey = (ip[1] * cx + ip[5] * cy) + (ip[9] * cz + ip[13]);
    float ez = (ip[2] * cx + ip[6] * cy) + (ip[10] * cz + ip[14]);
    if (ez + r < kMinDistance || ez - r >= kMaxDistance) return false;
    float fx = a.K[0], fy = a.K[4], px = a.K[6], py = a.K[7];
    float W = (float)a.w, H = (float)a.h;
    // pixel.x >= 0  <=>  fx*x + px*z >= 0 (z > 0);  pixel.x < W  <=>  (W-px)*z - fx*x > 0
    if ((fx * ex + px * ez) < -r * sqrtf(fx * fx + px * px)) return false;
    if (((W - px) * ez - fx * ex) < -r * sqrtf(fx * fx + (W - px) * (W - px))) return false;
    if ((fy * ey + py * ez) < -r * sqrtf(fy * fy + py * py)) return false;
    if (((H - py) * ez - fy * ey) < -r * sqrtf(fy * fy + (H - py) * (H - py))) return false;
    return true;
}

// tsdf.cu:20-75 for one voxel; returns true when the voxel was updated.
template <int FMT>
__device__ __forceinline__ bool fuse_voxel(const FuseArgs& a, const float* ip, v3 col /* ip0*x + ip1*y per row */,
                                           float z, int64_t q)
{
    // pe = (m0*x + m1*y) + (m2*z + m3*1): the first pair is shared by the column
    v3 pe = mk3(col.x + (ip[8] * z + ip[12] * 1.0f), col.y + (ip[9] * z + ip[13] * 1.0f),
                col.z + (ip[10] * z + ip[14] * 1.0f));
    if (!(pe.z >= kMinDistance && pe.z < kMaxDistance)) return false;
    v3 pix = m3v(a.K, pe);
    float px = pix.x / pix.z, py = pix.y / pix.z;
    if (!(px >= 0.0f && px < (float)a.w && py >= 0.0f && py < (float)a.h)) return false;
    float d = a.depth[(int64_t)(int)py * a.w + (int)px];
    if (!is_depth_valid(d)) return false;
    float diff = d - pe.z;
    if (!(diff >= -a.mue)) return false;
    float f = diff / a.mue;
    f = (f < 1.0f) ? f : 1.0f;
    if (FMT == KF_VOXEL_F32) {
        float* F = static_cast<float*>(a.f);
        float* Wt = static_cast<float*>(a.wt);
        float wt = 1.0f / d;
        float Fv = F[q], Wv = Wt[q];
        Fv = (Fv * Wv + f * wt) / (Wv + wt);
        Wv = Wv + wt;
        F[q] = Fv;
        Wt[q] = Wv;
    } else {
        // precision form: unit weight per observation, count capped at 128
        int16_t* F = static_cast<int16_t*>(a.f);
        uint8_t* Wt = static_cast<uint8_t*>(a.wt);
        float Wv = (float)Wt[q];
        float Fv = snorm_decode(F[q]);
        Fv = (Fv * Wv + f) / (Wv + 1.0f);
        F[q] = snorm_encode(Fv);
        Wt[q] = (uint8_t)(Wv < 128.0f ? Wv + 1.0f : 128.0f);
    }
    return true;
}

template <int FMT>
__global__ __launch_bounds__(256) void fuse_kernel(FuseArgs a)
{
    if (a.dyn && a.dyn->frame_skip_fuse) return;
    __shared__ float ip[16];
    if (threadIdx.x < 16) ip[threadIdx.x] = a.dyn ? a.dyn->inv_pose[threadIdx.x] : a.inv_pose[threadIdx.x];
    __syncthreads();
    const int64_t brick = blockIdx.x;
    const int bi = (int)(brick % a.nb), bj = (int)((brick / a.nb) % a.nb), bk = (int)(brick / ((int64_t)a.nb * a.nb));
    if (a.cull && !brick_visible(a, ip, bi, bj, bk)) return;
    const int lane = threadIdx.x;
    const int i = bi * 8 + (lane & 7), j = bj * 8 + ((lane >> 3) & 7), kl = lane >> 6;
    const float x = a.half + (float)i * a.res, y = a.half + (float)j * a.res;
    const v3 col = mk3(ip[0] * x + ip[4] * y, ip[1] * x + ip[5] * y, ip[2] * x + ip[6] * y);
    int cnt = 0;
#pragma unroll
    for (int hpart = 0; hpart < 2; ++hpart) {
        int k = bk * 8 + kl + 4 * hpart;
        cnt += fuse_voxel<FMT>(a, ip, col, a.zk[k], brick * kBrickVoxels + lane + 256 * hpart) ? 1 : 0;
    }
    // wave sum of the touched counts, one atomic per wave
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(a.touched, (unsigned long long)cnt);
}

template <int FMT>
__global__ __launch_bounds__(256) void init_kernel(void* f, void* wt, int64_t count, float f0, float w0)
{
    int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * 256;
    for (; q < count; q += stride) {
        if (FMT == KF_VOXEL_F32) {
            static_cast<float*>(f)[q] = f0;
            static_cast<float*>(wt)[q] = w0;
        } else {
            static_cast<int16_t*>(f)[q] = snorm_encode(f0);
            float wr = rintf(fminf(fmaxf(w0, 0.0f), 255.0f));
            static_cast<uint8_t*>(wt)[q] = (uint8_t)wr;
        }
    }
}

// Raw interchange <-> bricked planes, for the 8 z-slices of brick layer bk.
// slab: interleaved (f, w) floats, x fastest, n*n*8 voxels.
template <int FMT>
__global__ __launch_bounds__(256) void slab_to_volume(const float* __restrict__ slab, void* f, void* wt, int n, int nb,
                                                      int bk)
{
    int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t per = (int64_t)n * n * 8;
    if (q >= per) return;
    int i = (int)(q % n), j = (int)((q / n) % n), k = bk * 8 + (int)(q / ((int64_t)n * n));
    int64_t d = brick_index(nb, i, j, k);
    float fv = slab[2 * q], wv = slab[2 * q + 1];
    if (FMT == KF_VOXEL_F32) {
        static_cast<float*>(f)[d] = fv;
        static_cast<float*>(wt)[d] = wv;
    } else {
        static_cast<int16_t*>(f)[d] = snorm_encode(fv);
        static_cast<uint8_t*>(wt)[d] = (uint8_t)rintf(fminf(fmaxf(wv, 0.0f), 255.0f));
    }
}
template <int FMT>
__global__ __launch_bounds__(256) void volume_to_slab(float* __restrict__ slab, const void* f, const void* wt, int n,
                                                      int nb, int bk)
{
    int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t per = (int64_t)n * n * 8;
    if (q >= per) return;
    int i = (int)(q % n), j = (int)((q / n) % n), k = bk * 8 + (int)(q / ((int64_t)n * n));
    int64_t d = brick_index(nb, i, j, k);
    if (FMT == KF_VOXEL_F32) {
        slab[2 * q] = static_cast<const float*>(f)[d];
        slab[2 * q + 1] = static_cast<const float*>(wt)[d];
    } else {
        slab[2 * q] = snorm_decode(static_cast<const int16_t*>(f)[d]);
        slab[2 * q + 1] = (float)static_cast<const uint8_t*>(wt)[d];
    }
}

bool is_pinhole(const float* K)
{
    return K[1] == 0.0f && K[2] == 0.0f && K[3] == 0.0f && K[5] == 0.0f && K[8] == 1.0f && K[0] > 0.0f &&
           K[4] > 0.0f;
}

// Moves the volume through host memory one brick layer (8 z-slices) at a time.
template <class Fn>
int for_each_slab(const kf_volume* v, bool to_volume, Fn&& host_side)
{
    const int64_t per = (int64_t)v->n * v->n * 8;
    float* dslab = nullptr;
    std::vector<float> hslab((size_t)per * 2);
    KF_HIP(hipMalloc(&dslab, sizeof(float) * per * 2));
    int st = KF_OK;
    const unsigned blocks = (unsigned)((per + 255) / 256);
    for (int bk = 0; bk < v->nb && st == KF_OK; ++bk) {
        if (to_volume) {
            st = host_side(bk, hslab.data());
            if (st != KF_OK) break;
            hipError_t e = hipMemcpy(dslab, hslab.data(), sizeof(float) * per * 2, hipMemcpyHostToDevice);
            if (e != hipSuccess) { st = fail_hip(e, "volume upload copy"); break; }
            if (v->format == KF_VOXEL_F32)
                slab_to_volume<KF_VOXEL_F32><<<blocks, 256>>>(dslab, v->f, v->w, v->n, v->nb, bk);
            else
                slab_to_volume<KF_VOXEL_SNORM16_U8><<<blocks, 256>>>(dslab, v->f, v->w, v->n, v->nb, bk);
            e = hipDeviceSynchronize();
            if (e != hipSuccess) { st = fail_hip(e, "slab_to_volume"); break; }
        } else {
            if (v->format == KF_VOXEL_F32)
                volume_to_slab<KF_VOXEL_F32><<<blocks, 256>>>(dslab, v->f, v->w, v->n, v->nb, bk);
            else
                volume_to_slab<KF_VOXEL_SNORM16_U8><<<blocks, 256>>>(dslab, v->f, v->w, v->n, v->nb, bk);
            hipError_t e = hipMemcpy(hslab.data(), dslab, sizeof(float) * per * 2, hipMemcpyDeviceToHost);
            if (e != hipSuccess) { st = fail_hip(e, "volume download copy"); break; }
            st = host_side(bk, hslab.data());
        }
    }
    (void)hipFree(dslab);
    return st;
}

}  // namespace

int launch_fuse(const float* depth, int w, int h, kf_volume* v, const float* K, const float* inv_pose,
                const PoseState* dyn_state, hipStream_t s)
{
    FuseArgs a{};
    a.depth = depth;
    a.w = w;
    a.h = h;
    memcpy(a.K, K, sizeof(a.K));
    if (inv_pose) memcpy(a.inv_pose, inv_pose, sizeof(a.inv_pose));
    a.dyn = dyn_state;
    a.n = v->n;
    a.nb = v->nb;
    a.half = -v->width * 0.5f;
    a.res = v->res;
    a.mue = v->mue;
    a.zk = v->zk;
    a.f = v->f;
    a.wt = v->w;
    a.touched = v->touched;
    a.cull = is_pinhole(K) ? 1 : 0;
    const unsigned blocks = (unsigned)((int64_t)v->nb * v->nb * v->nb);
    if (v->format == KF_VOXEL_F32)
        fuse_kernel<KF_VOXEL_F32><<<blocks, 256, 0, s>>>(a);
    else
        fuse_kernel<KF_VOXEL_SNORM16_U8><<<blocks, 256, 0, s>>>(a);
    KF_LAUNCHED("fuse_kernel");
    return KF_OK;
}

int launch_initialize(kf_volume* v, float f, float wt, hipStream_t s)
{
    int64_t count = (int64_t)v->n * v->n * v->n;
    unsigned blocks = (unsigned)std::min<int64_t>((count + 255) / 256, 65536);
    if (v->format == KF_VOXEL_F32)
        init_kernel<KF_VOXEL_F32><<<blocks, 256, 0, s>>>(v->f, v->w, count, f, wt);
    else
        init_kernel<KF_VOXEL_SNORM16_U8><<<blocks, 256, 0, s>>>(v->f, v->w, count, f, wt);
    KF_LAUNCHED("init_kernel");
    return KF_OK;
}

}  // namespace kf

// ------------------------------------------------------------------ C ABI
extern "C" {

int kf_volume_create(int n, float width_m, float mue_m, int format, kf_volume** out)
{
    if (!out) return kf::fail(KF_ERR_ARG, "kf_volume_create: null out");
    *out = nullptr;
    if (n < 8 || n % 8 != 0) return kf::fail(KF_ERR_ARG, "kf_volume_create: n must be a positive multiple of 8");
    if (!(width_m > 0.0f) || !(width_m < 1e4f))
        return kf::fail(KF_ERR_ARG, "kf_volume_create: width must be in (0, 1e4) m");
    // the raycast marches at >= 0.1 * mue; 1e-4 m keeps every march finite and short
    if (!(mue_m >= 1e-4f)) return kf::fail(KF_ERR_ARG, "kf_volume_create: mue must be >= 1e-4 m");
    if (format != KF_VOXEL_F32 && format != KF_VOXEL_SNORM16_U8)
        return kf::fail(KF_ERR_ARG, "kf_volume_create: unknown voxel format");
    kf_volume* v = new kf_volume();
    v->n = n;
    v->nb = n / 8;
    v->format = format;
    v->width = width_m;
    v->res = width_m / (float)(n - 1);  // voxel_grid.h:43
    v->mue = mue_m;
    const int64_t count = (int64_t)n * n * n;
    const size_t fb = (format == KF_VOXEL_F32) ? 4 : 2, wb = (format == KF_VOXEL_F32) ? 4 : 1;
    v->bytes = (size_t)count * (fb + wb);
    hipError_t e;
    if ((e = hipMalloc(&v->f, (size_t)count * fb)) != hipSuccess || (e = hipMalloc(&v->w, (size_t)count * wb)) != hipSuccess ||
        (e = hipMalloc(&v->zk, sizeof(float) * n)) != hipSuccess ||
        (e = hipMalloc(&v->touched, sizeof(unsigned long long))) != hipSuccess) {
        kf_volume_destroy(v);
        return kf::fail_hip(e, "kf_volume_create");
    }
    // tsdf.cu:27-31: z starts at -width/2 - res and advances by repeated addition
    std::vector<float> zk(n);
    float z = -width_m * 0.5f + (-v->res);
    for (int k = 0; k < n; ++k) {
        z = z + v->res;
        zk[k] = z;
    }
    unsigned long long zero = 0;
    if ((e = hipMemcpy(v->zk, zk.data(), sizeof(float) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(v->touched, &zero, sizeof(zero), hipMemcpyHostToDevice)) != hipSuccess) {
        kf_volume_destroy(v);
        return kf::fail_hip(e, "kf_volume_create");
    }
    // Voxel{} = (2, 0) (voxel_grid.h:10-16), as KinectFusion's constructor initialises it
    int st = kf::launch_initialize(v, 2.0f, 0.0f, nullptr);
    if (st == KF_OK && (e = hipDeviceSynchronize()) != hipSuccess) st = kf::fail_hip(e, "initialize");
    if (st != KF_OK) {
        kf_volume_destroy(v);
        return st;
    }
    *out = v;
    return KF_OK;
}

int kf_volume_create_scaled(int n, float width_m, float mue_scale, int format, kf_volume** out)
{
    if (n < 2) return kf::fail(KF_ERR_ARG, "kf_volume_create_scaled: n too small");
    float res = width_m / (float)(n - 1);
    return kf_volume_create(n, width_m, mue_scale * res, format, out);  // voxel_grid.h:44
}

int kf_volume_destroy(kf_volume* v)
{
    if (!v) return KF_OK;
    (void)hipFree(v->f);
    (void)hipFree(v->w);
    (void)hipFree(v->zk);
    (void)hipFree(v->touched);
    delete v;
    return KF_OK;
}

int kf_volume_info(const kf_volume* v, int* n, float* width_m, float* resolution, float* mue_m, int* format)
{
    if (!v) return kf::fail(KF_ERR_ARG, "kf_volume_info: null volume");
    if (n) *n = v->n;
    if (width_m) *width_m = v->width;
    if (resolution) *resolution = v->res;
    if (mue_m) *mue_m = v->mue;
    if (format) *format = v->format;
    return KF_OK;
}

int kf_volume_bytes(const kf_volume* v, size_t* bytes)
{
    if (!v || !bytes) return kf::fail(KF_ERR_ARG, "kf_volume_bytes: null argument");
    *bytes = v->bytes;
    return KF_OK;
}

int kf_volume_touched_total(const kf_volume* v, int64_t* count)
{
    if (!v || !count) return kf::fail(KF_ERR_ARG, "kf_volume_touched_total: null argument");
    unsigned long long c = 0;
    KF_HIP(hipDeviceSynchronize());
    KF_HIP(hipMemcpy(&c, v->touched, sizeof(c), hipMemcpyDeviceToHost));
    *count = (int64_t)c;
    return KF_OK;
}

int kf_initialize_grid(kf_volume* v, float f, float w, kf_stream_t stream, float* elapsed_ms)
{
    if (!v) return kf::fail(KF_ERR_ARG, "initializeGrid: null volume");
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    int st;
    if ((st = t.start()) || (st = kf::launch_initialize(v, f, w, s))) return st;
    return t.finish();
}

int kf_fuse(const float* depth_m, int w, int h, kf_volume* v, const kf_sensor* sensor, kf_stream_t stream,
            float* elapsed_ms)
{
    if (!depth_m || !v || !sensor || w <= 0 || h <= 0) return kf::fail(KF_ERR_ARG, "fuse: bad argument");
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    int st;
    if ((st = t.start()) || (st = kf::launch_fuse(depth_m, w, h, v, sensor->intr[0], sensor->inv_pose, nullptr, s)))
        return st;
    return t.finish();
}

int kf_volume_download(const kf_volume* v, float* fw_host)
{
    if (!v || !fw_host) return kf::fail(KF_ERR_ARG, "kf_volume_download: null argument");
    KF_HIP(hipDeviceSynchronize());
    const int64_t per = (int64_t)v->n * v->n * 8;
    return kf::for_each_slab(v, false, [&](int bk, float* slab) {
        memcpy(fw_host + (size_t)bk * per * 2, slab, sizeof(float) * per * 2);
        return (int)KF_OK;
    });
}

int kf_volume_upload(kf_volume* v, const float* fw_host)
{
    if (!v || !fw_host) return kf::fail(KF_ERR_ARG, "kf_volume_upload: null argument");
    KF_HIP(hipDeviceSynchronize());
    const int64_t per = (int64_t)v->n * v->n * 8;
    return kf::for_each_slab(v, true, [&](int bk, float* slab) {
        memcpy(slab, fw_host + (size_t)bk * per * 2, sizeof(float) * per * 2);
        return (int)KF_OK;
    });
}

// voxel_grid.cpp:43-60
int kf_volume_save(const kf_volume* v, const char* path)
{
    if (!v || !path) return kf::fail(KF_ERR_ARG, "kf_volume_save: null argument");
    FILE* fp = fopen(path, "wb");
    if (!fp) return kf::fail(KF_ERR_IO, std::string("kf_volume_save: cannot open ") + path);
    KF_HIP(hipDeviceSynchronize());
    const int64_t per = (int64_t)v->n * v->n * 8;
    int st = kf::for_each_slab(v, false, [&](int, float* slab) {
        size_t want = (size_t)per * 2;
        return fwrite(slab, sizeof(float), want, fp) == want ? (int)KF_OK
                                                              : kf::fail(KF_ERR_IO, "kf_volume_save: short write");
    });
    if (fclose(fp) != 0 && st == KF_OK) st = kf::fail(KF_ERR_IO, "kf_volume_save: close failed");
    return st;
}

// voxel_grid.cpp:62-80 (the reference reads without checking; a short file is an error here)
int kf_volume_load(kf_volume* v, const char* path)
{
    if (!v || !path) return kf::fail(KF_ERR_ARG, "kf_volume_load: null argument");
    FILE* fp = fopen(path, "rb");
    if (!fp) return kf::fail(KF_ERR_IO, std::string("kf_volume_load: cannot open ") + path);
    KF_HIP(hipDeviceSynchronize());
    const int64_t per = (int64_t)v->n * v->n * 8;
    int st = kf::for_each_slab(v, true, [&](int, float* slab) {
        size_t want = (size_t)per * 2;
        return fread(slab, sizeof(float), want, fp) == want ? (int)KF_OK
                                                             : kf::fail(KF_ERR_IO, "kf_volume_load: short file");
    });
    fclose(fp);
    return st;
}

}  // extern "C"
