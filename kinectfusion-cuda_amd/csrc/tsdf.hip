// TSDF volume: allocation, initialisation, projective fusion, and the raw
// interchange format (project/src/tsdf.cu, voxel_grid.h, voxel_grid.cpp).
//
// HBM layout: two planes (f, w) of n^3 voxels in 8^3 bricks (kf_device.hpp).
//   KF_VOXEL_F32        : f float, w float          (8 B / voxel)
//   KF_VOXEL_SNORM16_U8 : f snorm16, w uint8 count  (3 B / voxel)
// A cull pass lists the bricks whose bounding sphere meets the view frustum;
// then one wave fuses one listed brick at a time, one brick column (8 voxels
// along z, contiguous in the plane) per lane, so a wave reads and writes a
// brick's planes as contiguous 16-byte-per-lane accesses, and the listed
// bricks are spread evenly over the waves.  Bricks whose bounding sphere lies outside the view frustum
// are skipped before any voxel is visited; the test is conservative, so the
// result is identical to visiting every voxel.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <vector>

#include "kf_device.hpp"
#include "kf_kernels.hpp"

namespace kf {

namespace {

// List entries name a brick by its coordinates, bi | bj << 10 | bk << 20
// (volumes are at most 1024^3 voxels: 128 bricks per axis), so the sweep
// forms its linear index with two multiply-adds instead of two divisions.
__host__ __device__ __forceinline__ unsigned pack_brick(int bi, int bj, int bk)
{
    return (unsigned)bi | (unsigned)bj << 10 | (unsigned)bk << 20;
}

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
    if ((fx * ex + px * ez) < -r * a.plane_norm[0]) return false;
    if (((W - px) * ez - fx * ex) < -r * a.plane_norm[1]) return false;
    if ((fy * ey + py * ez) < -r * a.plane_norm[2]) return false;
    if (((H - py) * ez - fy * ey) < -r * a.plane_norm[3]) return false;
    // Occlusion: a voxel is updated only if d - z >= -mue for the valid depth d
    // of its pixel.  With the sphere beyond the near plane its pixels lie in the
    // box below (x/z and y/z at the corners, one pixel of margin); if even the
    // largest valid depth there is more than mue (+1 mm for rounding) in front
    // of the sphere's nearest point, no voxel of the brick is updated.  The
    // quotients come from the hardware reciprocal: within a few 1e-7 of x/z,
    // which moves the box by under 1e-2 pixel (|x/z| is at most a few tens with
    // z > kMinDistance), far inside its one pixel of margin.
    const float zlo = ez - r;
    if (!a.dmax || !(zlo > kMinDistance)) return true;
    const float zhi = ez + r, xa = ex - r, xb = ex + r, ya = ey - r, yb = ey + r;
    const float r0 = __builtin_amdgcn_rcpf(zlo), r1 = __builtin_amdgcn_rcpf(zhi);
    const float u0 = fx * fminf(xa * r0, xa * r1) + px - 1.0f, u1 = fx * fmaxf(xb * r0, xb * r1) + px + 1.0f;
    const float v0 = fy * fminf(ya * r0, ya * r1) + py - 1.0f, v1 = fy * fmaxf(yb * r0, yb * r1) + py + 1.0f;
    const int i0 = (int)fmaxf(u0, 0.0f), i1 = (int)fminf(u1, W - 1.0f);
    const int j0 = (int)fmaxf(v0, 0.0f), j1 = (int)fminf(v1, H - 1.0f);
    if (i1 < i0 || j1 < j0) return a.mue + 1e-3f >= zlo;  // no pixel: no valid depth
    const float* m;
    int shift, tx;
    if (i1 - i0 < 48 && j1 - j0 < 48) {
        m = a.dmax;
        shift = 4;
        tx = a.tx16;
    } else if (i1 - i0 < 192 && j1 - j0 < 192) {
        m = a.dmax + a.tx16 * a.ty16;
        shift = 6;
        tx = a.tx64;
    } else {
        return true;
    }
    // a box under 48 (192) pixels spans at most 4 x 4 tiles of 16 (64): all 16
    // reads are issued together, those past the box's last tile row or column
    // repeating it (the maximum is the same)
    const int ti0 = i0 >> shift, tj0 = j0 >> shift, ti1 = i1 >> shift, tj1 = j1 >> shift;
    float t[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) t[q] = m[min(tj0 + (q >> 2), tj1) * tx + min(ti0 + (q & 3), ti1)];
    float dm = 0.0f;
#pragma unroll
    for (int q = 0; q < 16; ++q) dm = fmaxf(dm, t[q]);
    return dm + a.mue + 1e-3f >= zlo;
}

// Tile-max maps of a depth map (FuseArgs::dmax): one workgroup per 64x64
// pixel region; lane group t/16 takes one 16x16 tile, each lane a 4x4 block.
// Invalid depths (isDepthValid false) count as 0.
__global__ __launch_bounds__(256) void depth_max_kernel(const float* __restrict__ depth, int w, int h, float* dmax,
                                                        int tx16, int ty16, int tx64)
{
    __shared__ float part[16];
    const int q = threadIdx.x >> 4, sub = threadIdx.x & 15;
    const int tx = blockIdx.x * 4 + (q & 3), ty = blockIdx.y * 4 + (q >> 2);
    const int x0 = tx * 16 + (sub & 3) * 4, y0 = ty * 16 + (sub >> 2) * 4;
    float m = 0.0f;
    for (int y = y0; y < y0 + 4; ++y)
        for (int x = x0; x < x0 + 4; ++x)
            if (x < w && y < h) {
                const float d = depth[(int64_t)y * w + x];
                if (is_depth_valid(d)) m = fmaxf(m, d);
            }
    // max over the 16 lanes of the tile (one 16-lane row)
    m = fmaxf(m, dpp_partner<kDppRowMirror>(m));
    m = fmaxf(m, dpp_partner<kDppRowHalfMirror>(m));
    m = fmaxf(m, dpp_partner<kDppXor2>(m));
    m = fmaxf(m, dpp_partner<kDppXor1>(m));
    if (sub == 0) {
        part[q] = m;
        if (tx < tx16 && ty < ty16) dmax[ty * tx16 + tx] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.0f;
        for (int k = 0; k < 16; ++k) r = fmaxf(r, part[k]);
        dmax[tx16 * ty16 + blockIdx.y * tx64 + blockIdx.x] = r;
    }
}

// x / b rounded to nearest from y = RN(1/b): q = RN(x*y), r = x - b*q (exact by
// the fused multiply-add), RN(q + r*y) is RN(x/b) for every x and b whose
// quotient and residual neither overflow nor underflow (Markstein's theorem;
// checked exhaustively on the CPU for every x in [-mue, mue] at several mue
// and for every divisor 1..129 of the count update).  The fuse's operands
// (|x| >= 2^-25 or 0, b in [1e-4, 129]) stay far from both ranges.
__device__ __forceinline__ float div_by_exact_reciprocal(float x, float b, float y)
{
    const float q = x * y;
    const float r = fmaf(-q, b, x);
    return fmaf(r, y, q);
}

// True when x may lie within a few ulp of an integer (0 and the image size
// included): x*rcp(z) is within 2^-22 |x/z| of the rounded quotient x/z
// (v_rcp_f32 is within 1 ulp), so outside a band of half-width band =
// 2^-21 max(w, h) both have the same sign, the same integer part and the same
// side of every integer bound wherever |x| <= 2 max(w, h); beyond that both
// lie outside the image, and the integer part is not used.
__device__ __forceinline__ bool near_integer(float x, float band)
{
    return fabsf(x - rintf(x)) <= band;
}

// A brick column's planes in registers: the 8 voxels (i, j, 0..7) are
// contiguous, so f and w move as one or two 16-byte accesses each (FP32: two
// float4 per plane; snorm16: 8 codes in a uint4, 8 counts in a uint2).
template <int FMT>
struct ColumnPlanes;

template <>
struct ColumnPlanes<KF_VOXEL_F32> {
    float f[8], w[8];
    __device__ __forceinline__ void clear()
    {
#pragma unroll
        for (int k = 0; k < 8; ++k) f[k] = w[k] = 0.0f;
    }
    __device__ __forceinline__ void load(const FuseArgs& a, int64_t base)
    {
        const float4* F = reinterpret_cast<const float4*>(static_cast<const float*>(a.f) + base);
        const float4* W = reinterpret_cast<const float4*>(static_cast<const float*>(a.wt) + base);
        const float4 f0 = F[0], f1 = F[1], w0 = W[0], w1 = W[1];
        f[0] = f0.x, f[1] = f0.y, f[2] = f0.z, f[3] = f0.w, f[4] = f1.x, f[5] = f1.y, f[6] = f1.z, f[7] = f1.w;
        w[0] = w0.x, w[1] = w0.y, w[2] = w0.z, w[3] = w0.w, w[4] = w1.x, w[5] = w1.y, w[6] = w1.z, w[7] = w1.w;
    }
    __device__ __forceinline__ void store(const FuseArgs& a, int64_t base) const
    {
        float4* F = reinterpret_cast<float4*>(static_cast<float*>(a.f) + base);
        float4* W = reinterpret_cast<float4*>(static_cast<float*>(a.wt) + base);
        F[0] = make_float4(f[0], f[1], f[2], f[3]);
        F[1] = make_float4(f[4], f[5], f[6], f[7]);
        W[0] = make_float4(w[0], w[1], w[2], w[3]);
        W[1] = make_float4(w[4], w[5], w[6], w[7]);
    }
};

template <>
struct ColumnPlanes<KF_VOXEL_SNORM16_U8> {
    uint4 f;  // codes 2m (low half) and 2m+1 (high half) in word m
    uint2 w;  // count k in byte k % 4 of word k / 4
    __device__ __forceinline__ void clear()
    {
        f = make_uint4(0u, 0u, 0u, 0u);
        w = make_uint2(0u, 0u);
    }
    __device__ __forceinline__ void load(const FuseArgs& a, int64_t base)
    {
        f = *reinterpret_cast<const uint4*>(static_cast<const int16_t*>(a.f) + base);
        w = *reinterpret_cast<const uint2*>(static_cast<const uint8_t*>(a.wt) + base);
    }
    __device__ __forceinline__ void store(const FuseArgs& a, int64_t base) const
    {
        *reinterpret_cast<uint4*>(static_cast<int16_t*>(a.f) + base) = f;
        *reinterpret_cast<uint2*>(static_cast<uint8_t*>(a.wt) + base) = w;
    }
    __device__ __forceinline__ static uint32_t word(const uint4& v, int m)
    {
        return m == 0 ? v.x : (m == 1 ? v.y : (m == 2 ? v.z : v.w));
    }
    __device__ __forceinline__ int16_t code(int k) const
    {
        return (int16_t)(uint16_t)(word(f, k >> 1) >> (16 * (k & 1)));
    }
    __device__ __forceinline__ int count(int k) const
    {
        return (int)(((k < 4 ? w.x : w.y) >> (8 * (k & 3))) & 0xffu);
    }
    __device__ __forceinline__ void set(int k, int16_t q, int c)
    {
        const uint32_t sh = 16u * (uint32_t)(k & 1), m = 0xffffu << sh, v = (uint32_t)(uint16_t)q << sh;
        uint32_t& fw = (k >> 1) == 0 ? f.x : ((k >> 1) == 1 ? f.y : ((k >> 1) == 2 ? f.z : f.w));
        fw = (fw & ~m) | v;
        const uint32_t bs = 8u * (uint32_t)(k & 3), bm = 0xffu << bs;
        uint32_t& ww = k < 4 ? w.x : w.y;
        ww = (ww & ~bm) | ((uint32_t)c << bs);
    }
};

// tsdf.cu:20-75 for the 8 voxels (one per z-slice kl) of one lane's column of
// a brick (lane = (j, i) of the brick); returns the number of updated voxels.
// Projection (tsdf.cu:34-45): pe = inv_pose * (x, y, z, 1), pixel = K pe / pe.z;
// only the pixel's integer part and the range test use the quotient, so it is
// formed from the hardware reciprocal and recomputed by the correctly rounded
// division only near an integer (the decisions are the exact quotient's).
// Update (tsdf.cu:47-69): f = min(1, diff / mue) from the exact reciprocal of
// mue; the precision form divides by the observation count + 1 with the exact
// reciprocal of 1..129 from inv_count (LDS); every value written is the one the
// plain quotients give.  fmin takes the smallest value written, fk0 the value
// written to voxel 0 (the brick's k = 0 plane).  The column's planes are read
// when any of its voxels passes the depth test (after the depth gathers) and
// written back whole when any was updated (the others keep the bits they were
// read with; one wave owns the brick).
template <int FMT>
__device__ __forceinline__ int fuse_column(const FuseArgs& a, v3 col, const float4* zt, int64_t base, float& fmin,
                                           float& fk0, const float* inv_count)
{
    float dep[8], pz[8];
    unsigned pix[8];
    bool hit[8];
    const float fw = (float)a.w, fh = (float)a.h;
#pragma unroll
    for (int kl = 0; kl < 8; ++kl) {
        // pe = (m0*x + m1*y) + (m2*z + m3*1): the first pair is shared by the
        // column, the second by the slice (zt, formed once per fuse)
        v3 pe = mk3(col.x + zt[kl].x, col.y + zt[kl].y, col.z + zt[kl].z);
        bool ok = pe.z >= kMinDistance && pe.z < kMaxDistance;
        unsigned q = 0u;
        if (ok) {
            float px, py;
            if (a.cull) {
                // K pe for a pinhole K: the zero entries add exact zeros, so the product is
                // (fx x + cx z, fy y + cy z, z)
                const float vx = a.K[0] * pe.x + a.K[6] * pe.z, vy = a.K[4] * pe.y + a.K[7] * pe.z;
                const float r = __builtin_amdgcn_rcpf(pe.z);
                px = vx * r;
                py = vy * r;
                if (near_integer(px, a.band) || near_integer(py, a.band)) {
                    px = vx / pe.z;
                    py = vy / pe.z;
                }
            } else {
                const v3 pv = m3v(a.K, pe);
                px = pv.x / pv.z;
                py = pv.y / pv.z;
            }
            ok = px >= 0.0f && px < fw && py >= 0.0f && py < fh;
            // 0 <= px < w, 0 <= py < h: a 24-bit product (images are below 2^24 pixels a side)
            if (ok) q = __umul24((unsigned)py, (unsigned)a.w) + (unsigned)px;
        }
        hit[kl] = ok;
        pix[kl] = q;
        pz[kl] = pe.z;
    }
#pragma unroll
    for (int kl = 0; kl < 8; ++kl) {
        dep[kl] = 0.0f;
        if (hit[kl]) dep[kl] = a.depth[pix[kl]];
    }
    // the depth test first: the planes of a column none of whose voxels is
    // updated (behind the surface or on invalid depth) are not read
    bool upd[8], any = false;
#pragma unroll
    for (int kl = 0; kl < 8; ++kl) {
        upd[kl] = hit[kl] && is_depth_valid(dep[kl]) && dep[kl] - pz[kl] >= -a.mue;
        any = any || upd[kl];
    }
    ColumnPlanes<FMT> v;
    v.clear();
    if (any) v.load(a, base);
    int cnt = 0;
#pragma unroll
    for (int kl = 0; kl < 8; ++kl) {
        const float d = dep[kl];
        const float diff = d - pz[kl];
        if (!upd[kl]) continue;
        // f = min(1, diff / mue): diff >= mue makes the rounded quotient >= 1, so the
        // division (most updated voxels lie in front of the surface) is skipped there
        float f = 1.0f;
        if (diff < a.mue) {
            f = div_by_exact_reciprocal(diff, a.mue, a.inv_mue);
            f = (f < 1.0f) ? f : 1.0f;
        }
        if constexpr (FMT == KF_VOXEL_F32) {
            const float wt = 1.0f / d;
            float Fv = v.f[kl], Wv = v.w[kl];
            Fv = (Fv * Wv + f * wt) / (Wv + wt);
            Wv = Wv + wt;
            v.f[kl] = Fv;
            v.w[kl] = Wv;
            fmin = fminf(fmin, Fv);
            if (kl == 0) fk0 = Fv;
        } else {
            // precision form: unit weight per observation, count capped at 128
            const int wc = v.count(kl);
            const float Wv = (float)wc;
            float Fv = snorm_decode(v.code(kl));
            Fv = div_by_exact_reciprocal(Fv * Wv + f, Wv + 1.0f, inv_count[wc + 1]);
            const int16_t qf = snorm_encode(Fv);
            v.set(kl, qf, wc < 128 ? wc + 1 : 128);
            const float written = snorm_decode(qf);
            fmin = fminf(fmin, written);
            if (kl == 0) fk0 = written;
        }
        ++cnt;
    }
    if (cnt) v.store(a, base);
    return cnt;
}

// Exact reciprocals 1/k, k = 0..129 (entry 0 unused), for the count update.
constexpr int kInvCount = 130;
__device__ __forceinline__ void fill_inv_count(float* t)
{
    for (int k = threadIdx.x; k < kInvCount; k += blockDim.x) t[k] = k ? 1.0f / (float)k : 0.0f;
    __syncthreads();
}

__device__ __forceinline__ void load_inv_pose(const FuseArgs& a, float* ip)
{
    if (threadIdx.x < 16) ip[threadIdx.x] = a.dyn ? a.dyn->inv_pose[threadIdx.x] : a.inv_pose[threadIdx.x];
    __syncthreads();
}

// The z terms of slice k's projection, (m2 z + m3 * 1) per row of the inverse
// pose (the second pair of the reference's (m0 x + m1 y) + (m2 z + m3 * 1)),
// formed by the cull for every slice once per fuse and read by the sweep.
__device__ __forceinline__ void z_terms(const float* ip, const float* zk, float4* zt, int k)
{
    const float z = zk[k];
    zt[k] = make_float4(ip[8] * z + ip[12] * 1.0f, ip[9] * z + ip[13] * 1.0f, ip[10] * z + ip[14] * 1.0f, 0.0f);
}

// Pass 1: one lane per brick; bricks whose bounding sphere meets the view
// frustum are appended to `list`, one atomic per workgroup.  The list length
// lives in count[parity]; this pass also clears count[1 - parity], which the
// previous launch used, for the next launch.
constexpr int kCullThreads = 1024;

__global__ __launch_bounds__(kCullThreads) void fuse_cull_kernel(FuseArgs a, int64_t nbricks, int* list,
                                                                 unsigned int* count, int parity)
{
    if (a.dyn && a.dyn->frame_skip_fuse) return;
    __shared__ float ip[16];
    __shared__ unsigned int wave_base[kCullThreads / 64 + 1];
    load_inv_pose(a, ip);
    if (blockIdx.x == 0 && threadIdx.x == 0) count[1 - parity] = 0u;
    for (int k = blockIdx.x * kCullThreads + threadIdx.x; k < a.n; k += gridDim.x * kCullThreads) z_terms(ip, a.zk, a.zt, k);
    const int64_t brick = (int64_t)blockIdx.x * kCullThreads + threadIdx.x;
    bool vis = false;
    unsigned packed = 0u;
    if (brick < nbricks) {
        // at most 128^3 bricks: 32-bit index arithmetic
        const unsigned ub = (unsigned)brick, nb = (unsigned)a.nb;
        const int bi = (int)(ub % nb), bj = (int)((ub / nb) % nb), bk = (int)(ub / (nb * nb));
        vis = !a.cull || brick_visible(a, ip, bi, bj, bk);
        packed = pack_brick(bi, bj, bk);
    }
    const uint64_t m = __ballot(vis);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wave_base[wave] = (unsigned int)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int total = 0;
        for (int w = 0; w < kCullThreads / 64; ++w) {
            unsigned int c = wave_base[w];
            wave_base[w] = total;
            total += c;
        }
        wave_base[kCullThreads / 64] = total ? atomicAdd(&count[parity], total) : 0u;
    }
    __syncthreads();
    if (vis)
        list[wave_base[kCullThreads / 64] + wave_base[wave] + (unsigned int)__popcll(m & ((1ull << lane) - 1ull))] =
            (int)packed;
}

constexpr int kFuseWaves = 4;

// The bounds a fused brick may lower (kf_device.hpp): bound l < 8 belongs to
// brick (bi - (l & 1), bj - (l >> 1 & 1), bk - (l >> 2)), the brick itself and
// the 7 below it whose halos reach into it; bound 8 is the super-brick's.
// Lane l <= 8 loads bound l at the brick's start (-inf for a brick outside
// the volume: never lowered), so the load shares the column loads' round trip.
__device__ __forceinline__ float load_bound(const FuseArgs& a, int bi, int bj, int bk, int lane)
{
    float b = -__int_as_float(0x7f800000);
    if (lane < 8) {
        const int ni = bi - (lane & 1), nj = bj - ((lane >> 1) & 1), nk = bk - (lane >> 2);
        if (ni >= 0 && nj >= 0 && nk >= 0) b = order_value(a.brick_key[((int64_t)nk * a.nb + nj) * a.nb + ni]);
    } else if (lane == 8) {
        b = order_value(a.super_key[((bk >> 3) * a.ns + (bj >> 3)) * a.ns + (bi >> 3)]);
    }
    return b;
}

// Lowers the bounds the brick's written values fall below: every value lowers
// its brick's and super-brick's bound, values on the local i = 0 plane (lanes
// with i == 0) the -x neighbour's, on j = 0 (lanes with j == 0) the -y
// neighbour's, on k = 0 (each lane's first voxel) the -z neighbour's, and the
// values on two or three of these planes the diagonal neighbours'.  A lane
// issues an atomicMin only where its own smallest value on the set is below
// the bound loaded at the start; bounds are only ever lowered, so a value not
// below the loaded bound is not below the current one either.  Once the model
// has settled, no value is below its bounds and no atomic is issued.
__device__ __forceinline__ void lower_bounds(const FuseArgs& a, int bi, int bj, int bk, int lane, float bound,
                                             float lane_min, float lane_k0)
{
    float b[9];
#pragma unroll
    for (int l = 0; l < 9; ++l) b[l] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bound), l));
    const bool i0 = (lane & 7) == 0, j0 = (lane >> 3) == 0;
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        const int dx = l & 1, dy = (l >> 1) & 1, dz = l >> 2;
        const float v = dz ? lane_k0 : lane_min;
        if ((!dx || i0) && (!dy || j0) && v < b[l])
            atomicMin(&a.brick_key[((int64_t)(bk - dz) * a.nb + (bj - dy)) * a.nb + (bi - dx)], order_key(v));
    }
    if (lane_min < b[8]) atomicMin(&a.super_key[((bk >> 3) * a.ns + (bj >> 3)) * a.ns + (bi >> 3)], order_key(lane_min));
}

// One wave fuses brick (bi, bj, bk) (lane = (j, i) inside the brick, its
// column of 8 voxels along z) and lowers the bounds of F it reaches.  Returns
// the lane's count of updated voxels.
//
// The brick's z terms are uniform: they are read through the constant address
// space (scalar loads, issued together at the start).
typedef __attribute__((address_space(4))) const float ConstFloat;
typedef __attribute__((address_space(4))) const int ConstInt;
typedef __attribute__((address_space(4))) const unsigned int ConstUint;

template <int FMT>
__device__ __forceinline__ int fuse_brick(const FuseArgs& a, const float* ip, unsigned packed, int lane,
                                          const float* inv_count)
{
    const int bi = (int)(packed & 1023u), bj = (int)((packed >> 10) & 1023u), bk = (int)(packed >> 20);
    const int64_t brick = ((int64_t)bk * a.nb + bj) * a.nb + bi;
    float4 zt[8];
#pragma unroll
    for (int kl = 0; kl < 8; ++kl) {
        ConstFloat* z = (ConstFloat*)a.zt + (bk * 8 + kl) * 4;
        zt[kl] = make_float4(z[0], z[1], z[2], 0.0f);
    }
    const int i = bi * 8 + (lane & 7), j = bj * 8 + (lane >> 3);
    const float x = a.half + (float)i * a.res, y = a.half + (float)j * a.res;
    const v3 col = mk3(ip[0] * x + ip[4] * y, ip[1] * x + ip[5] * y, ip[2] * x + ip[6] * y);
    const float bound = load_bound(a, bi, bj, bk, lane);
    float fmin = __int_as_float(0x7f800000), fk0 = fmin;
    const int cnt = fuse_column<FMT>(a, col, zt, brick * kBrickVoxels + lane * 8, fmin, fk0, inv_count);
    lower_bounds(a, bi, bj, bk, lane, bound, fmin, fk0);
    return cnt;
}

// Sums the lanes' counts of a workgroup into its own slot (slot b belongs to
// workgroup b; the add returns nothing, so the wave does not wait for it).
__device__ __forceinline__ void add_touched(unsigned long long* touched, int cnt, int* wave_cnt)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    cnt = wave_sum(cnt);
    if (lane == 0) wave_cnt[wave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int c = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) c += wave_cnt[w];
        if (c) atomicAdd(&touched[blockIdx.x], (unsigned long long)c);
    }
}

// Pass 2: waves stride over the visible bricks.  The skip flag, the pose, the
// list length and the wave's first list entry are uniform loads issued
// together (the list buffer holds an entry per brick of the volume, so reading
// an entry past the length stays inside it); the next entry is fetched while
// a brick is fused.

template <int FMT>
__global__ __launch_bounds__(64 * kFuseWaves, 7) void fuse_list_kernel(FuseArgs a, const int* list,
                                                                    const unsigned int* count)
{
    __shared__ int wave_cnt[kFuseWaves];
    __shared__ float inv_count[kInvCount];
    if (FMT != KF_VOXEL_F32) fill_inv_count(inv_count);
    // constant address space: no kernel writes the PoseState, the list or its
    // length while this one runs
    typedef __attribute__((address_space(4))) const PoseState ConstState;
    const int lane = threadIdx.x & 63;
    const unsigned int wave = (unsigned)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned int cap = (unsigned)a.nb * (unsigned)a.nb * (unsigned)a.nb;
    const unsigned int waves = gridDim.x * kFuseWaves;
    unsigned int item = blockIdx.x * kFuseWaves + wave;
    const int skip = a.dyn ? ((ConstState*)a.dyn)->frame_skip_fuse : 0;
    float ip[16];
    if (a.dyn) {
        ConstState* cs = (ConstState*)a.dyn;
#pragma unroll
        for (int q = 0; q < 16; ++q) ip[q] = cs->inv_pose[q];
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) ip[q] = a.inv_pose[q];
    }
    const unsigned int total = *(ConstUint*)count;
    unsigned brick = item < cap ? ((ConstUint*)list)[item] : 0u;
    if (skip) return;
    int cnt = 0;
    while (item < total) {
        const unsigned int next = item + waves;
        const unsigned next_brick = next < cap ? ((ConstUint*)list)[next] : 0u;
        cnt += fuse_brick<FMT>(a, ip, brick, lane, inv_count);
        item = next;
        brick = next_brick;
    }
    add_touched(a.touched, cnt, wave_cnt);
}

// ---- many sessions per launch (batch ticks).  Jobs are sorted by brick0;
// global brick g of the tick belongs to the job whose range holds it.  The
// cull writes list entries job << 32 | pack_brick(bi, bj, bk) (the brick's
// coordinates in the job's volume), so the sweep needs no search.
constexpr int kBatchCullThreads = 256;
// Bricks a lane tests between two appends to the list: each test is a chain of
// dependent loads (the job search, the session's pose, the depth tile maxima),
// so a lane runs kBatchCullPer independent chains at once, and the workgroup
// takes one list slot range (one atomic on the tick's single counter) per
// kBatchCullPer * kBatchCullThreads bricks.
constexpr int kBatchCullPer = 4;

// Brick g of the tick: its list entry and whether it is listed.
__device__ __forceinline__ bool batch_cull_one(const FuseJob* jobs, int njobs, unsigned g, unsigned total,
                                               unsigned long long& entry)
{
    entry = 0;
    if (g >= total) return false;
    const int j = job_of(jobs, njobs, g, [](const FuseJob& x) { return x.brick0; });
    const FuseJob& J = jobs[j];
    const FuseArgs& a = J.a;
    if (a.dyn->frame_skip_fuse) return false;
    const float* ip = a.dyn->inv_pose;
    const unsigned brick = g - J.brick0, nb = (unsigned)a.nb;
    for (unsigned k = brick; k < (unsigned)a.n; k += nb * nb * nb) z_terms(ip, a.zk, a.zt, (int)k);
    const int bi = (int)(brick % nb), bj = (int)((brick / nb) % nb), bk = (int)(brick / (nb * nb));
    entry = (unsigned long long)j << 32 | (unsigned long long)pack_brick(bi, bj, bk);
    return !a.cull || brick_visible(a, ip, bi, bj, bk);
}

__global__ __launch_bounds__(kBatchCullThreads) void fuse_batch_cull_kernel(const BatchHeader* hd,
                                                                           const FuseJob* jobs)
{
    constexpr int kWaves = kBatchCullThreads / 64;
    __shared__ unsigned int wave_base[kBatchCullPer * kWaves + 1];
    const unsigned total = hd->fuse_bricks;
    const int njobs = hd->n_fuse;
    const unsigned span = kBatchCullThreads * kBatchCullPer;
    const unsigned stride = gridDim.x * span;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // uniform trip count so every lane reaches the barriers
    for (unsigned base = blockIdx.x * span; base < total; base += stride) {
        bool vis[kBatchCullPer];
        unsigned long long entry[kBatchCullPer];
        uint64_t m[kBatchCullPer];
#pragma unroll
        for (int q = 0; q < kBatchCullPer; ++q)
            vis[q] = batch_cull_one(jobs, njobs, base + q * kBatchCullThreads + threadIdx.x, total, entry[q]);
#pragma unroll
        for (int q = 0; q < kBatchCullPer; ++q) {
            m[q] = __ballot(vis[q]);
            if (lane == 0) wave_base[q * kWaves + wave] = (unsigned int)__popcll(m[q]);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // slots in brick order: q-th group of kBatchCullThreads bricks, then wave
            unsigned int t = 0;
            for (int w = 0; w < kBatchCullPer * kWaves; ++w) {
                unsigned int c = wave_base[w];
                wave_base[w] = t;
                t += c;
            }
            wave_base[kBatchCullPer * kWaves] = t ? atomicAdd(hd->fuse_count, t) : 0u;
        }
        __syncthreads();
        const unsigned int slot0 = wave_base[kBatchCullPer * kWaves];
#pragma unroll
        for (int q = 0; q < kBatchCullPer; ++q)
            if (vis[q])
                hd->fuse_list[slot0 + wave_base[q * kWaves + wave] + (unsigned int)__popcll(m[q] & ((1ull << lane) - 1ull))] =
                    entry[q];
        __syncthreads();
    }
}

// Field-wise copy of a job's arguments read through the constant address
// space (scalar loads; the inverse pose comes from the session's PoseState).
template <class CJ>
__device__ __forceinline__ void load_fuse_args(const CJ& j, FuseArgs& a)
{
    a.depth = j.a.depth;
    a.w = j.a.w;
    a.h = j.a.h;
#pragma unroll
    for (int q = 0; q < 9; ++q) a.K[q] = j.a.K[q];
    a.dyn = j.a.dyn;
    a.n = j.a.n;
    a.nb = j.a.nb;
    a.half = j.a.half;
    a.res = j.a.res;
    a.mue = j.a.mue;
    a.inv_mue = j.a.inv_mue;
    a.zk = j.a.zk;
    a.zt = j.a.zt;
    a.band = j.a.band;
    a.f = j.a.f;
    a.wt = j.a.wt;
    a.brick_key = j.a.brick_key;
    a.super_key = j.a.super_key;
    a.ns = j.a.ns;
    a.cull = j.a.cull;
    a.format = j.a.format;
}

// Each wave sweeps a contiguous run of the list: consecutive entries mostly
// belong to one session (the cull appends runs of up to 256 bricks of a job),
// so the job's arguments and pose are reloaded only when the job changes, and
// the next entry is fetched while a brick is fused. Eight waves per SIMD: the
// job arguments otherwise hold 106 SGPRs (seven waves); at eight, 40 of them
// move to VGPR lanes, no scratch, and a 64-session tick measured 0.6 % faster.
__global__ __launch_bounds__(64 * kFuseWaves, 8) void fuse_batch_list_kernel(const BatchHeader* hd, const FuseJob* jobs)
{
    __shared__ int wave_cnt[kFuseWaves];
    __shared__ float inv_count[kInvCount];
    fill_inv_count(inv_count);
    // the tick's table, the list and its length are never written while the
    // kernel runs: read them with scalar loads through the constant address space
    typedef __attribute__((address_space(4))) const FuseJob ConstJob;
    typedef __attribute__((address_space(4))) const BatchHeader ConstHeader;
    typedef __attribute__((address_space(4))) const PoseState ConstState;
    typedef __attribute__((address_space(4))) const unsigned long long ConstEntry;
    ConstJob* cj = (ConstJob*)jobs;
    ConstHeader* ch = (ConstHeader*)hd;
    const int lane = threadIdx.x & 63;
    const unsigned int wave = (unsigned)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned int total = *(ConstUint*)ch->fuse_count;
    ConstEntry* list = (ConstEntry*)ch->fuse_list;
    const unsigned int waves = gridDim.x * kFuseWaves;
    const unsigned int per = (total + waves - 1) / waves;
    const unsigned long long first = (unsigned long long)(blockIdx.x * kFuseWaves + wave) * per;
    unsigned int item = first < total ? (unsigned)first : total;
    const unsigned int end = item + per < total ? item + per : total;
    unsigned long long e = item < end ? list[item] : 0ull;
    int cur = -1;
    FuseArgs a;
    float ip[16];
    int cnt = 0;
    while (item < end) {
        const unsigned long long ne = item + 1 < end ? list[item + 1] : 0ull;
        const int job = (int)(e >> 32);
        const unsigned brick = (unsigned)(e & 0xffffffffull);
        if (job != cur) {
            cur = job;
            load_fuse_args(cj[job], a);
            ConstState* cs = (ConstState*)a.dyn;
#pragma unroll
            for (int q = 0; q < 16; ++q) ip[q] = cs->inv_pose[q];
        }
        if (a.format == KF_VOXEL_F32) cnt += fuse_brick<KF_VOXEL_F32>(a, ip, brick, lane, inv_count);
        else cnt += fuse_brick<KF_VOXEL_SNORM16_U8>(a, ip, brick, lane, inv_count);
        e = ne;
        ++item;
    }
    add_touched(hd->touched, cnt, wave_cnt);
}

template <int FMT>
__global__ __launch_bounds__(256) void init_kernel(void* f, void* wt, int64_t count, float f0, float w0,
                                                   int* brick_key, int64_t nbricks, int* super_key, int nsuper)
{
    int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * 256;
    const float fb = (FMT == KF_VOXEL_F32) ? f0 : snorm_decode(snorm_encode(f0));
    for (int64_t b = q; b < nbricks; b += stride) brick_key[b] = order_key(fb);
    for (int64_t b = q; b < nsuper; b += stride) super_key[b] = order_key(fb);
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

__global__ void super_reset_kernel(int* super_key, int nsuper)
{
    int q = blockIdx.x * 256 + threadIdx.x;
    if (q < nsuper) super_key[q] = 0x7f800000;  // order_key(+inf)
}

// Exact halo minimum of F per brick (after an upload): one wave per brick
// takes the 9^3 voxels [8b, 8b + 8] per axis that lie in the volume; the
// super-brick's bound takes the smallest of its bricks'.
template <int FMT>
__global__ __launch_bounds__(256) void brick_bound_kernel(const void* f, int* brick_key, int64_t nbricks, int n,
                                                          int nb, int ns, int* super_key)
{
    const int64_t brick = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (brick >= nbricks) return;
    const int lane = threadIdx.x & 63;
    const int bi = (int)(brick % nb), bj = (int)((brick / nb) % nb), bk = (int)(brick / ((int64_t)nb * nb));
    float m = __int_as_float(0x7f800000);
    for (int q = lane; q < 9 * 9 * 9; q += 64) {
        const int i = bi * 8 + q % 9, j = bj * 8 + (q / 9) % 9, k = bk * 8 + q / 81;
        if (i >= n || j >= n || k >= n) continue;
        const int64_t idx = brick_index(nb, i, j, k);
        const float v = (FMT == KF_VOXEL_F32) ? static_cast<const float*>(f)[idx]
                                              : snorm_decode(static_cast<const int16_t*>(f)[idx]);
        m = fminf(m, v);
    }
    m = wave_min(m);
    if (lane == 0) {
        brick_key[brick] = order_key(m);
        atomicMin(&super_key[((bk >> 3) * ns + (bj >> 3)) * ns + (bi >> 3)], order_key(m));
    }
}

// z of each k slice: one lane, the host loop's repeated additions in order.
__global__ void z_table_kernel(float* zk, int n, float half, float res)
{
    float z = half + (-res);
    for (int k = 0; k < n; ++k) {
        z = z + res;
        zk[k] = z;
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
    if (st == KF_OK && to_volume) {
        const int64_t nbricks = (int64_t)v->nb * v->nb * v->nb;
        const unsigned b = (unsigned)((nbricks + 3) / 4);
        const int nsuper = v->ns * v->ns * v->ns;
        super_reset_kernel<<<(nsuper + 255) / 256, 256>>>(v->super_key, nsuper);
        if (v->format == KF_VOXEL_F32)
            brick_bound_kernel<KF_VOXEL_F32><<<b, 256>>>(v->f, v->brick_key, nbricks, v->n, v->nb, v->ns, v->super_key);
        else
            brick_bound_kernel<KF_VOXEL_SNORM16_U8><<<b, 256>>>(v->f, v->brick_key, nbricks, v->n, v->nb, v->ns,
                                                                v->super_key);
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) st = fail_hip(e, "brick_bound_kernel");
    }
    return st;
}

}  // namespace

FuseArgs fuse_args(const float* depth, int w, int h, const kf_volume* v, const float* K, const float* inv_pose,
                   const PoseState* dyn_state)
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
    a.inv_mue = 1.0f / v->mue;
    a.zk = v->zk;
    a.zt = v->zt;
    a.band = (float)std::max(w, h) * 0x1p-21f;
    a.f = v->f;
    a.wt = v->w;
    a.touched = v->touched;
    a.brick_key = v->brick_key;
    a.super_key = v->super_key;
    a.ns = v->ns;
    a.cull = is_pinhole(K) ? 1 : 0;
    const double W = (double)w, H = (double)h;
    const double pn[4] = {std::hypot((double)K[0], (double)K[6]), std::hypot((double)K[0], W - (double)K[6]),
                          std::hypot((double)K[4], (double)K[7]), std::hypot((double)K[4], H - (double)K[7])};
    for (int q = 0; q < 4; ++q) a.plane_norm[q] = std::nextafter((float)pn[q], INFINITY);
    a.format = v->format;
    a.dmax = nullptr;
    a.tx16 = (w + 15) / 16;
    a.ty16 = (h + 15) / 16;
    a.tx64 = (w + 63) / 64;
    a.ty64 = (h + 63) / 64;
    return a;
}

int launch_fuse_batch(const BatchHeader* hd, const FuseJob* jobs, unsigned cull_blocks, unsigned list_blocks,
                      hipStream_t s, hipEvent_t mid)
{
    fuse_batch_cull_kernel<<<cull_blocks, kBatchCullThreads, 0, s>>>(hd, jobs);
    KF_LAUNCHED("fuse_batch_cull_kernel");
    if (mid) KF_HIP(hipEventRecord(mid, s));
    fuse_batch_list_kernel<<<list_blocks, 64 * kFuseWaves, 0, s>>>(hd, jobs);
    KF_LAUNCHED("fuse_batch_list_kernel");
    return KF_OK;
}

// Workgroups of the list sweep that are resident at once on the current
// device (at most kFuseBlocks, the size of the per-workgroup counters).
static int fuse_list_blocks(int format)
{
    static int cached[2] = {0, 0};
    int& c = cached[format == KF_VOXEL_F32 ? 0 : 1];
    if (!c) {
        int dev = 0, cus = 0, per_cu = 0;
        const void* k = format == KF_VOXEL_F32 ? reinterpret_cast<const void*>(&fuse_list_kernel<KF_VOXEL_F32>)
                                               : reinterpret_cast<const void*>(&fuse_list_kernel<KF_VOXEL_SNORM16_U8>);
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64 * kFuseWaves, 0) == hipSuccess && per_cu > 0 &&
            cus > 0)
            c = std::min<int>(kFuseBlocks, per_cu * cus);
        else
            c = kFuseBlocks;
    }
    return c;
}

float* volume_depth_max(kf_volume* v, int w, int h, hipStream_t s)
{
    const size_t need = depth_max_floats(w, h);
    if (v->dmax_cap < need) {
        // grown outside the stream order: wait for work that may still read the old map
        if (hipStreamSynchronize(s) != hipSuccess) return nullptr;
        (void)hipFree(v->dmax);
        v->dmax = nullptr;
        v->dmax_cap = 0;
        if (hipMalloc(&v->dmax, sizeof(float) * need) != hipSuccess) return nullptr;
        v->dmax_cap = need;
    }
    return v->dmax;
}

int launch_fuse(const float* depth, int w, int h, kf_volume* v, const float* K, const float* inv_pose,
                const PoseState* dyn_state, hipStream_t s, hipEvent_t* list_events, bool dmax_ready)
{
    float* dmax = volume_depth_max(v, w, h, s);
    if (!dmax) return fail(KF_ERR_OOM, "fuse: depth tile-max maps");
    FuseArgs a = fuse_args(depth, w, h, v, K, inv_pose, dyn_state);
    a.dmax = dmax;
    if (!dmax_ready) {
        depth_max_kernel<<<dim3((unsigned)a.tx64, (unsigned)a.ty64), 256, 0, s>>>(depth, w, h, dmax, a.tx16, a.ty16,
                                                                                a.tx64);
        KF_LAUNCHED("depth_max_kernel");
    }
    const int64_t nbricks = (int64_t)v->nb * v->nb * v->nb;
    const int parity = (int)(v->fuse_launches & 1);
    fuse_cull_kernel<<<(unsigned)((nbricks + kCullThreads - 1) / kCullThreads), kCullThreads, 0, s>>>(
        a, nbricks, v->brick_list, v->list_count, parity);
    KF_LAUNCHED("fuse_cull_kernel");
    // every workgroup resident at once (no second round); fewer when the volume is small
    const unsigned blocks =
        (unsigned)std::min<int64_t>((nbricks + kFuseWaves - 1) / kFuseWaves, fuse_list_blocks(v->format));
    if (list_events) KF_HIP(hipEventRecord(list_events[0], s));
    if (v->format == KF_VOXEL_F32)
        fuse_list_kernel<KF_VOXEL_F32><<<blocks, 64 * kFuseWaves, 0, s>>>(a, v->brick_list, v->list_count + parity);
    else
        fuse_list_kernel<KF_VOXEL_SNORM16_U8><<<blocks, 64 * kFuseWaves, 0, s>>>(a, v->brick_list,
                                                                                 v->list_count + parity);
    KF_LAUNCHED("fuse_list_kernel");
    if (list_events) KF_HIP(hipEventRecord(list_events[1], s));
    v->fuse_launches += 1;
    return KF_OK;
}

int launch_initialize(kf_volume* v, float f, float wt, hipStream_t s)
{
    int64_t count = (int64_t)v->n * v->n * v->n;
    unsigned blocks = (unsigned)std::min<int64_t>((count + 255) / 256, 65536);
    const int64_t nbricks = (int64_t)v->nb * v->nb * v->nb;
    const int nsuper = v->ns * v->ns * v->ns;
    if (v->format == KF_VOXEL_F32)
        init_kernel<KF_VOXEL_F32><<<blocks, 256, 0, s>>>(v->f, v->w, count, f, wt, v->brick_key, nbricks, v->super_key,
                                                         nsuper);
    else
        init_kernel<KF_VOXEL_SNORM16_U8><<<blocks, 256, 0, s>>>(v->f, v->w, count, f, wt, v->brick_key, nbricks,
                                                                v->super_key, nsuper);
    KF_LAUNCHED("init_kernel");
    return KF_OK;
}

// Allocates a volume and enqueues its initialisation on `s` (no host wait):
// the z table by repeated addition and every voxel at (2, 0).
int volume_create(int n, float width_m, float mue_m, int format, hipStream_t s, kf_volume** out)
{
    if (!out) return fail(KF_ERR_ARG, "kf_volume_create: null out");
    *out = nullptr;
    if (n < 8 || n % 8 != 0) return fail(KF_ERR_ARG, "kf_volume_create: n must be a positive multiple of 8");
    if (n > 1024) return fail(KF_ERR_ARG, "kf_volume_create: n must be at most 1024");
    if (!(width_m > 0.0f) || !(width_m < 1e4f)) return fail(KF_ERR_ARG, "kf_volume_create: width must be in (0, 1e4) m");
    // the raycast marches at >= 0.1 * mue; 1e-4 m keeps every march finite and short
    if (!(mue_m >= 1e-4f)) return fail(KF_ERR_ARG, "kf_volume_create: mue must be >= 1e-4 m");
    if (format != KF_VOXEL_F32 && format != KF_VOXEL_SNORM16_U8)
        return fail(KF_ERR_ARG, "kf_volume_create: unknown voxel format");
    kf_volume* v = new kf_volume();
    v->n = n;
    v->nb = n / 8;
    v->ns = (v->nb + 7) / 8;
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
        (e = hipMalloc(&v->zt, sizeof(float4) * n)) != hipSuccess ||
        (e = hipMalloc(&v->touched, sizeof(unsigned long long) * kFuseBlocks)) != hipSuccess ||
        (e = hipMemsetAsync(v->touched, 0, sizeof(unsigned long long) * kFuseBlocks, s)) != hipSuccess ||
        (e = hipMalloc(&v->brick_list, sizeof(int) * (size_t)v->nb * v->nb * v->nb)) != hipSuccess ||
        (e = hipMalloc(&v->brick_key, sizeof(int) * (size_t)v->nb * v->nb * v->nb)) != hipSuccess ||
        (e = hipMalloc(&v->super_key, sizeof(int) * (size_t)v->ns * v->ns * v->ns)) != hipSuccess ||
        (e = hipMalloc(&v->list_count, 2 * sizeof(unsigned int))) != hipSuccess ||
        (e = hipMemsetAsync(v->list_count, 0, 2 * sizeof(unsigned int), s)) != hipSuccess) {
        kf_volume_destroy(v);
        return fail_hip(e, "kf_volume_create");
    }
    // tsdf.cu:27-31: z starts at -width/2 - res and advances by repeated addition
    z_table_kernel<<<1, 1, 0, s>>>(v->zk, n, -width_m * 0.5f, v->res);
    e = hipGetLastError();
    if (e != hipSuccess) {
        kf_volume_destroy(v);
        return fail_hip(e, "z_table_kernel");
    }
    // Voxel{} = (2, 0) (voxel_grid.h:10-16), as KinectFusion's constructor initialises it
    int st = launch_initialize(v, 2.0f, 0.0f, s);
    if (st != KF_OK) {
        kf_volume_destroy(v);
        return st;
    }
    *out = v;
    return KF_OK;
}

}  // namespace kf

// ------------------------------------------------------------------ C ABI
extern "C" {

int kf_volume_create(int n, float width_m, float mue_m, int format, kf_volume** out)
{
    int st = kf::volume_create(n, width_m, mue_m, format, nullptr, out);
    if (st) return st;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        kf_volume_destroy(*out);
        *out = nullptr;
        return kf::fail_hip(e, "initialize");
    }
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
    (void)hipFree(v->zt);
    (void)hipFree(v->touched);
    (void)hipFree(v->brick_list);
    (void)hipFree(v->brick_key);
    (void)hipFree(v->super_key);
    (void)hipFree(v->list_count);
    (void)hipFree(v->dmax);
    (void)hipFree(v->ray_list);
    (void)hipFree(v->ray_count);
    (void)hipFree(v->tile_order);
    (void)hipFree(v->order_state);
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
    std::vector<unsigned long long> part(kf::kFuseBlocks);
    KF_HIP(hipDeviceSynchronize());
    KF_HIP(hipMemcpy(part.data(), v->touched, sizeof(unsigned long long) * part.size(), hipMemcpyDeviceToHost));
    unsigned long long c = 0;
    for (auto x : part) c += x;
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
    if ((st = t.start()) || (st = kf::launch_fuse(depth_m, w, h, v, sensor->intr[0], sensor->inv_pose, nullptr, s, nullptr, false)))
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
