// Raycaster (project/src/raycast.cu): per pixel of a pyramid level, march the
// TSDF from the sensor pose, find the zero crossing and write the world-space
// vertex and the normal (-grad F) of the surface.
//
// The reference samples the volume through a CUDA texture (normalized
// coordinates, linear filter, clamp; voxel_grid.cpp:24-33).  Here the same
// texel-centre rule is evaluated in software on the bricked f plane:
//   x = u*n - 1/2, i = floor(x), a = x - i, indices clamped to [0, n-1],
// with exact FP32 weights (the texture unit keeps 8 fractional bits).
//
// Launch: one lane per pixel, 8x8 pixel tiles per 64-lane wave (rays of a wave
// are coherent, so their trilinear stencils share bricks and cache lines);
// tiles are assigned so that neighbouring tiles run on the same XCD (the 8
// XCDs each have their own L2 and receive workgroups round-robin).
#include <string.h>

#include "kf_device.hpp"
#include "kf_kernels.hpp"

namespace kf {

namespace {

constexpr int kMaxMarchSteps = 1 << 20;
// A sample whose 8 trilinear corners lie in one brick whose lower bound of F is
// at least this value is >= 0.99 (convex weights; 1e-3 of margin covers the
// rounding of the three lerp levels), so it can neither end the march nor
// shorten the step.
constexpr float kSkipBound = 0.991f;


// The fast skip test works on texel coordinates computed with a reciprocal
// instead of the exact division (a few 1e-5 voxel apart): a sample counts as
// inside a region only when its coordinate is at least kMargin voxel inside it,
// so its exact stencil is inside that region too.
constexpr float kMargin = 0.01f;


struct RaycastArgs {
    RaycastVolume vol;
    float pose[16];
    const PoseState* dyn;  // when set, the pose comes from dyn->pose
    RaycastLevel lv[3];
    int levels, tiles;
};

template <int FMT>
__device__ __forceinline__ float voxel_f(const void* f, int64_t q)
{
    if (FMT == KF_VOXEL_F32) return static_cast<const float*>(f)[q];
    return snorm_decode(static_cast<const int16_t*>(f)[q]);
}

// Texel coordinates of a trilinear sample at normalized (u, v, s):
// x = u*n - 1/2, i0 = floor(x) clamped to [0, n-1], i1 = i0 + 1 clamped.
struct Stencil {
    int i0, j0, k0, i1, j1, k1;
    float a, b, c;
};
__device__ __forceinline__ Stencil stencil(int n, float u, float v, float s)
{
    Stencil st;
    float x = u * (float)n - 0.5f, y = v * (float)n - 0.5f, z = s * (float)n - 0.5f;
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    st.a = x - fx;
    st.b = y - fy;
    st.c = z - fz;
    const float lo = -1.0f, hi = (float)n;
    fx = fx < lo ? lo : (fx > hi ? hi : fx);
    fy = fy < lo ? lo : (fy > hi ? hi : fy);
    fz = fz < lo ? lo : (fz > hi ? hi : fz);
    int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
    st.i1 = clampi(i0 + 1, 0, n - 1);
    st.j1 = clampi(j0 + 1, 0, n - 1);
    st.k1 = clampi(k0 + 1, 0, n - 1);
    st.i0 = clampi(i0, 0, n - 1);
    st.j0 = clampi(j0, 0, n - 1);
    st.k0 = clampi(k0, 0, n - 1);
    return st;
}

template <int FMT>
__device__ __forceinline__ float trilinear(const void* F, int nb, const Stencil& t)
{
    float c000 = voxel_f<FMT>(F, brick_index(nb, t.i0, t.j0, t.k0)), c100 = voxel_f<FMT>(F, brick_index(nb, t.i1, t.j0, t.k0));
    float c010 = voxel_f<FMT>(F, brick_index(nb, t.i0, t.j1, t.k0)), c110 = voxel_f<FMT>(F, brick_index(nb, t.i1, t.j1, t.k0));
    float c001 = voxel_f<FMT>(F, brick_index(nb, t.i0, t.j0, t.k1)), c101 = voxel_f<FMT>(F, brick_index(nb, t.i1, t.j0, t.k1));
    float c011 = voxel_f<FMT>(F, brick_index(nb, t.i0, t.j1, t.k1)), c111 = voxel_f<FMT>(F, brick_index(nb, t.i1, t.j1, t.k1));
    const float a = t.a, b = t.b, c = t.c;
    float e00 = (1.0f - a) * c000 + a * c100, e10 = (1.0f - a) * c010 + a * c110;
    float e01 = (1.0f - a) * c001 + a * c101, e11 = (1.0f - a) * c011 + a * c111;
    float g0 = (1.0f - b) * e00 + b * e10, g1 = (1.0f - b) * e01 + b * e11;
    return (1.0f - c) * g0 + c * g1;
}

template <int FMT>
__device__ __forceinline__ float sample_tsdf(const void* F, int n, int nb, float u, float v, float s)
{
    return trilinear<FMT>(F, nb, stencil(n, u, v, s));
}

// True when the sample is known to be >= 0.99 without reading the voxels.
__device__ __forceinline__ bool skippable(const RaycastVolume& a, const Stencil& t)
{
    if (!a.brick_min) return false;
    const int b0 = ((t.k0 >> 3) * a.nb + (t.j0 >> 3)) * a.nb + (t.i0 >> 3);
    const int b1 = ((t.k1 >> 3) * a.nb + (t.j1 >> 3)) * a.nb + (t.i1 >> 3);
    return b0 == b1 && a.brick_min[b0] >= kSkipBound;
}

// Conservative skip test on approximate texel coordinates (xa, ya, za).
// Returns 1 when the exact stencil lies inside one super-brick (or one brick)
// whose lower bound of F is >= kSkipBound, 0 when it lies inside one brick
// whose bound is below it (so skippable() would say no), -1 when the margins
// leave it open.  The last super-brick's and brick's bounds are cached in
// (cs, cb) and (ck, cbk), so a march through free space, or along a brick, reads
// no memory.
__device__ __forceinline__ int fast_skip(const RaycastVolume& a, float xa, float ya, float za, int& cs, float& cb,
                                         int& ck, float& cbk)
{
    const float hi = (float)a.n - 1.0f - kMargin;
    if (!(xa >= kMargin && ya >= kMargin && za >= kMargin && xa < hi && ya < hi && za < hi)) return -1;
    const int xi = (int)xa, yi = (int)ya, zi = (int)za;
    const int sx = xi >> 6, sy = yi >> 6, sz = zi >> 6;
    const float ux = xa - (float)(sx * kSuper), uy = ya - (float)(sy * kSuper), uz = za - (float)(sz * kSuper);
    const float shi = (float)(kSuper - 1) - kMargin;
    if (ux >= kMargin && uy >= kMargin && uz >= kMargin && ux < shi && uy < shi && uz < shi) {
        const int sidx = (sz * a.ns + sy) * a.ns + sx;
        if (sidx != cs) {
            cs = sidx;
            cb = order_value(a.super_key[sidx]);
        }
        if (cb >= kSkipBound) return 1;
    }
    const int bx = xi >> 3, by = yi >> 3, bz = zi >> 3;
    const float vx = xa - (float)(bx * kBrick), vy = ya - (float)(by * kBrick), vz = za - (float)(bz * kBrick);
    const float bhi = (float)(kBrick - 1) - kMargin;
    if (vx >= kMargin && vy >= kMargin && vz >= kMargin && vx < bhi && vy < bhi && vz < bhi) {
        const int bidx = (bz * a.nb + by) * a.nb + bx;
        if (bidx != ck) {
            ck = bidx;
            cbk = a.brick_min[bidx];
        }
        return cbk >= kSkipBound ? 1 : 0;
    }
    return -1;
}

// The corner values of the last trilinear cell a march read: consecutive small
// steps (0.1 mue) often stay in one cell, and then need no loads.
struct CellCache {
    int ki = -1, kj = -1, kk = -1;  // (i0 | i1 << 16) per axis
    float c[8];
};

template <int FMT>
__device__ __forceinline__ float trilinear_cached(const void* F, int nb, const Stencil& t, CellCache& cc)
{
    const int ki = t.i0 | (t.i1 << 16), kj = t.j0 | (t.j1 << 16), kk = t.k0 | (t.k1 << 16);
    if (ki != cc.ki || kj != cc.kj || kk != cc.kk) {
        cc.ki = ki;
        cc.kj = kj;
        cc.kk = kk;
        cc.c[0] = voxel_f<FMT>(F, brick_index(nb, t.i0, t.j0, t.k0));
        cc.c[1] = voxel_f<FMT>(F, brick_index(nb, t.i1, t.j0, t.k0));
        cc.c[2] = voxel_f<FMT>(F, brick_index(nb, t.i0, t.j1, t.k0));
        cc.c[3] = voxel_f<FMT>(F, brick_index(nb, t.i1, t.j1, t.k0));
        cc.c[4] = voxel_f<FMT>(F, brick_index(nb, t.i0, t.j0, t.k1));
        cc.c[5] = voxel_f<FMT>(F, brick_index(nb, t.i1, t.j0, t.k1));
        cc.c[6] = voxel_f<FMT>(F, brick_index(nb, t.i0, t.j1, t.k1));
        cc.c[7] = voxel_f<FMT>(F, brick_index(nb, t.i1, t.j1, t.k1));
    }
    const float a = t.a, b = t.b, c = t.c;
    float e00 = (1.0f - a) * cc.c[0] + a * cc.c[1], e10 = (1.0f - a) * cc.c[2] + a * cc.c[3];
    float e01 = (1.0f - a) * cc.c[4] + a * cc.c[5], e11 = (1.0f - a) * cc.c[6] + a * cc.c[7];
    float g0 = (1.0f - b) * e00 + b * e10, g1 = (1.0f - b) * e01 + b * e11;
    return (1.0f - c) * g0 + c * g1;
}

// A ray that has taken kSmallBudget small steps is handed to the whole wave
// once no more than kCoopLanes lanes of the wave are still marching (about 1
// ray in 40 grazes a surface and takes hundreds of small steps; while many
// lanes march, the wave is busy anyway).
constexpr int kSmallBudget = 32;
constexpr int kCoopLanes = 16;

// One trilinear evaluation of the march at t: 1 = skipped (the sample is known
// to be >= 0.99), 0 = evaluated into f.
template <int FMT>
__device__ __forceinline__ int march_sample(const RaycastVolume& a, v3 o, v3 dir, float t, float width, float invw,
                                            float nf, int& cs, float& cb, int& ck, float& cbk, CellCache& cc,
                                            float& f)
{
    const v3 p = add3(o, mul3s(dir, t));
    const int fs = a.brick_min ? fast_skip(a, (p.x * invw + 0.5f) * nf - 0.5f, (p.y * invw + 0.5f) * nf - 0.5f,
                                           (p.z * invw + 0.5f) * nf - 0.5f, cs, cb, ck, cbk)
                               : 0;
    if (fs == 1) return 1;
    const Stencil st = stencil(a.n, p.x / width + 0.5f, p.y / width + 0.5f, p.z / width + 0.5f);
    if (fs < 0 && skippable(a, st)) return 1;
    f = trilinear_cached<FMT>(a.f, a.nb, st, cc);
    return 0;
}

__device__ __forceinline__ float lane_bcast(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// raycast.cu:44-145 for pixel (i, j) of one level map.  Called by all 64
// lanes of a wave together (in_image false for lanes past the map's edge).
// Phase 1: each lane marches its own ray; a long ray is deferred once few
// lanes are left marching (kSmallBudget, kCoopLanes).  Phase 2: the wave marches each
// deferred ray 64 samples at a time, lane k taking sample k of the chunk (its
// position by the same repeated additions t = t + step), and a ballot finds
// the first negative sample.  Every sample position, value and decision is the
// sequential march's, so the output is the same.
template <int FMT>
__device__ __forceinline__ void raycast_pixel(const RaycastVolume& a, const float* pose, const RaycastLevel& L, int i,
                                              int j, bool in_image)
{
    const int lane = threadIdx.x & 63;
    float Rm[9] = {pose[0], pose[1], pose[2], pose[4], pose[5], pose[6], pose[8], pose[9], pose[10]};
    float M[9];
    m3m3(Rm, L.Kinv, M);
    const v3 o = mk3(pose[12], pose[13], pose[14]);
    const float width = a.width, mue = a.mue, hw = width * 0.5f;

    v3 dir = m3v(M, mk3((float)i + 0.5f, (float)j + 0.5f, 1.0f));
    v3 inv = mk3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    // raycast.cu:7-27
    float t0x = (-hw - o.x) * inv.x, t1x = (hw - o.x) * inv.x;
    float t0y = (-hw - o.y) * inv.y, t1y = (hw - o.y) * inv.y;
    float t0z = (-hw - o.z) * inv.z, t1z = (hw - o.z) * inv.z;
    float tmin = fminf(t0x, t1x), tmax = fmaxf(t0x, t1x);
    tmin = fmaxf(tmin, fminf(t0y, t1y));
    tmax = fminf(tmax, fmaxf(t0y, t1y));
    tmin = fmaxf(tmin, fminf(t0z, t1z));
    tmax = fminf(tmax, fmaxf(t0z, t1z));
    if (tmax < tmin) {
        tmin = -1.0f;
        tmax = -1.0f;
    }
    const float nearv = (kMinDistance < tmin) ? tmin : kMinDistance;
    const float farv = (tmax < kMaxDistance) ? tmax : kMaxDistance;
    // raycast.cu:61-73: a ray behind the camera or with near >= far leaves the pixel untouched
    const bool active = in_image && !(tmax < 0.0f) && nearv < farv;

    // raycast.cu:75-117.  A skipped sample only sets prev; its value is
    // computed if and when the zero-crossing formula needs it.
    const float large = mue * 1.99f, small = mue * 0.1f;
    const float invw = 1.0f / width, nf = (float)a.n;
    float step = large, prev = 0.0f, tstar = 0.0f, prev_t = 0.0f, t = nearv;
    bool prev_skipped = false, deferred = false;
    int cs = -1, ck = -1;
    float cb = 0.0f, cbk = 0.0f;
    CellCache cc;
    int guard = 0;  // never binds for mue >= 1e-4 m (kf_volume_create enforces it)
    if (active) {
        int smalls = 0;
        for (; t < farv; t = t + step) {
            if (++guard > kMaxMarchSteps) break;
            const int marching = __popcll(__ballot(1));  // lanes of the wave still in this loop
            if (step == small && ++smalls > kSmallBudget && marching <= kCoopLanes) {
                deferred = true;  // this sample (guard already counted) is the wave's first
                break;
            }
            float f = 0.0f;
            if (march_sample<FMT>(a, o, dir, t, width, invw, nf, cs, cb, ck, cbk, cc, f)) {
                prev_skipped = true;
                prev_t = t;
                continue;
            }
            if (f < 0.0f) {
                if (t != nearv) {
                    if (step == large) {
                        t = t - mue;
                        step = small;
                        continue;
                    }
                    if (prev_skipped) {
                        v3 pp = add3(o, mul3s(dir, prev_t));
                        prev = sample_tsdf<FMT>(a.f, a.n, a.nb, pp.x / width + 0.5f, pp.y / width + 0.5f,
                                                pp.z / width + 0.5f);
                    }
                    tstar = t - step * prev / (f - prev);
                }
                break;
            } else if (f < 0.99f) {
                step = small;
            }
            prev = f;
            prev_skipped = false;
        }
    }

    // phase 2: the deferred rays of the wave, one at a time, 64 samples per chunk
    uint64_t pending = __ballot(deferred);
    while (pending) {
        const int r = __ffsll((unsigned long long)pending) - 1;
        pending &= pending - 1;
        const v3 dr = mk3(lane_bcast(dir.x, r), lane_bcast(dir.y, r), lane_bcast(dir.z, r));
        const float fr = lane_bcast(farv, r);
        float tb = lane_bcast(t, r);  // position of sample 0 of the chunk
        float pv = lane_bcast(prev, r), pt = lane_bcast(prev_t, r);
        int psk = __builtin_amdgcn_readlane(prev_skipped ? 1 : 0, r);
        int gb = __builtin_amdgcn_readlane(guard, r);  // guard count of sample 0
        float ts = 0.0f;
        for (;;) {
            float tk = tb;
            for (int q = 0; q < lane; ++q) tk = tk + small;
            // sample k exists if every sample up to it passes the loop test and the guard
            const bool end = !(tk < fr) || gb + lane > kMaxMarchSteps;
            float f = 0.0f;
            int sk = 1;
            if (!end) sk = march_sample<FMT>(a, o, dr, tk, width, invw, nf, cs, cb, ck, cbk, cc, f);
            const bool neg = !end && !sk && f < 0.0f;
            const uint64_t ev = __ballot(end || neg);
            if (ev) {
                const int k = __ffsll((unsigned long long)ev) - 1;
                const bool kneg = __builtin_amdgcn_readlane(neg ? 1 : 0, k) != 0;
                if (kneg) {
                    // prev = the value of sample k - 1 (exactly, also if it was skipped)
                    float pk;
                    if (k > 0) {
                        float fe = f;
                        if (lane == k - 1 && sk) {
                            v3 pp = add3(o, mul3s(dr, tk));
                            fe = sample_tsdf<FMT>(a.f, a.n, a.nb, pp.x / width + 0.5f, pp.y / width + 0.5f,
                                                  pp.z / width + 0.5f);
                        }
                        pk = lane_bcast(fe, k - 1);
                    } else if (psk) {
                        v3 pp = add3(o, mul3s(dr, pt));
                        pk = sample_tsdf<FMT>(a.f, a.n, a.nb, pp.x / width + 0.5f, pp.y / width + 0.5f,
                                              pp.z / width + 0.5f);
                    } else {
                        pk = pv;
                    }
                    const float tsk = tk - small * pk / (f - pk);
                    ts = lane_bcast(tsk, k);
                }
                break;
            }
            // no event: the chunk's last sample becomes prev, the next chunk follows it
            const float t63 = lane_bcast(tk, 63);
            psk = __builtin_amdgcn_readlane(sk, 63);
            pv = lane_bcast(f, 63);
            pt = t63;
            tb = t63 + small;
            gb += 64;
        }
        if (lane == r) tstar = ts;
    }

    if (!active) return;
    const int64_t q = (int64_t)j * L.w + i;
    if (tstar != 0.0f) {
        v3 vtx = add3(o, mul3s(dir, tstar));
        // raycast.cu:29-42
        float r = mue * 0.5f / width;
        float u = vtx.x / width + 0.5f, v = vtx.y / width + 0.5f, s = vtx.z / width + 0.5f;
        // one axis at a time (16 loads in flight, not 48): fewer live registers,
        // more waves per SIMD for the march.  x + 0 == x, so the samples are the
        // six of raycast.cu:33-38.
        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll 1
        for (int ax = 0; ax < 3; ++ax) {
            const float du = ax == 0 ? r : 0.0f, dv = ax == 1 ? r : 0.0f, ds = ax == 2 ? r : 0.0f;
            const float fp = sample_tsdf<FMT>(a.f, a.n, a.nb, u + du, v + dv, s + ds);
            const float fm = sample_tsdf<FMT>(a.f, a.n, a.nb, u - du, v - dv, s - ds);
            const float d = fm - fp;
            gx = ax == 0 ? d : gx;
            gy = ax == 1 ? d : gy;
            gz = ax == 2 ? d : gz;
        }
        v3 g = div3s(mk3(gx, gy, gz), r);
        float gn = len3(g);
        if (is_depth_valid(gn)) {
            v3 nn = div3s(g, gn);
            L.vout[q] = make_float4(vtx.x, vtx.y, vtx.z, kValid);
            L.nout[q] = make_float4(nn.x, nn.y, nn.z, kValid);
            return;
        }
    }
    L.vout[q] = make_float4(0.0f, 0.0f, 0.0f, kInvalid);
    L.nout[q] = make_float4(0.0f, 0.0f, 0.0f, kInvalid);
}

// XCD-aware tile order: workgroup g runs on XCD g % 8; give each XCD a
// contiguous band of tiles.
__device__ __forceinline__ int xcd_tile(int tiles)
{
    const int g = blockIdx.x;
    const int per_xcd = (tiles + 7) / 8;
    return (g % 8) * per_xcd + g / 8;
}

template <int FMT>
__global__ __launch_bounds__(64, 7) void raycast_kernel(RaycastArgs a)
{
    const int tile = xcd_tile(a.tiles);
    if (tile >= a.tiles) return;
    int l = 0;
    while (l + 1 < a.levels && tile >= a.lv[l + 1].tile0) ++l;
    const RaycastLevel& L = a.lv[l];
    const int lt = tile - L.tile0;
    const int i = (lt % L.tiles_x) * 8 + (threadIdx.x & 7);
    const int j = (lt / L.tiles_x) * 8 + (threadIdx.x >> 3);
    raycast_pixel<FMT>(a.vol, a.dyn ? a.dyn->pose : a.pose, L, i, j, i < L.w && j < L.h);
}

// Field-wise copy of a job read through the constant address space.
template <class CJ>
__device__ __forceinline__ void copy_job(const CJ& j, RaycastVolume& vol, RaycastLevel& L)
{
    vol.f = j.vol.f;
    vol.brick_min = j.vol.brick_min;
    vol.super_key = j.vol.super_key;
    vol.n = j.vol.n;
    vol.nb = j.vol.nb;
    vol.ns = j.vol.ns;
    vol.format = j.vol.format;
    vol.width = j.vol.width;
    vol.mue = j.vol.mue;
#pragma unroll
    for (int q = 0; q < 9; ++q) L.Kinv[q] = j.L.Kinv[q];
    L.vout = j.L.vout;
    L.nout = j.L.nout;
    L.w = j.L.w;
    L.h = j.L.h;
    L.tiles_x = j.L.tiles_x;
    L.tile0 = j.L.tile0;
}

// Many sessions and levels per persistent launch, one launch per voxel format:
// the jobs of format FMT are jobs[j0, j1), covering tiles [t0, t1) of the tick
// (jobs[k] covers [jobs[k].L.tile0, jobs[k+1].L.tile0)).  March lengths vary
// by an order of magnitude between tiles (rays grazing a surface take hundreds
// of small steps), so tiles are claimed dynamically: the range is cut into 8
// bands, workgroups on XCD x claim pairs of tiles of band x from its counter
// (locality in that XCD's L2) and then help the other bands until all are done.
constexpr int kClaim = 2;  // tiles per claim

template <int FMT>
__global__ __launch_bounds__(64, 7) void raycast_batch_kernel(const BatchHeader* hd, const RaycastJob* jobs)
{
    const int nf = hd->n_ray_f32, tf = hd->ray_tiles_f32;
    const int j0 = FMT == KF_VOXEL_F32 ? 0 : nf, j1 = FMT == KF_VOXEL_F32 ? nf : hd->n_ray;
    const int t0 = FMT == KF_VOXEL_F32 ? 0 : tf, t1 = FMT == KF_VOXEL_F32 ? tf : hd->ray_tiles;
    if (t1 <= t0) return;
    unsigned* next = hd->ray_next + (FMT == KF_VOXEL_F32 ? 0 : 8);
    const int per = (t1 - t0 + 7) / 8;
    const int xcd = (int)(blockIdx.x & 7);
    // The table is read through the constant address space: it is never written
    // while the kernel runs, so its fields load with scalar loads into scalar
    // registers even after the kernel's stores (the vertex and normal maps).
    typedef __attribute__((address_space(4))) const RaycastJob ConstJob;
    ConstJob* cj = (ConstJob*)jobs;
    for (int bi = 0; bi < 8; ++bi) {
        const int band = (xcd + bi) & 7;
        const int b0 = t0 + band * per, b1 = min(t1, b0 + per);
        for (;;) {
            int c = 0;
            if (threadIdx.x == 0) c = (int)atomicAdd(&next[band], 1u);
            c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, 64));
            const int tc = b0 + c * kClaim;
            if (tc >= b1) break;
            for (int t = tc; t < min(b1, tc + kClaim); ++t) {
                int lo = j0, hi = j1 - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (cj[mid].L.tile0 <= t) lo = mid;
                    else hi = mid - 1;
                }
                const int jb = __builtin_amdgcn_readfirstlane(lo);
                RaycastVolume vol;
                RaycastLevel L;
                copy_job(cj[jb], vol, L);
                const float* pose = cj[jb].pose;
                const int lt = t - L.tile0;
                const int i = (lt % L.tiles_x) * 8 + (threadIdx.x & 7);
                const int j = (lt / L.tiles_x) * 8 + (threadIdx.x >> 3);
                raycast_pixel<FMT>(vol, pose, L, i, j, i < L.w && j < L.h);
            }
        }
    }
}

}  // namespace

RaycastVolume raycast_volume(const kf_volume* v)
{
    RaycastVolume r{};
    r.f = v->f;
    r.brick_min = v->brick_min;
    r.super_key = v->super_key;
    r.n = v->n;
    r.nb = v->nb;
    r.ns = v->ns;
    r.format = v->format;
    r.width = v->width;
    r.mue = v->mue;
    return r;
}

int launch_raycast_batch(const BatchHeader* hd, const RaycastJob* jobs, unsigned blocks, hipStream_t s)
{
    raycast_batch_kernel<KF_VOXEL_F32><<<blocks, 64, 0, s>>>(hd, jobs);
    KF_LAUNCHED("raycast_batch_kernel");
    raycast_batch_kernel<KF_VOXEL_SNORM16_U8><<<blocks, 64, 0, s>>>(hd, jobs);
    KF_LAUNCHED("raycast_batch_kernel");
    return KF_OK;
}

int launch_raycast_levels(const kf_volume* v, const float* pose, const PoseState* dyn_state, int levels,
                          const float (*inv_k)[9], float4* const* vout, float4* const* nout, int w0, int h0,
                          hipStream_t s)
{
    RaycastArgs a{};
    a.vol = raycast_volume(v);
    if (pose) memcpy(a.pose, pose, sizeof(a.pose));
    a.dyn = dyn_state;
    a.levels = levels;
    int tiles = 0;
    for (int l = 0; l < levels; ++l) {
        RaycastLevel& L = a.lv[l];
        memcpy(L.Kinv, inv_k[l], sizeof(L.Kinv));
        L.vout = vout[l];
        L.nout = nout[l];
        L.w = w0 >> l;
        L.h = h0 >> l;
        L.tiles_x = (L.w + 7) / 8;
        L.tile0 = tiles;
        tiles += L.tiles_x * ((L.h + 7) / 8);
    }
    a.tiles = tiles;
    const unsigned blocks = (unsigned)(((tiles + 7) / 8) * 8);
    if (v->format == KF_VOXEL_F32)
        raycast_kernel<KF_VOXEL_F32><<<blocks, 64, 0, s>>>(a);
    else
        raycast_kernel<KF_VOXEL_SNORM16_U8><<<blocks, 64, 0, s>>>(a);
    KF_LAUNCHED("raycast_kernel");
    return KF_OK;
}

int launch_raycast(const kf_volume* v, const float* pose, const PoseState* dyn_state, const float* inv_k,
                   float4* vout, float4* nout, int w, int h, hipStream_t s)
{
    float ki[1][9];
    memcpy(ki[0], inv_k, sizeof(ki[0]));
    return launch_raycast_levels(v, pose, dyn_state, 1, ki, &vout, &nout, w, h, s);
}

}  // namespace kf

extern "C" int kf_raycast(const kf_volume* v, const kf_sensor* sensor, float* vertex_map, float* normal_map, int w,
                          int h, int level, kf_stream_t stream, float* elapsed_ms)
{
    if (!v || !sensor || !vertex_map || !normal_map || w <= 0 || h <= 0)
        return kf::fail(KF_ERR_ARG, "raycast: bad argument");
    if (level < 0 || level > 2) return kf::fail(KF_ERR_ARG, "raycast: level must be 0..2 (sensor.h:51-52)");
    hipStream_t s = kf::stream_of(stream);
    kf::CallTimer t(elapsed_ms, s);
    int st;
    if ((st = t.start()) ||
        (st = kf::launch_raycast(v, sensor->pose, nullptr, sensor->inv_intr[level], reinterpret_cast<float4*>(vertex_map),
                                 reinterpret_cast<float4*>(normal_map), w, h, s)))
        return st;
    return t.finish();
}
