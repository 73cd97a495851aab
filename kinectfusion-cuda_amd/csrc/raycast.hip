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
// Skipping: a brick's bound covers every stencil whose first corner lies in it
// (its halo, kf_device.hpp), a super-brick's the stencils inside it.  On a
// large step a sample whose bound is at least kSkipBound is >= 0.99 (convex
// weights; 1e-3 of margin covers the rounding of the three lerp levels), so
// it can neither end the march nor shorten the step.  On a small step only a
// negative sample changes the march (the step is already small), and a sample
// whose bound is >= 0 is >= 0 (non-negative corners and weights give a
// non-negative value in any rounding), so kSmallSkipBound = 0 applies there.
// A skipped sample only sets prev_t; its value is computed if and when the
// zero-crossing formula needs it.
constexpr float kSkipBound = 0.991f;
constexpr float kSmallSkipBound = 0.0f;


// The fast skip test works on texel coordinates computed with a reciprocal
// instead of the exact division (a few 1e-5 voxel apart): a sample counts as
// inside a region only when its coordinate is at least kMargin voxel inside it,
// so its exact stencil is inside that region too.
constexpr float kMargin = 0.01f;


// Dispatch order of the first pass.  Workgroup g of a level runs on XCD
// g % 8 and takes the (g / 8)-th tile of its level's band on that XCD; which
// tile that is comes from the previous launch: there, every workgroup timed
// its wave and appended its tile to the next launch's order of its band, at
// the front when the wave took longer than the level's mean, else at the back,
// so the tiles that marched longest start first (they made the kernel's tail
// when they started late).  The mean is estimated from one wave in eight,
// summed per level and XCD (one shared sum per level made every wave wait on
// the same contended atomic).  Only the order changes, never a result.
struct TileOrderState {
    unsigned ctr[2][3][8][2];         // [parity][level][XCD][front, back] appends
    unsigned long long dur[3][3][8];  // [launch % 3][level][XCD] sum of sampled wave times
    unsigned waves[3][3][8];          // [launch % 3][level][XCD] waves sampled
};
struct RaycastArgs {
    RaycastVolume vol;
    float pose[16];
    const PoseState* dyn;  // when set, the pose comes from dyn->pose
    RaycastLevel lv[3];
    int levels, tiles;
    DeferredRay* rays;  // rays handed to the finishing pass (at most cap)
    unsigned* count;    // their count for this launch
    unsigned* clear;    // the count of the other launch parity, cleared for the next launch
    unsigned cap;
    const int* order_in;   // band slot of each workgroup, from the previous launch (null: g / 8)
    int* order_out;        // the same for the next launch (null: not recorded)
    TileOrderState* ost;
    int parity, tri;       // launch index mod 2 and mod 3
};


// Normalized texture coordinate x / width + 1/2 of a world coordinate.  For a
// power-of-two width the quotient is the product with the exact reciprocal
// (the same real number, so the same rounding), which skips the division; the
// product is exact (or far below the half's ulp), so fusing it with the
// addition rounds the same way.
__device__ __forceinline__ float tex(const RaycastVolume& a, float x)
{
    return a.inv_width != 0.0f ? fmaf(x, a.inv_width, 0.5f) : x / a.width + 0.5f;
}

// Approximate texel coordinate of a world coordinate for the skip tests only
// (they allow kMargin of error; this is within a few 1e-5 texel).
__device__ __forceinline__ float approx_texel(float x, float invw, float nf)
{
    return fmaf(fmaf(x, invw, 0.5f), nf, -0.5f);
}

// Texel coordinates of a trilinear sample at normalized (u, v, s):
// x = u*n - 1/2, i0 = floor(x) clamped to [0, n-1], i1 = i0 + 1 clamped.
struct Stencil {
    int i0, j0, k0, i1, j1, k1;
    float a, b, c;
};
// The floors are integers (sample positions are finite), so clamping them, and
// them plus one (exact below 2^24; beyond, both clamp to the same end), to
// [0, n-1] in float and converting gives the clamped indices: one median of
// three and a conversion per index.
__device__ __forceinline__ Stencil stencil_texel(int n, float x, float y, float z)
{
    Stencil st;
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    st.a = x - fx;
    st.b = y - fy;
    st.c = z - fz;
    const float top = (float)(n - 1);
    st.i0 = (int)__builtin_amdgcn_fmed3f(fx, 0.0f, top);
    st.j0 = (int)__builtin_amdgcn_fmed3f(fy, 0.0f, top);
    st.k0 = (int)__builtin_amdgcn_fmed3f(fz, 0.0f, top);
    st.i1 = (int)__builtin_amdgcn_fmed3f(fx + 1.0f, 0.0f, top);
    st.j1 = (int)__builtin_amdgcn_fmed3f(fy + 1.0f, 0.0f, top);
    st.k1 = (int)__builtin_amdgcn_fmed3f(fz + 1.0f, 0.0f, top);
    return st;
}
// From normalized coordinates: x = u*n - 1/2 rounded twice; for a power-of-two
// n the product is exact, so one fused multiply-add rounds the same way (P2).
template <bool P2>
__device__ __forceinline__ Stencil stencil(int n, float u, float v, float s)
{
    const float nf = (float)n;
    if (P2) return stencil_texel(n, fmaf(u, nf, -0.5f), fmaf(v, nf, -0.5f), fmaf(s, nf, -0.5f));
    return stencil_texel(n, u * nf - 0.5f, v * nf - 0.5f, s * nf - 0.5f);
}

// Voxel q of the f plane (32-bit element offset from the plane's base).
template <int FMT>
__device__ __forceinline__ float voxel_at(const void* f, uint32_t q)
{
    const char* base = static_cast<const char*>(f);
    if (FMT == KF_VOXEL_F32) return *reinterpret_cast<const float*>(base + (size_t)(q * 4u));
    return snorm_decode(*reinterpret_cast<const int16_t*>(base + (size_t)(q * 2u)));
}

// Element offsets of the 8 corners of a stencil's cell, in the order
// c000 c100 c010 c110 c001 c101 c011 c111 (i fastest).
__device__ __forceinline__ void cell_offsets(int nb, const Stencil& t, uint32_t* o)
{
    const uint32_t x0 = brick_ox(t.i0), x1 = brick_ox(t.i1);
    const uint32_t y0 = brick_oy(nb, t.j0), y1 = brick_oy(nb, t.j1);
    const uint32_t z0 = brick_oz(nb, t.k0), z1 = brick_oz(nb, t.k1);
    const uint32_t y0z0 = y0 + z0, y1z0 = y1 + z0, y0z1 = y0 + z1, y1z1 = y1 + z1;
    o[0] = x0 + y0z0;
    o[1] = x1 + y0z0;
    o[2] = x0 + y1z0;
    o[3] = x1 + y1z0;
    o[4] = x0 + y0z1;
    o[5] = x1 + y0z1;
    o[6] = x0 + y1z1;
    o[7] = x1 + y1z1;
}

template <int FMT>
__device__ __forceinline__ float trilinear(const void* F, int nb, const Stencil& t)
{
    uint32_t o[8];
    cell_offsets(nb, t, o);
    float c000 = voxel_at<FMT>(F, o[0]), c100 = voxel_at<FMT>(F, o[1]);
    float c010 = voxel_at<FMT>(F, o[2]), c110 = voxel_at<FMT>(F, o[3]);
    float c001 = voxel_at<FMT>(F, o[4]), c101 = voxel_at<FMT>(F, o[5]);
    float c011 = voxel_at<FMT>(F, o[6]), c111 = voxel_at<FMT>(F, o[7]);
    const float a = t.a, b = t.b, c = t.c;
    float e00 = (1.0f - a) * c000 + a * c100, e10 = (1.0f - a) * c010 + a * c110;
    float e01 = (1.0f - a) * c001 + a * c101, e11 = (1.0f - a) * c011 + a * c111;
    float g0 = (1.0f - b) * e00 + b * e10, g1 = (1.0f - b) * e01 + b * e11;
    return (1.0f - c) * g0 + c * g1;
}

template <int FMT, bool P2>
__device__ __forceinline__ float sample_tsdf(const void* F, int n, int nb, float u, float v, float s)
{
    return trilinear<FMT>(F, nb, stencil<P2>(n, u, v, s));
}

// Brick holding the stencil's first corner (its halo bound covers the stencil).
__device__ __forceinline__ int stencil_brick(const RaycastVolume& a, const Stencil& t)
{
    return grid_index(t.i0 >> 3, t.j0 >> 3, t.k0 >> 3, a.nb);
}

// True when the sample's bound is >= thr (the exact stencil), through the
// brick cache (ck, cbk).
__device__ __forceinline__ bool skippable(const RaycastVolume& a, const Stencil& t, float thr, int& ck, float& cbk)
{
    if (!a.brick_key) return false;
    const int b0 = stencil_brick(a, t);
    if (b0 != ck) {
        ck = b0;
        cbk = order_value(a.brick_key[b0]);
    }
    return cbk >= thr;
}

// Conservative skip test on approximate texel coordinates (xa, ya, za).
// Returns 1 when the exact stencil lies inside one super-brick, or has its
// first corner in one brick, whose bound of F is >= thr, 0 when its first
// corner is in one brick whose bound is below it (so skippable() would say
// no), -1 when the margins leave it open.  The last super-brick's and brick's
// bounds are cached in (cs, cb) and (ck, cbk), so a march through free space,
// or along a brick, reads no memory.  (A stencil inside a free super-brick whose
// brick's halo reaches into a neighbour with a lower bound is then evaluated
// when that brick is the cached one: conservative.)
__device__ __forceinline__ int fast_skip(const RaycastVolume& a, float xa, float ya, float za, float thr, int& cs,
                                         float& cb, int& ck, float& cbk)
{
    const float hi = (float)a.n - 1.0f - kMargin;
    if (!(xa >= kMargin && ya >= kMargin && za >= kMargin && xa < hi && ya < hi && za < hi)) return -1;
    const int xi = (int)xa, yi = (int)ya, zi = (int)za;
    // the exact first corner is in brick (bx, by, bz) when the coordinate is
    // kMargin inside the brick's [0, 8) span on every axis
    const int bx = xi >> 3, by = yi >> 3, bz = zi >> 3;
    const float vx = xa - (float)(bx * kBrick), vy = ya - (float)(by * kBrick), vz = za - (float)(bz * kBrick);
    const float bhi = (float)kBrick - kMargin;
    const bool in_brick = vx >= kMargin && vy >= kMargin && vz >= kMargin && vx < bhi && vy < bhi && vz < bhi;
    const int bidx = grid_index(bx, by, bz, a.nb);
    // the cached brick decides first (along a brick or near a surface, where the
    // super-brick test rarely passes): no memory read and no super-brick test
    if (in_brick && bidx == ck) return cbk >= thr ? 1 : 0;
    const int sx = xi >> 6, sy = yi >> 6, sz = zi >> 6;
    const float ux = xa - (float)(sx * kSuper), uy = ya - (float)(sy * kSuper), uz = za - (float)(sz * kSuper);
    const float shi = (float)(kSuper - 1) - kMargin;
    if (ux >= kMargin && uy >= kMargin && uz >= kMargin && ux < shi && uy < shi && uz < shi) {
        const int sidx = grid_index(sx, sy, sz, a.ns);
        if (sidx != cs) {
            cs = sidx;
            cb = order_value(a.super_key[sidx]);
        }
        if (cb >= thr) return 1;
    }
    if (in_brick) {
        ck = bidx;
        cbk = order_value(a.brick_key[bidx]);
        return cbk >= thr ? 1 : 0;
    }
    return -1;
}

// Samples after the one at t (same step, t + k*step by repeated addition)
// whose stencils certainly stay in the texel region [lo, hi] per axis (the
// first corner at most hi - 1).  The march's approximate texel coordinates are
// linear in t up to rounding: t stays below 6 m, so each addition is off by at
// most 2^-22 m, and k of them by k * 2^-22 m * (texels per metre); k is capped
// so that this stays below kMargin / 2, the coordinates' own rounding is a few
// 1e-4 texel.  A sample whose linear coordinate is 2 kMargin inside the region
// thus has a computed coordinate kMargin inside it, and fast_skip's tests hold
// for it.  The step count is formed with the hardware reciprocal: a relative
// error of a few 1e-7 on at most 64 texels of distance moves the last counted
// sample by under 1e-4 texel, far inside the kMargin / 2 left for it.  -1 when
// the sample at t itself is not 2 kMargin inside.
__device__ __forceinline__ int region_run(v3 o, v3 dir, float t, float step, float invw, float nf, const int* lo,
                                          const int* hi)
{
    const v3 p = add3(o, mul3s(dir, t));
    const float pos[3] = {approx_texel(p.x, invw, nf), approx_texel(p.y, invw, nf), approx_texel(p.z, invw, nf)};
    const float d[3] = {dir.x, dir.y, dir.z};
    const float m2 = 2.0f * kMargin;
    float kmax = 20000.0f / (nf * invw);  // k * 2^-22 * texels per metre <= kMargin / 2
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float l = (float)lo[ax] + m2, h = (float)hi[ax] - m2;
        if (!(pos[ax] >= l && pos[ax] <= h)) return -1;
        const float rate = d[ax] * invw * nf * step;  // texels per step
        const float inv = __builtin_amdgcn_rcpf(rate);
        if (rate > 0.0f) kmax = fminf(kmax, (h - pos[ax]) * inv);
        else if (rate < 0.0f) kmax = fminf(kmax, (l - pos[ax]) * inv);
    }
    return (int)kmax;
}

// Samples after the one at t inside the cached super-brick, which covers
// stencils inside it; 0 when its bound is below thr or the sample at t is
// near its faces.  The super-brick's coordinates are those of the sample's
// position when its index is the cached one (region_run checks the margins).
__device__ __forceinline__ int super_run(const RaycastVolume& a, v3 o, v3 dir, float t, float step, float invw,
                                         float nf, float thr, int cs, float cb)
{
    if (cs < 0 || !(cb >= thr)) return 0;
    const v3 p = add3(o, mul3s(dir, t));
    const int sc[3] = {(int)approx_texel(p.x, invw, nf) / kSuper, (int)approx_texel(p.y, invw, nf) / kSuper,
                       (int)approx_texel(p.z, invw, nf) / kSuper};
    if (sc[0] < 0 || sc[1] < 0 || sc[2] < 0 || grid_index(sc[0], sc[1], sc[2], a.ns) != cs) return 0;
    int lo[3], hi[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        lo[ax] = sc[ax] * kSuper;
        hi[ax] = min(sc[ax] * kSuper + kSuper - 1, a.n - 1);
    }
    const int k = region_run(o, dir, t, step, invw, nf, lo, hi);
    return k > 0 ? k : 0;
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
        uint32_t o[8];
        cell_offsets(nb, t, o);
#pragma unroll
        for (int q = 0; q < 8; ++q) cc.c[q] = voxel_at<FMT>(F, o[q]);
    }
    const float a = t.a, b = t.b, c = t.c;
    float e00 = (1.0f - a) * cc.c[0] + a * cc.c[1], e10 = (1.0f - a) * cc.c[2] + a * cc.c[3];
    float e01 = (1.0f - a) * cc.c[4] + a * cc.c[5], e11 = (1.0f - a) * cc.c[6] + a * cc.c[7];
    float g0 = (1.0f - b) * e00 + b * e10, g1 = (1.0f - b) * e01 + b * e11;
    return (1.0f - c) * g0 + c * g1;
}

// A lane marches at most kLaneBudget samples of its ray unless it took at most
// kWorkFloor trips of the march loop for them (a ray crossing free space skips
// many samples per trip and stays); a longer ray (about 1 in 40 grazes a
// surface and takes hundreds of small steps) is appended to a list that a second pass
// finishes with a whole wave per ray, so no wave of the first pass waits on
// one long ray.
// The single-session first pass (its waves' tail bounds the frame) hands rays
// over after 40 samples; the batch's persistent pass (throughput-bound) after 48.
constexpr int kLaneBudget = 40;
constexpr int kBatchLaneBudget = 48;
constexpr int kWorkFloor = 16;

// One trilinear evaluation of the march at t: 1 = skipped (the sample's bound
// is >= thr: kSkipBound on a large step, kSmallSkipBound on a small one),
// 0 = evaluated into f.  With P2 (power-of-two width and n: invw is the exact
// reciprocal, and tex then x = u*n - 1/2 round like the two fused
// multiply-adds of approx_texel) the skip test's texel coordinates are the
// stencil's.
template <int FMT, bool P2>
__device__ __forceinline__ int march_sample(const RaycastVolume& a, v3 o, v3 dir, float t, float width, float invw,
                                            float nf, float thr, int& cs, float& cb, int& ck, float& cbk,
                                            CellCache& cc, float& f)
{
    const v3 p = add3(o, mul3s(dir, t));
    const float xa = approx_texel(p.x, invw, nf), ya = approx_texel(p.y, invw, nf), za = approx_texel(p.z, invw, nf);
    const int fs = a.brick_key ? fast_skip(a, xa, ya, za, thr, cs, cb, ck, cbk) : 0;
    if (fs == 1) return 1;
    const Stencil st = P2 ? stencil_texel(a.n, xa, ya, za) : stencil<false>(a.n, tex(a, p.x), tex(a, p.y), tex(a, p.z));
    if (fs < 0 && skippable(a, st, thr, ck, cbk)) return 1;
    f = trilinear_cached<FMT>(a.f, a.nb, st, cc);
    return 0;
}

__device__ __forceinline__ float lane_bcast(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Camera ray of pixel (i, j) and its march interval (raycast.cu:7-27, 44-73).
struct RayGeom {
    v3 o, dir;
    float nearv, farv;
    bool active;  // false: the pixel is left untouched (raycast.cu:61-73)
};
__device__ __forceinline__ RayGeom ray_geom(const RaycastVolume& a, const float* pose, const RaycastLevel& L, int i,
                                            int j, bool in_image)
{
    float Rm[9] = {pose[0], pose[1], pose[2], pose[4], pose[5], pose[6], pose[8], pose[9], pose[10]};
    float M[9];
    m3m3(Rm, L.Kinv, M);
    RayGeom g;
    g.o = mk3(pose[12], pose[13], pose[14]);
    const float hw = a.width * 0.5f;
    g.dir = m3v(M, mk3((float)i + 0.5f, (float)j + 0.5f, 1.0f));
    const v3 inv = mk3(1.0f / g.dir.x, 1.0f / g.dir.y, 1.0f / g.dir.z);
    float t0x = (-hw - g.o.x) * inv.x, t1x = (hw - g.o.x) * inv.x;
    float t0y = (-hw - g.o.y) * inv.y, t1y = (hw - g.o.y) * inv.y;
    float t0z = (-hw - g.o.z) * inv.z, t1z = (hw - g.o.z) * inv.z;
    float tmin = fminf(t0x, t1x), tmax = fmaxf(t0x, t1x);
    tmin = fmaxf(tmin, fminf(t0y, t1y));
    tmax = fminf(tmax, fmaxf(t0y, t1y));
    tmin = fmaxf(tmin, fminf(t0z, t1z));
    tmax = fminf(tmax, fmaxf(t0z, t1z));
    if (tmax < tmin) {
        tmin = -1.0f;
        tmax = -1.0f;
    }
    g.nearv = (kMinDistance < tmin) ? tmin : kMinDistance;
    g.farv = (tmax < kMaxDistance) ? tmax : kMaxDistance;
    g.active = in_image && !(tmax < 0.0f) && g.nearv < g.farv;
    return g;
}

// raycast.cu:29-42 and 119-145: the surface point at tstar (tstar == 0: no
// crossing) and its normal -grad F, written by one lane.  The 6 gradient
// samples' corner loads are issued together; x + 0 == x, so the samples are
// those of raycast.cu:33-38.
template <int FMT, bool P2>
__device__ __forceinline__ void write_surface(const RaycastVolume& a, const RaycastLevel& L, int i, int j,
                                              const RayGeom& g, float tstar)
{
    const int64_t q = (int64_t)j * L.w + i;
    if (tstar != 0.0f) {
        const float width = a.width;
        const v3 vtx = add3(g.o, mul3s(g.dir, tstar));
        const float r = a.mue * 0.5f / width;
        const float u = tex(a, vtx.x), v = tex(a, vtx.y), s = tex(a, vtx.z);
        // the 48 corner loads of the 6 samples are issued together (one round trip)
        const float gx = sample_tsdf<FMT, P2>(a.f, a.n, a.nb, u - r, v, s) - sample_tsdf<FMT, P2>(a.f, a.n, a.nb, u + r, v, s);
        const float gy = sample_tsdf<FMT, P2>(a.f, a.n, a.nb, u, v - r, s) - sample_tsdf<FMT, P2>(a.f, a.n, a.nb, u, v + r, s);
        const float gz = sample_tsdf<FMT, P2>(a.f, a.n, a.nb, u, v, s - r) - sample_tsdf<FMT, P2>(a.f, a.n, a.nb, u, v, s + r);
        const v3 gg = div3s(mk3(gx, gy, gz), r);
        const float gn = len3(gg);
        if (is_depth_valid(gn)) {
            const v3 nn = div3s(gg, gn);
            L.vout[q] = make_float4(vtx.x, vtx.y, vtx.z, kValid);
            L.nout[q] = make_float4(nn.x, nn.y, nn.z, kValid);
            return;
        }
    }
    L.vout[q] = make_float4(0.0f, 0.0f, 0.0f, kInvalid);
    L.nout[q] = make_float4(0.0f, 0.0f, 0.0f, kInvalid);
}

// The same, with the 6 gradient samples taken by lanes 0-5 of the wave and the
// result written by lane 0 (the finishing pass: one ray per wave).
template <int FMT, bool P2>
__device__ __forceinline__ void write_surface_wave(const RaycastVolume& a, const RaycastLevel& L, int i, int j,
                                                   const RayGeom& g, float tstar)
{
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)j * L.w + i;
    if (tstar != 0.0f) {
        const float width = a.width;
        const v3 vtx = add3(g.o, mul3s(g.dir, tstar));
        const float r = a.mue * 0.5f / width;
        const float u = tex(a, vtx.x), v = tex(a, vtx.y), s = tex(a, vtx.z);
        float val = 0.0f;
        if (lane < 6) {
            const int ax = lane >> 1;
            const float du = ax == 0 ? r : 0.0f, dv = ax == 1 ? r : 0.0f, ds = ax == 2 ? r : 0.0f;
            if (lane & 1) val = sample_tsdf<FMT, P2>(a.f, a.n, a.nb, u - du, v - dv, s - ds);
            else val = sample_tsdf<FMT, P2>(a.f, a.n, a.nb, u + du, v + dv, s + ds);
        }
        const float gx = lane_bcast(val, 1) - lane_bcast(val, 0);
        const float gy = lane_bcast(val, 3) - lane_bcast(val, 2);
        const float gz = lane_bcast(val, 5) - lane_bcast(val, 4);
        const v3 gg = div3s(mk3(gx, gy, gz), r);
        const float gn = len3(gg);
        if (is_depth_valid(gn)) {
            if (lane == 0) {
                const v3 nn = div3s(gg, gn);
                L.vout[q] = make_float4(vtx.x, vtx.y, vtx.z, kValid);
                L.nout[q] = make_float4(nn.x, nn.y, nn.z, kValid);
            }
            return;
        }
    }
    if (lane == 0) {
        L.vout[q] = make_float4(0.0f, 0.0f, 0.0f, kInvalid);
        L.nout[q] = make_float4(0.0f, 0.0f, 0.0f, kInvalid);
    }
}

// Where the first pass appends the rays it hands over.
struct RayQueue {
    DeferredRay* rays;
    unsigned* count;
    unsigned cap;
    int budget;  // samples a lane marches before it may hand its ray over
};

// raycast.cu:44-145 for pixel (i, j) of one level map, one lane per pixel,
// called by all 64 lanes of a wave together (in_image false past the map's
// edge).  A ray still marching after its budget of samples is appended to q with
// its exact march state and left to the finishing pass (when q is full it
// marches on here).  A skipped sample only sets prev_t; its value is computed
// if and when the zero-crossing formula needs it.
template <int FMT, bool P2>
__device__ __forceinline__ void raycast_pixel(const RaycastVolume& a, const float* pose, const RaycastLevel& L, int i,
                                              int j, bool in_image, const RayQueue& rq, int job)
{
    const RayGeom g = ray_geom(a, pose, L, i, j, in_image);
    const v3 o = g.o, dir = g.dir;
    const float nearv = g.nearv, farv = g.farv;
    const float width = a.width, mue = a.mue;
    const float large = mue * 1.99f, small = mue * 0.1f;
    const float invw = 1.0f / width, nf = (float)a.n;
    float step = large, prev = 0.0f, tstar = 0.0f, prev_t = 0.0f, t = nearv;
    bool prev_skipped = false, can_defer = true;
    int work = 0;  // trips of the march loop so far (each a skip test or an evaluation)
    int cs = -1, ck = -1;
    float cb = 0.0f, cbk = 0.0f;
    CellCache cc;
    int guard = 0;  // never binds for mue >= 1e-4 m (kf_volume_create enforces it)
    if (!g.active) return;
    for (;;) {
        if (!(t < farv)) break;
        if (++guard > kMaxMarchSteps) break;
        // the lanes handing their rays over in this trip take their slots with one
        // atomic per wave (one counter for the whole launch: per-lane atomics on it
        // serialise)
        const bool defer = can_defer && ++work > kWorkFloor && guard > rq.budget;
        const uint64_t dm = __ballot(defer);
        if (dm) {
            const int leader = __ffsll((unsigned long long)dm) - 1;
            unsigned base = 0u;
            if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(rq.count, (unsigned)__popcll(dm));
            base = (unsigned)__builtin_amdgcn_readlane((int)base, leader);
            const unsigned slot = base + (unsigned)__popcll(dm & ((1ull << (threadIdx.x & 63)) - 1ull));
            if (defer && slot < rq.cap) {
                DeferredRay d;
                d.job = job;
                d.px = j * L.w + i;
                d.guard = guard;
                d.flags = (prev_skipped ? 1 : 0) | (step == small ? 2 : 0);
                d.t = t;
                d.prev = prev;
                d.prev_t = prev_t;
                d.pad = 0.0f;
                rq.rays[slot] = d;
                return;
            }
            if (defer) can_defer = false;
        }
        // one sample per trip on either step: large and small steps share the code,
        // so lanes of a wave in different phases of their march do not run two paths
        const bool lg = step == large;
        float f = 0.0f;
        if (march_sample<FMT, P2>(a, o, dir, t, width, invw, nf, lg ? kSkipBound : kSmallSkipBound, cs, cb, ck, cbk, cc,
                              f)) {
            // a skipped large step, and the following ones inside the same free
            // super-brick: positions by the same additions, no coordinates computed
            int run = lg ? super_run(a, o, dir, t, step, invw, nf, kSkipBound, cs, cb) : 0;
            prev_skipped = true;
            prev_t = t;
            t = t + step;
            for (; run > 0 && t < farv && guard < kMaxMarchSteps; --run) {
                ++guard;
                prev_t = t;
                t = t + step;
            }
            continue;
        }
        if (f < 0.0f) {
            if (t == nearv) break;
            if (lg) {
                t = t - mue;  // back off and march on with small steps
                step = small;
                t = t + step;
                continue;
            }
            if (prev_skipped) {
                const v3 pp = add3(o, mul3s(dir, prev_t));
                prev = sample_tsdf<FMT, P2>(a.f, a.n, a.nb, tex(a, pp.x), tex(a, pp.y), tex(a, pp.z));
            }
            tstar = t - small * prev / (f - prev);
            break;
        }
        if (lg && f < 0.99f) step = small;
        prev = f;
        prev_skipped = false;
        t = t + step;
    }
    write_surface<FMT, P2>(a, L, i, j, g, tstar);
}

// Position of sample `lane` of a chunk of 64: t_0 = t and t_{k+1} = t_k + step,
// one rounded addition per sample as the march takes them.  The additions form
// a chain, but inside one binade of t they nearly always add the same rounded
// increment: each lane guesses t_{k-1} on the line t_b + (k - 1 - b) d through
// an exact t_b, takes one rounded step from the guess, and checks that the
// result is the line's next point (so the next lane's guess is exact).  The
// first lane whose check fails still got the exact value (its own guess was
// exact) and becomes the next t_b.  By induction every lane ends with the
// chained value; a chunk takes one pass per change of the increment (a binade
// crossing, the settling of a rounding tie) instead of 63 dependent additions.
__device__ __forceinline__ float chunk_position(float t, float step, int lane)
{
    float res = t;
    int b = 0;
    float tb = t;
    for (;;) {
        const float d = (tb + step) - tb;
        const float c = fmaf((float)(lane - b - 1), d, tb);  // t_{lane-1} when the guesses so far held
        const float tk = c + step;
        const bool on = tk == fmaf((float)(lane - b), d, tb);
        if (lane > b) res = tk;
        const uint64_t off = __ballot(lane > b && !on);
        if (!off) break;
        b = __ffsll((unsigned long long)off) - 1;
        tb = lane_bcast(res, b);
    }
    return res;
}

// The finishing pass for one handed-over ray: the wave marches it 64 samples
// at a time, lane k taking sample k of the chunk (its position by the same
// repeated additions t = t + step, chunk_position), and a ballot finds the first sample that
// ends the ray or changes the march (end of the interval or the guard, f < 0,
// or f < 0.99 on a large step).  The samples before it are consumed as the
// single-lane march would (prev from the last of them), the ones after it are
// discarded, so every sample position, value and decision is the sequential
// march's and the output is the same.
template <int FMT, bool P2>
__device__ __forceinline__ void finish_ray(const RaycastVolume& a, const float* pose, const RaycastLevel& L,
                                           const DeferredRay& r)
{
    const int lane = threadIdx.x & 63;
    const int i = r.px % L.w, j = r.px / L.w;
    const RayGeom g = ray_geom(a, pose, L, i, j, true);
    const float width = a.width, mue = a.mue;
    const float large = mue * 1.99f, small = mue * 0.1f;
    const float invw = 1.0f / width, nf = (float)a.n;
    float step = (r.flags & 2) ? small : large;
    float t = r.t, pv = r.prev, pt = r.prev_t;
    bool psk = (r.flags & 1) != 0;
    int gb = r.guard;
    int cs = -1, ck = -1;
    float cb = 0.0f, cbk = 0.0f;
    CellCache cc;
    float tstar = 0.0f;
    for (;;) {
        const float tk = chunk_position(t, step, lane);
        const bool end = !(tk < g.farv) || gb + lane > kMaxMarchSteps;
        float f = 0.0f;
        int sk = 1;
        if (!end)
            sk = march_sample<FMT, P2>(a, g.o, g.dir, tk, width, invw, nf, step == large ? kSkipBound : kSmallSkipBound,
                                   cs, cb, ck, cbk, cc, f);
        const bool neg = !end && !sk && f < 0.0f;
        const bool slow = !end && !sk && step == large && f < 0.99f;
        const uint64_t evm = __ballot(end || neg || slow);
        const int k = evm ? __ffsll((unsigned long long)evm) - 1 : 64;
        if (k > 0) {
            // sample k - 1 is the last one consumed: it sets prev
            if (__builtin_amdgcn_readlane(sk, k - 1)) {
                psk = true;
                pt = lane_bcast(tk, k - 1);
            } else {
                psk = false;
                pv = lane_bcast(f, k - 1);
            }
        }
        if (k == 64) {
            t = lane_bcast(tk, 63) + step;
            gb += 64;
            continue;
        }
        const float tkk = lane_bcast(tk, k), fk = lane_bcast(f, k);
        if (__builtin_amdgcn_readlane(end ? 1 : 0, k)) break;
        if (__builtin_amdgcn_readlane(neg ? 1 : 0, k)) {
            if (tkk == g.nearv) break;
            if (step == large) {
                t = tkk - mue;  // back off and march on with small steps
                step = small;
                t = t + step;
                gb = gb + k + 1;
                continue;
            }
            float pk = pv;
            if (psk) {
                const v3 pp = add3(g.o, mul3s(g.dir, pt));
                pk = sample_tsdf<FMT, P2>(a.f, a.n, a.nb, tex(a, pp.x), tex(a, pp.y), tex(a, pp.z));
            }
            tstar = tkk - step * pk / (fk - pk);
            break;
        }
        // a large step found f < 0.99: small steps from the next sample on
        step = small;
        pv = fk;
        psk = false;
        t = tkk + step;
        gb = gb + k + 1;
    }
    write_surface_wave<FMT, P2>(a, L, i, j, g, tstar);
}

// Workgroups of the first pass per level: its tile count rounded up to 8.
__host__ __device__ __forceinline__ int level_blocks(const RaycastArgs& a, int l)
{
    const int n = (l + 1 < a.levels ? a.lv[l + 1].tile0 : a.tiles) - a.lv[l].tile0;
    return (n + 7) & ~7;
}

// Dispatch order of the first pass: the coarsest level first (a tile of a
// coarse level covers more of the scene and its waves march longest; started
// last they made the kernel's tail), each level's tiles in 8 XCD-contiguous
// bands (workgroup g runs on XCD g % 8), so every XCD gets its share of every
// level.
// Appends band slot `slot` of (level l, XCD x) to the next launch's order
// (lane 0): front when the wave took longer than the level's mean wave time
// of the previous launch (always, without one), else back.
__device__ __forceinline__ void record_order(const RaycastArgs& a, int l, int x, int per, int base, int slot,
                                             unsigned long long dur)
{
    TileOrderState* o = a.ost;
    const int prev = (a.tri + 2) % 3;
    unsigned long long n = 0, sum = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        n += o->waves[prev][l][q];
        sum += o->dur[prev][l][q];
    }
    const bool front = n == 0 || dur * n > sum;
    unsigned* c = &o->ctr[1 - a.parity][l][x][front ? 0 : 1];
    const int k = (int)atomicAdd(c, 1u);
    if (k < per) a.order_out[base + (front ? k : per - 1 - k) * 8 + x] = slot;
    if (dur != 0 && ((int)blockIdx.x / 8) % 8 == 0) {  // one workgroup in eight
        atomicAdd(&o->dur[a.tri][l][x], dur);
        atomicAdd(&o->waves[a.tri][l][x], 1u);
    }
}

// At least five waves per SIMD: the FP32 form compiles to 68 VGPRs (seven waves, no spill).
// A bound of eight (64 VGPRs, 20 bytes of spill per lane) made the three-level raycast 4 %
// slower.
template <int FMT>
__global__ __launch_bounds__(64, 5) void raycast_kernel(RaycastArgs a)
{
    const unsigned long long t0 = wall_clock64();
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) *a.clear = 0u;
        if (a.ost) {
            // counters the next launch appends with, and the wave times of the launch after it
            unsigned* c = &a.ost->ctr[a.parity][0][0][0];
            if (threadIdx.x < 3 * 8 * 2) c[threadIdx.x] = 0u;
            if (threadIdx.x < 3 * 8) {
                (&a.ost->dur[(a.tri + 1) % 3][0][0])[threadIdx.x] = 0ull;
                (&a.ost->waves[(a.tri + 1) % 3][0][0])[threadIdx.x] = 0u;
            }
        }
    }
    int g = blockIdx.x, l = a.levels - 1, base = 0;
    for (; l >= 0; --l) {
        const int n8 = level_blocks(a, l);
        if (g < n8) break;
        g -= n8;
        base += n8;
    }
    if (l < 0) return;
    const int n8 = level_blocks(a, l), per = n8 / 8, x = g % 8;
    const int slot = a.order_in ? a.order_in[blockIdx.x] : g / 8;
    const int lt = x * per + slot;
    const bool live = slot >= 0 && slot < per && lt < (l + 1 < a.levels ? a.lv[l + 1].tile0 : a.tiles) - a.lv[l].tile0;
    if (live) {
        const RaycastLevel& L = a.lv[l];
        const int i = (lt % L.tiles_x) * 8 + (threadIdx.x & 7);
        const int j = (lt / L.tiles_x) * 8 + (threadIdx.x >> 3);
        const RayQueue rq{a.rays, a.count, a.cap, kLaneBudget};
        const float* pose = a.dyn ? a.dyn->pose : a.pose;
        if (a.vol.p2) raycast_pixel<FMT, true>(a.vol, pose, L, i, j, i < L.w && j < L.h, rq, l);
        else raycast_pixel<FMT, false>(a.vol, pose, L, i, j, i < L.w && j < L.h, rq, l);
    }
    if (a.order_out && threadIdx.x == 0)
        record_order(a, l, x, per, base, slot, live ? wall_clock64() - t0 : 0ull);
}

// Waves of the finishing pass; each takes every kFinishWaves-th handed-over
// ray.  The batch's pass (hundreds of thousands of rays per tick) takes every
// kBatchFinishWaves-th: the rays' lengths vary, and the finer split shortens
// the tail (per 64-session tick, one wave per workgroup: 0.69 ms at 8192
// waves, 0.60 at 16384, 0.56 at 32768, 0.54 at 65536).  A single frame's few
// thousand rays gain nothing from more than 8192.  The waves are independent
// (no LDS, no barrier) and go kFinishPer to a workgroup: fewer workgroups to
// dispatch took the batch's pass from 0.54 to 0.51 ms (8 per workgroup gave
// 0.55, 16 gave 0.66 and slowed the single-frame pass).
constexpr unsigned kFinishWaves = 8192;
constexpr unsigned kBatchFinishWaves = 65536;
constexpr unsigned kFinishPer = 4;

template <int FMT>
__global__ __launch_bounds__(64 * kFinishPer) void raycast_finish_kernel(RaycastArgs a)
{
    typedef __attribute__((address_space(4))) const unsigned ConstCount;
    typedef __attribute__((address_space(4))) const DeferredRay ConstRay;
    const unsigned total = *(ConstCount*)a.count;
    const unsigned n = total < a.cap ? total : a.cap;
    const float* pose = a.dyn ? a.dyn->pose : a.pose;
    const unsigned w0 = blockIdx.x * kFinishPer + (unsigned)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (unsigned r = w0; r < n; r += gridDim.x * kFinishPer) {
        ConstRay& cr = ((ConstRay*)a.rays)[r];
        DeferredRay d;
        d.job = cr.job;
        d.px = cr.px;
        d.guard = cr.guard;
        d.flags = cr.flags;
        d.t = cr.t;
        d.prev = cr.prev;
        d.prev_t = cr.prev_t;
        if (a.vol.p2) finish_ray<FMT, true>(a.vol, pose, a.lv[d.job], d);
        else finish_ray<FMT, false>(a.vol, pose, a.lv[d.job], d);
    }
}

// Field-wise copy of a job read through the constant address space.
template <class CJ>
__device__ __forceinline__ void copy_job(const CJ& j, RaycastVolume& vol, RaycastLevel& L)
{
    vol.f = j.vol.f;
    vol.brick_key = j.vol.brick_key;
    vol.super_key = j.vol.super_key;
    vol.n = j.vol.n;
    vol.nb = j.vol.nb;
    vol.ns = j.vol.ns;
    vol.format = j.vol.format;
    vol.width = j.vol.width;
    vol.mue = j.vol.mue;
    vol.inv_width = j.vol.inv_width;
    vol.p2 = j.vol.p2;
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
__global__ __launch_bounds__(64) void raycast_batch_kernel(const BatchHeader* hd, const RaycastJob* jobs)
{
    const RayQueue rq{hd->ray_list + (FMT == KF_VOXEL_F32 ? 0u : hd->ray_list_cap),
                      hd->ray_next + (FMT == KF_VOXEL_F32 ? 16 : 17), hd->ray_list_cap, kBatchLaneBudget};
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
            c = __builtin_amdgcn_readlane(c, 0);
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
                if (vol.p2) raycast_pixel<FMT, true>(vol, pose, L, i, j, i < L.w && j < L.h, rq, jb);
                else raycast_pixel<FMT, false>(vol, pose, L, i, j, i < L.w && j < L.h, rq, jb);
            }
        }
    }
}

// The finishing pass of one voxel format over the rays its first pass handed
// over (each record names its job).
template <int FMT>
__global__ __launch_bounds__(64 * kFinishPer) void raycast_batch_finish_kernel(const BatchHeader* hd,
                                                                               const RaycastJob* jobs)
{
    typedef __attribute__((address_space(4))) const RaycastJob ConstJob;
    typedef __attribute__((address_space(4))) const BatchHeader ConstHeader;
    typedef __attribute__((address_space(4))) const unsigned ConstCount;
    typedef __attribute__((address_space(4))) const DeferredRay ConstRay;
    ConstJob* cj = (ConstJob*)jobs;
    ConstHeader* ch = (ConstHeader*)hd;
    const unsigned cap = ch->ray_list_cap;
    const unsigned total = *(ConstCount*)(ch->ray_next + (FMT == KF_VOXEL_F32 ? 16 : 17));
    const unsigned n = total < cap ? total : cap;
    ConstRay* rays = (ConstRay*)(ch->ray_list + (FMT == KF_VOXEL_F32 ? 0u : cap));
    const unsigned w0 = blockIdx.x * kFinishPer + (unsigned)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (unsigned r = w0; r < n; r += gridDim.x * kFinishPer) {
        DeferredRay d;
        d.job = rays[r].job;
        d.px = rays[r].px;
        d.guard = rays[r].guard;
        d.flags = rays[r].flags;
        d.t = rays[r].t;
        d.prev = rays[r].prev;
        d.prev_t = rays[r].prev_t;
        RaycastVolume vol;
        RaycastLevel L;
        copy_job(cj[d.job], vol, L);
        if (vol.p2) finish_ray<FMT, true>(vol, cj[d.job].pose, L, d);
        else finish_ray<FMT, false>(vol, cj[d.job].pose, L, d);
    }
}

}  // namespace

RaycastVolume raycast_volume(const kf_volume* v)
{
    RaycastVolume r{};
    r.f = v->f;
    r.brick_key = v->brick_key;
    r.super_key = v->super_key;
    r.n = v->n;
    r.nb = v->nb;
    r.ns = v->ns;
    r.format = v->format;
    r.width = v->width;
    r.mue = v->mue;
    int e = 0;
    r.inv_width = frexpf(v->width, &e) == 0.5f ? 1.0f / v->width : 0.0f;
    r.p2 = r.inv_width != 0.0f && (v->n & (v->n - 1)) == 0 ? 1 : 0;
    return r;
}

// The two voxel formats use disjoint tile counters, ray-list halves and maps, so
// both first passes run before both finishing passes.
int launch_raycast_batch(const BatchHeader* hd, const RaycastJob* jobs, unsigned blocks, hipStream_t s, hipEvent_t mid)
{
    raycast_batch_kernel<KF_VOXEL_F32><<<blocks, 64, 0, s>>>(hd, jobs);
    KF_LAUNCHED("raycast_batch_kernel");
    raycast_batch_kernel<KF_VOXEL_SNORM16_U8><<<blocks, 64, 0, s>>>(hd, jobs);
    KF_LAUNCHED("raycast_batch_kernel");
    if (mid) KF_HIP(hipEventRecord(mid, s));
    raycast_batch_finish_kernel<KF_VOXEL_F32><<<kBatchFinishWaves / kFinishPer, 64 * kFinishPer, 0, s>>>(hd, jobs);
    KF_LAUNCHED("raycast_batch_finish_kernel");
    raycast_batch_finish_kernel<KF_VOXEL_SNORM16_U8><<<kBatchFinishWaves / kFinishPer, 64 * kFinishPer, 0, s>>>(
        hd, jobs);
    KF_LAUNCHED("raycast_batch_finish_kernel");
    return KF_OK;
}

// The volume's deferred-ray list, grown to hold every ray of a launch of
// `rays` rays (grown outside the stream order: waits for work that may still
// use the old list).
static int ensure_ray_list(const kf_volume* v, size_t rays, hipStream_t s)
{
    if (!v->ray_count) {
        KF_HIP(hipMalloc(&v->ray_count, 2 * sizeof(unsigned)));
        KF_HIP(hipMemset(v->ray_count, 0, 2 * sizeof(unsigned)));
    }
    if (v->ray_cap >= rays) return KF_OK;
    KF_HIP(hipStreamSynchronize(s));
    (void)hipFree(v->ray_list);
    v->ray_list = nullptr;
    v->ray_cap = 0;
    KF_HIP(hipMalloc(&v->ray_list, sizeof(DeferredRay) * rays));
    v->ray_cap = rays;
    return KF_OK;
}

// The first pass's dispatch-order state of a volume, for `blocks` workgroups.
static int ensure_tile_order(const kf_volume* v, size_t blocks, hipStream_t s)
{
    if (!v->order_state) {
        KF_HIP(hipMalloc(&v->order_state, sizeof(TileOrderState)));
        KF_HIP(hipMemset(v->order_state, 0, sizeof(TileOrderState)));
    }
    if (v->order_cap >= blocks) return KF_OK;
    KF_HIP(hipStreamSynchronize(s));
    (void)hipFree(v->tile_order);
    v->tile_order = nullptr;
    v->order_cap = 0;
    v->order_key = -1;
    KF_HIP(hipMalloc(&v->tile_order, 2 * sizeof(int) * blocks));
    v->order_cap = blocks;
    return KF_OK;
}

int launch_raycast_levels(const kf_volume* v, const float* pose, const PoseState* dyn_state, int levels,
                          const float (*inv_k)[9], float4* const* vout, float4* const* nout, int w0, int h0,
                          hipStream_t s)
{
    size_t rays = 0;
    for (int l = 0; l < levels; ++l) rays += (size_t)(w0 >> l) * (size_t)(h0 >> l);
    int st = ensure_ray_list(v, rays, s);
    if (st) return st;
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
    unsigned blocks = 0;
    for (int l = 0; l < levels; ++l) blocks += (unsigned)level_blocks(a, l);
    st = ensure_tile_order(v, blocks, s);
    if (st) return st;
    const int64_t launch = v->ray_launches++;
    const int parity = (int)(launch & 1);
    a.rays = static_cast<DeferredRay*>(v->ray_list);
    a.count = v->ray_count + parity;
    a.clear = v->ray_count + (1 - parity);
    a.cap = (unsigned)v->ray_cap;
    // the order recorded by the previous launch is used when it was made for the same maps
    const int64_t key = ((int64_t)levels << 42) | ((int64_t)w0 << 21) | (int64_t)h0;
    a.order_in = v->order_key == key ? v->tile_order + (size_t)parity * v->order_cap : nullptr;
    a.order_out = v->tile_order + (size_t)(1 - parity) * v->order_cap;
    a.ost = static_cast<TileOrderState*>(v->order_state);
    a.parity = parity;
    a.tri = (int)(launch % 3);
    v->order_key = key;
    if (v->format == KF_VOXEL_F32) {
        raycast_kernel<KF_VOXEL_F32><<<blocks, 64, 0, s>>>(a);
        KF_LAUNCHED("raycast_kernel");
        raycast_finish_kernel<KF_VOXEL_F32><<<kFinishWaves / kFinishPer, 64 * kFinishPer, 0, s>>>(a);
    } else {
        raycast_kernel<KF_VOXEL_SNORM16_U8><<<blocks, 64, 0, s>>>(a);
        KF_LAUNCHED("raycast_kernel");
        raycast_finish_kernel<KF_VOXEL_SNORM16_U8><<<kFinishWaves / kFinishPer, 64 * kFinishPer, 0, s>>>(a);
    }
    KF_LAUNCHED("raycast_finish_kernel");
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
