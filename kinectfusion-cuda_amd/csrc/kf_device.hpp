// Device-side math shared by every kernel of the per-frame path.
//
// Every expression below is evaluated in a fixed operation order with no FMA
// contraction (the library is compiled with -ffp-contract=off) and with
// correctly rounded FP32 division and square root (hipcc's default for HIP
// device code), so a kernel's result is a deterministic function of its inputs.
// The orders are glm's, as the reference uses them (project/src/*.cu include
// glm with GLM_FORCE_PURE, project/src/glm_macro.h:5-12):
//   mat3*vec3 : r_i = (m[0][i]*v.x + m[1][i]*v.y) + m[2][i]*v.z
//   mat4*vec4 : r   = (m[0]*v.x + m[1]*v.y) + (m[2]*v.z + m[3]*v.w)
//   mat3*mat3 : C[c][r] = (A[0][r]*B[c][0] + A[1][r]*B[c][1]) + A[2][r]*B[c][2]
//   dot       : (a.x*b.x + a.y*b.y) + a.z*b.z
//   normalize : v * (1 / sqrt(dot(v, v)))
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kf {

constexpr float kValid = -1.0f;     // device_helper.cuh:9
constexpr float kInvalid = -2.0f;   // device_helper.cuh:10
constexpr float kMinDepth = 0.01f;  // device_helper.cuh:11
constexpr float kMinDistance = 0.4f;  // tsdf.cuh:11
constexpr float kMaxDistance = 6.0f;  // tsdf.cuh:12

struct v3 {
    float x, y, z;
};

__device__ __forceinline__ v3 mk3(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 mul3s(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 div3s(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float len3(v3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ v3 normalize3(v3 a)
{
    float inv = 1.0f / sqrtf(dot3(a, a));
    return mul3s(a, inv);
}
__device__ __forceinline__ v3 cross3(v3 x, v3 y)
{
    return mk3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
// m: column-major 3x3
__device__ __forceinline__ v3 m3v(const float* m, v3 v)
{
    return mk3((m[0] * v.x + m[3] * v.y) + m[6] * v.z,
               (m[1] * v.x + m[4] * v.y) + m[7] * v.z,
               (m[2] * v.x + m[5] * v.y) + m[8] * v.z);
}
__device__ __forceinline__ void m3m3(const float* A, const float* B, float* C)
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            C[c * 3 + r] = (A[0 * 3 + r] * B[c * 3 + 0] + A[1 * 3 + r] * B[c * 3 + 1]) + A[2 * 3 + r] * B[c * 3 + 2];
}
__device__ __forceinline__ void m3t(const float* A, float* T)
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) T[c * 3 + r] = A[r * 3 + c];
}
// glm mat4 * vec4(p, 1), xyz only
__device__ __forceinline__ v3 m4p(const float* m, v3 p)
{
    float a0 = m[0] * p.x, a1 = m[4] * p.y, a2 = m[8] * p.z, a3 = m[12] * 1.0f;
    float b0 = m[1] * p.x, b1 = m[5] * p.y, b2 = m[9] * p.z, b3 = m[13] * 1.0f;
    float c0 = m[2] * p.x, c1 = m[6] * p.y, c2 = m[10] * p.z, c3 = m[14] * 1.0f;
    return mk3((a0 + a1) + (a2 + a3), (b0 + b1) + (b2 + b3), (c0 + c1) + (c2 + c3));
}

__device__ __forceinline__ bool is_depth_valid(float d) { return d > kMinDepth && isfinite(d); }
// v clamped to [lo, hi]; lo <= hi (every caller clamps to [0, size - 1] of a
// non-empty range), which lets it be a max and a min (no compare and select)
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// float4 map access; out-of-range validity reads are invalid (device_helper.cuh:43-50)
__device__ __forceinline__ bool map_valid(const float4* m, int w, int h, int i, int j)
{
    if (i < 0 || j < 0 || i >= w || j >= h) return false;
    return m[(int64_t)j * w + i].w == kValid;
}
__device__ __forceinline__ v3 xyz(float4 e) { return mk3(e.x, e.y, e.z); }

// Raw depth code -> binary16 (RNE) -> float, integer restatement of the IEEE
// rule (measurement.cu:19-21); codes >= 65520 overflow to +inf.
__device__ __forceinline__ float half_rne(uint32_t c)
{
    if (c <= 2048u) return (float)c;
    int bits = 32 - __clz(c);
    int drop = bits - 11;
    uint32_t q = c >> drop, r = c & ((1u << drop) - 1u), half = 1u << (drop - 1);
    if (r > half || (r == half && (q & 1u))) q += 1u;
    uint32_t v = q << drop;
    if (v > 65504u) return __int_as_float(0x7f800000);
    return (float)v;
}

// Single-precision exp: Cody-Waite reduction by ln2 with the Cephes expf
// polynomial (S. L. Moshier) in fused multiply-adds (the oracle's kfo_expf
// makes the same fmaf calls), then exact scaling by 2^k with one final rounding.
__device__ __forceinline__ float kf_expf(float x)
{
    if (!(x == x)) return x;
    if (x < -104.0f) return 0.0f;
    if (x > 88.72283f) return __int_as_float(0x7f800000);
    float fk = floorf(x * 1.44269504088896341f + 0.5f);
    int k = (int)fk;
    float r = x - fk * 0.693359375f;  // exact product (9-bit constant)
    r = fmaf(fk, 2.12194440e-4f, r);
    float z = r * r;
    float p = 1.9875691500e-4f;
    p = fmaf(p, r, 1.3981999507e-3f);
    p = fmaf(p, r, 8.3334519073e-3f);
    p = fmaf(p, r, 4.1665795894e-2f);
    p = fmaf(p, r, 1.6666665459e-1f);
    p = fmaf(p, r, 5.0000001201e-1f);
    float y = fmaf(p, z, r);
    y = y + 1.0f;
    if (k > 127) { y = y * 1.7014118346046923e38f; k -= 127; }
    if (k < -126) { y = y * 5.421010862427522e-20f; k += 64; }
    return y * __int_as_float((k + 127) << 23);
}
// kf_expf for x <= 0 or NaN (the bilateral weights, -a - b with a, b >= 0):
// the overflow branches of kf_expf are never taken there, so the same
// operations without them give the same bits.  Its scaling by 2^k is one
// rounding of y * 2^k (for k < -126 the first step, by 2^-64, is exact, and
// y in [0.7, 1.5] keeps it normal), which is what v_ldexp_f32 computes.
__device__ __forceinline__ float kf_expf_nonpos(float x)
{
    if (!(x == x)) return x;
    if (x < -104.0f) return 0.0f;
    float fk = floorf(x * 1.44269504088896341f + 0.5f);
    int k = (int)fk;
    float r = x - fk * 0.693359375f;  // exact product (9-bit constant)
    r = fmaf(fk, 2.12194440e-4f, r);
    float z = r * r;
    float p = 1.9875691500e-4f;
    p = fmaf(p, r, 1.3981999507e-3f);
    p = fmaf(p, r, 8.3334519073e-3f);
    p = fmaf(p, r, 4.1665795894e-2f);
    p = fmaf(p, r, 1.6666665459e-1f);
    p = fmaf(p, r, 5.0000001201e-1f);
    float y = fmaf(p, z, r);
    y = y + 1.0f;
    return __builtin_amdgcn_ldexpf(y, k);
}

// Live normal at (i,j) of a vertex map (device_helper.cuh:81-110).
// The three vertices a normal needs: the pixel, its right and its lower
// neighbour (out-of-range entries read as invalid).  The three loads are
// independent, so they are in flight together; an out-of-range entry loads
// element 0 and is then replaced (a value select, not an address select).
struct NormalStencil {
    float4 c, r, d;
};
__device__ __forceinline__ float4 load_or_invalid(const float4* v, bool ok, int64_t q)
{
    float4 e = v[ok ? q : 0];
    if (!ok) e.w = kInvalid;
    return e;
}
__device__ __forceinline__ NormalStencil load_normal_stencil(const float4* v, int w, int h, int i, int j)
{
    const bool ic = i >= 0 && j >= 0 && i < w && j < h;
    const int64_t q = (int64_t)j * w + i;
    NormalStencil s;
    s.c = load_or_invalid(v, ic, q);
    s.r = load_or_invalid(v, ic && i + 1 < w, q + 1);
    s.d = load_or_invalid(v, ic && j + 1 < h, q + w);
    return s;
}
__device__ __forceinline__ float4 normal_of(const NormalStencil& s)
{
    if (s.c.w == kValid && s.r.w == kValid && s.d.w == kValid) {
        v3 c = xyz(s.c), r = xyz(s.r), d = xyz(s.d);
        v3 n = normalize3(cross3(sub3(r, c), sub3(d, c)));
        return make_float4(n.x, n.y, n.z, kValid);
    }
    return make_float4(0.0f, 0.0f, 0.0f, kInvalid);
}
__device__ __forceinline__ float4 compute_normal(const float4* v, int w, int h, int i, int j)
{
    return normal_of(load_normal_stencil(v, w, h, i, j));
}

// ---------------------------------------------------------------------------
// Lane exchanges within a wave in VALU instructions only (no LDS round trip):
// the gfx950 permlane swaps for lane bits 5 and 4, and DPP moves for bits 3
// to 0 (row_mirror pairs lane r of a 16-lane row with 15 - r, row_half_mirror
// r of an 8-lane half-row with 7 - r, quad_perm the lanes two and one apart).
// Every pattern pairs each lane with one partner whose index differs in the
// step's bit, so six steps combine all 64 lanes.
constexpr int kDppRowMirror = 0x140, kDppRowHalfMirror = 0x141, kDppXor2 = 0x4E, kDppXor1 = 0xB1;
template <int CTRL>
__device__ __forceinline__ float dpp_partner(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ int dpp_partner(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
// Lanes with bit log2(OFF) clear get (a, partner's a), the others (partner's
// b, b): in every lane `lo` is the operand of the lower lane of the pair.
template <int OFF>
__device__ __forceinline__ void swap_pair(float a, float b, float& lo, float& hi)
{
    static_assert(OFF == 32 || OFF == 16, "permlane swaps exchange 32- or 16-lane blocks");
    const auto r = OFF == 32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false)
                             : __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    lo = __uint_as_float(r[0]);
    hi = __uint_as_float(r[1]);
}

// Wave-wide min, max and integer sum (order-free), in every lane.
template <class Op>
__device__ __forceinline__ float wave_reduce(float v, Op op)
{
    float lo, hi;
    swap_pair<32>(v, v, lo, hi);
    v = op(lo, hi);
    swap_pair<16>(v, v, lo, hi);
    v = op(lo, hi);
    v = op(v, dpp_partner<kDppRowMirror>(v));
    v = op(v, dpp_partner<kDppRowHalfMirror>(v));
    v = op(v, dpp_partner<kDppXor2>(v));
    return op(v, dpp_partner<kDppXor1>(v));
}
__device__ __forceinline__ float wave_min(float v)
{
    return wave_reduce(v, [](float a, float b) { return fminf(a, b); });
}
__device__ __forceinline__ float wave_max(float v)
{
    return wave_reduce(v, [](float a, float b) { return fmaxf(a, b); });
}
__device__ __forceinline__ int wave_sum(int v)
{
    const auto r32 = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    v = (int)r32[0] + (int)r32[1];
    const auto r16 = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
    v = (int)r16[0] + (int)r16[1];
    v += dpp_partner<kDppRowMirror>(v);
    v += dpp_partner<kDppRowHalfMirror>(v);
    v += dpp_partner<kDppXor2>(v);
    v += dpp_partner<kDppXor1>(v);
    return v;
}

// ---------------------------------------------------------------------------
// TSDF volume layout in HBM: 8x8x8 bricks of 512 voxels, bricks ordered x then
// y then z, voxels inside a brick ordered z then x then y (z fastest).  The 8
// voxels of a brick column (i, j, 0..7) are contiguous, so the fusion sweep
// (one lane per column) reads and writes a column with one or two wide accesses
// and a wave covers a brick's planes contiguously; a trilinear stencil usually
// stays inside one brick.
constexpr int kBrick = 8;
constexpr int kBrickVoxels = 512;

__device__ __forceinline__ int64_t brick_index(int nb, int i, int j, int k)
{
    int64_t b = ((int64_t)(k >> 3) * nb + (j >> 3)) * nb + (i >> 3);
    return b * kBrickVoxels + (((j & 7) * 8 + (i & 7)) * 8 + (k & 7));
}
// The same index as a sum of per-axis terms, brick_index = ox(i) + oy(j) + oz(k),
// in 32 bits (volumes are at most 1024^3 voxels): the 8 corners of a trilinear
// cell then cost 6 terms and 12 additions, and a load is a 32-bit offset from
// the plane's base.
__device__ __forceinline__ uint32_t brick_ox(int i) { return (uint32_t)(i >> 3) * 512u + (uint32_t)(i & 7) * 8u; }
// The row and slab strides of the brick grid, 512 nb and 512 nb^2, stay below
// 2^24 (nb <= 128), so the products use the 24-bit multiply (full rate; a
// 32-bit multiply issues in four passes).
__device__ __forceinline__ uint32_t brick_oy(int nb, int j)
{
    return __umul24((uint32_t)(j >> 3), (uint32_t)nb << 9) + (uint32_t)(j & 7) * 64u;
}
__device__ __forceinline__ uint32_t brick_oz(int nb, int k)
{
    return __umul24((uint32_t)(k >> 3), ((uint32_t)nb * (uint32_t)nb) << 9) + (uint32_t)(k & 7);
}
// Linear index of cell (x, y, z) of an m^3 grid (bricks or super-bricks:
// m <= 128), in 24-bit multiplies.
__device__ __forceinline__ int grid_index(int x, int y, int z, int m)
{
    return (int)(__umul24(__umul24((uint32_t)z, (uint32_t)m) + (uint32_t)y, (uint32_t)m) + (uint32_t)x);
}

// Brick bounds (raycast skipping): brick (bi, bj, bk) keeps a lower bound of F
// over its halo, the voxels [8bi, 8bi + 8] x [8bj, 8bj + 8] x [8bk, 8bk + 8]
// clamped to the volume: its own 512 voxels and the first voxel plane of its
// +x, +y and +z neighbours.  A trilinear stencil spans (i0, i0 + 1) per axis
// (clamped), so the bound of the brick holding its first corner (i0, j0, k0)
// covers all 8 corners, also where the stencil crosses into the next brick.
// Several waves lower one bound in a fuse (the brick's own and those of the
// bricks whose halo reaches into it), so it is kept as an ordered integer key
// (order_key) and lowered with an integer atomicMin.
// Super-bricks: 8x8x8 bricks (64^3 voxels) carry a lower bound of F over their
// own voxels too, in the same key form.
constexpr int kSuper = 64;
__device__ __forceinline__ int order_key(float f)
{
    int b = __float_as_int(f);
    return b >= 0 ? b : (b ^ 0x7fffffff);
}
__device__ __forceinline__ float order_value(int k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff)); }

// snorm16 TSDF codes (precision form): f = q / 32767 with q in [-32767, 32767];
// the code snorm16 leaves unused, -32768, holds the unobserved value 2 of
// tsdf.cu:7-18 (any f > 1 is stored as it), so samples next to unobserved
// space interpolate with 2 exactly as the FP32 volume does.
constexpr float kSnormScale = 32767.0f;
constexpr int kSnormUnobserved = -32768;
__device__ __forceinline__ float snorm_decode(int16_t q)
{
    return q == kSnormUnobserved ? 2.0f : (float)q * (1.0f / kSnormScale);
}
__device__ __forceinline__ int16_t snorm_encode(float f)
{
    if (f > 1.0f) return (int16_t)kSnormUnobserved;
    float s = f * kSnormScale;
    s = fminf(fmaxf(s, -kSnormScale), kSnormScale);
    return (int16_t)rintf(s);
}

// Rank of workgroup g among G such that the ranks on one XCD (workgroups are
// dealt to the 8 XCDs round-robin) are contiguous; identity unless G % 8 == 0.
__device__ __forceinline__ int xcd_rank(int g, int G)
{
    return (G & 7) ? g : (g & 7) * (G >> 3) + (g >> 3);
}

// Index of the job whose range [key(job), key(job + 1)) holds `item`; the
// jobs are sorted by key and key(jobs[0]) == 0.
template <class J, class K>
__device__ __forceinline__ int job_of(const J* jobs, int njobs, long long item, K key)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((long long)key(jobs[mid]) <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace kf
