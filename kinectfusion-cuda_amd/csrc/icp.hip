// Projective point-to-plane ICP (project/src/icp.cu, icp.cuh, icp.cpp,
// linear_least_squares.cpp), kept on the device for a whole frame.
//
// One iteration = two launches:
//   reduce   : 256 lanes x 4 pixels per workgroup.  Each lane associates its
//              pixels (icp.cu:7-82) and accumulates the 27 unique terms of the
//              normal equations (21 of A^T A, 6 of A^T b) in registers, either
//              plainly in FP32 or compensated (TwoSum, Kahan-Babuska-Neumaier);
//              then a 64-lane shuffle tree, an LDS pass over the 4 waves, and
//              one partial row per workgroup.  A and b are never materialised.
//   finalize : one workgroup sums the partial rows in a fixed order (so the
//              result is deterministic), solves the 6x6 system with an FP64
//              Cholesky, applies the pose update (icp.cpp:65-87) and the early
//              stop test (icp.cpp:45-53) to the device-resident PoseState.
// The kernel boundary between the two launches is what makes every
// workgroup's partial row visible to the finalize across the 8 XCDs' L2s.
// The reference instead writes A (N x 6) and b to HBM, runs cublasSgemm and
// cublasSgemv, copies 42 floats to the host and solves with Eigen's FP32 LLT,
// with a host synchronisation per iteration.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "kf_device.hpp"
#include "kf_kernels.hpp"

namespace kf {

namespace {

constexpr int kThreads = 256;
constexpr int kPixPerThread = 4;  // pixels per lane, at most (see icp_pix_for)
constexpr int kTerms = 27;
// partial layout (component-major): [0,27) sums, [27,54) compensations, 54 count
constexpr int kCountSlot = 54;


// icp.cu:7-82 split in two: what a pixel needs that does not depend on the
// pose (its vertex and its live normal, device_helper.cuh:81-110) and the
// pose-dependent association, so the loads of the first part can be issued
// before the pose of the iteration is known.
struct PixelIn {
    v3 vc;       // camera-space vertex
    v3 cr;       // live normal before normalization (device_helper.cuh:81-110's cross product)
    bool ok;     // vertex and normal valid
};
__device__ __forceinline__ PixelIn icp_pixel_from(const NormalStencil& ns)
{
    PixelIn in;
    in.ok = ns.c.w == kValid;
    in.vc = in.ok ? xyz(ns.c) : mk3(0.0f, 0.0f, 0.0f);
    const v3 c = xyz(ns.c);
    in.cr = cross3(sub3(xyz(ns.r), c), sub3(xyz(ns.d), c));  // normal_of before its normalize3
    in.ok = in.ok && ns.r.w == kValid && ns.d.w == kValid;
    return in;
}
__device__ __forceinline__ PixelIn icp_pixel_in(const IcpArgs& a, int u, int v)
{
    return icp_pixel_from(load_normal_stencil(a.V, a.w, a.h, u, v));
}

// false = zero row (rejection reasons as icp.cu:29-70, evaluated in an order
// that does not change the result: every rejection writes the same zero row).
// Split in two so that the target loads of all of a lane's pixels are in
// flight together: icp_pixel_gather transforms and projects the pixel and
// issues the loads of its target vertex and normal (at an in-map index even
// when the projection is off the map, so no branch surrounds the loads);
// icp_pixel_finish applies the tests and builds the row.
// The two rejection tests decide a boolean only, so each is first decided
// from a cheap approximation and evaluated exactly (the reference arithmetic,
// correctly rounded sqrt and division) only when the approximation lies too
// close to the threshold to decide: the decision is the exact one always.
//
// Distance, icp.cuh:57-60: reject iff RN(sqrt(D)) > T, D = dot(v, v) (the
// same bits on both paths).  D <= T^2 gives sqrt(D) <= T; D >= (T + ulp(T))^2
// gives RN(sqrt(D)) > T.  lo and hi below bound those squares from inside
// (lo <= T^2, hi >= T^2 (1 + 2^-21) > (T + ulp(T))^2); NaN takes the exact path.
__device__ __forceinline__ bool distance_rejects(const IcpArgs& a, v3 v)
{
    const float D = dot3(v, v);
    const float T2 = a.dist_thresh * a.dist_thresh;
    const float lo = T2 * (1.0f - 0x1p-22f), hi = T2 * (1.0f + 0x1p-20f);
    if (D >= hi) return true;
    if (D <= lo) return false;
    return len3(v) > a.dist_thresh;
}
// Angle, icp.cuh:63-70 in its cosine form (acos(d) > th <=> d >= -1 && d < cos
// th), d = dot(normalize(Rc normalize(cr)), n).  The approximation takes both
// normalizations with the hardware reciprocal square root (1 ulp); its d is
// within 2^-19 |n|_1 of the exact d (measured worst case over 8M random
// inputs with every rsq off by one ulp either way: 2^-21.8), so a margin of
// 2^-16 |n|_1 decides; squared lengths outside [2^-100, 2^100] (where the
// hardware routine may flush or overflow) and NaN take the exact path.
__device__ __forceinline__ bool angle_rejects(const IcpArgs& a, v3 cr, const float* Rc, v3 n)
{
    const float q = dot3(cr, cr);
    const v3 m = m3v(Rc, mul3s(cr, __builtin_amdgcn_rsqf(q)));
    const float mq = dot3(m, m);
    const float da = dot3(mul3s(m, __builtin_amdgcn_rsqf(mq)), n);
    const float del = 0x1p-16f * (fabsf(n.x) + fabsf(n.y) + fabsf(n.z));
    const bool ranged = q >= 0x1p-100f && q <= 0x1p100f && mq >= 0x1p-100f && mq <= 0x1p100f;
    const float c = a.cos_thresh;
    if (ranged) {
        if (da >= c + del || da < -1.0f - del) return false;
        if (da >= -1.0f + del && da < c - del) return true;
    }
    const float dd = dot3(normalize3(m3v(Rc, normalize3(cr))), n);
    return dd >= -1.0f && dd < c;
}

struct PixelGather {
    v3 s;             // the vertex in the frame of the current estimate
    float4 tv, tn;    // target vertex and normal at the projection
    bool ok;          // the pixel is valid and projects into the map
};
__device__ __forceinline__ PixelGather icp_pixel_gather(const IcpArgs& a, const PixelIn& in, const float* Rc,
                                                       const float* tc, const float* M, const float* tp)
{
    PixelGather g;
    g.s = add3(m3v(Rc, in.vc), mk3(tc[0], tc[1], tc[2]));
    // icp.cuh:40-49: (K * R_prev^T) * (s - t_prev), floor of the perspective division
    v3 p = m3v(M, sub3(g.s, mk3(tp[0], tp[1], tp[2])));
    float fu = floorf(p.x / p.z), fv = floorf(p.y / p.z);
    g.ok = in.ok && fu >= 0.0f && fv >= 0.0f && fu < (float)a.w && fv < (float)a.h;
    const int64_t q = g.ok ? (int64_t)(int)fv * a.w + (int)fu : 0;
    g.tv = a.TV[q];
    g.tn = a.TN[q];
    return g;
}
__device__ __forceinline__ bool icp_pixel_finish(const IcpArgs& a, const PixelIn& in, const PixelGather& g,
                                                 const float* Rc, float* row, float& b)
{
    if (!g.ok) return false;
    const float4 tv = g.tv, tn = g.tn;
    const v3 s = g.s;
    if (tv.w != kValid || tn.w != kValid) return false;
    v3 d = xyz(tv);
    if (distance_rejects(a, sub3(d, s))) return false;  // icp.cuh:57-60
    v3 n = xyz(tn);
    if (angle_rejects(a, in.cr, Rc, n)) return false;  // icp.cuh:63-70
    // icp.cuh:72-95
    row[0] = n.y * s.z - n.z * s.y;
    row[1] = n.z * s.x - n.x * s.z;
    row[2] = n.x * s.y - n.y * s.x;
    row[3] = n.x;
    row[4] = n.y;
    row[5] = n.z;
    b = ((((n.x * d.x + n.y * d.y) + n.z * d.z) - n.x * s.x) - n.y * s.y) - n.z * s.z;
    return true;
}
__device__ __forceinline__ bool icp_pixel_row(const IcpArgs& a, const PixelIn& in, const float* Rc, const float* tc,
                                              const float* M, const float* tp, float* row, float& b)
{
    return icp_pixel_finish(a, in, icp_pixel_gather(a, in, Rc, tc, M, tp), Rc, row, b);
}

__device__ __forceinline__ bool icp_row(const IcpArgs& a, int u, int v, const float* Rc, const float* tc,
                                        const float* M, const float* tp, float* row, float& b)
{
    return icp_pixel_row(a, icp_pixel_in(a, u, v), Rc, tc, M, tp, row, b);
}

// (s1, c1) += (s2, c2)
__device__ __forceinline__ void comp_merge(float& s1, float& c1, float s2, float c2)
{
    float t = s1 + s2;
    float bp = t - s1;
    float e = (s1 - (t - bp)) + (s2 - bp);
    s1 = t;
    c1 = (c1 + c2) + e;
}

template <bool COMP>
// Adds a pixel's row to the lane's sums of A^T A (upper triangle) and A^T b,
// each product fused into its sum (one rounding instead of two).  A lane sums
// at most kPixPerThread rows, so its few roundings are small against the
// total; the compensated form carries the rounding errors of every addition
// after that (lanes, waves, workgroups: block_reduce_store and the partial-row
// sums), where nearly all of the additions happen.
__device__ __forceinline__ void accumulate(float* s, float* c, const float* row, float b)
{
    (void)c;
    int q = 0;
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int r = p; r < 6; ++r, ++q) s[q] = fmaf(row[p], row[r], s[q]);
#pragma unroll
    for (int p = 0; p < 6; ++p) s[21 + p] = fmaf(row[p], b, s[21 + p]);
}

// One halving step of the lane reduction: the first n/2 entries of s (and c)
// become, on the lane with the step's bit clear, its first half plus the
// partner's copy and, on the other lane, its second half plus the partner's
// copy; the compensated merge takes the lower lane's operand first.
template <bool COMP, int N, int OFF>
__device__ __forceinline__ void halve_swap(float* s, float* c)
{
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
        float lo, hi;
        swap_pair<OFF>(s[i], s[N / 2 + i], lo, hi);
        if (COMP) {
            float clo, chi;
            swap_pair<OFF>(c[i], c[N / 2 + i], clo, chi);
            comp_merge(lo, clo, hi, chi);
            s[i] = lo;
            c[i] = clo;
        } else {
            s[i] = lo + hi;
        }
    }
}
template <bool COMP, int N, int CTRL>
__device__ __forceinline__ void halve_dpp(float* s, float* c, bool upper)
{
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
        const float send = upper ? s[i] : s[N / 2 + i];
        const float keep = upper ? s[N / 2 + i] : s[i];
        const float recv = dpp_partner<CTRL>(send);
        if (COMP) {
            const float csend = upper ? c[i] : c[N / 2 + i];
            const float ckeep = upper ? c[N / 2 + i] : c[i];
            const float crecv = dpp_partner<CTRL>(csend);
            float ls = upper ? recv : keep, lc = upper ? crecv : ckeep;
            comp_merge(ls, lc, upper ? keep : recv, upper ? ckeep : crecv);
            s[i] = ls;
            c[i] = lc;
        } else {
            s[i] = keep + recv;
        }
    }
}
// The compensated form: lane 0 holds the total (the lower lane's operand first
// in every merge).
__device__ __forceinline__ void wave_comp_sum(float& s, float& c)
{
    const int lane = threadIdx.x & 63;
    float lo, hi, clo, chi;
    swap_pair<32>(s, s, lo, hi);
    swap_pair<32>(c, c, clo, chi);
    comp_merge(lo, clo, hi, chi);
    swap_pair<16>(lo, lo, s, hi);
    swap_pair<16>(clo, clo, c, chi);
    comp_merge(s, c, hi, chi);
#define KF_DPP_MERGE(CTRL, BIT)                                                  \
    {                                                                            \
        const bool up = (lane & (BIT)) != 0;                                     \
        const float rs = dpp_partner<CTRL>(s), rc = dpp_partner<CTRL>(c);        \
        float ls = up ? rs : s, lc = up ? rc : c;                                \
        comp_merge(ls, lc, up ? s : rs, up ? c : rc);                            \
        s = ls;                                                                  \
        c = lc;                                                                  \
    }
    KF_DPP_MERGE(kDppRowMirror, 8)
    KF_DPP_MERGE(kDppRowHalfMirror, 4)
    KF_DPP_MERGE(kDppXor2, 2)
    KF_DPP_MERGE(kDppXor1, 1)
#undef KF_DPP_MERGE
}

// Reduces the lane values of a workgroup and writes one partial row.  The 27
// terms are halved across the lanes (28 -> 14 -> 7 -> 4 -> 2 -> 1, one zero
// pad at 27 and one at the 7 -> 8 step): lane L ends with term
// 14*b5 + 7*b4 + 4*b3 + 2*b2 + b1 of its bits, every pair (L, L^1) holding it
// (each step keeps the first half on the lane whose step bit is clear).
template <bool COMP>
__device__ __forceinline__ void block_reduce_store(float* s_in, float* c_in, int cnt, float* partials, int stride,
                                                   int row)
{
    __shared__ float red[kThreads / 64][2 * kTerms + 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float s[28], c[28];
#pragma unroll
    for (int q = 0; q < kTerms; ++q) {
        s[q] = s_in[q];
        c[q] = COMP ? c_in[q] : 0.0f;
    }
    s[27] = c[27] = 0.0f;
    halve_swap<COMP, 28, 32>(s, c);
    halve_swap<COMP, 14, 16>(s, c);
    s[7] = c[7] = 0.0f;
    halve_dpp<COMP, 8, kDppRowMirror>(s, c, (lane & 8) != 0);
    halve_dpp<COMP, 4, kDppRowHalfMirror>(s, c, (lane & 4) != 0);
    halve_dpp<COMP, 2, kDppXor2>(s, c, (lane & 2) != 0);
    {
        // last step: both lanes of the pair add the same two operands, lower first
        const bool upper = (lane & 1) != 0;
        const float recv = dpp_partner<kDppXor1>(s[0]);
        if (COMP) {
            const float crecv = dpp_partner<kDppXor1>(c[0]);
            float ls = upper ? recv : s[0], lc = upper ? crecv : c[0];
            comp_merge(ls, lc, upper ? s[0] : recv, upper ? c[0] : crecv);
            s[0] = ls;
            c[0] = lc;
        } else {
            s[0] = s[0] + recv;
        }
    }
    cnt = wave_sum(cnt);
    const int sub = 4 * ((lane >> 3) & 1) + 2 * ((lane >> 2) & 1) + ((lane >> 1) & 1);
    const int q = 14 * ((lane >> 5) & 1) + 7 * ((lane >> 4) & 1) + sub;
    if ((lane & 1) == 0 && sub < 7 && q < kTerms) {
        red[wave][q] = s[0];
        red[wave][kTerms + q] = COMP ? c[0] : 0.0f;
    }
    if (lane == 0) red[wave][2 * kTerms] = (float)cnt;
    __syncthreads();
    if (threadIdx.x < kTerms) {
        const int q = threadIdx.x;
        float ss = red[0][q], cc = red[0][kTerms + q];
        for (int wv = 1; wv < kThreads / 64; ++wv) {
            if (COMP) comp_merge(ss, cc, red[wv][q], red[wv][kTerms + q]);
            else ss = ss + red[wv][q];
        }
        partials[(int64_t)q * stride + row] = ss;
        partials[(int64_t)(kTerms + q) * stride + row] = cc;
    } else if (threadIdx.x == kTerms) {
        float n = 0.0f;
        for (int wv = 0; wv < kThreads / 64; ++wv) n = n + red[wv][2 * kTerms];
        partials[(int64_t)kCountSlot * stride + row] = n;
    }
}

// The pixels of lane (block, k, lane), k < pix <= PIX, and their
// pose-independent inputs: pixel (block * pix + k) * kThreads + lane.
template <int PIX = kPixPerThread>
struct LanePixels {
    PixelIn in[PIX];
};
// Their vertex loads only (a pixel past the map reads as invalid), so the
// loads can be in flight while the workgroup finishes the previous iteration.
template <int PIX = kPixPerThread>
struct LaneLoads {
    NormalStencil ns[PIX];
};
// Column and row of pixel px of a w-wide map, px < kMaxMapPixels: px is exact
// in FP32 and px * (1/w) is within one of px / w, so one correction each way
// gives the quotient.
constexpr int kMaxMapPixels = 1 << 24;
__device__ __forceinline__ void pixel_uv(int px, int w, float inv_w, int& u, int& v)
{
    v = (int)((float)px * inv_w);
    u = px - v * w;
    if (u < 0) {
        u += w;
        --v;
    } else if (u >= w) {
        u -= w;
        ++v;
    }
}
template <int PIX>
__device__ __forceinline__ void issue_lane_loads(const IcpArgs& a, LaneLoads<PIX>& ll, int block, int pix)
{
    const int npx = a.w * a.h;  // <= kMaxMapPixels (checked on the host)
    const float inv_w = 1.0f / (float)a.w;
#pragma unroll
    for (int k = 0; k < PIX; ++k) {
        if (k >= pix) break;  // uniform
        const int px = (block * pix + k) * kThreads + (int)threadIdx.x;
        int u = -1, v = -1;
        if (px < npx) pixel_uv(px, a.w, inv_w, u, v);
        ll.ns[k] = load_normal_stencil(a.V, a.w, a.h, u, v);
    }
}
template <int PIX>
__device__ __forceinline__ void finish_lane_pixels(const LaneLoads<PIX>& ll, LanePixels<PIX>& lp, int pix)
{
#pragma unroll
    for (int k = 0; k < PIX; ++k) {
        if (k >= pix) break;
        lp.in[k] = icp_pixel_from(ll.ns[k]);
    }
}
template <int PIX>
__device__ __forceinline__ void load_lane_pixels(const IcpArgs& a, LanePixels<PIX>& lp, int block, int pix)
{
    LaneLoads<PIX> ll;
    issue_lane_loads(a, ll, block, pix);
    finish_lane_pixels(ll, lp, pix);
}

// pose: R (9), t (3), Rp (9), tp (3) of the iteration, in LDS.  Writes partial
// row `block` of a.partials.
template <bool COMP, int PIX>
__device__ void associate_reduce(const IcpArgs& a, const LanePixels<PIX>& lp, const float* pose, int block, int pix)
{
    const float* Rc = pose;
    const float* tc = pose + 9;
    float Rt[9], M[9];
    m3t(pose + 12, Rt);
    m3m3(a.K, Rt, M);
    const float* tp = pose + 21;
    float s[kTerms], c[kTerms];
#pragma unroll
    for (int q = 0; q < kTerms; ++q) s[q] = c[q] = 0.0f;
    int cnt = 0;
    PixelGather g[PIX];
#pragma unroll
    for (int k = 0; k < PIX; ++k) {
        if (k >= pix) break;
        g[k] = icp_pixel_gather(a, lp.in[k], Rc, tc, M, tp);
    }
#pragma unroll
    for (int k = 0; k < PIX; ++k) {
        if (k >= pix) break;
        float row[6], b;
        if (icp_pixel_finish(a, lp.in[k], g[k], Rc, row, b)) {
            accumulate<COMP>(s, c, row, b);
            ++cnt;
        }
    }
    block_reduce_store<COMP>(s, c, cnt, a.partials, a.stride, block);
}

template <bool COMP>
__device__ void reduce_body(const IcpArgs& a)
{
    PoseState* st = a.st;
    __shared__ float pose[24];
    LanePixels<> lp;
    load_lane_pixels(a, lp, blockIdx.x, kPixPerThread);
    if (threadIdx.x < 9) pose[threadIdx.x] = st->R[threadIdx.x];
    else if (threadIdx.x < 12) pose[threadIdx.x] = st->t[threadIdx.x - 9];
    else if (threadIdx.x < 21) pose[threadIdx.x] = st->Rp[threadIdx.x - 12];
    else if (threadIdx.x < 24) pose[threadIdx.x] = st->tp[threadIdx.x - 21];
    __syncthreads();
    associate_reduce<COMP>(a, lp, pose, blockIdx.x, kPixPerThread);
}

template <bool COMP>
__global__ __launch_bounds__(kThreads) void icp_reduce_kernel(IcpArgs a)
{
    if (a.st->status != 0 || a.st->level_done[a.level]) return;
    reduce_body<COMP>(a);
}

// Rows of a given A (n x 6 row-major) and b (LinearLeastSquares::solve).
template <bool COMP>
__global__ __launch_bounds__(kThreads) void ab_reduce_kernel(const float* A, const float* B, int64_t n,
                                                             float* partials, int stride)
{
    float s[kTerms], c[kTerms];
#pragma unroll
    for (int q = 0; q < kTerms; ++q) s[q] = c[q] = 0.0f;
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < kPixPerThread; ++k) {
        int64_t r = ((int64_t)blockIdx.x * kPixPerThread + k) * kThreads + threadIdx.x;
        if (r >= n) break;
        float row[6];
#pragma unroll
        for (int p = 0; p < 6; ++p) row[p] = A[r * 6 + p];
        float b = B[r];
        accumulate<COMP>(s, c, row, b);
        bool nz = b != 0.0f;
#pragma unroll
        for (int p = 0; p < 6; ++p) nz = nz || row[p] != 0.0f;
        cnt += nz ? 1 : 0;
    }
    block_reduce_store<COMP>(s, c, cnt, partials, stride, blockIdx.x);
}

// icp.cu:7-82, writing the rows (kernel::constructIcpResiduals).
__global__ __launch_bounds__(kThreads) void icp_rows_kernel(IcpArgs a, float* A, float* B)
{
    const float* pose = reinterpret_cast<const float*>(a.partials);  // 24 floats by value (R, t, Rp, tp)
    float Rt[9], M[9];
    m3t(pose + 12, Rt);
    m3m3(a.K, Rt, M);
    int64_t px = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (px >= (int64_t)a.w * a.h) return;
    int u = (int)(px % a.w), v = (int)(px / a.w);
    float row[6] = {0, 0, 0, 0, 0, 0}, b = 0.0f;
    if (!icp_row(a, u, v, pose, pose + 9, M, pose + 21, row, b)) {
#pragma unroll
        for (int p = 0; p < 6; ++p) row[p] = 0.0f;
        b = 0.0f;
    }
#pragma unroll
    for (int p = 0; p < 6; ++p) A[px * 6 + p] = row[p];
    B[px] = b;
}

// ---- FP64 6x6 Cholesky solve, unblocked, in the column order of Eigen's LLT
// (3rd/eigen/Eigen/src/Cholesky/LLT.h:306-326), by one wave: lane r < 6
// holds row r of A (from the 27 sums) and of L; column k is finished by its
// lanes and its values are broadcast with readlane:
//   D = A(k,k) - sum_j<k L(k,j)^2,   i(k) = 1 / sqrt(D),   L(k,k) = D i(k),
//   L(r,k) = (A(r,k) - sum_j<k L(r,j) L(k,j)) * i(k),
//   y(r) = (b(r) - sum_j<r L(r,j) y(j)) * i(r),
//   x(r) = (y(r) - sum_j>r L(j,r) x(j)) * i(r)
// (a reciprocal square root and multiplications instead of a square root and
// divisions by L(r,r), and fused multiply-adds in the sums: each value moves
// by a few ulp of FP64, far below the FP32 pose it feeds, and the FP64 square
// roots and divisions leave the dependent chain).  No lane holds the whole matrix (a one-lane FP64 solve
// would set the register budget of every kernel that embeds it).
// Called by all 64 lanes of a wave; returns 0 or the failing column + 1 and
// the solution in x (the same values in every lane).
__device__ __forceinline__ double bcast(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// index of A(p, q), p <= q, among the 21 sums: rows of the upper triangle
__device__ __forceinline__ int sym_index(int p, int q) { return p * 6 - (p * (p - 1)) / 2 + (q - p); }

__device__ int chol6_wave(const double* tot, double* x)
{
    const int lane = threadIdx.x & 63;
    const int r = lane < 6 ? lane : 5;  // lanes >= 6 shadow row 5; their values are never read
    double a[6], Lr[6], inv[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        a[c] = tot[sym_index(min(r, c), max(r, c))];
        Lr[c] = 0.0;
    }
    // each diagonal entry's reciprocal square root is taken once per column;
    // the column below it and both substitutions multiply by it, so the
    // substitutions' twelve dependent steps are multiplications, not divisions
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double d = a[k];
#pragma unroll
        for (int j = 0; j < k; ++j) d = fma(-Lr[j], Lr[j], d);
        d = bcast(d, k);
        if (!(d > 0.0)) return k + 1;
        inv[k] = rsqrt(d);  // 1/sqrt, within 2 ulp of FP64
        d = d * inv[k];
        double Lk[6];
#pragma unroll
        for (int j = 0; j < k; ++j) Lk[j] = bcast(Lr[j], k);
        if (r == k) {
            Lr[k] = d;
        } else if (r > k) {
            double v = a[k];
#pragma unroll
            for (int j = 0; j < k; ++j) v = fma(-Lr[j], Lk[j], v);
            Lr[k] = v * inv[k];
        }
    }
    double y[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        double v = tot[21 + r];
#pragma unroll
        for (int j = 0; j < q; ++j) v = fma(-Lr[j], y[j], v);
        y[q] = bcast(v, q) * inv[q];
    }
#pragma unroll
    for (int q = 5; q >= 0; --q) {
        double v = y[q];
#pragma unroll
        for (int j = q + 1; j < 6; ++j) v = fma(-bcast(Lr[q], j), x[j], v);
        x[q] = v * inv[q];
    }
    return 0;
}

// icp.cpp:65-87 in FP64: R_inc = R_y(x1) R_x(x0) R_z(x2) (glm orientate3(x0, x2, x1)),
// R <- R_inc^T R, t <- R_inc^T t + x[3..5].
//
// sin and cos in FP64 of an angle x: Cody-Waite reduction r = x - k pi/2 with
// pi/2 split in a 33-bit head (k * head exact for |k| < 2^20) and a tail,
// then the Taylor series of sin and cos on |r| <= pi/4 (terms through r^17
// and r^18: the next terms are below 2^-60 there; Horner steps fused).  Accurate to a few ulp of
// double for |x| < 2^20 pi/2, far beyond any ICP update; the library routines
// carry a huge-argument reduction whose registers every lane of the kernel
// would pay for.
__device__ __forceinline__ void sincos_small(double x, double* sn, double* cs)
{
    const double kHead = 1.57079632673412561417e+00;  // leading 33 bits of pi/2
    const double kTail = 6.07710050650619224932e-11;  // pi/2 - kHead
    const double k = rint(x * 0.636619772367581343076);
    const double r = (x - k * kHead) - k * kTail;
    const double z = r * r;
    double ps = 1.0 / 355687428096000.0;            // 1/17!
    ps = fma(ps, z, -1.0 / 1307674368000.0);        // 1/15!
    ps = fma(ps, z, 1.0 / 6227020800.0);            // 1/13!
    ps = fma(ps, z, -1.0 / 39916800.0);             // 1/11!
    ps = fma(ps, z, 1.0 / 362880.0);                // 1/9!
    ps = fma(ps, z, -1.0 / 5040.0);                 // 1/7!
    ps = fma(ps, z, 1.0 / 120.0);                   // 1/5!
    ps = fma(ps, z, -1.0 / 6.0);                    // 1/3!
    const double s = fma(r, z * ps, r);
    double pc = 1.0 / 6402373705728000.0;           // 1/18!
    pc = fma(pc, z, -1.0 / 20922789888000.0);       // 1/16!
    pc = fma(pc, z, 1.0 / 87178291200.0);           // 1/14!
    pc = fma(pc, z, -1.0 / 479001600.0);            // 1/12!
    pc = fma(pc, z, 1.0 / 3628800.0);               // 1/10!
    pc = fma(pc, z, -1.0 / 40320.0);                // 1/8!
    pc = fma(pc, z, 1.0 / 720.0);                   // 1/6!
    pc = fma(pc, z, -1.0 / 24.0);                   // 1/4!
    pc = fma(pc, z, 0.5);                           // 1/2!
    const double c = fma(-z, pc, 1.0);
    const int q = (int)((long long)k & 3);
    *sn = q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
    *cs = q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}

// sin and cos of the three angles of an update, as update_pose uses them:
// yaw x[1] (sh, ch), pitch x[0] (sp, cp), roll x[2] (sb, cb).
struct SinCos3 {
    double sh, ch, sp, cp, sb, cb;
};
__device__ __forceinline__ SinCos3 angle_sincos(const float* x)
{
    SinCos3 a;
    sincos_small(x[1], &a.sh, &a.ch);
    sincos_small(x[0], &a.sp, &a.cp);
    sincos_small(x[2], &a.sb, &a.cb);
    return a;
}
// The same by lanes 0, 1, 2 of a wave, from the FP64 solution xd (every lane
// of the wave calls it and receives all six values).
__device__ __forceinline__ SinCos3 angle_sincos_wave(const double* xd)
{
    const int lane = threadIdx.x & 63;
    const float ang = lane == 0 ? (float)xd[1] : (lane == 1 ? (float)xd[0] : (float)xd[2]);
    double sn, cs;
    sincos_small(ang, &sn, &cs);
    SinCos3 a;
    a.sh = bcast(sn, 0);
    a.ch = bcast(cs, 0);
    a.sp = bcast(sn, 1);
    a.cp = bcast(cs, 1);
    a.sb = bcast(sn, 2);
    a.cb = bcast(cs, 2);
    return a;
}

__device__ void update_pose(float* Rm, float* t, const float* x, const SinCos3& a)
{
    const double sh = a.sh, ch = a.ch, sp = a.sp, cp = a.cp, sb = a.sb, cb = a.cb;
    double Ri[9];
    Ri[0] = ch * cb + sh * sp * sb;  Ri[1] = sb * cp;  Ri[2] = -sh * cb + ch * sp * sb;
    Ri[3] = -ch * sb + sh * sp * cb; Ri[4] = cb * cp;  Ri[5] = sb * sh + ch * sp * cb;
    Ri[6] = sh * cp;                 Ri[7] = -sp;      Ri[8] = ch * cp;
    double Rn[9], tn[3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double s = 0;
            for (int q = 0; q < 3; ++q) s += Ri[r * 3 + q] * (double)Rm[c * 3 + q];
            Rn[c * 3 + r] = s;
        }
    for (int r = 0; r < 3; ++r) {
        double s = 0;
        for (int q = 0; q < 3; ++q) s += Ri[r * 3 + q] * (double)t[q];
        tn[r] = s + (double)x[3 + r];
    }
    for (int q = 0; q < 9; ++q) Rm[q] = (float)Rn[q];
    for (int q = 0; q < 3; ++q) t[q] = (float)tn[q];
}

// general_helper.cpp:11-21
__device__ void pose_error(const float* R1, const float* t1, const float* R2, const float* t2, float* out)
{
    v3 v = normalize3(mk3(1.0f, 1.0f, 1.0f));
    v3 a = m3v(R1, v), b = m3v(R2, v);
    out[0] = len3(sub3(b, a));
    out[1] = len3(sub3(mk3(t2[0], t2[1], t2[2]), mk3(t1[0], t1[1], t1[2])));
}

// Sums the partial rows (fixed order) into 27 doubles + count in LDS.  Wave w
// owns rows q = w, w + 4, ... (7 of the 28); lane l adds blocks l, l + 64, ...
// in that order, then the lanes combine (wave_reduce / wave_comp_sum: lane 0
// ends with the total, the lower lane's operand first).  The loads of all of a
// wave's rows for a chunk of kRowHalves 64-block groups (RowChunk) are issued
// before any is added, so a row set costs one memory round trip per 320 blocks
// (a 640x480 level is 300); the first chunk can be loaded by the caller ahead
// of other loads (sum_loaded_partials).
constexpr int kRowsPerWave = (kTerms + 1 + 3) / 4;
constexpr int kRowHalves = 5;
struct RowChunk {
    float vs[kRowHalves][kRowsPerWave], vc[kRowHalves][kRowsPerWave];
};
// Needs nblocks > 0.  Lanes past the last block load the last block's entry
// (no divergent branch around the loads, so none waits for another);
// add_row_chunk skips them.  Groups wholly past the last block are not loaded
// (a wave-uniform test).
template <bool COMP>
__device__ __forceinline__ void load_row_chunk(const float* partials, int stride, int nblocks, int b0, RowChunk& ch)
{
    static_assert(kThreads == 256 && 3 + 4 * (kRowsPerWave - 1) == kTerms, "four waves own rows 0..kTerms");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int h = 0; h < kRowHalves; ++h) {
        if (h > 0 && b0 + 64 * h >= nblocks) break;
        const int b = min(b0 + 64 * h + lane, nblocks - 1);
#pragma unroll
        for (int k = 0; k < kRowsPerWave; ++k) {
            const int q = wave + 4 * k;
            const bool is_count = (q == kTerms);  // wave-uniform
            ch.vs[h][k] = partials[(int64_t)(is_count ? kCountSlot : q) * stride + b];
            ch.vc[h][k] = (COMP && !is_count) ? partials[(int64_t)(kTerms + q) * stride + b] : 0.0f;
        }
    }
}
template <bool COMP>
__device__ __forceinline__ void add_row_chunk(int nblocks, int b0, const RowChunk& ch, float* s, float* c)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int h = 0; h < kRowHalves; ++h) {
        if (h > 0 && b0 + 64 * h >= nblocks) break;
        const int b = b0 + 64 * h + lane;
#pragma unroll
        for (int k = 0; k < kRowsPerWave; ++k) {
            const int q = wave + 4 * k;
            if (b < nblocks) {
                if (COMP && q != kTerms) comp_merge(s[k], c[k], ch.vs[h][k], ch.vc[h][k]);
                else s[k] = s[k] + ch.vs[h][k];
            }
        }
    }
}
// `first` = load_row_chunk(partials, stride, nblocks, 0, first); nblocks > 0.
template <bool COMP>
__device__ void sum_loaded_partials(const float* partials, int stride, int nblocks, const RowChunk& first,
                                    double* tot, double* count)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float s[kRowsPerWave], c[kRowsPerWave];
#pragma unroll
    for (int k = 0; k < kRowsPerWave; ++k) s[k] = c[k] = 0.0f;
    add_row_chunk<COMP>(nblocks, 0, first, s, c);
    for (int b0 = 64 * kRowHalves; b0 < nblocks; b0 += 64 * kRowHalves) {
        RowChunk ch;
        load_row_chunk<COMP>(partials, stride, nblocks, b0, ch);
        add_row_chunk<COMP>(nblocks, b0, ch, s, c);
    }
#pragma unroll
    for (int k = 0; k < kRowsPerWave; ++k) {
        const int q = wave + 4 * k;
        const bool is_count = (q == kTerms);
        float sk = s[k], ck = c[k];
        if (COMP && !is_count) wave_comp_sum(sk, ck);
        else sk = wave_reduce(sk, [](float x, float y) { return x + y; });  // both lanes of a pair: same sum
        if (lane == 0) {
            if (is_count) *count = (double)sk;
            else tot[q] = COMP ? (double)sk + (double)ck : (double)sk;
        }
    }
    __syncthreads();
}
template <bool COMP>
__device__ void sum_partials(const float* partials, int stride, int nblocks, double* tot, double* count)
{
    RowChunk first = {};
    if (nblocks > 0) load_row_chunk<COMP>(partials, stride, nblocks, 0, first);
    sum_loaded_partials<COMP>(partials, stride, nblocks, first, tot, count);
}

// Pose update of one iteration from its FP64 solution (one lane): the float
// update x, the pose update (icp.cpp:65-87) and the early-stop test
// (icp.cpp:45-53).  Returns 0 or KF_ERR_NONFINITE; *done = early stop.
__device__ int pose_step(const double* xd, const SinCos3& sc, float* R, float* t, float stop_angle,
                         float stop_dist, float* x, bool* done)
{
    bool finite = true;
    for (int q = 0; q < 6; ++q) {
        x[q] = (float)xd[q];
        finite = finite && isfinite(x[q]);
    }
    *done = false;
    if (!finite) return KF_ERR_NONFINITE;
    float Rprev[9], tprev[3], err[2];
    for (int q = 0; q < 9; ++q) Rprev[q] = R[q];
    for (int q = 0; q < 3; ++q) tprev[q] = t[q];
    update_pose(R, t, x, sc);
    pose_error(Rprev, tprev, R, t, err);
    *done = err[0] < stop_angle && err[1] < stop_dist;
    return KF_OK;
}

// mode 0: ICP iteration (solve + pose update + early stop); mode 1: solve only.
// Executed by every lane of one workgroup.
template <bool COMP>
__device__ void finalize_body(PoseState* st, const float* partials, int stride, int nblocks, int level,
                              float stop_angle, float stop_dist, int mode)
{
    __shared__ double tot[kTerms];
    __shared__ double count;
    sum_partials<COMP>(partials, stride, nblocks, tot, &count);
    if (threadIdx.x >= 64) return;
    double xd[6];
    const int rc = chol6_wave(tot, xd);
    if (threadIdx.x != 0) return;
    for (int q = 0; q < kTerms; ++q) st->last_sums[q] = tot[q];
    st->last_inliers = (long long)count;
    if (rc != 0) {
        if (st->status == 0) st->status = KF_ERR_NOT_PD;
        if (st->sticky_status == 0) st->sticky_status = KF_ERR_NOT_PD;
        st->frame_skip_fuse = 1;
        return;
    }
    float x[6];
    bool finite = true;
    for (int q = 0; q < 6; ++q) {
        x[q] = (float)xd[q];
        finite = finite && isfinite(x[q]);
        st->last_x[q] = x[q];
    }
    if (!finite) {
        if (st->status == 0) st->status = KF_ERR_NONFINITE;
        if (st->sticky_status == 0) st->sticky_status = KF_ERR_NONFINITE;
        st->frame_skip_fuse = 1;
        return;
    }
    if (mode != 0) return;
    float Rprev[9], tprev[3], err[2];
    for (int q = 0; q < 9; ++q) Rprev[q] = st->R[q];
    for (int q = 0; q < 3; ++q) tprev[q] = st->t[q];
    update_pose(st->R, st->t, x, angle_sincos(x));
    pose_error(Rprev, tprev, st->R, st->t, err);
    st->iterations = st->iterations + 1;
    if (err[0] < stop_angle && err[1] < stop_dist) st->level_done[level] = 1;
}

template <bool COMP>
__global__ __launch_bounds__(kThreads) void icp_finalize_kernel(PoseState* st, const float* partials, int stride,
                                                                int nblocks, int level, float stop_angle,
                                                                float stop_dist, int mode)
{
    if (mode == 0 && (st->status != 0 || st->level_done[level])) return;
    finalize_body<COMP>(st, partials, stride, nblocks, level, stop_angle, stop_dist, mode);
}

// ---- the device-resident coarse-to-fine loop: one launch per iteration.
// Launch k first finishes iteration k-1 (every workgroup sums launch k-1's
// partial rows in the same fixed order and solves, so all of them hold the same
// pose), then associates and reduces iteration k.  The carried state (IcpCarry)
// and the partial rows are double-buffered by launch parity, so no workgroup
// reads what another workgroup of the same launch writes.

// Applies the solve of the pending iteration of `c` (thread 0): rc and xd from
// chol6_wave, tot/count the sums it was solved from.
__device__ void apply_solution(IcpCarry& c, int rc, const double* xd, const SinCos3& sc, const double* tot,
                               double count,
                               PoseState* st, bool record, float stop_angle, float stop_dist)
{
    if (!c.pending || c.status != 0) return;
    float x[6] = {0, 0, 0, 0, 0, 0};
    bool done = false;
    if (rc) rc = KF_ERR_NOT_PD;
    else rc = pose_step(xd, sc, c.R, c.t, stop_angle, stop_dist, x, &done);
    if (rc) {
        c.status = rc;
    } else {
        c.iterations += 1;
        // c.level_done[c.pending_level] = 1 without a dynamic index (keeps c in registers)
#pragma unroll
        for (int l = 0; l < 4; ++l)
            if (done && l == c.pending_level) c.level_done[l] = 1;
    }
    if (record) {
        for (int q = 0; q < kTerms; ++q) st->last_sums[q] = tot[q];
        st->last_inliers = (long long)count;
        for (int q = 0; q < 6; ++q) st->last_x[q] = x[q];
    }
}

// Wave 0 solves the pending iteration of cin (if any) from tot; every lane of
// the wave gets rc and xd.
__device__ __forceinline__ int solve_pending(const IcpCarry* cin, const double* tot, double* xd)
{
    if (!(cin->pending && cin->status == 0)) return 0;
    return chol6_wave(tot, xd);
}

__device__ __forceinline__ int level_flag(const int* v, int level)
{
    int r = 0;
#pragma unroll
    for (int l = 0; l < 4; ++l) r = (l == level) ? v[l] : r;
    return r;
}

// Phase B of a launch for one session: finish the previous iteration from
// (cin, pin) and put the pose of this iteration in LDS.  `first` = this
// workgroup holds the session's partial row 0: it records the solve and writes
// the carried state for the next launch.  Returns (via run) whether the
// session iterates at `level` in this launch.  Ends with a barrier.
// Loads of phase B that do not depend on this launch's pose: the carried
// state (one word per lane of wave 0), the previous pose (lanes 64..75) and
// the first chunk of the previous launch's partial rows.  Issued before the
// pixel loads, so waiting for them does not wait for those (loads complete in
// issue order).
struct BeginLoads {
    int cw;
    float pw;
    RowChunk rows;
};
template <bool COMP>
__device__ __forceinline__ void issue_begin_loads(const PoseState* st, const IcpCarry* cin, const float* pin,
                                                  int prev_blocks, BeginLoads& bl)
{
    // every lane loads (indexes clamped; begin_iteration keeps the wanted
    // lanes' words), so no branch makes a load wait for the loads before it
    constexpr int kWords = (int)(sizeof(IcpCarry) / sizeof(int));
    bl.cw = reinterpret_cast<const int*>(cin)[min((int)threadIdx.x & 63, kWords - 1)];
    const int pi = min((int)threadIdx.x & 63, 11);
    bl.pw = *(pi < 9 ? st->Rp + pi : st->tp + (pi - 9));
    // no branch around the row loads either: with no previous launch
    // (prev_blocks = 0) every lane reads pin[0], and nothing is summed
    load_row_chunk<COMP>(pin, prev_blocks, max(prev_blocks, 1), 0, bl.rows);
}

// Phase B of a launch for one session: finish the previous iteration from
// (cin, pin) and put the pose of this iteration in LDS.  `first` = this
// workgroup holds the session's partial row 0: it records the solve and writes
// the carried state for the next launch.  Returns (via run) whether the
// session iterates at `level` in this launch.  Ends with a barrier.  The rows
// are summed whether or not the previous launch left a pending iteration
// (prev_blocks is that launch's row count, known to the host; a sum nobody
// needs is discarded), so no load waits for the carry.
template <bool COMP>
__device__ void begin_iteration(const BeginLoads& bl, PoseState* st, IcpCarry* cout, const float* pin,
                                int prev_blocks, int level, int blocks, bool first, float stop_angle, float stop_dist,
                                float* pose, int* run, double* tot, double* count)
{
    __shared__ IcpCarry carry;
    constexpr int kWords = (int)(sizeof(IcpCarry) / sizeof(int));
    if (prev_blocks > 0) sum_loaded_partials<COMP>(pin, prev_blocks, prev_blocks, bl.rows, tot, count);
    if (threadIdx.x < kWords) reinterpret_cast<int*>(&carry)[threadIdx.x] = bl.cw;
    else if (threadIdx.x >= 64 && threadIdx.x < 76) pose[12 + (threadIdx.x - 64)] = bl.pw;
    __syncthreads();
    if (threadIdx.x < 64) {
        double xd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const int rc = (carry.pending && carry.status == 0) ? chol6_wave(tot, xd) : 0;
        const SinCos3 sc = angle_sincos_wave(xd);
        if (threadIdx.x == 0) {
        IcpCarry c = carry;
        apply_solution(c, rc, xd, sc, tot, *count, st, first, stop_angle, stop_dist);
        const int go = (c.status == 0 && !level_flag(c.level_done, level)) ? 1 : 0;
        c.pending = go;
        c.pending_level = level;
        c.pending_blocks = blocks;
        for (int q = 0; q < 9; ++q) pose[q] = c.R[q];
        for (int q = 0; q < 3; ++q) pose[9 + q] = c.t[q];
        *run = go;
        if (first) *cout = c;
        }
    }
    __syncthreads();
}

template <bool COMP, int PIX>
__global__ __launch_bounds__(kThreads) void icp_step_kernel(IcpArgs a, const IcpCarry* cin, IcpCarry* cout,
                                                            const float* pin, int prev_blocks, float stop_angle,
                                                            float stop_dist)
{
    __shared__ float pose[24];
    __shared__ int run;
    __shared__ double tot[kTerms];
    __shared__ double count;
    // phase A: the loads of phase B, then the pose-independent loads of this
    // lane's pixels (in flight during B)
    BeginLoads bl;
    issue_begin_loads<COMP>(a.st, cin, pin, prev_blocks, bl);
    LaneLoads<PIX> ll;
    issue_lane_loads(a, ll, blockIdx.x, PIX);
    __builtin_amdgcn_sched_barrier(0);  // nothing of phase B moves above the pixel loads
    // phase B: finish the previous iteration
    begin_iteration<COMP>(bl, a.st, cout, pin, prev_blocks, a.level, (int)gridDim.x, blockIdx.x == 0, stop_angle,
                          stop_dist, pose, &run, tot, &count);
    if (!run) return;
    // phase C: live normals, association + reduction of this iteration
    LanePixels<PIX> lp;
    finish_lane_pixels(ll, lp, PIX);
    associate_reduce<COMP>(a, lp, pose, blockIdx.x, PIX);
}

// Starts a frame's ICP: the working and previous poses are the sensor pose
// (icp.cpp:27-29); from `pose` when given by value, else from st->pose.
struct Pose16 {
    float m[16];
};
__device__ void begin_state(PoseState* st, IcpCarry* c0)
{
    const float* P = st->pose;
    float Rm[9] = {P[0], P[1], P[2], P[4], P[5], P[6], P[8], P[9], P[10]};
    IcpCarry c;
    for (int q = 0; q < 9; ++q) st->R[q] = st->Rp[q] = c.R[q] = Rm[q];
    for (int q = 0; q < 3; ++q) st->t[q] = st->tp[q] = c.t[q] = P[12 + q];
    st->status = 0;
    for (int l = 0; l < 4; ++l) st->level_done[l] = c.level_done[l] = 0;
    st->iterations = 0;
    st->frame_skip_fuse = 0;
    c.status = 0;
    c.iterations = 0;
    c.pending = 0;
    c.pending_level = 0;
    c.pending_blocks = 0;
    c.pad[0] = c.pad[1] = 0;
    *c0 = c;
}
__global__ void icp_begin_kernel(PoseState* st, IcpCarry* c0, Pose16 pose, int from_value)
{
    if (threadIdx.x != 0) return;
    if (from_value)
        for (int q = 0; q < 16; ++q) st->pose[q] = pose.m[q];
    begin_state(st, c0);
}

// Commits the carried estimate as the sensor pose (Sensor::setPose,
// kinect_fusion.cpp:146-161, sensor.h:30-34); on failure the previous pose is
// kept and the frame is not fused.  One lane.
__device__ void commit_state(PoseState* st, const IcpCarry& c)
{
    for (int q = 0; q < 9; ++q) st->R[q] = c.R[q];
    for (int q = 0; q < 3; ++q) st->t[q] = c.t[q];
    for (int l = 0; l < 4; ++l) st->level_done[l] = c.level_done[l];
    st->iterations = c.iterations;
    st->status = c.status;
    if (c.status != 0) {
        if (st->sticky_status == 0) st->sticky_status = c.status;
        st->frame_skip_fuse = 1;
        return;
    }
    float* P = st->pose;
    for (int cc = 0; cc < 3; ++cc)
        for (int r = 0; r < 3; ++r) P[cc * 4 + r] = c.R[cc * 3 + r];
    P[3] = P[7] = P[11] = 0.0f;
    for (int r = 0; r < 3; ++r) P[12 + r] = c.t[r];
    P[15] = 1.0f;
    float* I = st->inv_pose;
    for (int cc = 0; cc < 3; ++cc)
        for (int r = 0; r < 3; ++r) I[cc * 4 + r] = c.R[r * 3 + cc];
    I[3] = I[7] = I[11] = 0.0f;
    for (int r = 0; r < 3; ++r) {
        double sum = 0;
        for (int q = 0; q < 3; ++q) sum += (double)c.R[r * 3 + q] * (double)c.t[q];
        I[12 + r] = (float)(-sum);
    }
    I[15] = 1.0f;
}

// Finishes the last iteration and commits the estimate.
template <bool COMP>
__device__ void commit_body(PoseState* st, const IcpCarry* cin, const float* pin, float stop_angle, float stop_dist,
                            double* tot, double* count)
{
    if (cin->pending && cin->status == 0) sum_partials<COMP>(pin, cin->pending_blocks, cin->pending_blocks, tot, count);
    if (threadIdx.x >= 64) return;
    double xd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int rc = solve_pending(cin, tot, xd);
    const SinCos3 sc = angle_sincos_wave(xd);
    if (threadIdx.x != 0) return;
    IcpCarry c = *cin;
    apply_solution(c, rc, xd, sc, tot, *count, st, true, stop_angle, stop_dist);
    commit_state(st, c);
}

template <bool COMP>
__global__ __launch_bounds__(kThreads) void icp_commit_kernel(PoseState* st, const IcpCarry* cin, const float* pin,
                                                              float stop_angle, float stop_dist)
{
    __shared__ double tot[kTerms];
    __shared__ double count;
    commit_body<COMP>(st, cin, pin, stop_angle, stop_dist, tot, &count);
}

// ---- the same loop for many sessions per launch (batch ticks).  Each
// iteration is two launches: a solve launch (one workgroup per session)
// finishes the previous iteration from the session's partial rows and puts the
// pose of this iteration into its carry; an association launch reduces every
// block of every session (workgroup ranks own contiguous, XCD-contiguous block
// ranges).  Splitting keeps the FP64 solve out of the registers of the wide
// launch.  Partial rows are indexed by the session's own block number and
// summed in that order, so a session's result does not depend on the mix and
// equals the single-session loop's.
__global__ void icp_batch_begin_kernel(const BatchHeader* hd, const IcpJob* jobs)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < hd->n_icp) begin_state(jobs[q].a.st, jobs[q].carry);
}

// final = 0: finish the pending iteration and open `level`; final = 1: finish
// the last iteration and commit the estimate (as icp_commit_kernel).
template <bool COMP>
__global__ __launch_bounds__(kThreads) void icp_batch_solve_kernel(const BatchHeader* hd, const IcpJob* jobs,
                                                                   int level, int final_, float stop_angle,
                                                                   float stop_dist)
{
    __shared__ double tot[kTerms];
    __shared__ double count;
    for (int q = blockIdx.x; q < hd->n_icp; q += gridDim.x) {
        const IcpJob& J = jobs[q];
        IcpCarry* c = J.carry;
        PoseState* st = J.a.st;
        const bool pend = c->pending && c->status == 0;
        __syncthreads();  // every lane has read the carry before lane 0 rewrites it
        if (pend) sum_partials<COMP>(J.partials, c->pending_blocks, c->pending_blocks, tot, &count);
        if (threadIdx.x < 64) {
            double xd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            const int rc = pend ? chol6_wave(tot, xd) : 0;
            const SinCos3 sc = angle_sincos_wave(xd);
            if (threadIdx.x == 0) {
                IcpCarry cc = *c;
                apply_solution(cc, rc, xd, sc, tot, count, st, true, stop_angle, stop_dist);
                if (!final_) {
                    cc.pending = (cc.status == 0 && !level_flag(cc.level_done, level)) ? 1 : 0;
                    cc.pending_level = level;
                    cc.pending_blocks = J.a.stride;
                    *c = cc;
                } else {
                    cc.pending = 0;
                    *c = cc;
                    commit_state(st, cc);
                }
            }
        }
        __syncthreads();
    }
}

template <bool COMP>
__global__ __launch_bounds__(kThreads) void icp_batch_assoc_kernel(const BatchHeader* hd, const IcpJob* jobs,
                                                                   int level)
{
    __shared__ float pose[24];
    __shared__ int run;
    const int njobs = hd->n_icp, total = hd->icp_blocks[level];
    const int G = (int)gridDim.x;
    const int rank = xcd_rank((int)blockIdx.x, G);
    const int chunk = (total + G - 1) / G;
    int b = rank * chunk;
    const int end = min(total, b + chunk);
    if (b >= end) return;
    // the table is never written while the kernel runs: scalar loads through
    // the constant address space keep the job's fields in scalar registers
    typedef __attribute__((address_space(4))) const IcpJob ConstJob;
    ConstJob* cj = (ConstJob*)jobs;
    int j = job_of(jobs, njobs, b, [](const IcpJob& J) { return J.block0; });
    int cur = -1;
    for (; b < end; ++b) {
        while (j + 1 < njobs && cj[j + 1].block0 <= b) ++j;
        j = __builtin_amdgcn_readfirstlane(j);
        IcpArgs a;
        a.V = cj[j].a.V;
        a.TV = cj[j].a.TV;
        a.TN = cj[j].a.TN;
        a.w = cj[j].a.w;
        a.h = cj[j].a.h;
#pragma unroll
        for (int q = 0; q < 9; ++q) a.K[q] = cj[j].a.K[q];
        a.dist_thresh = cj[j].a.dist_thresh;
        a.cos_thresh = cj[j].a.cos_thresh;
        a.st = cj[j].a.st;
        a.level = level;
        a.partials = cj[j].partials;
        a.stride = cj[j].a.stride;
        a.pix = cj[j].a.pix;
        const int lb = b - cj[j].block0;
        LanePixels<> lp;
        load_lane_pixels(a, lp, lb, a.pix);
        if (j != cur) {
            cur = j;
            __syncthreads();  // the previous session's pose is no longer read
            if (threadIdx.x == 0) {
                const IcpCarry* c = cj[j].carry;
                const PoseState* st = a.st;
                for (int q = 0; q < 9; ++q) pose[q] = c->R[q];
                for (int q = 0; q < 3; ++q) pose[9 + q] = c->t[q];
                for (int q = 0; q < 9; ++q) pose[12 + q] = st->Rp[q];
                for (int q = 0; q < 3; ++q) pose[21 + q] = st->tp[q];
                run = c->pending;
            }
            __syncthreads();
        } else {
            __syncthreads();  // the previous block's reduction LDS is free again
        }
        if (!run) continue;
        associate_reduce<COMP>(a, lp, pose, lb, a.pix);
    }
}

inline int blocks_for(int64_t n, int pix = kPixPerThread)
{
    return (int)((n + (int64_t)kThreads * pix - 1) / ((int64_t)kThreads * pix));
}

}  // namespace

// Pixels per lane of an iteration launch over an n-pixel map: the most (4, 2
// or 1) that still gives kFillBlocks workgroups, else 1.  An iteration's
// association is one lane's pixels in sequence, so a coarse level spread over
// more workgroups ends sooner.  A function of n alone: the batch and the
// single-session loop choose alike and sum the same partial rows.
int icp_pix_for(int64_t n)
{
    constexpr int kFillBlocks = 256;
    for (int pix = kPixPerThread; pix > 1; pix >>= 1)
        if (blocks_for(n, pix) >= kFillBlocks) return pix;
    return 1;
}
int64_t icp_max_map_pixels() { return kMaxMapPixels; }
int icp_level_blocks(int w, int h)
{
    const int64_t n = (int64_t)w * h;
    return blocks_for(n, icp_pix_for(n));
}
int icp_blocks_for(int w, int h)
{
    int m = 0;
    for (int l = 0; l < 3; ++l) m = std::max(m, icp_level_blocks(std::max(w >> l, 1), std::max(h >> l, 1)));
    return m;
}

static IcpCarry* carry_of(kf_icp* icp, int k) { return reinterpret_cast<IcpCarry*>(icp->carry) + (k & 1); }
static float* partials_of(kf_icp* icp, int k)
{
    return icp->partials + (size_t)(k & 1) * kPartialStride * (size_t)icp->max_blocks;
}

int launch_icp_begin(kf_icp* icp, const float* pose16, hipStream_t s)
{
    Pose16 p{};
    if (pose16) memcpy(p.m, pose16, sizeof(p.m));
    icp_begin_kernel<<<1, 64, 0, s>>>(icp->state, carry_of(icp, 0), p, pose16 ? 1 : 0);
    KF_LAUNCHED("icp_begin_kernel");
    return KF_OK;
}

static IcpArgs level_args(const kf_icp* icp, const IcpLevel& L, int l)
{
    IcpArgs a{};
    a.V = L.V;
    a.TV = L.TV;
    a.TN = L.TN;
    a.w = L.w;
    a.h = L.h;
    memcpy(a.K, L.K, sizeof(a.K));
    a.dist_thresh = icp->cfg.distance_thresh;
    a.cos_thresh = icp->cos_thresh;
    a.st = icp->state;
    a.level = l;
    a.pix = icp_pix_for((int64_t)L.w * L.h);
    a.stride = blocks_for((int64_t)L.w * L.h, a.pix);
    return a;
}

// Enqueues the iteration launches; `events` (optional) receives one event
// after every launch.  Returns the number of launches in *launches.
static int enqueue_steps(kf_icp* icp, const IcpLevel* levels, int n_levels, hipStream_t s, hipEvent_t* events,
                         int* launches)
{
    const float sa = icp->cfg.iteration_stop_thresh_angle, sd = icp->cfg.iteration_stop_thresh_distance;
    int k = 0, prev_blocks = 0;  // row count of the previous launch (0: none this frame)
    for (int l = n_levels - 1; l >= 0; --l)
        for (int it = 0; it < icp->cfg.iters_per_layer[l]; ++it, ++k) {
            IcpArgs a = level_args(icp, levels[l], l);
            a.partials = partials_of(icp, k);
            const int blocks = a.stride;
            if (blocks > icp->max_blocks || (int64_t)a.w * a.h > kMaxMapPixels)
                return fail(KF_ERR_DIMS, "ICP: map larger than the configured width x height");
            using Step = void (*)(IcpArgs, const IcpCarry*, IcpCarry*, const float*, int, float, float);
            const Step steps[2][3] = {{icp_step_kernel<false, 1>, icp_step_kernel<false, 2>, icp_step_kernel<false, 4>},
                                      {icp_step_kernel<true, 1>, icp_step_kernel<true, 2>, icp_step_kernel<true, 4>}};
            const Step step = steps[icp->cfg.compensated ? 1 : 0][a.pix == 4 ? 2 : a.pix == 2 ? 1 : 0];
            step<<<blocks, kThreads, 0, s>>>(a, carry_of(icp, k), carry_of(icp, k + 1), partials_of(icp, k + 1),
                                             prev_blocks, sa, sd);
            KF_LAUNCHED("icp_step_kernel");
            if (events) KF_HIP(hipEventRecord(events[k], s));
            prev_blocks = blocks;
        }
    if (icp->cfg.compensated)
        icp_commit_kernel<true><<<1, kThreads, 0, s>>>(icp->state, carry_of(icp, k), partials_of(icp, k + 1), sa, sd);
    else
        icp_commit_kernel<false><<<1, kThreads, 0, s>>>(icp->state, carry_of(icp, k), partials_of(icp, k + 1), sa, sd);
    KF_LAUNCHED("icp_commit_kernel");
    if (launches) *launches = k;
    return KF_OK;
}

int launch_icp_loop(kf_icp* icp, const IcpLevel* levels, int n_levels, hipStream_t s)
{
    return enqueue_steps(icp, levels, n_levels, s, nullptr, nullptr);
}

int launch_icp_batch(const BatchHeader* hd, const IcpJob* jobs, const int* iters, int levels, int max_sessions,
                     int compensated, float stop_angle, float stop_dist, unsigned step_blocks, hipStream_t s)
{
    if (max_sessions <= 0) return KF_OK;
    icp_batch_begin_kernel<<<(unsigned)((max_sessions + 63) / 64), 64, 0, s>>>(hd, jobs);
    KF_LAUNCHED("icp_batch_begin_kernel");
    const IcpJob* level_jobs[3] = {jobs, jobs + max_sessions, jobs + 2 * max_sessions};
    const unsigned sb = (unsigned)std::min(max_sessions, 1024);
    for (int l = levels - 1; l >= 0; --l)
        for (int it = 0; it < iters[l]; ++it) {
            if (compensated) {
                icp_batch_solve_kernel<true><<<sb, kThreads, 0, s>>>(hd, level_jobs[l], l, 0, stop_angle, stop_dist);
                icp_batch_assoc_kernel<true><<<step_blocks, kThreads, 0, s>>>(hd, level_jobs[l], l);
            } else {
                icp_batch_solve_kernel<false><<<sb, kThreads, 0, s>>>(hd, level_jobs[l], l, 0, stop_angle, stop_dist);
                icp_batch_assoc_kernel<false><<<step_blocks, kThreads, 0, s>>>(hd, level_jobs[l], l);
            }
            KF_LAUNCHED("icp_batch_assoc_kernel");
        }
    if (compensated)
        icp_batch_solve_kernel<true><<<sb, kThreads, 0, s>>>(hd, jobs, 0, 1, stop_angle, stop_dist);
    else
        icp_batch_solve_kernel<false><<<sb, kThreads, 0, s>>>(hd, jobs, 0, 1, stop_angle, stop_dist);
    KF_LAUNCHED("icp_batch_solve_kernel");
    return KF_OK;
}

// ---- LinearLeastSquares::solve without allocation (kf_solve_normal_equations_noalloc).
// Its working memory -- the PoseState the solve writes and the per-block partial
// rows -- lives in one file-scope device array, `scratch`, of kNoAllocBytes on
// every GPU that runs it (held until the process ends).  Host code reaches a
// device's copy only through the address hipGetSymbolAddress returns for it
// (noalloc_scratch); the array's name is a host-side stand-in, not a device
// address.  Every call on a GPU shares that GPU's copy.
constexpr unsigned kNoAllocMaxEquations = 1u << 22;
constexpr int kNoAllocMaxBlocks = (int)((kNoAllocMaxEquations + kThreads * kPixPerThread - 1) /
                                        (kThreads * kPixPerThread));
constexpr size_t kNoAllocStateBytes = (sizeof(PoseState) + 255) & ~(size_t)255;
constexpr size_t kNoAllocBytes = kNoAllocStateBytes + sizeof(float) * kPartialStride * kNoAllocMaxBlocks;

__global__ void noalloc_clear_state_kernel(PoseState* st)
{
    int* w = reinterpret_cast<int*>(st);
    for (size_t q = threadIdx.x; q < sizeof(PoseState) / sizeof(int); q += blockDim.x) w[q] = 0;
}

__global__ void noalloc_result_kernel(const PoseState* st, float* x, int* status)
{
    if (threadIdx.x < 6) x[threadIdx.x] = st->last_x[threadIdx.x];
    if (threadIdx.x == 0 && status) *status = st->status;
}

}  // namespace kf

static __device__ __align__(256) unsigned char scratch[kf::kNoAllocBytes];

namespace {

// The device address of this GPU's `scratch`, looked up once per device.
int noalloc_scratch(char** out)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static char* ptr[kMaxDevices] = {};
    int dev = 0;
    KF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return kf::fail(KF_ERR_ARG, "solve (no allocation): device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return kf::fail_hip(e, "solve (no allocation): scratch address");
        ptr[dev] = static_cast<char*>(p);
    }
    *out = ptr[dev];
    return KF_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

int kf_construct_icp_residuals(const float* vertex_map, const float* target_vertex_map, const float* target_normal_map,
                               int w, int h, const float prev_rot[9], const float prev_transl[3],
                               const float cur_rot[9], const float cur_transl[3], const float intrinsics[9],
                               float distance_thresh, float angle_thresh, float* mat_a, float* vec_b,
                               kf_stream_t stream, float* elapsed_ms)
{
    if (!vertex_map || !target_vertex_map || !target_normal_map || !mat_a || !vec_b || w <= 0 || h <= 0 || !prev_rot ||
        !prev_transl || !cur_rot || !cur_transl || !intrinsics)
        return kf::fail(KF_ERR_ARG, "constructIcpResiduals: bad argument");
    hipStream_t s = kf::stream_of(stream);
    // the pose travels by value through a 24-float staging block
    float pose[24];
    memcpy(pose, cur_rot, 9 * sizeof(float));
    memcpy(pose + 9, cur_transl, 3 * sizeof(float));
    memcpy(pose + 12, prev_rot, 9 * sizeof(float));
    memcpy(pose + 21, prev_transl, 3 * sizeof(float));
    float* dpose = nullptr;
    KF_HIP(hipMalloc(&dpose, sizeof(pose)));
    kf::IcpArgs a{};
    a.V = reinterpret_cast<const float4*>(vertex_map);
    a.TV = reinterpret_cast<const float4*>(target_vertex_map);
    a.TN = reinterpret_cast<const float4*>(target_normal_map);
    a.w = w;
    a.h = h;
    memcpy(a.K, intrinsics, sizeof(a.K));
    a.dist_thresh = distance_thresh;
    a.cos_thresh = (float)cos((double)angle_thresh);
    a.partials = dpose;
    kf::CallTimer t(elapsed_ms, s);
    int st = KF_OK;
    hipError_t e = hipMemcpyAsync(dpose, pose, sizeof(pose), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) st = kf::fail_hip(e, "constructIcpResiduals pose copy");
    if (st == KF_OK) st = t.start();
    if (st == KF_OK) {
        const unsigned blocks = (unsigned)(((int64_t)w * h + kf::kThreads - 1) / kf::kThreads);
        kf::icp_rows_kernel<<<blocks, kf::kThreads, 0, s>>>(a, mat_a, vec_b);
        e = hipGetLastError();
        if (e != hipSuccess) st = kf::fail_hip(e, "icp_rows_kernel");
    }
    if (st == KF_OK) st = t.finish();
    e = hipStreamSynchronize(s);
    if (st == KF_OK && e != hipSuccess) st = kf::fail_hip(e, "constructIcpResiduals");
    (void)hipFree(dpose);
    return st;
}

// Shared by kf_solve_normal_equations and kf_icp_reduce: partial rows -> sums/solve.
static int reduce_to_state(kf::PoseState** st_out, float** partials_out, int blocks)
{
    KF_HIP(hipMalloc(st_out, sizeof(kf::PoseState)));
    KF_HIP(hipMemset(*st_out, 0, sizeof(kf::PoseState)));
    KF_HIP(hipMalloc(partials_out, sizeof(float) * kf::kPartialStride * (size_t)(blocks > 0 ? blocks : 1)));
    return KF_OK;
}

int kf_solve_normal_equations(const float* mat_a, const float* vec_b, unsigned int n_equations, float result[6],
                              kf_stream_t stream, float* elapsed_ms)
{
    if (!mat_a || !vec_b || !result || n_equations == 0)
        return kf::fail(KF_ERR_ARG, "LinearLeastSquares::solve: bad argument");
    hipStream_t s = kf::stream_of(stream);
    const int blocks = kf::blocks_for(n_equations);
    kf::PoseState* st = nullptr;
    float* partials = nullptr;
    int rc = reduce_to_state(&st, &partials, blocks);
    kf::CallTimer t(elapsed_ms, s);
    if (rc == KF_OK) rc = t.start();
    if (rc == KF_OK) {
        kf::ab_reduce_kernel<false><<<blocks, kf::kThreads, 0, s>>>(mat_a, vec_b, n_equations, partials, blocks);
        kf::icp_finalize_kernel<false><<<1, kf::kThreads, 0, s>>>(st, partials, blocks, blocks, 0, 0.0f, 0.0f, 1);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = kf::fail_hip(e, "solve kernels");
    }
    if (rc == KF_OK) rc = t.finish();
    kf::PoseState hs;
    if (rc == KF_OK) {
        hipError_t e = hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = kf::fail_hip(e, "solve readback");
    }
    if (rc == KF_OK) {
        for (int q = 0; q < 6; ++q) result[q] = hs.last_x[q];
        if (hs.status) rc = kf::fail(hs.status, kf_status_string(hs.status));
    }
    (void)hipFree(st);
    (void)hipFree(partials);
    return rc;
}

int kf_solve_normal_equations_noalloc(const float* mat_a, const float* vec_b, unsigned int n_equations,
                                      float* result_dev, int* status_dev, kf_stream_t stream)
{
    if (!mat_a || !vec_b || !result_dev || n_equations == 0)
        return kf::fail(KF_ERR_ARG, "solve (no allocation): bad argument");
    if (n_equations > kf::kNoAllocMaxEquations)
        return kf::fail(KF_ERR_ARG, "solve (no allocation): more than 2^22 equations");
    char* base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    auto* st = reinterpret_cast<kf::PoseState*>(base);
    auto* partials = reinterpret_cast<float*>(base + kf::kNoAllocStateBytes);
    hipStream_t s = kf::stream_of(stream);
    const int blocks = kf::blocks_for(n_equations);
    kf::noalloc_clear_state_kernel<<<1, 256, 0, s>>>(st);
    kf::ab_reduce_kernel<false><<<blocks, kf::kThreads, 0, s>>>(mat_a, vec_b, n_equations, partials, blocks);
    kf::icp_finalize_kernel<false><<<1, kf::kThreads, 0, s>>>(st, partials, blocks, blocks, 0, 0.0f, 0.0f, 1);
    kf::noalloc_result_kernel<<<1, 64, 0, s>>>(st, result_dev, status_dev);
    KF_HIP(hipGetLastError());
    return KF_OK;
}

int kf_icp_reduce(const float* vertex_map, const float* target_vertex_map, const float* target_normal_map, int w,
                  int h, const float prev_rot[9], const float prev_transl[3], const float cur_rot[9],
                  const float cur_transl[3], const float intrinsics[9], float distance_thresh, float angle_thresh,
                  int compensated, double sums_host[27], int64_t* inliers, kf_stream_t stream, float* elapsed_ms)
{
    if (!vertex_map || !target_vertex_map || !target_normal_map || !sums_host || w <= 0 || h <= 0)
        return kf::fail(KF_ERR_ARG, "kf_icp_reduce: bad argument");
    if ((int64_t)w * h > kf::kMaxMapPixels) return kf::fail(KF_ERR_DIMS, "kf_icp_reduce: map above 2^24 pixels");
    hipStream_t s = kf::stream_of(stream);
    const int blocks = kf::blocks_for((int64_t)w * h);
    kf::PoseState* st = nullptr;
    float* partials = nullptr;
    int rc = reduce_to_state(&st, &partials, blocks);
    kf::PoseState hs;
    memset(&hs, 0, sizeof(hs));
    memcpy(hs.R, cur_rot, sizeof(hs.R));
    memcpy(hs.t, cur_transl, sizeof(hs.t));
    memcpy(hs.Rp, prev_rot, sizeof(hs.Rp));
    memcpy(hs.tp, prev_transl, sizeof(hs.tp));
    if (rc == KF_OK) {
        hipError_t e = hipMemcpy(st, &hs, sizeof(hs), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = kf::fail_hip(e, "kf_icp_reduce state");
    }
    kf::IcpArgs a{};
    a.V = reinterpret_cast<const float4*>(vertex_map);
    a.TV = reinterpret_cast<const float4*>(target_vertex_map);
    a.TN = reinterpret_cast<const float4*>(target_normal_map);
    a.w = w;
    a.h = h;
    memcpy(a.K, intrinsics, sizeof(a.K));
    a.dist_thresh = distance_thresh;
    a.cos_thresh = (float)cos((double)angle_thresh);
    a.st = st;
    a.level = 0;
    a.partials = partials;
    a.stride = blocks;
    kf::CallTimer t(elapsed_ms, s);
    if (rc == KF_OK) rc = t.start();
    if (rc == KF_OK) {
        if (compensated) {
            kf::icp_reduce_kernel<true><<<blocks, kf::kThreads, 0, s>>>(a);
            kf::icp_finalize_kernel<true><<<1, kf::kThreads, 0, s>>>(st, partials, blocks, blocks, 0, 0, 0, 1);
        } else {
            kf::icp_reduce_kernel<false><<<blocks, kf::kThreads, 0, s>>>(a);
            kf::icp_finalize_kernel<false><<<1, kf::kThreads, 0, s>>>(st, partials, blocks, blocks, 0, 0, 0, 1);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = kf::fail_hip(e, "kf_icp_reduce kernels");
    }
    if (rc == KF_OK) rc = t.finish();
    if (rc == KF_OK) {
        hipError_t e = hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = kf::fail_hip(e, "kf_icp_reduce readback");
    }
    if (rc == KF_OK) {
        for (int q = 0; q < 27; ++q) sums_host[q] = hs.last_sums[q];
        if (inliers) *inliers = hs.last_inliers;
    }
    (void)hipFree(st);
    (void)hipFree(partials);
    return rc;
}

int kf_icp_create(const kf_icp_config* config, kf_icp** out)
{
    if (!config || !out) return kf::fail(KF_ERR_ARG, "kf_icp_create: null argument");
    *out = nullptr;
    if (config->levels < 1 || config->levels > 3 || config->width <= 0 || config->height <= 0)
        return kf::fail(KF_ERR_ARG, "kf_icp_create: levels must be 1..3 and sizes positive");
    if ((int64_t)config->width * config->height > kf::icp_max_map_pixels())
        return kf::fail(KF_ERR_DIMS, "kf_icp_create: width x height above 2^24 pixels");
    for (int l = 0; l < config->levels; ++l)
        if (config->iters_per_layer[l] < 0) return kf::fail(KF_ERR_ARG, "kf_icp_create: negative iteration count");
    kf_icp* icp = new kf_icp();
    icp->cfg = *config;
    icp->cos_thresh = (float)cos((double)config->angle_thresh);
    icp->max_blocks = kf::icp_blocks_for(config->width, config->height);
    hipError_t e;
    if ((e = hipMalloc(&icp->state, sizeof(kf::PoseState))) != hipSuccess ||
        (e = hipMemset(icp->state, 0, sizeof(kf::PoseState))) != hipSuccess ||
        (e = hipMalloc(&icp->carry, 2 * sizeof(kf::IcpCarry))) != hipSuccess ||
        (e = hipMemset(icp->carry, 0, 2 * sizeof(kf::IcpCarry))) != hipSuccess ||
        (e = hipMalloc(&icp->partials, 2 * sizeof(float) * kf::kPartialStride * (size_t)icp->max_blocks)) != hipSuccess) {
        kf_icp_destroy(icp);
        return kf::fail_hip(e, "kf_icp_create");
    }
    *out = icp;
    return KF_OK;
}

int kf_icp_destroy(kf_icp* icp)
{
    if (!icp) return KF_OK;
    (void)hipFree(icp->state);
    (void)hipFree(icp->carry);
    (void)hipFree(icp->partials);
    delete icp;
    return KF_OK;
}

int kf_icp_compute_pose(kf_icp* icp, const float* const vertex_pyr[], const float* const target_vertex_pyr[],
                        const float* const target_normal_pyr[], const kf_sensor* sensor, float pose_out[16],
                        kf_stream_t stream, float* elapsed_ms)
{
    if (!icp || !vertex_pyr || !target_vertex_pyr || !target_normal_pyr || !sensor || !pose_out)
        return kf::fail(KF_ERR_ARG, "ICP::computePose: null argument");
    hipStream_t s = kf::stream_of(stream);
    kf::IcpLevel lv[3];
    for (int l = 0; l < icp->cfg.levels; ++l) {
        lv[l].V = reinterpret_cast<const float4*>(vertex_pyr[l]);
        lv[l].TV = reinterpret_cast<const float4*>(target_vertex_pyr[l]);
        lv[l].TN = reinterpret_cast<const float4*>(target_normal_pyr[l]);
        lv[l].w = icp->cfg.width >> l;
        lv[l].h = icp->cfg.height >> l;
        memcpy(lv[l].K, sensor->intr[l], sizeof(lv[l].K));
    }
    // Events after every launch (ICP::getExecutionTimes): times[0] = the iteration
    // launches (association, reduction and the solve of the previous iteration),
    // times[1] = the final solve and commit.
    int total = 0;
    for (int l = 0; l < icp->cfg.levels; ++l) total += icp->cfg.iters_per_layer[l];
    std::vector<hipEvent_t> ev((size_t)total + 2, nullptr);
    int rc = KF_OK;
    for (auto& e : ev)
        if (rc == KF_OK && hipEventCreate(&e) != hipSuccess) rc = kf::fail(KF_ERR_HIP, "ICP event create");
    if (rc == KF_OK) rc = kf::launch_icp_begin(icp, sensor->pose, s);
    if (rc == KF_OK && hipEventRecord(ev[0], s) != hipSuccess) rc = kf::fail(KF_ERR_HIP, "ICP event record");
    int k = 0;
    if (rc == KF_OK) rc = kf::enqueue_steps(icp, lv, icp->cfg.levels, s, ev.data() + 1, &k);
    if (rc == KF_OK && hipEventRecord(ev.back(), s) != hipSuccess) rc = kf::fail(KF_ERR_HIP, "ICP event record");
    kf::PoseState hs;
    if (rc == KF_OK) {
        hipError_t e = hipMemcpyAsync(&hs, icp->state, sizeof(hs), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = kf::fail_hip(e, "ICP readback");
    }
    if (rc == KF_OK) {
        icp->times[0] = icp->times[1] = 0.0f;
        if (k > 0) (void)hipEventElapsedTime(&icp->times[0], ev[0], ev[k]);
        (void)hipEventElapsedTime(&icp->times[1], ev[k], ev.back());
        if (elapsed_ms) (void)hipEventElapsedTime(elapsed_ms, ev[0], ev.back());
        icp->last_iterations = hs.iterations;
        // the last good estimate (the commit does not run on failure)
        float P[16] = {0};
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) P[c * 4 + r] = hs.R[c * 3 + r];
        for (int r = 0; r < 3; ++r) P[12 + r] = hs.t[r];
        P[15] = 1.0f;
        memcpy(pose_out, P, sizeof(P));
        if (hs.status) rc = kf::fail(hs.status, kf_status_string(hs.status));
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int kf_icp_execution_times(const kf_icp* icp, float times[2])
{
    if (!icp || !times) return kf::fail(KF_ERR_ARG, "kf_icp_execution_times: null argument");
    times[0] = icp->times[0];
    times[1] = icp->times[1];
    return KF_OK;
}

int kf_icp_iterations(const kf_icp* icp, int* iterations)
{
    if (!icp || !iterations) return kf::fail(KF_ERR_ARG, "kf_icp_iterations: null argument");
    *iterations = icp->last_iterations;
    return KF_OK;
}

}  // extern "C"
