/*
 * kfref.c — CPU restatement of the reference's per-frame KinectFusion path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker or
 * the CPU baseline, never as the thing measured or shipped.
 *
 * Compiled twice (see oracle/Makefile):
 *   default       -> kfo_*   FP32 ("real" = float)
 *   -DKFREF_F64   -> kfo64_* FP64 ("real" = double), same algorithm
 * Both with -ffp-contract=off and no fast-math, so every + - * / sqrt is one
 * IEEE-754 correctly rounded operation, in the order written here.  The HIP
 * kernels evaluate the same expressions in the same order, which is what makes
 * masks, touched voxels and weights bit-exact between the two.
 *
 * Operation order follows glm as the reference uses it (glm is an empty
 * submodule in the reference, 3rd/glm; the orders restated here are glm's
 * documented operator definitions):
 *   mat3*vec3 : r_i = (m[0][i]*v.x + m[1][i]*v.y) + m[2][i]*v.z
 *   mat4*vec4 : r   = (m[0]*v.x + m[1]*v.y) + (m[2]*v.z + m[3]*v.w)
 *   mat3*mat3 : C[c][r] = (A[0][r]*B[c][0] + A[1][r]*B[c][1]) + A[2][r]*B[c][2]
 *   dot(a,b)  : (a.x*b.x + a.y*b.y) + a.z*b.z
 *   normalize : v * (1 / sqrt(dot(v,v)))
 *   cross     : (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y)
 *
 * Parity pins (what checks each function besides the GPU comparison) are noted
 * per function as "pinned by:"; the tests live in tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <time.h>

#ifdef KFREF_F64
typedef double real;
typedef struct { double x, y, z, w; } real4;
#define KFO(name) kfo64_##name
#define R(c) ((double)(c))
#define SQRT sqrt
#define FLOOR floor
#define FABS fabs
#define FMIN fmin
#define FMAX fmax
#else
#include "kfref.h"
typedef float real;
typedef kfo_float4 real4;
#define KFO(name) kfo_##name
#define R(c) ((float)(c))
#define SQRT sqrtf
#define FLOOR floorf
#define FABS fabsf
#define FMIN fminf
#define FMAX fmaxf
#endif

#define KF_VALID   R(-1.0)   /* device_helper.cuh:9  */
#define KF_INVALID R(-2.0)   /* device_helper.cuh:10 */
#define KF_MIN_DEPTH R(0.01) /* device_helper.cuh:11 */
#define KF_MIN_DIST R(0.4)   /* tsdf.cuh:11 */
#define KF_MAX_DIST R(6.0)   /* tsdf.cuh:12 */

typedef struct { real x, y, z; } v3;

static inline v3 mk3(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3s(v3 a, real s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline v3 div3s(v3 a, real s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline real dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline real len3(v3 a) { return SQRT(dot3(a, a)); }
static inline v3 normalize3(v3 a) { real inv = R(1.0) / SQRT(dot3(a, a)); return mul3s(a, inv); }
static inline v3 cross3(v3 x, v3 y)
{
    return mk3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
/* m column-major 3x3 */
static inline v3 m3v(const real* m, v3 v)
{
    return mk3((m[0] * v.x + m[3] * v.y) + m[6] * v.z,
               (m[1] * v.x + m[4] * v.y) + m[7] * v.z,
               (m[2] * v.x + m[5] * v.y) + m[8] * v.z);
}
static inline void m3m3(const real* A, const real* B, real* C)
{
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            C[c * 3 + r] = (A[0 * 3 + r] * B[c * 3 + 0] + A[1 * 3 + r] * B[c * 3 + 1]) + A[2 * 3 + r] * B[c * 3 + 2];
}
static inline void m3t(const real* A, real* T)
{
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) T[c * 3 + r] = A[r * 3 + c];
}
/* glm mat4*vec4 with w given; returns xyz */
static inline v3 m4v(const real* m, v3 p, real w)
{
    real a0 = m[0] * p.x, a1 = m[4] * p.y, a2 = m[8] * p.z, a3 = m[12] * w;
    real b0 = m[1] * p.x, b1 = m[5] * p.y, b2 = m[9] * p.z, b3 = m[13] * w;
    real c0 = m[2] * p.x, c1 = m[6] * p.y, c2 = m[10] * p.z, c3 = m[14] * w;
    return mk3((a0 + a1) + (a2 + a3), (b0 + b1) + (b2 + b3), (c0 + c1) + (c2 + c3));
}
static inline int is_depth_valid(real d) { return d > KF_MIN_DEPTH && isfinite(d); }
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int map_valid(const real4* m, int w, int h, int i, int j)
{
    /* device_helper.cuh:43-50: reads outside the map return 0 (cudaBoundaryModeZero) -> invalid */
    if (i < 0 || j < 0 || i >= w || j >= h) return 0;
    return m[(int64_t)j * w + i].w == KF_VALID;
}
static inline v3 map_xyz(const real4* m, int w, int i, int j)
{
    const real4* e = &m[(int64_t)j * w + i];
    return mk3(e->x, e->y, e->z);
}
static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ------------------------------------------------------------------------- */
/* Raw depth code -> binary16 (round to nearest even) -> float.
 * measurement.cu:19-21 reads the 16-bit cell as unsigned short and converts it
 * numerically through __half (IEEE-754 binary16, RNE); codes >= 65520 overflow
 * to +inf.  Pinned by: exhaustive comparison over all 65536 codes with numpy's
 * float16 conversion (IEEE RNE) in tests/test_oracle_measurement.py. */
static float half_rne(uint16_t code)
{
    uint32_t c = code;
    if (c <= 2048u) return (float)c;             /* 11 significant bits: exact */
    int bits = 32 - __builtin_clz(c);            /* 12..16 */
    int drop = bits - 11;
    uint32_t q = c >> drop, r = c & ((1u << drop) - 1u), half = 1u << (drop - 1);
    if (r > half || (r == half && (q & 1u))) q += 1u;
    uint32_t v = q << drop;
    if (v > 65504u) return INFINITY;
    return (float)v;
}
float KFO(u16_to_half_float)(uint16_t code) { return half_rne(code); }

/* exp for the bilateral weights.  The reference calls expf (measurement.cu:58),
 * built with --use_fast_math (project.vcxproj:139); its result is not
 * reproducible here.  Both this oracle and the HIP kernel use the same
 * project-written single-precision exp: Cody-Waite reduction by ln2 and the
 * degree-6 polynomial of the Cephes library's expf (S. L. Moshier), evaluated
 * with fused multiply-adds (fmaf: one rounding per step, on both sides), followed
 * by an exact power-of-two scaling with a single final rounding (denormals kept).
 * Pinned by: max relative error vs FP64 exp over a dense sweep of [-104, 88]
 * (tests/test_oracle_measurement.py). */
#ifndef KFREF_F64
static float kf_expf_impl(float x)
{
    if (!(x == x)) return x;
    if (x < -104.0f) return 0.0f;
    if (x > 88.72283f) return INFINITY;
    float fk = floorf(x * 1.44269504088896341f + 0.5f);
    int k = (int)fk;
    float r = x - fk * 0.693359375f;  /* exact product (9-bit constant) */
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
    if (k > 127) { y = y * 1.7014118346046923e38f; k -= 127; }     /* * 2^127, exact */
    if (k < -126) { y = y * 5.421010862427522e-20f; k += 64; }      /* * 2^-64, exact */
    union { uint32_t u; float f; } s;
    s.u = (uint32_t)(k + 127) << 23;
    return y * s.f;
}
float kfo_expf(float x) { return kf_expf_impl(x); }
#endif

#ifndef KFREF_F64
#ifdef _OPENMP
#include <omp.h>
#endif
/* Threads the oracle's parallel loops use (the CPU baseline's core count). */
int kfo_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
/* Sets the threads of the following parallel loops (the CPU baseline's single-thread figure). */
void kfo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
#endif
static inline real kf_exp(real x)
{
#ifdef KFREF_F64
    return exp(x);
#else
    return kf_expf_impl(x);
#endif
}

/* Intrinsics at pyramid level L.  sensor.h:10-20 hard-codes c = (319.5,239.5),
 * (159.5,119.5), (79.5,59.5) for f, f/2, f/4; the generalised rule used here,
 * c_L = (c_0 + 0.5) / 2^L - 0.5 and f_L = f / 2^L, reproduces those values for
 * 640x480 and extends them to other frame sizes.  K_inv is glm::inverse of the
 * 3x3 (cofactor form, glm func_matrix), evaluated in the same order.
 * Pinned by: K*K_inv == I within 1 ulp and the reference's own level constants
 * (tests/test_oracle_icp.py). */
void KFO(intrinsics)(real fx, real fy, real cx, real cy, int level, real* K, real* Ki)
{
    real s = R(1.0);
    for (int l = 0; l < level; ++l) s = s * R(0.5);
    real fxl = fx * s, fyl = fy * s;
    real cxl = (cx + R(0.5)) * s - R(0.5), cyl = (cy + R(0.5)) * s - R(0.5);
    real m[9] = {fxl, 0, 0, 0, fyl, 0, cxl, cyl, 1};
    memcpy(K, m, sizeof(m));
#define M(c, r) m[(c) * 3 + (r)]
    real det = (+M(0, 0) * (M(1, 1) * M(2, 2) - M(2, 1) * M(1, 2))
                - M(1, 0) * (M(0, 1) * M(2, 2) - M(2, 1) * M(0, 2)))
               + M(2, 0) * (M(0, 1) * M(1, 2) - M(1, 1) * M(0, 2));
    real ood = R(1.0) / det;
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

/* ------------------------------------------------------------------------- */
/* measurement.cu:14-31.  d = float(half(code)) * scale, optional mirror of the
 * column (the reference hard-codes 639 - i; generalised to w-1-i). */
void KFO(convert_to_depth_meters)(const uint16_t* raw, real* out, int w, int h, real scale, int unmirror)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < h; ++j)
        for (int i = 0; i < w; ++i) {
            real d = (real)half_rne(raw[(int64_t)j * w + i]);
            int io = unmirror ? (w - 1 - i) : i;
            out[(int64_t)j * w + io] = d * scale;
        }
}

/* measurement.cu:33-66.  7x7 bilateral, sigma_s = 4 px, sigma_r = 0.25 m,
 * x outer / y inner, clamp-to-edge reads, invalid zeros take part.
 * Pinned by: constant image -> identity; FP64 brute force on small images. */
void KFO(apply_bilateral_filter)(const real* in, real* out, int w, int h)
{
    const real inv_s = R(1.0) / (R(4.0) * R(4.0));    /* measurement.cu:42 */
    const real inv_r = R(1.0) / (R(0.25) * R(0.25));  /* measurement.cu:43 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < h; ++j)
        for (int i = 0; i < w; ++i) {
            real center = in[(int64_t)j * w + i];
            real norm = 0, acc = 0;
            for (int x = -3; x <= 3; ++x)
                for (int y = -3; y <= 3; ++y) {
                    real cur = in[(int64_t)clampi(j + y, 0, h - 1) * w + clampi(i + x, 0, w - 1)];
                    real sd = (real)(x * x + y * y);
                    real id = center - cur;
                    id = id * id;
                    real f = kf_exp(-sd * inv_s - id * inv_r);
                    norm = norm + f;
                    acc = acc + f * cur;
                }
            out[(int64_t)j * w + i] = acc / norm;
        }
}

/* measurement.cu:68-102.  2x2 -> 1: average of the taps within 3*sigma_r of the
 * top-left tap, in the order f2, f3, f4.  in is w x h, out is (w/2) x (h/2).
 * Pinned by: constant -> identity and each 3-sigma branch (tests). */
void KFO(down_sample)(const real* in, int w, int h, real* out)
{
    const real tsd = R(3.0) * R(0.25);
    int wo = w / 2, ho = h / 2;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ho; ++j)
        for (int i = 0; i < wo; ++i) {
            real f1 = in[(int64_t)(2 * j) * w + 2 * i], f2 = in[(int64_t)(2 * j) * w + 2 * i + 1];
            real f3 = in[(int64_t)(2 * j + 1) * w + 2 * i], f4 = in[(int64_t)(2 * j + 1) * w + 2 * i + 1];
            real acc = f1;
            int cnt = 1;
            if (FABS(f1 - f2) <= tsd) { acc = acc + f2; ++cnt; }
            if (FABS(f1 - f3) <= tsd) { acc = acc + f3; ++cnt; }
            if (FABS(f1 - f4) <= tsd) { acc = acc + f4; ++cnt; }
            out[(int64_t)j * wo + i] = acc / (real)cnt;
        }
}

/* measurement.cu:104-126.  v = (K^-1 (i+1/2, j+1/2, 1)) * d, w = -1; invalid ->
 * (0,0,0,-2).  Pinned by: closed form (i+1/2-cx)/fx*d (tests). */
void KFO(create_vertex_map)(const real* depth, int w, int h, const real* Ki, real4* out)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < h; ++j)
        for (int i = 0; i < w; ++i) {
            real d = depth[(int64_t)j * w + i];
            real4 o = {0, 0, 0, KF_INVALID};
            if (is_depth_valid(d)) {
                v3 p = m3v(Ki, mk3((real)i + R(0.5), (real)j + R(0.5), R(1.0)));
                p = mul3s(p, d);
                o.x = p.x; o.y = p.y; o.z = p.z; o.w = KF_VALID;
            }
            out[(int64_t)j * w + i] = o;
        }
}

/* device_helper.cuh:81-110.  normalize(cross(v(i+1,j)-v, v(i,j+1)-v)) when the
 * three vertices are valid (out of range = invalid), else (0,0,0,-2).
 * Pinned by: the reference's TestComputeNormal with its validity flag corrected
 * (transformation_tests.cpp:41-59 -> (0,0,1)), and plane closed forms. */
void KFO(compute_normal)(const real4* v, int w, int h, int i, int j, real4* out)
{
    real4 o = {0, 0, 0, KF_INVALID};
    if (map_valid(v, w, h, i, j) && map_valid(v, w, h, i + 1, j) && map_valid(v, w, h, i, j + 1)) {
        v3 c = map_xyz(v, w, i, j), r = map_xyz(v, w, i + 1, j), d = map_xyz(v, w, i, j + 1);
        v3 n = normalize3(cross3(sub3(r, c), sub3(d, c)));
        o.x = n.x; o.y = n.y; o.z = n.z; o.w = KF_VALID;
    }
    *out = o;
}
void KFO(create_normal_map)(const real4* v, int w, int h, real4* out)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < h; ++j)
        for (int i = 0; i < w; ++i) KFO(compute_normal)(v, w, h, i, j, &out[(int64_t)j * w + i]);
}

/* ------------------------------------------------------------------------- */
/* tsdf.cu:7-18 */
void KFO(initialize_grid)(real* F, real* W, int n, real f0, real w0)
{
    int64_t cnt = (int64_t)n * n * n;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < cnt; ++q) { F[q] = f0; W[q] = w0; }
}

/* tsdf.cu:20-75.  Per voxel (i,j,k):
 *   x = -width/2 + i*res, y likewise, z by repeated addition from -width/2 - res
 *   (tsdf.cu:27-31: point.z += res once per k, before use);
 *   pe = inv_pose * (x,y,z,1); pix = K * pe; pix /= pix.z;
 *   skip unless 0.4 <= pe.z < 6 and 0 <= pix < (w, h) (reference: 640x480);
 *   d = depth[int(pix.y)][int(pix.x)], skip unless valid;
 *   sdf = d - pe.z; if sdf >= -mue: f = min(1, sdf/mue), wt = 1/d,
 *   F = (F*W + f*wt) / (W + wt), W += wt.
 * Returns the number of updated ("touched") voxels.
 * Pinned by: the fronto-parallel wall closed form, repeat-invariance of F and
 * W = k/D after k identical frames (tests/test_oracle_tsdf.py). */
/* Precision form (snorm16 F, uint8 W): F is stored as q = rint(clamp(f, -1, 1) * 32767)
 * and read as q / 32767; the one code snorm16 leaves unused, -32768, holds the
 * unobserved value 2 of tsdf.cu:7-18 (any f > 1 is stored as it), so trilinear
 * samples next to unobserved space mix in 2 exactly as the FP32 volume does.
 * W counts observations: F = (F*W + f) / (W + 1), W = min(W + 1, 128).  Same
 * projection and tests as the FP32 form.
 * Count-weight form (format 2, both builds): the precision form's update rule on
 * unquantised F and W planes of `real` -- compiled as FP64 it is the error
 * reference of the precision form, which then differs from it only by the
 * snorm16 rounding of F and FP32 arithmetic. */
#define KF_SNORM R(32767.0)
#define KF_SNORM_UNOBSERVED (-32768)
static inline real snorm_decode(int16_t q)
{
    return q == KF_SNORM_UNOBSERVED ? R(2.0) : (real)q * (R(1.0) / KF_SNORM);
}
static inline int16_t snorm_encode(real f)
{
    if (f > R(1.0)) return (int16_t)KF_SNORM_UNOBSERVED;
    real s = f * KF_SNORM;
    s = FMIN(FMAX(s, -KF_SNORM), KF_SNORM);
#ifdef KFREF_F64
    return (int16_t)nearbyint(s);
#else
    return (int16_t)nearbyintf(s);
#endif
}

static int64_t fuse_impl(void* Fp, void* Wp, int fmt, int n, real width, real mue,
                         const real* depth, int w, int h, const real* inv_pose, const real* K)
{
    real res = width / (real)(n - 1);
    real half = -width * R(0.5);
    real* zk = (real*)malloc(sizeof(real) * (size_t)n);
    real z = half + (-res);
    for (int k = 0; k < n; ++k) { z = z + res; zk[k] = z; }
    const real fw = (real)w, fh = (real)h;
    int64_t touched = 0;
#pragma omp parallel for schedule(static) reduction(+ : touched) collapse(2)
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j) {
            real y = half + (real)j * res;
            for (int i = 0; i < n; ++i) {
                real x = half + (real)i * res;
                v3 pe = m4v(inv_pose, mk3(x, y, zk[k]), R(1.0));
                v3 pix = m3v(K, pe);
                pix = div3s(pix, pix.z);
                if (!(pe.z >= KF_MIN_DIST && pe.z < KF_MAX_DIST && pix.x >= R(0.0) && pix.x < fw &&
                      pix.y >= R(0.0) && pix.y < fh))
                    continue;
                real d = depth[(int64_t)(int)pix.y * w + (int)pix.x];
                if (!is_depth_valid(d)) continue;
                real diff = d - pe.z;
                if (diff >= -mue) {
                    real f = diff / mue;
                    f = (f < R(1.0)) ? f : R(1.0);   /* glm::min(1, f) = (f < 1) ? f : 1 */
                    int64_t q = ((int64_t)k * n + j) * n + i;
                    if (fmt == 2) {
                        real* F = (real*)Fp;
                        real* W = (real*)Wp;
                        real Wv = W[q];
                        F[q] = (F[q] * Wv + f) / (Wv + R(1.0));
                        W[q] = Wv < R(128.0) ? Wv + R(1.0) : R(128.0);
                    } else if (fmt == 0) {
                        real* F = (real*)Fp;
                        real* W = (real*)Wp;
                        real wt = R(1.0) / d;
                        real Fv = F[q], Wv = W[q];
                        Fv = (Fv * Wv + f * wt) / (Wv + wt);
                        Wv = Wv + wt;
                        F[q] = Fv;
                        W[q] = Wv;
                    } else {
                        int16_t* F = (int16_t*)Fp;
                        uint8_t* W = (uint8_t*)Wp;
                        real Wv = (real)W[q];
                        real Fv = snorm_decode(F[q]);
                        Fv = (Fv * Wv + f) / (Wv + R(1.0));
                        F[q] = snorm_encode(Fv);
                        W[q] = (uint8_t)(Wv < R(128.0) ? Wv + R(1.0) : R(128.0));
                    }
                    ++touched;
                }
            }
        }
    free(zk);
    return touched;
}
int64_t KFO(fuse)(real* F, real* W, int n, real width, real mue, const real* depth, int w, int h,
                  const real* inv_pose, const real* K)
{
    return fuse_impl(F, W, 0, n, width, mue, depth, w, h, inv_pose, K);
}
#ifndef KFREF_F64
int64_t kfo_fuse_q(int16_t* F, uint8_t* W, int n, float width, float mue, const float* depth, int w, int h,
                   const float* inv_pose, const float* K)
{
    return fuse_impl(F, W, 1, n, width, mue, depth, w, h, inv_pose, K);
}
void kfo_initialize_grid_q(int16_t* F, uint8_t* W, int n, float f0, float w0)
{
    int64_t cnt = (int64_t)n * n * n;
    int16_t fq = snorm_encode(f0);
    float wr = nearbyintf(fminf(fmaxf(w0, 0.0f), 255.0f));
    for (int64_t q = 0; q < cnt; ++q) { F[q] = fq; W[q] = (uint8_t)wr; }
}
#endif

/* ------------------------------------------------------------------------- */
/* Trilinear sample of the f plane at normalized coordinates (u,v,s), texel
 * centres at (i+1/2)/n, indices clamped to [0,n-1] — the CUDA texture
 * "linear filter, normalized coords, clamp" rule of voxel_grid.cpp:24-33 as the
 * CUDA C++ Programming Guide ("Texture Fetching") states it:
 *   x = u*n - 0.5, i = floor(x), a = x - i,
 *   T = (1-a)(1-b)(1-c) T[i,j,k] + a(1-b)(1-c) T[i+1,j,k] + ... .
 * The hardware holds a, b, c in 1.8 fixed point; this restatement uses exact
 * FP32 weights (parity with the texture unit itself is unpinned).  Evaluated
 * as nested (1-t)*p + t*q along x, then y, then z. */
static inline real lerp_(real p, real q, real t) { return (R(1.0) - t) * p + t * q; }
static inline int64_t vidx(int n, int i, int j, int k) { return ((int64_t)k * n + j) * n + i; }
static inline real voxel_value(const void* F, int fmt, int64_t q)
{
    return fmt != 1 ? ((const real*)F)[q] : snorm_decode(((const int16_t*)F)[q]);
}
/* Footprint recording (measurement only, kfo_raycast_footprint): when set, every
 * trilinear sample marks its 8 corner voxels in this n^3 byte map and counts
 * itself.  Concurrent marks all store 1, so the map needs no ordering. */
static uint8_t* g_footprint = NULL;
static int64_t g_footprint_samples = 0;
static real sample_any(const void* F, int fmt, int n, real u, real v, real s)
{
    real x = u * (real)n - R(0.5), y = v * (real)n - R(0.5), z = s * (real)n - R(0.5);
    real fx = FLOOR(x), fy = FLOOR(y), fz = FLOOR(z);
    real a = x - fx, b = y - fy, c = z - fz;
    /* floor of huge/NaN coordinates is clamped before the int conversion */
    real lo = R(-1.0), hi = (real)n;
    fx = fx < lo ? lo : (fx > hi ? hi : fx);
    fy = fy < lo ? lo : (fy > hi ? hi : fy);
    fz = fz < lo ? lo : (fz > hi ? hi : fz);
    int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
    int i1 = clampi(i0 + 1, 0, n - 1), j1 = clampi(j0 + 1, 0, n - 1), k1 = clampi(k0 + 1, 0, n - 1);
    i0 = clampi(i0, 0, n - 1); j0 = clampi(j0, 0, n - 1); k0 = clampi(k0, 0, n - 1);
    if (g_footprint) {
        g_footprint[vidx(n, i0, j0, k0)] = 1; g_footprint[vidx(n, i1, j0, k0)] = 1;
        g_footprint[vidx(n, i0, j1, k0)] = 1; g_footprint[vidx(n, i1, j1, k0)] = 1;
        g_footprint[vidx(n, i0, j0, k1)] = 1; g_footprint[vidx(n, i1, j0, k1)] = 1;
        g_footprint[vidx(n, i0, j1, k1)] = 1; g_footprint[vidx(n, i1, j1, k1)] = 1;
#pragma omp atomic
        g_footprint_samples += 1;
    }
    real c000 = voxel_value(F, fmt, vidx(n, i0, j0, k0)), c100 = voxel_value(F, fmt, vidx(n, i1, j0, k0));
    real c010 = voxel_value(F, fmt, vidx(n, i0, j1, k0)), c110 = voxel_value(F, fmt, vidx(n, i1, j1, k0));
    real c001 = voxel_value(F, fmt, vidx(n, i0, j0, k1)), c101 = voxel_value(F, fmt, vidx(n, i1, j0, k1));
    real c011 = voxel_value(F, fmt, vidx(n, i0, j1, k1)), c111 = voxel_value(F, fmt, vidx(n, i1, j1, k1));
    real e00 = lerp_(c000, c100, a), e10 = lerp_(c010, c110, a);
    real e01 = lerp_(c001, c101, a), e11 = lerp_(c011, c111, a);
    real g0 = lerp_(e00, e10, b), g1 = lerp_(e01, e11, b);
    return lerp_(g0, g1, c);
}
real KFO(sample_tsdf)(const real* F, int n, real u, real v, real s) { return sample_any(F, 0, n, u, v, s); }

/* raycast.cu:7-27 slab test with 1/dir. */
static inline void intersect_bbox(v3 o, v3 inv, real lo, real hi, real* tmin_out, real* tmax_out)
{
    real t0x = (lo - o.x) * inv.x, t1x = (hi - o.x) * inv.x;
    real t0y = (lo - o.y) * inv.y, t1y = (hi - o.y) * inv.y;
    real t0z = (lo - o.z) * inv.z, t1z = (hi - o.z) * inv.z;
    real tmin = FMIN(t0x, t1x), tmax = FMAX(t0x, t1x);
    tmin = FMAX(tmin, FMIN(t0y, t1y));
    tmax = FMIN(tmax, FMAX(t0y, t1y));
    tmin = FMAX(tmin, FMIN(t0z, t1z));
    tmax = FMIN(tmax, FMAX(t0z, t1z));
    if (tmax < tmin) { tmin = R(-1.0); tmax = R(-1.0); }
    *tmin_out = tmin;
    *tmax_out = tmax;
}

/* raycast.cu:29-42: central difference at +-mue/2 (normalized offset
 * mue*0.5/width), returns (F(-) - F(+)) / offset, i.e. -grad F scaled. */
static inline v3 compute_gradient(const void* F, int fmt, int n, real width, real mue, v3 p)
{
    real r = mue * R(0.5) / width;
    real u = p.x / width + R(0.5), v = p.y / width + R(0.5), s = p.z / width + R(0.5);
    real fx0 = sample_any(F, fmt, n, u + r, v, s);
    real fy0 = sample_any(F, fmt, n, u, v + r, s);
    real fz0 = sample_any(F, fmt, n, u, v, s + r);
    real fx = sample_any(F, fmt, n, u - r, v, s);
    real fy = sample_any(F, fmt, n, u, v - r, s);
    real fz = sample_any(F, fmt, n, u, v, s - r);
    return div3s(mk3(fx - fx0, fy - fy0, fz - fz0), r);
}

/* raycast.cu:44-145.  Per pixel of a level-L map:
 *   o = pose[3].xyz, dir = (mat3(pose) * K_inv) * (i+1/2, j+1/2, 1) (not
 *   normalized, so t is camera depth); slab test; return WITHOUT writing when
 *   tmax < 0 or near >= far (stale pixels, raycast.cu:61-73);
 *   march t from near by 1.99*mue, switching to 0.1*mue once F < 0.99; on the
 *   first negative sample after a large step back off mue and continue with
 *   small steps (Fp not updated); otherwise t* = t - s*Fp/(F-Fp);
 *   at t* != 0 write vertex o+dir*t* and -grad/|grad| when |grad| is a valid
 *   "depth" (> 0.01, finite), else (0,0,0,-2) to both maps.
 * Pinned by: analytic plane and sphere TSDFs (hit depth, +-axis normals) and
 * the stale-pixel rule (tests/test_oracle_raycast.py). */
static void raycast_impl(const void* F, int fmt, int n, real width, real mue, const real* pose, const real* Ki,
                         int w, int h, real4* vout, real4* nout)
{
    real Rm[9] = {pose[0], pose[1], pose[2], pose[4], pose[5], pose[6], pose[8], pose[9], pose[10]};
    real M[9];
    m3m3(Rm, Ki, M);
    v3 o = mk3(pose[12], pose[13], pose[14]);
    real hw = width * R(0.5);
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < h; ++j)
        for (int i = 0; i < w; ++i) {
            v3 dir = m3v(M, mk3((real)i + R(0.5), (real)j + R(0.5), R(1.0)));
            v3 inv = mk3(R(1.0) / dir.x, R(1.0) / dir.y, R(1.0) / dir.z);
            real tmin, tmax;
            intersect_bbox(o, inv, -hw, hw, &tmin, &tmax);
            if (tmax < R(0.0)) continue;
            real nearv = (KF_MIN_DIST < tmin) ? tmin : KF_MIN_DIST;  /* glm::max(0.4, tmin) */
            real farv = (tmax < KF_MAX_DIST) ? tmax : KF_MAX_DIST;   /* glm::min(6.0, tmax) */
            if (nearv >= farv) continue;
            const real large = mue * R(1.99), small = mue * R(0.1);
            real step = large, prev = R(0.0), tstar = R(0.0);
            int guard = 0; /* the library's march bound; never binds for mue >= 1e-4 m */
            for (real t = nearv; t < farv; t = t + step) {
                if (++guard > (1 << 20)) break;
                v3 p = add3(o, mul3s(dir, t));
                real f = sample_any(F, fmt, n, p.x / width + R(0.5), p.y / width + R(0.5), p.z / width + R(0.5));
                if (f < R(0.0)) {
                    if (t != nearv) {
                        if (step == large) {
                            t = t - mue;
                            step = small;
                            continue;
                        }
                        tstar = t - step * prev / (f - prev);
                    }
                    break;
                } else if (f < R(0.99)) {
                    step = small;
                }
                prev = f;
            }
            int64_t q = (int64_t)j * w + i;
            real4 inval = {0, 0, 0, KF_INVALID};
            if (tstar != R(0.0)) {
                v3 vtx = add3(o, mul3s(dir, tstar));
                v3 g = compute_gradient(F, fmt, n, width, mue, vtx);
                real gn = len3(g);
                if (is_depth_valid(gn)) {
                    real4 a = {vtx.x, vtx.y, vtx.z, KF_VALID};
                    v3 nn = div3s(g, gn);
                    real4 b = {nn.x, nn.y, nn.z, KF_VALID};
                    vout[q] = a;
                    nout[q] = b;
                    continue;
                }
            }
            vout[q] = inval;
            nout[q] = inval;
        }
}
void KFO(raycast)(const real* F, int n, real width, real mue, const real* pose, const real* Ki, int w, int h,
                  real4* vout, real4* nout)
{
    raycast_impl(F, 0, n, width, mue, pose, Ki, w, h, vout, nout);
}
#ifndef KFREF_F64
void kfo_raycast_q(const int16_t* F, int n, float width, float mue, const float* pose, const float* Ki, int w, int h,
                   kfo_float4* vout, kfo_float4* nout)
{
    raycast_impl(F, 1, n, width, mue, pose, Ki, w, h, vout, nout);
}
/* The raycast's footprint (measurement only): runs the march of raycast_impl
 * (fmt 0: float F, 1: snorm16 codes) with every sample -- march samples and
 * the 6 gradient samples of each hit -- marking its trilinear corners in
 * mark (n^3 bytes, zeroed by the caller, accumulated over calls).  Returns the
 * number of samples taken.  Not reentrant. */
int64_t kfo_raycast_footprint(const void* F, int fmt, int n, float width, float mue, const float* pose, const float* Ki,
                              int w, int h, uint8_t* mark)
{
    kfo_float4* maps = (kfo_float4*)malloc(sizeof(kfo_float4) * 2 * (size_t)w * h);
    if (!maps) return -1;
    g_footprint = mark;
    g_footprint_samples = 0;
    raycast_impl(F, fmt, n, width, mue, pose, Ki, w, h, maps, maps + (size_t)w * h);
    g_footprint = NULL;
    free(maps);
    return g_footprint_samples;
}
/* Number of non-zero bytes of a footprint map. */
int64_t kfo_count_marked(const uint8_t* mark, int64_t count)
{
    int64_t c = 0;
#pragma omp parallel for schedule(static) reduction(+ : c)
    for (int64_t q = 0; q < count; ++q) c += mark[q] != 0;
    return c;
}
#endif

/* ------------------------------------------------------------------------- */
/* icp.cuh:40-49: pixel of s in the previous frame, (K * R_prev^T) * (s - t_prev),
 * floor of the perspective division.  fuv returns the floored floats (the range
 * test is done on them, so huge or NaN values never reach an int conversion).
 * Pinned by: the reference's TestComputeCorrespondence (icp_tests.cpp:55-80). */
static inline void correspondence(v3 s, const real* Rp, const real* tp, const real* K, real* fu, real* fv)
{
    real Rt[9], M[9];
    m3t(Rp, Rt);
    m3m3(K, Rt, M);
    v3 p = m3v(M, sub3(s, mk3(tp[0], tp[1], tp[2])));
    *fu = FLOOR(p.x / p.z);
    *fv = FLOOR(p.y / p.z);
}
#ifndef KFREF_F64
void kfo_correspondence(const float s[3], const float Rp[9], const float tp[3], const float K[9], int uv[2], float fuv[2])
{
    float fu, fv;
    correspondence(mk3(s[0], s[1], s[2]), Rp, tp, K, &fu, &fv);
    fuv[0] = fu;
    fuv[1] = fv;
    uv[0] = (fu > -2e9f && fu < 2e9f) ? (int)fu : 0;
    uv[1] = (fv > -2e9f && fv < 2e9f) ? (int)fv : 0;
}
#endif

/* Cosine of the ICP normal-angle threshold, evaluated once per call on the
 * host in FP64 and rounded to float (the HIP library's host code does the same). */
float KFO(cos_threshold)(float angle) { return (float)cos((double)angle); }

/* icp.cuh:63-70 rejects when acos(dot(normalize(R*n), tn)) > angle.  Restated
 * without acos: since acos is decreasing on [-1,1] and NaN compares false, the
 * test is "d >= -1 && d < cos(angle)" (d > 1 gives NaN in acos and is kept, as
 * here).  cos(angle) is evaluated once on the host in double and rounded.
 * Pinned by: the reference's TestNormalsAreTooDifferent (icp_tests.cpp:82-98)
 * and agreement with the acos form on random inputs away from the boundary. */
static inline int too_different(v3 n, v3 tn, const real* Rm, real cos_thresh)
{
    v3 nn = normalize3(m3v(Rm, n));
    real d = dot3(nn, tn);
    return d >= R(-1.0) && d < cos_thresh;
}
#ifndef KFREF_F64
int kfo_normals_too_different(const float n[3], const float tn[3], const float Rm[9], float cos_thresh)
{
    return too_different(mk3(n[0], n[1], n[2]), mk3(tn[0], tn[1], tn[2]), Rm, cos_thresh);
}
int kfo_normals_too_different_acos(const float n[3], const float tn[3], const float Rm[9], float angle)
{
    v3 nn = normalize3(m3v(Rm, mk3(n[0], n[1], n[2])));
    return acosf(dot3(nn, mk3(tn[0], tn[1], tn[2]))) > angle;
}
#endif

/* icp.cu:7-82 and icp.cuh:51-95: one row A_i = (n x s, n), b_i = n.d - n.s per
 * pixel, zero row when rejected.  Pinned by: TestComputeAndFillA/B and
 * TestConstructIcpResiduals (icp_tests.cpp:100-127,159-231). */
static inline int icp_row(const real4* V, const real4* TV, const real4* TN, int w, int h, int u, int v,
                          const real* Rp, const real* tp, const real* Rc, const real* tc, const real* K,
                          real dist_thresh, real cos_thresh, real* a, real* b)
{
    if (!map_valid(V, w, h, u, v)) return 0;
    v3 vc = map_xyz(V, w, u, v);
    v3 s = add3(m3v(Rc, vc), mk3(tc[0], tc[1], tc[2]));
    real fu, fv;
    correspondence(s, Rp, tp, K, &fu, &fv);
    if (!(fu >= R(0.0) && fv >= R(0.0) && fu < (real)w && fv < (real)h)) return 0;
    int uc = (int)fu, vcr = (int)fv;
    if (!map_valid(TV, w, h, uc, vcr) || !map_valid(TN, w, h, uc, vcr)) return 0;
    v3 d = map_xyz(TV, w, uc, vcr);
    if (len3(sub3(d, s)) > dist_thresh) return 0;
    v3 n = map_xyz(TN, w, uc, vcr);
    real4 nl;
    KFO(compute_normal)(V, w, h, u, v, &nl);
    if (nl.w == KF_INVALID) return 0;
    if (too_different(mk3(nl.x, nl.y, nl.z), n, Rc, cos_thresh)) return 0;
    a[0] = n.y * s.z - n.z * s.y;
    a[1] = n.z * s.x - n.x * s.z;
    a[2] = n.x * s.y - n.y * s.x;
    a[3] = n.x;
    a[4] = n.y;
    a[5] = n.z;
    *b = ((((n.x * d.x + n.y * d.y) + n.z * d.z) - n.x * s.x) - n.y * s.y) - n.z * s.z;
    return 1;
}

void KFO(construct_icp_residuals)(const real4* V, const real4* TV, const real4* TN, int w, int h,
                                  const real* Rp, const real* tp, const real* Rc, const real* tc,
                                  const real* K, real dist_thresh, real cos_thresh, real* A, real* b)
{
#pragma omp parallel for schedule(static)
    for (int v = 0; v < h; ++v)
        for (int u = 0; u < w; ++u) {
            int64_t q = (int64_t)v * w + u;
            real* a = &A[q * 6];
            if (!icp_row(V, TV, TN, w, h, u, v, Rp, tp, Rc, tc, K, dist_thresh, cos_thresh, a, &b[q])) {
                a[0] = a[1] = a[2] = a[3] = a[4] = a[5] = 0;
                b[q] = 0;
            }
        }
}

/* linear_least_squares.cpp:23-38: A^T A (cublasSgemm) and A^T b (cublasSgemv).
 * cuBLAS's FP32 summation order is unspecified; the oracle sums every product in
 * FP64 over fixed chunks combined in order (deterministic for any thread count).
 * ata is column-major 6x6 (symmetric).  Returns the number of non-zero rows. */
int64_t KFO(normal_equations)(const real* A, const real* b, int64_t n, double* ata, double* atb)
{
    enum { CH = 64 };
    double part[CH][27];
    int64_t nz[CH];
    int64_t per = (n + CH - 1) / CH;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < CH; ++c) {
        double s[27] = {0};
        int64_t cnt = 0;
        int64_t lo = c * per, hi = lo + per < n ? lo + per : n;
        for (int64_t r = lo; r < hi; ++r) {
            const real* a = &A[r * 6];
            int any = 0;
            for (int q = 0; q < 6; ++q) any |= (a[q] != 0);
            cnt += any || b[r] != 0;
            int t = 0;
            for (int p = 0; p < 6; ++p)
                for (int q = p; q < 6; ++q) s[t++] += (double)a[p] * (double)a[q];
            for (int p = 0; p < 6; ++p) s[21 + p] += (double)a[p] * (double)b[r];
        }
        memcpy(part[c], s, sizeof(s));
        nz[c] = cnt;
    }
    double s[27] = {0};
    int64_t cnt = 0;
    for (int c = 0; c < CH; ++c) {
        for (int q = 0; q < 27; ++q) s[q] += part[c][q];
        cnt += nz[c];
    }
    int t = 0;
    for (int p = 0; p < 6; ++p)
        for (int q = p; q < 6; ++q) { ata[q * 6 + p] = s[t]; ata[p * 6 + q] = s[t]; ++t; }
    for (int p = 0; p < 6; ++p) atb[p] = s[21 + p];
    return cnt;
}

#ifndef KFREF_F64
/* FP64 lower Cholesky and the two triangular solves, in the unblocked order of
 * Eigen's LLT (3rd/eigen/Eigen/src/Cholesky/LLT.h:306-326 for size < 32):
 * for each column k: x = A_kk - |L_k,0:k|^2; fail if x <= 0; L_kk = sqrt(x);
 * L_k+1:,k = (A_k+1:,k - L_k+1:,0:k L_k,0:k^T) / L_kk.  Returns 0, or k+1 at the
 * failing column (Eigen reports NumericalIssue and the reference ignores it).
 * Pinned by: the reference's TestSolveLinearSystem (icp_tests.cpp:129-157),
 * Eigen 3.3.7's own LLT built from the reference tree (oracle/_ref) and a
 * residual check on random SPD systems. */
static int chol6_d(const double* ata, const double* atb, double* x)
{
    double L[6][6];
    for (int c = 0; c < 6; ++c)
        for (int r = 0; r < 6; ++r) L[r][c] = ata[c * 6 + r];
    for (int k = 0; k < 6; ++k) {
        double d = L[k][k];
        for (int j = 0; j < k; ++j) d -= L[k][j] * L[k][j];
        if (!(d > 0.0)) return k + 1;
        d = sqrt(d);
        L[k][k] = d;
        for (int r = k + 1; r < 6; ++r) {
            double v = L[r][k];
            for (int j = 0; j < k; ++j) v -= L[r][j] * L[k][j];
            L[r][k] = v / d;
        }
    }
    double y[6];
    for (int r = 0; r < 6; ++r) {
        double v = atb[r];
        for (int j = 0; j < r; ++j) v -= L[r][j] * y[j];
        y[r] = v / L[r][r];
    }
    for (int r = 5; r >= 0; --r) {
        double v = y[r];
        for (int j = r + 1; j < 6; ++j) v -= L[j][r] * x[j];
        x[r] = v / L[r][r];
    }
    return 0;
}
int kfo_cholesky_solve6(const double ata[36], const double atb[6], double x[6]) { return chol6_d(ata, atb, x); }

/* The same unblocked LLT in FP32: the reference's Eigen::LLT<Matrix<float,6,6>>
 * (linear_least_squares.cpp:41), for the cross-check against Eigen itself. */
int kfo_llt_solve6_f32(const float ata[36], const float atb[6], float x[6])
{
    float L[6][6];
    for (int c = 0; c < 6; ++c)
        for (int r = 0; r < 6; ++r) L[r][c] = ata[c * 6 + r];
    for (int k = 0; k < 6; ++k) {
        float d = L[k][k];
        for (int j = 0; j < k; ++j) d -= L[k][j] * L[k][j];
        if (!(d > 0.0f)) return k + 1;
        d = sqrtf(d);
        L[k][k] = d;
        for (int r = k + 1; r < 6; ++r) {
            float v = L[r][k];
            for (int j = 0; j < k; ++j) v -= L[r][j] * L[k][j];
            L[r][k] = v / d;
        }
    }
    float y[6];
    for (int r = 0; r < 6; ++r) {
        float v = atb[r];
        for (int j = 0; j < r; ++j) v -= L[r][j] * y[j];
        y[r] = v / L[r][r];
    }
    for (int r = 5; r >= 0; --r) {
        float v = y[r];
        for (int j = r + 1; j < 6; ++j) v -= L[j][r] * x[j];
        x[r] = v / L[r][r];
    }
    return 0;
}
#else
int kfo_cholesky_solve6(const double ata[36], const double atb[6], double x[6]);
#endif

/* icp.cpp:65-87.  R_inc = orientate3(vec3(x0, x2, x1)) = yawPitchRoll(x1, x0, x2)
 * = R_y(x1) R_x(x0) R_z(x2) (glm gtx/euler_angles, restated from its definition);
 * R <- R_inc^T R, t <- R_inc^T t + (x3, x4, x5).  Evaluated in FP64 from the
 * FP32 pose and rounded back (the device does the same).  Pinned by: closed-form
 * single-axis rotations and first-order consistency R_inc^T ~ I - [x]_x. */
void KFO(update_pose)(real* Rm, real* t, const real* x)
{
    double pitch = x[0], yaw = x[1], roll = x[2];
    double ch = cos(yaw), sh = sin(yaw), cp = cos(pitch), sp = sin(pitch), cb = cos(roll), sb = sin(roll);
    double Ri[9]; /* column-major, glm yawPitchRoll */
    Ri[0] = ch * cb + sh * sp * sb;  Ri[1] = sb * cp;  Ri[2] = -sh * cb + ch * sp * sb;
    Ri[3] = -ch * sb + sh * sp * cb; Ri[4] = cb * cp;  Ri[5] = sb * sh + ch * sp * cb;
    Ri[6] = sh * cp;                 Ri[7] = -sp;      Ri[8] = ch * cp;
    /* RiT * R: (RiT)[c][r] = Ri[r][c] */
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
    for (int q = 0; q < 9; ++q) Rm[q] = (real)Rn[q];
    for (int q = 0; q < 3; ++q) t[q] = (real)tn[q];
}

/* general_helper.cpp:11-21: (|R1 v - R2 v|, |t1 - t2|), v = normalize(1,1,1). */
void KFO(pose_error)(const real* R1, const real* t1, const real* R2, const real* t2, real* out)
{
    v3 v = normalize3(mk3(1, 1, 1));
    v3 a = m3v(R1, v), b = m3v(R2, v);
    out[0] = len3(sub3(b, a));
    out[1] = len3(sub3(mk3(t2[0], t2[1], t2[2]), mk3(t1[0], t1[1], t1[2])));
}

/* Inverse of the rigid pose [R|t] as [R^T | -R^T t] (sensor.h:30-34 uses the
 * general glm::inverse; for a rigid transform both give the same matrix up to
 * rounding).  Evaluated in FP64 and rounded. */
void KFO(rigid_inverse)(const real* Rm, const real* t, real* inv)
{
    memset(inv, 0, sizeof(real) * 16);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) inv[c * 4 + r] = Rm[r * 3 + c];
    for (int r = 0; r < 3; ++r) {
        double s = 0;
        for (int q = 0; q < 3; ++q) s += (double)Rm[r * 3 + q] * (double)t[q];
        inv[12 + r] = (real)(-s);
    }
    inv[15] = R(1.0);
}

/* icp.cpp:25-58: coarse-to-fine (levels-1 .. 0), iters[l] iterations at level l;
 * residual rows -> normal equations -> FP64 Cholesky -> pose update -> early
 * stop when pose_error < (stop_angle, stop_dist).  Returns 0, or a negative
 * code: -1 non-finite update (icp.cpp:67-73 throws), -2 not positive definite.
 * sums_log (optional) receives the 27 sums of every executed iteration. */
int KFO(compute_pose)(const real4* const vpyr[], const real4* const tvpyr[], const real4* const tnpyr[],
                      int w0, int h0, int levels, const int iters[], const real (*Ks)[9],
                      real dist_thresh, real angle_thresh, real stop_angle, real stop_dist,
                      const real* R_in, const real* t_in, real* R_out, real* t_out,
                      double* sums_log, int* n_iters)
{
    real Rp[9], tp[3], Rc[9], tc[3];
    memcpy(Rp, R_in, sizeof(Rp)); memcpy(tp, t_in, sizeof(tp));
    memcpy(Rc, R_in, sizeof(Rc)); memcpy(tc, t_in, sizeof(tc));
    real cos_thresh = (real)KFO(cos_threshold)((float)angle_thresh);
    int done = 0, status = 0;
    real* A = (real*)malloc(sizeof(real) * 6 * (size_t)w0 * h0);
    real* b = (real*)malloc(sizeof(real) * (size_t)w0 * h0);
    for (int l = levels - 1; l >= 0 && status == 0; --l) {
        int w = w0 >> l, h = h0 >> l;
        for (int it = 0; it < iters[l]; ++it) {
            KFO(construct_icp_residuals)(vpyr[l], tvpyr[l], tnpyr[l], w, h, Rp, tp, Rc, tc, Ks[l],
                                         dist_thresh, cos_thresh, A, b);
            double ata[36], atb[6], xd[6];
            KFO(normal_equations)(A, b, (int64_t)w * h, ata, atb);
            if (sums_log) {
                int q = 0;
                for (int p = 0; p < 6; ++p)
                    for (int r = p; r < 6; ++r) sums_log[done * 27 + q++] = ata[r * 6 + p];
                for (int p = 0; p < 6; ++p) sums_log[done * 27 + 21 + p] = atb[p];
            }
            ++done;
            if (kfo_cholesky_solve6(ata, atb, xd) != 0) { status = -2; break; }
            real x[6];
            int finite = 1;
            for (int q = 0; q < 6; ++q) { x[q] = (real)xd[q]; finite &= isfinite(x[q]) ? 1 : 0; }
            if (!finite) { status = -1; break; }
            real Rprev[9], tprev[3], err[2];
            memcpy(Rprev, Rc, sizeof(Rprev)); memcpy(tprev, tc, sizeof(tprev));
            KFO(update_pose)(Rc, tc, x);
            KFO(pose_error)(Rprev, tprev, Rc, tc, err);
            if (err[0] < stop_angle && err[1] < stop_dist) break;
        }
    }
    free(A);
    free(b);
    memcpy(R_out, Rc, sizeof(Rp));
    memcpy(t_out, tc, sizeof(tp));
    if (n_iters) *n_iters = done;
    return status;
}

/* voxel_grid.cpp:43-80 raw interchange: n^3 x (f, w) interleaved, x fastest. */
void KFO(interleave)(const real* F, const real* W, int64_t count, real* fw)
{
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < count; ++q) { fw[2 * q] = F[q]; fw[2 * q + 1] = W[q]; }
}

/* ------------------------------------------------------------------------- */
/* kinect_fusion.cpp:44-183 with one volume used for tracking and fusion.
 * First frame: convert + fuse at the initial pose (initializeTracking, :75-94).
 * Later frames: convert -> bilateral -> vertex L0 -> (downsample -> vertex) L1,
 * L2 -> raycast L0..2 at the previous pose -> ICP -> set pose -> fuse the
 * unfiltered metric depth at the new pose.  Raycast outputs start invalid and
 * keep stale pixels across frames as in the reference. */
struct KFO(pipeline) {
    int w, h, n, levels;
    real width, mue, scale, dist_thresh, angle_thresh, stop_angle, stop_dist;
    int iters[3];
    real K[3][9], Ki[3][9];
    real pose[16], inv_pose[16];
    int fmt;          /* 0: 1/d weights, 1: snorm16 F + uint8 W (FP32 build only), 2: count weights */
    int16_t* Fq;
    uint8_t* Wq;
    real *F, *W, *raw_m, *dpyr[3];
    real4 *vpyr[3], *pv[3], *pn[3];
    int frame;
    double stage_ms[8];
    int64_t touched;
};
typedef struct KFO(pipeline) pipe_t;

static void pose_from(real* pose, const real* Rm, const real* t)
{
    memset(pose, 0, sizeof(real) * 16);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) pose[c * 4 + r] = Rm[c * 3 + r];
    for (int r = 0; r < 3; ++r) pose[12 + r] = t[r];
    pose[15] = R(1.0);
}
static void set_pose(pipe_t* p, const real* Rm, const real* t)
{
    pose_from(p->pose, Rm, t);
    KFO(rigid_inverse)(Rm, t, p->inv_pose);
}

pipe_t* KFO(pipeline_create)(int w, int h, int n, real width, real mue, real fx, real fy, real cx, real cy,
                             real scale, const int iters[3], real dist_thresh, real angle_thresh,
                             real stop_angle, real stop_dist, const real* init_pose)
{
    pipe_t* p = (pipe_t*)calloc(1, sizeof(pipe_t));
    p->w = w; p->h = h; p->n = n; p->levels = 3;
    p->width = width; p->mue = mue; p->scale = scale;
    p->dist_thresh = dist_thresh; p->angle_thresh = angle_thresh;
    p->stop_angle = stop_angle; p->stop_dist = stop_dist;
    for (int l = 0; l < 3; ++l) {
        p->iters[l] = iters[l];
        KFO(intrinsics)(fx, fy, cx, cy, l, p->K[l], p->Ki[l]);
    }
    int64_t cnt = (int64_t)n * n * n;
    p->F = (real*)malloc(sizeof(real) * cnt);
    p->W = (real*)malloc(sizeof(real) * cnt);
    KFO(initialize_grid)(p->F, p->W, n, R(2.0), R(0.0));
    p->raw_m = (real*)malloc(sizeof(real) * (size_t)w * h);
    real4 inval = {0, 0, 0, KF_INVALID};
    for (int l = 0; l < 3; ++l) {
        size_t px = (size_t)(w >> l) * (h >> l);
        p->dpyr[l] = (real*)malloc(sizeof(real) * px);
        p->vpyr[l] = (real4*)malloc(sizeof(real4) * px);
        p->pv[l] = (real4*)malloc(sizeof(real4) * px);
        p->pn[l] = (real4*)malloc(sizeof(real4) * px);
        for (size_t q = 0; q < px; ++q) { p->pv[l][q] = inval; p->pn[l][q] = inval; }
    }
    real Rm[9] = {init_pose[0], init_pose[1], init_pose[2], init_pose[4], init_pose[5], init_pose[6],
                  init_pose[8], init_pose[9], init_pose[10]};
    real t[3] = {init_pose[12], init_pose[13], init_pose[14]};
    set_pose(p, Rm, t);
    return p;
}

void KFO(pipeline_destroy)(pipe_t* p)
{
    if (!p) return;
    free(p->F); free(p->W); free(p->Fq); free(p->Wq); free(p->raw_m);
    for (int l = 0; l < 3; ++l) { free(p->dpyr[l]); free(p->vpyr[l]); free(p->pv[l]); free(p->pn[l]); }
    free(p);
}

/* stage indices for pipeline_stage_ms: 0 preprocess, 1 raycast, 2 icp, 3 fuse */
int KFO(pipeline_frame)(pipe_t* p, const uint16_t* raw)
{
    int w = p->w, h = p->h, status = 0;
    double t0 = now_ms();
    KFO(convert_to_depth_meters)(raw, p->raw_m, w, h, p->scale, 0);
    if (p->frame > 0) {
        KFO(apply_bilateral_filter)(p->raw_m, p->dpyr[0], w, h);
        KFO(create_vertex_map)(p->dpyr[0], w, h, p->Ki[0], p->vpyr[0]);
        for (int l = 1; l < p->levels; ++l) {
            KFO(down_sample)(p->dpyr[l - 1], w >> (l - 1), h >> (l - 1), p->dpyr[l]);
            KFO(create_vertex_map)(p->dpyr[l], w >> l, h >> l, p->Ki[l], p->vpyr[l]);
        }
        double t1 = now_ms();
        p->stage_ms[0] += t1 - t0;
        for (int l = 0; l < p->levels; ++l)
            raycast_impl(p->fmt == 1 ? (const void*)p->Fq : (const void*)p->F, p->fmt, p->n, p->width, p->mue, p->pose,
                         p->Ki[l], w >> l, h >> l, p->pv[l], p->pn[l]);
        double t2 = now_ms();
        p->stage_ms[1] += t2 - t1;
        real Rin[9] = {p->pose[0], p->pose[1], p->pose[2], p->pose[4], p->pose[5], p->pose[6],
                       p->pose[8], p->pose[9], p->pose[10]};
        real tin[3] = {p->pose[12], p->pose[13], p->pose[14]};
        real Ro[9], to[3];
        const real4* vp[3] = {p->vpyr[0], p->vpyr[1], p->vpyr[2]};
        const real4* tv[3] = {p->pv[0], p->pv[1], p->pv[2]};
        const real4* tn[3] = {p->pn[0], p->pn[1], p->pn[2]};
        status = KFO(compute_pose)(vp, tv, tn, w, h, p->levels, p->iters, (const real(*)[9])p->K,
                                   p->dist_thresh, p->angle_thresh, p->stop_angle, p->stop_dist,
                                   Rin, tin, Ro, to, NULL, NULL);
        if (status == 0) set_pose(p, Ro, to);
        double t3 = now_ms();
        p->stage_ms[2] += t3 - t2;
        t0 = t3;
    } else {
        p->stage_ms[0] += now_ms() - t0;
        t0 = now_ms();
    }
    if (status == 0)
        p->touched = p->fmt == 1 ? fuse_impl(p->Fq, p->Wq, 1, p->n, p->width, p->mue, p->raw_m, w, h, p->inv_pose, p->K[0])
                                 : fuse_impl(p->F, p->W, p->fmt, p->n, p->width, p->mue, p->raw_m, w, h, p->inv_pose,
                                             p->K[0]);
    p->stage_ms[3] += now_ms() - t0;
    p->frame++;
    return status;
}

void KFO(pipeline_pose)(const pipe_t* p, real* pose) { memcpy(pose, p->pose, sizeof(real) * 16); }
const real* KFO(pipeline_tsdf)(const pipe_t* p) { return p->F; }
const real* KFO(pipeline_weight)(const pipe_t* p) { return p->W; }
/* which: 0 vertex pyramid (camera space), 1 predicted vertices, 2 predicted normals */
const real4* KFO(pipeline_map)(const pipe_t* p, int which, int level)
{
    return which == 0 ? p->vpyr[level] : (which == 1 ? p->pv[level] : p->pn[level]);
}
/* which: 0 unfiltered metric depth (level ignored), 1 filtered depth pyramid */
const real* KFO(pipeline_depth)(const pipe_t* p, int which, int level)
{
    return which == 0 ? p->raw_m : p->dpyr[level];
}
double KFO(pipeline_stage_ms)(const pipe_t* p, int stage) { return p->stage_ms[stage]; }
int64_t KFO(pipeline_touched)(const pipe_t* p) { return p->touched; }
/* Switches a fresh pipeline to the count-weight update (format 2) before its first frame. */
int KFO(pipeline_use_count_weights)(pipe_t* p)
{
    if (p->frame != 0 || p->fmt != 0) return -1;
    p->fmt = 2;
    return 0;
}
/* Fuse with the count-weight update into F and W planes of `real`. */
int64_t KFO(fuse_counts)(real* F, real* W, int n, real width, real mue, const real* depth, int w, int h,
                         const real* inv_pose, const real* K)
{
    return fuse_impl(F, W, 2, n, width, mue, depth, w, h, inv_pose, K);
}
#ifndef KFREF_F64
/* Switches a fresh pipeline to the precision form's volume (before the first frame). */
int kfo_pipeline_use_snorm16(pipe_t* p)
{
    if (p->frame != 0) return -1;
    int64_t cnt = (int64_t)p->n * p->n * p->n;
    p->Fq = (int16_t*)malloc(sizeof(int16_t) * cnt);
    p->Wq = (uint8_t*)malloc(sizeof(uint8_t) * cnt);
    if (!p->Fq || !p->Wq) return -1;
    kfo_initialize_grid_q(p->Fq, p->Wq, p->n, 2.0f, 0.0f);
    free(p->F); free(p->W);
    p->F = NULL; p->W = NULL;
    p->fmt = 1;
    return 0;
}
const int16_t* kfo_pipeline_tsdf_q(const pipe_t* p) { return p->Fq; }
const uint8_t* kfo_pipeline_weight_q(const pipe_t* p) { return p->Wq; }
#endif
