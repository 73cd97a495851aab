/*
 * kfsynth.c — seeded synthetic depth sequences for the per-frame path.
 *
 * Scene (world frame = volume frame, cube of side L centred at the origin,
 * camera x right / y down / z forward):
 *   floor plane y = +0.45L, walls x = -0.45L, x = +0.45L and z = +0.45L
 *   (infinite planes), and 8 spheres with r ~ U[0.1, 0.4] m, centres
 *   x ~ U[-0.3L, 0.3L], y ~ U[-0.1L, 0.45L - r], z ~ U[0.1L, 0.4L], rejected
 *   while any sphere surface comes closer than 0.45 m to the camera path.
 * Trajectory: each of the camera centre's 3 coordinates and 3 Euler angles is a
 *   sum of 3 sinusoids with seeded amplitudes, frequencies and phases, scaled so
 *   that |c| <= 0.25 m, |angle| <= 10 deg, and per frame |dc| <= 2 cm and the
 *   relative rotation <= 1 deg.  Camera-to-world R = Rz(c) Ry(b) Rx(a).
 * Depth: FP64 ray cast through the pixel centre (i+1/2, j+1/2) with intrinsics
 *   fx, fy, cx, cy, z = camera depth of the first hit, noise
 *   z' = z + N(0, (k z^2)^2) (k = 0.0012), u16 = clamp(round(1000 z'), 0, 65535);
 *   an optional dropout fraction zeroes pixels (robustness tests only).
 * Randomness: SplitMix64 of a counter keyed by (seed, stream, index); normals
 *   by Box-Muller.  Everything is deterministic for a given seed and size.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline uint64_t key3(uint64_t seed, uint64_t stream, uint64_t index)
{
    return mix64(mix64(mix64(seed) ^ (stream * 0xD1B54A32D192ED03ull)) ^ index);
}
/* uniform in (0, 1] */
static inline double u01(uint64_t seed, uint64_t stream, uint64_t index)
{
    return ((double)(key3(seed, stream, index) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}
static inline double uab(uint64_t seed, uint64_t stream, uint64_t index, double a, double b)
{
    return a + (b - a) * u01(seed, stream, index);
}

enum { S_TRAJ = 1, S_SPHERE = 2, S_NOISE = 3, S_DROP = 4 };

static void euler_to_R(double a, double b, double c, double* Rm /* col-major 3x3 */)
{
    double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cc = cos(c), sc = sin(c);
    /* Rz(c) Ry(b) Rx(a), row-major first */
    double M[3][3] = {
        {cc * cb, cc * sb * sa - sc * ca, cc * sb * ca + sc * sa},
        {sc * cb, sc * sb * sa + cc * ca, sc * sb * ca - cc * sa},
        {-sb, cb * sa, cb * ca},
    };
    for (int c0 = 0; c0 < 3; ++c0)
        for (int r = 0; r < 3; ++r) Rm[c0 * 3 + r] = M[r][c0];
}

static double rel_angle(const double* R1, const double* R2)
{
    /* angle of R1^T R2: acos((tr - 1)/2) */
    double tr = 0;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) tr += R1[c * 3 + r] * R2[c * 3 + r];
    double x = (tr - 1.0) * 0.5;
    x = x > 1 ? 1 : (x < -1 ? -1 : x);
    return acos(x);
}

/* poses: n_frames x 16 doubles, column-major camera-to-world. */
int kfs_trajectory(uint64_t seed, int n_frames, double* poses)
{
    if (n_frames <= 0) return -1;
    double amp[6][3], om[6][3], ph[6][3];
    for (int d = 0; d < 6; ++d)
        for (int q = 0; q < 3; ++q) {
            uint64_t idx = (uint64_t)(d * 3 + q) * 3;
            amp[d][q] = uab(seed, S_TRAJ, idx + 0, 0.3, 1.0);
            om[d][q] = uab(seed, S_TRAJ, idx + 1, 0.01, 0.06);
            ph[d][q] = uab(seed, S_TRAJ, idx + 2, 0.0, 2.0 * M_PI);
        }
    double* raw = (double*)malloc(sizeof(double) * 6 * (size_t)n_frames);
    double peak[6] = {0};
    for (int f = 0; f < n_frames; ++f)
        for (int d = 0; d < 6; ++d) {
            double v = 0;
            for (int q = 0; q < 3; ++q) v += amp[d][q] * sin(om[d][q] * f + ph[d][q]);
            raw[f * 6 + d] = v;
            if (fabs(v) > peak[d]) peak[d] = fabs(v);
        }
    /* scale to the position and angle envelopes */
    double scale[6];
    for (int d = 0; d < 3; ++d) scale[d] = peak[d] > 0 ? (0.25 / sqrt(3.0)) / peak[d] : 0;
    for (int d = 3; d < 6; ++d) scale[d] = peak[d] > 0 ? (10.0 * M_PI / 180.0) / peak[d] : 0;
    /* then shrink translation / rotation until the per-frame limits hold */
    for (int pass = 0; pass < 60; ++pass) {
        double max_dc = 0, max_da = 0;
        double Rprev[9], Rcur[9];
        for (int f = 0; f < n_frames; ++f) {
            double* v = &raw[f * 6];
            euler_to_R(v[3] * scale[3], v[4] * scale[4], v[5] * scale[5], Rcur);
            if (f > 0) {
                double* u = &raw[(f - 1) * 6];
                double dx = (v[0] - u[0]) * scale[0], dy = (v[1] - u[1]) * scale[1], dz = (v[2] - u[2]) * scale[2];
                double dc = sqrt(dx * dx + dy * dy + dz * dz);
                if (dc > max_dc) max_dc = dc;
                double da = rel_angle(Rprev, Rcur);
                if (da > max_da) max_da = da;
            }
            memcpy(Rprev, Rcur, sizeof(Rcur));
        }
        int ok = 1;
        if (max_dc > 0.02) { for (int d = 0; d < 3; ++d) scale[d] *= 0.98 * 0.02 / max_dc; ok = 0; }
        if (max_da > M_PI / 180.0) { for (int d = 3; d < 6; ++d) scale[d] *= 0.98 * (M_PI / 180.0) / max_da; ok = 0; }
        if (ok) break;
    }
    for (int f = 0; f < n_frames; ++f) {
        double* v = &raw[f * 6];
        double* P = &poses[(size_t)f * 16];
        double Rm[9];
        euler_to_R(v[3] * scale[3], v[4] * scale[4], v[5] * scale[5], Rm);
        memset(P, 0, sizeof(double) * 16);
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) P[c * 4 + r] = Rm[c * 3 + r];
        P[12] = v[0] * scale[0];
        P[13] = v[1] * scale[1];
        P[14] = v[2] * scale[2];
        P[15] = 1.0;
    }
    free(raw);
    return 0;
}

/* spheres: 8 x (cx, cy, cz, r).  Returns the number of rejected draws, or -1. */
int kfs_spheres(uint64_t seed, double L, const double* poses, int n_frames, int n_spheres, double* spheres)
{
    int rejected = 0;
    uint64_t draw = 0;
    for (int s = 0; s < n_spheres; ++s) {
        for (int tries = 0;; ++tries) {
            if (tries > 100000) return -1;
            double r = uab(seed, S_SPHERE, draw++, 0.1, 0.4);
            double x = uab(seed, S_SPHERE, draw++, -0.3 * L, 0.3 * L);
            double y = uab(seed, S_SPHERE, draw++, -0.1 * L, 0.45 * L - r);
            double z = uab(seed, S_SPHERE, draw++, 0.1 * L, 0.4 * L);
            int ok = 1;
            for (int f = 0; f < n_frames && ok; ++f) {
                const double* P = &poses[(size_t)f * 16];
                double dx = x - P[12], dy = y - P[13], dz = z - P[14];
                if (sqrt(dx * dx + dy * dy + dz * dz) - r < 0.45) ok = 0;
            }
            if (ok) {
                spheres[s * 4 + 0] = x; spheres[s * 4 + 1] = y;
                spheres[s * 4 + 2] = z; spheres[s * 4 + 3] = r;
                break;
            }
            ++rejected;
        }
    }
    return rejected;
}

/* First hit distance t > 0 (camera depth, since the camera-space ray has z = 1). */
static double cast(double L, const double* sph, int n_sph, const double o[3], const double d[3])
{
    double best = INFINITY;
    double p = 0.45 * L;
    if (d[1] > 0) { double t = (p - o[1]) / d[1]; if (t > 0 && t < best) best = t; }
    if (d[0] > 0) { double t = (p - o[0]) / d[0]; if (t > 0 && t < best) best = t; }
    if (d[0] < 0) { double t = (-p - o[0]) / d[0]; if (t > 0 && t < best) best = t; }
    if (d[2] > 0) { double t = (p - o[2]) / d[2]; if (t > 0 && t < best) best = t; }
    double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    for (int s = 0; s < n_sph; ++s) {
        double oc[3] = {o[0] - sph[s * 4], o[1] - sph[s * 4 + 1], o[2] - sph[s * 4 + 2]};
        double b = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
        double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - sph[s * 4 + 3] * sph[s * 4 + 3];
        double disc = b * b - dd * c;
        if (disc < 0) continue;
        double sq = sqrt(disc);
        double t = (-b - sq) / dd;
        if (t <= 0) t = (-b + sq) / dd;
        if (t > 0 && t < best) best = t;
    }
    return best;
}

/* Renders frames [first, first+count) of the sequence into out (count x h x w).
 * true_depth (optional) receives the noiseless FP64 depth in metres. */
void kfs_render(double L, const double* spheres, int n_sph, const double* poses, int first, int count,
                int w, int h, double fx, double fy, double cx, double cy, uint64_t seed, double noise_k,
                double dropout, uint16_t* out, double* true_depth)
{
#pragma omp parallel for schedule(static) collapse(2)
    for (int f = 0; f < count; ++f)
        for (int j = 0; j < h; ++j) {
            const double* P = &poses[(size_t)(first + f) * 16];
            double o[3] = {P[12], P[13], P[14]};
            for (int i = 0; i < w; ++i) {
                double dc[3] = {(i + 0.5 - cx) / fx, (j + 0.5 - cy) / fy, 1.0};
                double d[3];
                for (int r = 0; r < 3; ++r) d[r] = P[r] * dc[0] + P[4 + r] * dc[1] + P[8 + r] * dc[2];
                double z = cast(L, spheres, n_sph, o, d);
                size_t q = ((size_t)f * h + j) * w + i;
                uint64_t pix = ((uint64_t)(first + f) << 32) | (uint64_t)(j * w + i);
                if (true_depth) true_depth[q] = isfinite(z) ? z : 0.0;
                double zn = 0.0;
                if (isfinite(z)) {
                    double u1 = u01(seed, S_NOISE, pix * 2), u2 = u01(seed, S_NOISE, pix * 2 + 1);
                    double g = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
                    zn = z + g * noise_k * z * z;
                }
                if (dropout > 0 && u01(seed, S_DROP, pix) <= dropout) zn = 0.0;
                double mm = floor(zn * 1000.0 + 0.5);
                mm = mm < 0 ? 0 : (mm > 65535 ? 65535 : mm);
                out[q] = (uint16_t)mm;
            }
        }
}
