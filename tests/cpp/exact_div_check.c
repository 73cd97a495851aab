/* CPU check of the division the fuse kernels use in place of a correctly
 * rounded quotient (kinectfusion-cuda_amd/csrc/tsdf.hip,
 * div_by_exact_reciprocal): with y = RN(1/b),
 *     q = RN(x*y),  r = fma(-q, b, x),  result = fma(r, y, q)
 * must equal RN(x/b) for
 *   - the truncation quotient diff / mue: every float x in [-mue, mue] except
 *     the magnitudes below 2^-25 (a depth difference is 0 or at least 2^-25),
 *     for several mue;
 *   - the count update: numerators in [-130, 130] over divisors 1..129 (every
 *     float of [2^-30, 130] for b = 129, and a seeded sample for the others).
 * Prints the number of mismatches and exits non-zero if any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static inline float fdiv(float x, float b, float y)
{
    const float q = x * y;
    const float r = fmaf(-q, b, x);
    return fmaf(r, y, q);
}
static inline uint32_t bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float fl(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint64_t state = 0x9E3779B97F4A7C15ull;
static inline uint64_t next(void)
{
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int main(void)
{
    long bad = 0, total = 0;
    const float mues[] = {0.03f, 3.0f * 4.0f / 511.0f, 0.01f, 0.1f, 1e-4f};
    for (int m = 0; m < 5; ++m) {
        const float b = mues[m], y = 1.0f / b;
        for (uint32_t u = bits(0x1p-25f); u <= bits(b); ++u)
            for (int s = 0; s < 2; ++s) {
                const float x = s ? -fl(u) : fl(u);
                ++total;
                bad += bits(x / b) != bits(fdiv(x, b, y));
            }
    }
    for (int k = 1; k <= 129; ++k) {
        const float b = (float)k, y = 1.0f / b;
        if (k == 129) {
            for (uint32_t u = bits(0x1p-30f); u <= bits(130.0f); ++u)
                for (int s = 0; s < 2; ++s) {
                    const float x = s ? -fl(u) : fl(u);
                    ++total;
                    bad += bits(x / b) != bits(fdiv(x, b, y));
                }
        } else {
            for (int t = 0; t < 2000000; ++t) {
                const uint64_t v = next();
                float x = fl(bits(0x1p-30f) + (uint32_t)(v % (bits(130.0f) - bits(0x1p-30f))));
                if (v >> 63) x = -x;
                ++total;
                bad += bits(x / b) != bits(fdiv(x, b, y));
            }
        }
    }
    printf("%ld of %ld quotients differ\n", bad, total);
    return bad != 0;
}
