/* Host restatement of div3() in binarymeshfitting_amd/csrc/bmf_diff.hip
 * (x / d in three FP32 operations), checked against IEEE division over
 * ranges of float bit patterns.  Test infrastructure: built and called by
 * tests/test_exact_division.py only. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static float bits_float(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static uint32_t float_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* q0 = x zh, r = x - q0 d exactly (one fma), q0 + r zh (one fma); zeros and
 * infinities return q0. */
static float div_short(float x, float d, float zh) {
    const float q0 = x * zh;
    if (x == 0.0f || isinf(x)) return q0;
    return fmaf(fmaf(-q0, d, x), zh, q0);
}

/* Number of patterns u in [lo, hi) for which div_short(x, d, RN(1/d))
 * differs bitwise from x / d; NaN results compare equal.  (Every step is
 * odd in x under round-to-nearest-even, so positive patterns suffice.) */
long exact_division_mismatches(int d, uint32_t lo, uint32_t hi) {
    const float fd = (float)d, zh = 1.0f / fd;
    long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = (int64_t)lo; i < (int64_t)hi; i++) {
        const float x = bits_float((uint32_t)i);
        const float want = x / fd, got = div_short(x, fd, zh);
        bad += float_bits(got) != float_bits(want) && !(isnan(got) && isnan(want));
    }
    return bad;
}
