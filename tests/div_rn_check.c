/* The correctly rounded division of rt_device.hpp's div_rn (q = RN(a y),
 * r = a - b q exact by an FMA, q' = RN(q + r y), with y = RN(1/b)) against
 * IEEE division: random normal operands over 56 binades, and the
 * back-projection domain of the gradient pass, a = (x - c) z over pixel
 * columns x and every 16-bit depth z, divided by the focal length and then by
 * z, and the 256 8-bit gray values over 255. Prints the mismatch count;
 * exit status 1 on any mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float div_rn(float a, float b, float y) {
    float q = a * y;
    float r = fmaf(-b, q, a);
    return fmaf(r, y, q);
}
static uint64_t state = 88172645463325252ull;
static uint32_t next(void) {
    state ^= state << 13; state ^= state >> 7; state ^= state << 17;
    return (uint32_t)(state >> 11);
}
static int same(float x, float y) {
    uint32_t a, b;
    memcpy(&a, &x, 4); memcpy(&b, &y, 4);
    return a == b || (x == 0.0f && y == 0.0f);  /* the sign of a zero quotient may differ */
}
int main(int argc, char** argv) {
    long nrand = argc > 1 ? atol(argv[1]) : 10000000L, ncols = argc > 2 ? atol(argv[2]) : 640;
    long bad = 0, n = 0;
    for (long i = 0; i < nrand; i++) {
        uint32_t ua = (next() & 0x007fffffu) | ((uint32_t)(100 + next() % 56) << 23) | (next() & 1u) << 31;
        uint32_t ub = (next() & 0x007fffffu) | ((uint32_t)(100 + next() % 56) << 23) | (next() & 1u) << 31;
        float a, b;
        memcpy(&a, &ua, 4); memcpy(&b, &ub, 4);
        n++; bad += !same(div_rn(a, b, 1.0f / b), a / b);
    }
    const float cal[][2] = {{319.5f, -525.0f}, {239.5f, 525.0f}, {320.0f, -517.3f}, {0.3f, 1234.5678f}};
    for (int k = 0; k < 4; k++) {
        const float c = cal[k][0], f = cal[k][1], yf = 1.0f / f;
        for (long x = 0; x < ncols; x++)
            for (int d = 1; d < 65536; d++) {
                const float z = -(float)d, a = ((float)x - c) * z, p = a / f;
                n++; bad += !same(div_rn(a, f, yf), p);
                n++; bad += !same(div_rn(p, z, 1.0f / z), p / z);
            }
    }
    /* the 8-bit gray values over 255 (the pyramid's conversion, u8_unit) */
    for (int t = 0; t < 256; t++) { n++; bad += !same(div_rn((float)t, 255.0f, 1.0f / 255.0f), (float)t / 255.0f); }
    printf("%ld mismatches in %ld divisions\n", bad, n);
    return bad ? 1 : 0;
}
