/* Exhaustive check of the sweeps' two-operation division by 3,
 * q = fma(x, zh, x * zl) with zh = RN(1/3), zl = RN(1/3 - zh), against IEEE
 * division over every non-negative finite binary32 x.  The reference value is
 * (float)((double)x / 3.0): x / 3 is never within 2^-53 (relative) of a
 * binary32 midpoint, so rounding through binary64 first gives the correctly
 * rounded quotient.
 *
 * usage: div3_check <zl as a hex float> [stride] -> prints "<mismatches>
 * <largest x that mismatched, as bits>"; stride > 1 checks every stride-th
 * bit pattern only. */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NT 8
static float g_zh = 0x1.555556p-2f, g_zl;
static uint32_t g_stride = 1;
static uint64_t g_bad[NT];
static uint32_t g_last[NT];

static void* worker(void* arg) {
    const int t = (int)(intptr_t)arg;
    const uint64_t end = 0x7f800000u;  /* +inf */
    for (uint64_t v = (uint64_t)t * g_stride; v < end; v += (uint64_t)NT * g_stride) {
        const uint32_t u = (uint32_t)v;
        float x;
        memcpy(&x, &u, 4);
        const float ref = (float)((double)x / 3.0);
        const float q = fmaf(x, g_zh, x * g_zl);
        if (memcmp(&q, &ref, 4) != 0) {
            ++g_bad[t];
            g_last[t] = u;
        }
    }
    return NULL;
}

int main(int argc, char** argv) {
    g_zl = argc > 1 ? strtof(argv[1], NULL) : -0x1.555556p-27f;
    g_stride = argc > 2 ? (uint32_t)strtoul(argv[2], NULL, 10) : 1u;
    if (g_stride == 0) g_stride = 1;
    pthread_t th[NT];
    for (int t = 0; t < NT; ++t) pthread_create(&th[t], NULL, worker, (void*)(intptr_t)t);
    uint64_t bad = 0;
    uint32_t last = 0;
    for (int t = 0; t < NT; ++t) {
        pthread_join(th[t], NULL);
        bad += g_bad[t];
        if (g_bad[t] && g_last[t] > last) last = g_last[t];
    }
    printf("%llu %u\n", (unsigned long long)bad, last);
    return 0;
}
