/* Division of a float by a constant as a double product: for every float x
 * (all 2^32 bit patterns) and each constant c given on the command line,
 * checks that r = (float)((double)x * (1.0 / (double)c)), replaced by x / c
 * where r is subnormal (kun::div_rc), has the bits of the IEEE float
 * quotient x / c (NaNs: both NaN). With -bare first, r is checked without
 * the replacement. Prints one line per constant (mismatches of the checked
 * form) and exits 1 if any constant has one. Threads split the bit
 * patterns. */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NT 8

typedef struct {
    int bare;
    float c;
    double rc;
    uint64_t lo, hi;
    uint64_t bad;
    uint32_t first_bad;
} job_t;

static void *run(void *p) {
    job_t *j = (job_t *)p;
    volatile float cv = j->c; /* the division stays a division */
    const float c = cv;
    for (uint64_t u = j->lo; u < j->hi; u++) {
        float x;
        uint32_t b = (uint32_t)u;
        memcpy(&x, &b, 4);
        const float q = x / c;
        float r = (float)((double)x * j->rc);
        if (!j->bare && fpclassify(r) == FP_SUBNORMAL) r = x / c;
        uint32_t bq, br;
        memcpy(&bq, &q, 4);
        memcpy(&br, &r, 4);
        if (bq != br && !(isnan(q) && isnan(r))) {
            if (!j->bad) j->first_bad = b;
            j->bad++;
        }
    }
    return NULL;
}

int main(int argc, char **argv) {
    int rc_all = 0, bare = 0, a0 = 1;
    if (argc > 1 && strcmp(argv[1], "-bare") == 0) bare = 1, a0 = 2;
    for (int a = a0; a < argc; a++) {
        const float c = strtof(argv[a], NULL);
        job_t jobs[NT];
        pthread_t th[NT];
        const uint64_t n = 1ull << 32, step = n / NT;
        for (int t = 0; t < NT; t++) {
            jobs[t].bare = bare;
            jobs[t].c = c;
            jobs[t].rc = 1.0 / (double)c;
            jobs[t].lo = t * step;
            jobs[t].hi = t == NT - 1 ? n : (t + 1) * step;
            jobs[t].bad = 0;
            pthread_create(&th[t], NULL, run, &jobs[t]);
        }
        uint64_t bad = 0;
        uint32_t first = 0;
        for (int t = 0; t < NT; t++) {
            pthread_join(th[t], NULL);
            if (jobs[t].bad && !bad) first = jobs[t].first_bad;
            bad += jobs[t].bad;
        }
        printf("c=%.9g mismatches=%llu%s\n", (double)c, (unsigned long long)bad, bad ? " FIRST" : "");
        if (bad) {
            printf("  first x bits 0x%08x\n", first);
            rc_all = 1;
        }
    }
    return rc_all;
}
