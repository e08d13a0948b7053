/* Stress driver of the host copy pool (bridct_amd/csrc/copy_pool.c), CPU only.
 * usage: copy_pool_stress POOLS ROUNDS MAXBYTES
 * Creates POOLS pools one after another; each runs ROUNDS calls of one to four jobs of
 * random sizes below MAXBYTES right after creation (the helpers' start-up race), checks
 * every copied byte, and is destroyed.  Prints "ok <helpers>" on success. */
#include "copy_pool.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static unsigned next(unsigned *s)
{
    *s = *s * 1664525u + 1013904223u;
    return *s >> 8;
}

int main(int argc, char **argv)
{
    if (argc != 4)
        return 2;
    const int pools = atoi(argv[1]), rounds = atoi(argv[2]);
    const size_t cap = (size_t)atol(argv[3]);
    unsigned char *src[4], *dst[4];
    for (int j = 0; j < 4; j++) {
        src[j] = malloc(cap ? cap : 1);
        dst[j] = malloc(cap ? cap : 1);
        if (!src[j] || !dst[j])
            return 2;
    }
    unsigned seed = 12345;
    int helpers = -1;
    for (int i = 0; i < pools; i++) {
        copy_pool *p = pool_create();
        if (!p)
            return 3;
        helpers = pool_helpers(p);
        for (int r = 0; r < rounds; r++) {
            const int nj = 1 + (int)(next(&seed) & 3);
            copy_job jobs[4];
            for (int j = 0; j < nj; j++) {
                const size_t n = cap ? next(&seed) % (cap + 1) : 0;
                const unsigned char tag = (unsigned char)next(&seed);
                for (size_t b = 0; b < n; b++)
                    src[j][b] = (unsigned char)(tag + b * 7);
                memset(dst[j], 0, n);
                jobs[j] = (copy_job){dst[j], src[j], n};
            }
            pool_copy(p, jobs, nj);
            for (int j = 0; j < nj; j++)
                if (memcmp(dst[j], src[j], jobs[j].bytes)) {
                    fprintf(stderr, "pool %d round %d job %d: copy differs\n", i, r, j);
                    return 1;
                }
        }
        pool_destroy(p);
    }
    printf("ok %d\n", helpers);
    return 0;
}
