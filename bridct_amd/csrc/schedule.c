/* Schedule table: route and launch parameters per shape, chosen by the autotuner
 * (bridct_amd/autotune.py) and read when a plan is created.
 *
 * The reference keeps its per-shape policy in compiled-in tables (src/rect_policy.inc:1-46,
 * small_policy.json:14-21, the square-variant table of src/bridct.c:47-74).  Here the table
 * is a text file, one line per shape:
 *
 *     <height> <width> <route> <pack>
 *
 * route: butterfly | mfma_f16 | mfma_bf16; pack: arrays per workgroup of the whole-array
 * kernel (0 = automatic), or -1 (BRIDCT_PACK_TWO_PASS): the row and column pass kernels
 * instead of the whole-array kernel.  '#' starts a comment.  The file is $BRIDCT_SCHEDULE when that
 * variable is set (empty or "none": no table), else schedule_gfx950.txt in the directory
 * of this library.  A shape missing from the table, or a missing table, keeps the
 * defaults (butterfly, automatic packing).
 */
#define _GNU_SOURCE
#include "bridct_gpu.h"

#include "host_internal.h"

#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int route_id(const char *name)
{
    if (!strcmp(name, "butterfly"))
        return BRIDCT_ROUTE_BUTTERFLY;
    if (!strcmp(name, "mfma_f16"))
        return BRIDCT_ROUTE_MFMA_F16;
    if (!strcmp(name, "mfma_bf16"))
        return BRIDCT_ROUTE_MFMA_BF16;
    return -1;
}

int bdx_schedule_lookup_file(const char *path, int h, int w, int *route, int *pack)
{
    FILE *f = fopen(path, "r");
    if (!f)
        return -1;
    char line[256];
    int found = -1;
    while (fgets(line, sizeof(line), f)) {
        char *hash = strchr(line, '#');
        if (hash)
            *hash = 0;
        int lh, lw, lp;
        char name[32];
        if (sscanf(line, "%d %d %31s %d", &lh, &lw, name, &lp) != 4)
            continue;
        if (lh != h || lw != w)
            continue;
        const int r = route_id(name);
        if (r < 0 || lp < BRIDCT_PACK_TWO_PASS || lp > 4096)
            continue;
        *route = r;
        *pack = lp;
        found = 0;   /* the last matching line wins */
    }
    fclose(f);
    return found;
}

int bridct_schedule_lookup(int h, int w, int *route, int *pack)
{
    const char *env = getenv("BRIDCT_SCHEDULE");
    if (env)
        return *env && strcmp(env, "none") ? bdx_schedule_lookup_file(env, h, w, route, pack) : -1;
    Dl_info info;
    if (!dladdr((void *)&bridct_schedule_lookup, &info) || !info.dli_fname)
        return -1;
    char path[4096];
    const char *slash = strrchr(info.dli_fname, '/');
    const size_t dir = slash ? (size_t)(slash - info.dli_fname) + 1 : 0;
    static const char name[] = "schedule_gfx950.txt";
    if (dir + sizeof(name) > sizeof(path))
        return -1;
    memcpy(path, info.dli_fname, dir);
    memcpy(path + dir, name, sizeof(name));
    return bdx_schedule_lookup_file(path, h, w, route, pack);
}
