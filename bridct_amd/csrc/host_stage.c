/* Staged host <-> device copies: see host_stage.h. */
#include "host_stage.h"

#include "bridct_device.h"
#include "copy_pool.h"

#include <stdlib.h>

struct host_stage {
    size_t chunk;        /* bytes per pinned buffer */
    void *pin[2];
    void *ev[2];         /* last DMA of each buffer */
    int used[2];         /* ev[i] has been recorded */
    copy_pool *pool;
};

host_stage *stage_create(size_t chunk_bytes)
{
    if (!chunk_bytes)
        return NULL;
    host_stage *s = calloc(1, sizeof(*s));
    if (!s)
        return NULL;
    s->chunk = chunk_bytes;
    for (int i = 0; i < 2; i++)
        if (bdx_host_malloc(&s->pin[i], chunk_bytes) || bdx_event_create(&s->ev[i])) {
            stage_destroy(s);
            return NULL;
        }
    if (!(s->pool = pool_create())) {
        stage_destroy(s);
        return NULL;
    }
    return s;
}

/* Buffer i is free for the host once its last DMA has completed. */
static int slot_free(host_stage *s, int i)
{
    if (!s->used[i])
        return 0;
    s->used[i] = 0;
    return bdx_event_sync(s->ev[i]);
}

void stage_destroy(host_stage *s)
{
    if (!s)
        return;
    for (int i = 0; i < 2; i++) {
        slot_free(s, i);
        if (s->ev[i])
            bdx_event_destroy(s->ev[i]);
        if (s->pin[i])
            bdx_host_free(s->pin[i]);
    }
    pool_destroy(s->pool);
    free(s);
}

int stage_h2d(host_stage *s, void *d_dst, const void *h_src, size_t bytes, void *stream)
{
    for (size_t at = 0, k = 0; at < bytes; at += s->chunk, k++) {
        const int i = (int)(k & 1);
        const size_t n = bytes - at < s->chunk ? bytes - at : s->chunk;
        if (slot_free(s, i))
            return -1;
        const copy_job job = {s->pin[i], (const char *)h_src + at, n};
        pool_copy(s->pool, &job, 1);
        if (bdx_copy_h2d((char *)d_dst + at, s->pin[i], n, stream) || bdx_event_record(s->ev[i], stream))
            return -1;
        s->used[i] = 1;
    }
    return 0;
}

int stage_d2h(host_stage *s, void *h_dst, const void *d_src, size_t bytes, void *stream)
{
    const size_t nchunks = (bytes + s->chunk - 1) / s->chunk;
    /* chunk k + 1 is copied out of the device while the host copies chunk k out of its
     * buffer */
    for (size_t k = 0; k < nchunks + 1; k++) {
        if (k < nchunks) {
            const int i = (int)(k & 1);
            const size_t at = k * s->chunk, n = bytes - at < s->chunk ? bytes - at : s->chunk;
            if (slot_free(s, i) || bdx_copy_d2h(s->pin[i], (const char *)d_src + at, n, stream) ||
                bdx_event_record(s->ev[i], stream))
                return -1;
            s->used[i] = 1;
        }
        if (k >= 1) {
            const int i = (int)((k - 1) & 1);
            const size_t at = (k - 1) * s->chunk, n = bytes - at < s->chunk ? bytes - at : s->chunk;
            if (slot_free(s, i))
                return -1;
            const copy_job job = {(char *)h_dst + at, s->pin[i], n};
            pool_copy(s->pool, &job, 1);
        }
    }
    return 0;
}
