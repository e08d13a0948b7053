/* Host copy pool of the pinned pipeline: see copy_pool.h.
 *
 * A round: the caller publishes the jobs, bumps `round` and wakes the helpers; every
 * thread copies its piece of each job; the caller waits until `pending` (helpers not yet
 * done with the round) is zero.  The caller never starts a round before the previous one
 * has finished, so a helper cannot miss a round.  A helper counts rounds from zero, the
 * count when it is created: a helper that first takes the lock after the caller has
 * already started the first round must still take part in it. */
#include "copy_pool.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    copy_pool *pool;
    int id;
} pool_member;

struct copy_pool {
    pthread_t th[POOL_HELPERS];
    pool_member member[POOL_HELPERS];
    int started;          /* helpers running */
    pthread_mutex_t m;
    pthread_cond_t go, done;
    unsigned round;       /* rounds started so far */
    int pending;          /* helpers still copying in this round */
    int quit;
    copy_job jobs[2];
    int njobs;
};

static void pool_share(const copy_pool *p, int id, int parts)
{
    for (int j = 0; j < p->njobs; j++) {
        const copy_job *c = &p->jobs[j];
        /* ceil(bytes / parts), rounded up to 256 bytes: the pieces cover the whole job */
        const size_t piece = ((c->bytes + (size_t)parts - 1) / (size_t)parts + 255) & ~(size_t)255;
        const size_t at = (size_t)id * piece;
        if (at < c->bytes)
            memcpy((char *)c->dst + at, (const char *)c->src + at, c->bytes - at < piece ? c->bytes - at : piece);
    }
}

static void *pool_main(void *arg)
{
    const pool_member *me = arg;
    copy_pool *p = me->pool;
    unsigned seen = 0;   /* the round current at creation: none has started yet */
    pthread_mutex_lock(&p->m);
    for (;;) {
        while (p->round == seen && !p->quit)
            pthread_cond_wait(&p->go, &p->m);
        if (p->quit)
            break;
        seen = p->round;
        const int parts = p->started + 1;
        pthread_mutex_unlock(&p->m);
        pool_share(p, me->id, parts);
        pthread_mutex_lock(&p->m);
        if (--p->pending == 0)
            pthread_cond_signal(&p->done);
    }
    pthread_mutex_unlock(&p->m);
    return NULL;
}

copy_pool *pool_create(void)
{
    copy_pool *p = calloc(1, sizeof(*p));
    if (!p)
        return NULL;
    if (pthread_mutex_init(&p->m, NULL)) {
        free(p);
        return NULL;
    }
    if (pthread_cond_init(&p->go, NULL)) {
        pthread_mutex_destroy(&p->m);
        free(p);
        return NULL;
    }
    if (pthread_cond_init(&p->done, NULL)) {
        pthread_cond_destroy(&p->go);
        pthread_mutex_destroy(&p->m);
        free(p);
        return NULL;
    }
    pthread_mutex_lock(&p->m);
    for (int i = 0; i < POOL_HELPERS; i++) {
        p->member[i].pool = p;
        p->member[i].id = i + 1;
        if (pthread_create(&p->th[i], NULL, pool_main, &p->member[i]))
            break;
        p->started++;
    }
    pthread_mutex_unlock(&p->m);
    return p;
}

void pool_destroy(copy_pool *p)
{
    if (!p)
        return;
    pthread_mutex_lock(&p->m);
    p->quit = 1;
    pthread_cond_broadcast(&p->go);
    pthread_mutex_unlock(&p->m);
    for (int i = 0; i < p->started; i++)
        pthread_join(p->th[i], NULL);
    pthread_cond_destroy(&p->done);
    pthread_cond_destroy(&p->go);
    pthread_mutex_destroy(&p->m);
    free(p);
}

void pool_copy(copy_pool *p, const copy_job *jobs, int njobs)
{
    if (njobs <= 0)
        return;
    if (njobs > 2) {   /* one round per pair of jobs */
        for (int j = 0; j < njobs; j += 2)
            pool_copy(p, jobs + j, njobs - j < 2 ? njobs - j : 2);
        return;
    }
    pthread_mutex_lock(&p->m);
    for (int j = 0; j < njobs; j++)
        p->jobs[j] = jobs[j];
    p->njobs = njobs;
    p->pending = p->started;
    p->round++;
    pthread_cond_broadcast(&p->go);
    const int parts = p->started + 1;
    pthread_mutex_unlock(&p->m);
    pool_share(p, 0, parts);
    pthread_mutex_lock(&p->m);
    while (p->pending)
        pthread_cond_wait(&p->done, &p->m);
    pthread_mutex_unlock(&p->m);
}

int pool_helpers(const copy_pool *p)
{
    return p ? p->started : 0;
}
