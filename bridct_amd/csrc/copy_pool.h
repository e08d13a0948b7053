/* Host copy pool of the pinned pipeline (plan.c): a few helper threads that share large
 * memcpy jobs with the calling thread.  One pool serves one plan, from one thread at a
 * time (a plan must not be used concurrently). */
#pragma once
#include <stddef.h>

#ifndef POOL_HELPERS
#define POOL_HELPERS 5
#endif

typedef struct {
    void *dst;
    const void *src;
    size_t bytes;
} copy_job;

typedef struct copy_pool copy_pool;

/* Starts up to POOL_HELPERS helpers (with none, the caller copies alone); NULL if out of
 * memory. */
__attribute__((visibility("hidden"))) copy_pool *pool_create(void);

/* Stops and joins the helpers; NULL is a no-op. */
__attribute__((visibility("hidden"))) void pool_destroy(copy_pool *p);

/* Copies the jobs, each split into one contiguous piece per thread (two jobs per round),
 * and returns when every piece is done. */
__attribute__((visibility("hidden"))) void pool_copy(copy_pool *p, const copy_job *jobs, int njobs);

/* Helpers running (for tests). */
__attribute__((visibility("hidden"))) int pool_helpers(const copy_pool *p);
