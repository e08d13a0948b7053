/* BRiDCT plan API, MI355X implementation.
 *
 * Drop-in for the reference public header (include/bridct.h:1-55 of BRiDCT 0.2.0-dev.7):
 * the same seven functions, two enums, status values and BRIDCT_VERSION, with the same
 * argument meaning and error behaviour (docs/API.md).  The arithmetic runs on the GPU;
 * the host-pointer functions copy through device buffers owned by the plan.
 */
#ifndef BRIDCT_H
#define BRIDCT_H

#include <stddef.h>

#if defined(__GNUC__)
#define BRIDCT_API __attribute__((visibility("default")))
#else
#define BRIDCT_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Same version string as the reference interface it replaces (include/bridct.h:22). */
#define BRIDCT_VERSION "0.2.0-dev.7"

/* include/bridct.h:24-28 */
enum bridct_mode {
    BRIDCT_FORWARD = 0,
    BRIDCT_INVERSE = 1,
    BRIDCT_ROUNDTRIP = 2
};

/* include/bridct.h:30-36 */
enum bridct_status {
    BRIDCT_OK = 0,
    BRIDCT_INVALID_ARGUMENT = -1,
    BRIDCT_WORKSPACE_TOO_SMALL = -2,
    BRIDCT_BAD_ALIGNMENT = -3,
    BRIDCT_OVERLAPPING_BUFFERS = -4
};

typedef struct bridct_plan bridct_plan;

/* Replaces bridct_supported (include/bridct.h:40, src/plan.c:33-37): both axes powers of
 * two in 8..1024. */
BRIDCT_API int bridct_supported(int height, int width);

/* Replaces bridct_plan_create (include/bridct.h:41, src/plan.c:59-115).  NULL for an
 * unsupported shape, an allocation failure, or when no gfx950 GPU is available. */
BRIDCT_API bridct_plan *bridct_plan_create(int height, int width);

/* Replaces bridct_plan_destroy (include/bridct.h:42, src/plan.c:49-57); NULL is a no-op. */
BRIDCT_API void bridct_plan_destroy(bridct_plan *plan);

/* Replaces bridct_plan_route (include/bridct.h:43, src/plan.c:117-124): the static string
 * "paper8-square" (square up to 256) or "blocked-extension"; NULL for a NULL plan. */
BRIDCT_API const char *bridct_plan_route(const bridct_plan *plan);

/* Replaces bridct_plan_apply (include/bridct.h:44-45, src/plan.c:126-163).  Host
 * pointers to h*w floats each, disjoint.  Validates before any write. */
BRIDCT_API int bridct_plan_apply(bridct_plan *plan, enum bridct_mode mode,
                                 const float *input, float *output);

/* Replaces bridct_workspace_floats (include/bridct.h:47, src/bridct.c:13-21). */
BRIDCT_API size_t bridct_workspace_floats(int size);

/* Replaces bridct_apply (include/bridct.h:48-50, src/bridct.c:30-75): square 8..256 with
 * caller workspace, error precedence -1, -2, -3, -4.  The GPU state of each size (tables,
 * stream, staging) is made by a thread's first call of that size and kept until the
 * thread exits; BRIDCT_DEVICE_ERROR when it cannot be made. */
BRIDCT_API int bridct_apply(int size, enum bridct_mode mode,
                            const float *input, float *output,
                            float *workspace, size_t workspace_count);

#ifdef __cplusplus
}
#endif
#endif
