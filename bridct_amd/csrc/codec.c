/* 8x8 block codec step (C11 host side of bdx_codec.hip): level shift, 8x8 transform and
 * quantization fused in one kernel per direction.  No counterpart in the reference, whose
 * callers would run bridct_plan_apply on each 8x8 block (src/plan.c:126-163) and quantize
 * the result themselves.
 *
 * The length-8 core's output normalization alpha is constant along each line, so it
 * commutes with the other axis' core: the encode kernel applies alpha_u alpha_v / Q_uv
 * once per coefficient after both passes, and the decode kernel Q_uv alpha_u alpha_v once
 * before them.  Both factor tables are evaluated in FP64 and rounded once.
 *
 * The object owns its stream, the 2 x 64 factors, and staging buffers for the
 * host-pointer calls; it has no global state.
 */
#include "bridct_gpu.h"

#include "bridct_device.h"
#include "host_stage.h"
#include "host_internal.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

struct bridct_codec8 {
    void *stream;
    float *d_factors;         /* [0, 64): alpha alpha / Q (encode); [64, 128): Q alpha alpha */
    unsigned char *d_image;   /* host-call staging, grown on demand */
    short *d_coeffs;
    size_t image_cap, coeff_cap;
    host_stage *stage;        /* pinned staging of large host calls, on first use */
};

/* Host-call copies of at least this many bytes are staged through pinned memory. */
#define STAGE_MIN_BYTES ((size_t)1 << 20)
#define STAGE_CHUNK_BYTES ((size_t)8 << 20)

static int to_device(bridct_codec8 *c, void *d_dst, const void *src, size_t bytes)
{
    if (bytes >= STAGE_MIN_BYTES && !c->stage)
        c->stage = stage_create(STAGE_CHUNK_BYTES);
    return bytes >= STAGE_MIN_BYTES && c->stage ? stage_h2d(c->stage, d_dst, src, bytes, c->stream)
                                                 : bdx_copy_h2d(d_dst, src, bytes, c->stream);
}

/* after the work issued so far; the host buffer holds the bytes on return (success) */
static int to_host(bridct_codec8 *c, void *dst, const void *d_src, size_t bytes)
{
    if (bytes >= STAGE_MIN_BYTES && !c->stage)
        c->stage = stage_create(STAGE_CHUNK_BYTES);
    if (bytes >= STAGE_MIN_BYTES && c->stage)
        return stage_d2h(c->stage, dst, d_src, bytes, c->stream);
    return bdx_copy_d2h(dst, d_src, bytes, c->stream) || bdx_stream_sync(c->stream);
}

void bridct_codec8_destroy(bridct_codec8 *c)
{
    if (!c)
        return;
    if (c->stream)
        bdx_stream_sync(c->stream);
    stage_destroy(c->stage);
    void *bufs[] = {c->d_factors, c->d_image, c->d_coeffs};
    for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++)
        if (bufs[i])
            bdx_free(bufs[i]);
    if (c->stream)
        bdx_stream_destroy(c->stream);
    free(c);
}

bridct_codec8 *bridct_codec8_create(const float quant[64])
{
    if (!quant)
        return NULL;
    for (int i = 0; i < 64; i++)
        if (!isfinite(quant[i]) || !(quant[i] > 0.0f))
            return NULL;
    if (bdx_device_ready() != 0)
        return NULL;
    float factors[128];
    for (int u = 0; u < 8; u++)
        for (int v = 0; v < 8; v++) {
            const double a = bridct_axis_alpha(8, u) * bridct_axis_alpha(8, v), q = quant[u * 8 + v];
            factors[u * 8 + v] = (float)(a / q);
            factors[64 + u * 8 + v] = (float)(a * q);
        }
    bridct_codec8 *c = calloc(1, sizeof(*c));
    if (!c)
        return NULL;
    if (bdx_stream_create(&c->stream) || bdx_malloc((void **)&c->d_factors, sizeof(factors)) ||
        bdx_copy_h2d(c->d_factors, factors, sizeof(factors), c->stream) || bdx_stream_sync(c->stream)) {
        bridct_codec8_destroy(c);
        return NULL;
    }
    return c;
}

/* Shared validation of sizes; 0 when the geometry is usable. */
static int check_geometry(size_t height, size_t width, size_t stride)
{
    if (height == 0 || width == 0 || height % 8 || width % 8 || stride < width)
        return BRIDCT_INVALID_ARGUMENT;
    if (height > SIZE_MAX / stride || (height / 8) * (width / 8) > SIZE_MAX / 128)
        return BRIDCT_INVALID_ARGUMENT;
    return BRIDCT_OK;
}

static int run(bridct_codec8 *c, int decode, const void *src, void *dst, size_t height, size_t width,
               size_t stride)
{
    if (!c || !src || !dst)
        return BRIDCT_INVALID_ARGUMENT;
    const int g = check_geometry(height, width, stride);
    if (g)
        return g;
    const void *image = decode ? dst : src, *coeffs = decode ? src : dst;
    if ((uintptr_t)image % 8 || stride % 8 || (uintptr_t)coeffs % 16)
        return BRIDCT_BAD_ALIGNMENT;
    return bdx_codec8(decode, src, dst, height, width, stride, c->d_factors, c->stream)
               ? BRIDCT_DEVICE_ERROR
               : BRIDCT_OK;
}

int bridct_codec8_encode_device(bridct_codec8 *c, const unsigned char *d_image, size_t height, size_t width,
                                size_t stride, short *d_coeffs)
{
    return run(c, 0, d_image, d_coeffs, height, width, stride);
}

int bridct_codec8_decode_device(bridct_codec8 *c, const short *d_coeffs, size_t height, size_t width,
                                size_t stride, unsigned char *d_image)
{
    return run(c, 1, d_coeffs, d_image, height, width, stride);
}

static int grow(void **buf, size_t *cap, size_t bytes)
{
    if (*cap >= bytes)
        return 0;
    if (*buf) {
        bdx_free(*buf);
        *buf = NULL;
        *cap = 0;
    }
    if (bdx_malloc(buf, bytes))
        return -1;
    *cap = bytes;
    return 0;
}

/* Host calls: the device raster is compact (stride = width); rows of a strided host
 * image are packed or unpacked on the host. */
static int host_run(bridct_codec8 *c, int decode, const void *src, void *dst, size_t height, size_t width,
                    size_t stride)
{
    if (!c || !src || !dst)
        return BRIDCT_INVALID_ARGUMENT;
    const int g = check_geometry(height, width, stride);
    if (g)
        return g;
    const size_t image_bytes = height * width, coeff_bytes = (height / 8) * (width / 8) * 128;
    const size_t host_image_bytes = (height - 1) * stride + width;
    const uintptr_t im = (uintptr_t)(decode ? dst : src), co = (uintptr_t)(decode ? src : dst);
    if (im < co + coeff_bytes && co < im + host_image_bytes)
        return BRIDCT_OVERLAPPING_BUFFERS;
    if (grow((void **)&c->d_image, &c->image_cap, image_bytes) ||
        grow((void **)&c->d_coeffs, &c->coeff_cap, coeff_bytes))
        return BRIDCT_DEVICE_ERROR;
    unsigned char *packed = NULL;
    if (stride != width && !(packed = malloc(image_bytes)))
        return BRIDCT_DEVICE_ERROR;
    int rc = BRIDCT_OK;
    if (!decode) {
        const unsigned char *img = src;
        if (packed)
            for (size_t r = 0; r < height; r++)
                memcpy(packed + r * width, img + r * stride, width);
        if (to_device(c, c->d_image, packed ? packed : img, image_bytes) ||
            bdx_codec8(0, c->d_image, c->d_coeffs, height, width, width, c->d_factors, c->stream) ||
            to_host(c, dst, c->d_coeffs, coeff_bytes))
            rc = BRIDCT_DEVICE_ERROR;
    } else {
        unsigned char *img = dst;
        if (to_device(c, c->d_coeffs, src, coeff_bytes) ||
            bdx_codec8(1, c->d_coeffs, c->d_image, height, width, width, c->d_factors, c->stream) ||
            to_host(c, packed ? packed : img, c->d_image, image_bytes))
            rc = BRIDCT_DEVICE_ERROR;
        else if (packed)
            for (size_t r = 0; r < height; r++)
                memcpy(img + r * stride, packed + r * width, width);
    }
    free(packed);
    return rc;
}

int bridct_codec8_encode(bridct_codec8 *c, const unsigned char *image, size_t height, size_t width,
                         size_t stride, short *coeffs)
{
    return host_run(c, 0, image, coeffs, height, width, stride);
}

int bridct_codec8_decode(bridct_codec8 *c, const short *coeffs, size_t height, size_t width, size_t stride,
                         unsigned char *image)
{
    return host_run(c, 1, coeffs, image, height, width, stride);
}

int bridct_codec8_synchronize(bridct_codec8 *c)
{
    if (!c)
        return BRIDCT_INVALID_ARGUMENT;
    return bdx_stream_sync(c->stream) ? BRIDCT_DEVICE_ERROR : BRIDCT_OK;
}

void *bridct_codec8_stream(const bridct_codec8 *c)
{
    return c ? c->stream : NULL;
}
