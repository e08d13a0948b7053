// Row-pass and column-pass kernels for lines of 64 and more (templates in bdx_passes.h).
#include "bdx_passes.h"

namespace bdx {

// XCD count of the current device (hipDeviceAttributeNumberOfXccs), read once per device;
// 1 when it cannot be read, which gives the plain block order.
unsigned xcd_count()
{
    constexpr int kMaxDevices = 64;
    static int count[kMaxDevices];   // 0: not read yet (racing first reads store the same value)
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return 1;
    int c = __atomic_load_n(&count[dev], __ATOMIC_RELAXED);
    if (c == 0) {
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || c < 1)
            c = 1;
        __atomic_store_n(&count[dev], c, __ATOMIC_RELAXED);
    }
    return (unsigned)c;
}

int launch_rows(int w, const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s,
                const size_t *offs, int h, bool nt)
{
    switch (w) {
    case 64: return rows_w<64>(in, out, nrows, op, tab, s, offs, h, nt);
    case 128: return rows_w<128>(in, out, nrows, op, tab, s, offs, h, nt);
    case 256: return rows_w<256>(in, out, nrows, op, tab, s, offs, h, nt);
    case 512: return rows_w<512>(in, out, nrows, op, tab, s, offs, h, nt);
    case 1024: return rows_w<1024>(in, out, nrows, op, tab, s, offs, h, nt);
    default: return launch_rows_lines(w, in, out, nrows, op, tab, s, offs, h, nt);
    }
}

int launch_cols(int h, int w, const float *in, float *out, size_t batch, int mode, const float *tab,
                hipStream_t s, const size_t *offs, bool nt)
{
    switch (h) {
    case 64: return cols_h<64>(w, in, out, batch, mode, tab, s, offs, nt);
    case 128: return cols_h<128>(w, in, out, batch, mode, tab, s, offs, nt);
    case 256: return cols_h<256>(w, in, out, batch, mode, tab, s, offs, nt);
    case 512: return cols_h<512>(w, in, out, batch, mode, tab, s, offs, nt);
    case 1024: return cols_h<1024>(w, in, out, batch, mode, tab, s, offs, nt);
    default: return launch_cols_lines(h, w, in, out, batch, mode, tab, s, offs, nt);
    }
}

}  // namespace bdx
