// Row-pass and column-pass kernels for lines of 8, 16 and 32 (one lane per line; templates
// in bdx_passes.h), used where the schedule table sends a whole-array shape to the
// two-pass kernels.  A code object of their own (see bdx_passes.h).
#include "bdx_passes.h"

namespace bdx {

int launch_rows_lines(int w, const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s,
                      const size_t *offs, int h, bool nt)
{
    switch (w) {
    case 8: return rows_w<8>(in, out, nrows, op, tab, s, offs, h, nt);
    case 16: return rows_w<16>(in, out, nrows, op, tab, s, offs, h, nt);
    case 32: return rows_w<32>(in, out, nrows, op, tab, s, offs, h, nt);
    default: return -1;
    }
}

int launch_cols_lines(int h, int w, const float *in, float *out, size_t batch, int mode, const float *tab,
                      hipStream_t s, const size_t *offs, bool nt)
{
    switch (h) {
    case 8: return cols_h<8>(w, in, out, batch, mode, tab, s, offs, nt);
    case 16: return cols_h<16>(w, in, out, batch, mode, tab, s, offs, nt);
    case 32: return cols_h<32>(w, in, out, batch, mode, tab, s, offs, nt);
    default: return -1;
    }
}

}  // namespace bdx
