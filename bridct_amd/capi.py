"""ctypes binding of the in-tree C ABI (include/bridct.h, include/bridct_gpu.h).

This is how Python code that wants device-resident buffers (benchmarks, batch
pipelines) reaches the library; array users should prefer the NumPy extension
``bridct_amd.bridct_numpy``.  Loading fails loudly when the library has not been built:
there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libbridct_amd.so')

FORWARD, INVERSE, ROUNDTRIP, FORWARD_THEN_INVERSE = 0, 1, 2, 3
ROUTE_BUTTERFLY, ROUTE_MFMA_F16, ROUTE_MFMA_BF16 = 0, 1, 2
PACK_TWO_PASS = -1   # row and column pass kernels instead of the whole-array kernel
BATCH_AUTO, BATCH_PERSISTENT, BATCH_GROUPED = 0, 1, 2
OK, INVALID_ARGUMENT, WORKSPACE_TOO_SMALL, BAD_ALIGNMENT, OVERLAPPING_BUFFERS, DEVICE_ERROR = (
    0, -1, -2, -3, -4, -5)

_fp = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes); every function declared in include/bridct.h and bridct_gpu.h
SIGNATURES = {
    'bridct_supported': (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    'bridct_plan_create': (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int]),
    'bridct_plan_destroy': (None, [ctypes.c_void_p]),
    'bridct_plan_route': (ctypes.c_char_p, [ctypes.c_void_p]),
    'bridct_plan_apply': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    'bridct_workspace_floats': (ctypes.c_size_t, [ctypes.c_int]),
    'bridct_apply': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_size_t]),
    'bridct_gpu_available': (ctypes.c_int, []),
    'bridct_plan_apply_device': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_size_t]),
    'bridct_plan_apply_device_static': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                       ctypes.c_void_p, ctypes.c_size_t]),
    'bridct_plan_apply_batch': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_size_t]),
    'bridct_plan_apply_device_pass': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                     ctypes.c_void_p, ctypes.c_size_t]),
    'bridct_plan_synchronize': (ctypes.c_int, [ctypes.c_void_p]),
    'bridct_plan_set_route': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    'bridct_plan_get_route': (ctypes.c_int, [ctypes.c_void_p]),
    'bridct_plan_set_pack': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    'bridct_plan_get_pack': (ctypes.c_int, [ctypes.c_void_p]),
    'bridct_plan_stream': (ctypes.c_void_p, [ctypes.c_void_p]),
    'bridct_batch_create': (ctypes.c_void_p, [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]),
    'bridct_batch_destroy': (None, [ctypes.c_void_p]),
    'bridct_batch_elements': (ctypes.c_size_t, [ctypes.c_void_p]),
    'bridct_batch_offset': (ctypes.c_size_t, [ctypes.c_void_p, ctypes.c_size_t]),
    'bridct_batch_items': (ctypes.c_size_t, [ctypes.c_void_p, ctypes.c_int]),
    'bridct_batch_apply_device': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_void_p]),
    'bridct_batch_graph_apply': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    'bridct_batch_synchronize': (ctypes.c_int, [ctypes.c_void_p]),
    'bridct_batch_apply': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    'bridct_batch_stream': (ctypes.c_void_p, [ctypes.c_void_p]),
    'bridct_batch_set_spin_limit': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint]),
    'bridct_batch_set_strategy': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    'bridct_batch_get_strategy': (ctypes.c_int, [ctypes.c_void_p]),
    'bridct_codec8_create': (ctypes.c_void_p, [ctypes.c_void_p]),
    'bridct_codec8_destroy': (None, [ctypes.c_void_p]),
    'bridct_codec8_encode_device': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                   ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p]),
    'bridct_codec8_decode_device': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                   ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p]),
    'bridct_codec8_encode': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                            ctypes.c_size_t, ctypes.c_void_p]),
    'bridct_codec8_decode': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                            ctypes.c_size_t, ctypes.c_void_p]),
    'bridct_codec8_synchronize': (ctypes.c_int, [ctypes.c_void_p]),
    'bridct_codec8_stream': (ctypes.c_void_p, [ctypes.c_void_p]),
    'bridct_debug_alloc_faults': (ctypes.c_long, [ctypes.c_long]),
    'bridct_debug_grid_limit': (ctypes.c_long, [ctypes.c_long]),
    'bridct_gpu_malloc': (ctypes.c_void_p, [ctypes.c_size_t]),
    'bridct_gpu_free': (ctypes.c_int, [ctypes.c_void_p]),
    'bridct_gpu_upload': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    'bridct_gpu_download': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
}

_lib = None


def load():
    """Returns the loaded library with typed entry points (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError('libbridct_amd.so is not built; run make -C bridct_amd')
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def ptr(a):
    """Host pointer of a NumPy array as c_void_p."""
    return ctypes.c_void_p(a.ctypes.data)


class DeviceBuffer:
    """Owning device allocation of `nbytes` bytes."""

    def __init__(self, nbytes):
        self.lib = load()
        self.nbytes = int(nbytes)
        self.ptr = self.lib.bridct_gpu_malloc(self.nbytes)
        if not self.ptr:
            raise MemoryError('bridct_gpu_malloc(%d) failed' % self.nbytes)

    def upload(self, a):
        assert a.nbytes <= self.nbytes and a.flags.c_contiguous
        rc = self.lib.bridct_gpu_upload(self.ptr, ptr(a), a.nbytes)
        if rc:
            raise RuntimeError('upload failed: %d' % rc)

    def download(self, a):
        assert a.nbytes <= self.nbytes and a.flags.c_contiguous
        rc = self.lib.bridct_gpu_download(ptr(a), self.ptr, a.nbytes)
        if rc:
            raise RuntimeError('download failed: %d' % rc)

    def free(self):
        if self.ptr:
            self.lib.bridct_gpu_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Plan:
    """Owning wrapper of a bridct_plan for device-pointer calls."""

    def __init__(self, h, w):
        self.lib = load()
        self.h, self.w = h, w
        self.ptr = self.lib.bridct_plan_create(h, w)
        if not self.ptr:
            raise RuntimeError('bridct_plan_create(%d, %d) failed' % (h, w))

    def apply_device(self, mode, d_in, d_out, batch):
        return self.lib.bridct_plan_apply_device(self.ptr, mode, d_in, d_out, batch)

    def apply_device_pass(self, mode, pass_, d_in, d_out, batch):
        return self.lib.bridct_plan_apply_device_pass(self.ptr, mode, pass_, d_in, d_out, batch)

    def apply_batch(self, mode, x, y):
        batch = x.size // (self.h * self.w)
        return self.lib.bridct_plan_apply_batch(self.ptr, mode, ptr(x), ptr(y), batch)

    def apply(self, mode, x, y):
        return self.lib.bridct_plan_apply(self.ptr, mode, ptr(x), ptr(y))

    def synchronize(self):
        return self.lib.bridct_plan_synchronize(self.ptr)

    def set_route(self, route):
        return self.lib.bridct_plan_set_route(self.ptr, route)

    @property
    def route(self):
        return self.lib.bridct_plan_get_route(self.ptr)

    def set_pack(self, pack):
        return self.lib.bridct_plan_set_pack(self.ptr, pack)

    @property
    def pack(self):
        return self.lib.bridct_plan_get_pack(self.ptr)

    @property
    def stream(self):
        return self.lib.bridct_plan_stream(self.ptr)

    def close(self):
        if self.ptr:
            self.lib.bridct_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Owning wrapper of a bridct_batch: arrays of shapes `shapes` [(h, w), ...] stored back
    to back in one buffer."""

    def __init__(self, shapes):
        import numpy as np
        self.lib = load()
        self.shapes = [(int(h), int(w)) for h, w in shapes]
        hs = np.ascontiguousarray([h for h, _ in self.shapes], dtype=np.int32)
        ws = np.ascontiguousarray([w for _, w in self.shapes], dtype=np.int32)
        self.ptr = self.lib.bridct_batch_create(len(self.shapes), ptr(hs), ptr(ws))
        if not self.ptr:
            raise RuntimeError('bridct_batch_create failed')
        self.elements = self.lib.bridct_batch_elements(self.ptr)
        self.offsets = [self.lib.bridct_batch_offset(self.ptr, i) for i in range(len(self.shapes))]

    def items(self, mode):
        return self.lib.bridct_batch_items(self.ptr, mode)

    def apply_device(self, mode, d_in, d_out):
        return self.lib.bridct_batch_apply_device(self.ptr, mode, d_in, d_out)

    def graph_apply(self, mode, d_in, d_out):
        return self.lib.bridct_batch_graph_apply(self.ptr, mode, d_in, d_out)

    def apply(self, mode, x, y):
        assert x.size == self.elements and y.size == self.elements
        return self.lib.bridct_batch_apply(self.ptr, mode, ptr(x), ptr(y))

    def synchronize(self):
        return self.lib.bridct_batch_synchronize(self.ptr)

    def set_spin_limit(self, limit):
        return self.lib.bridct_batch_set_spin_limit(self.ptr, limit)

    def set_strategy(self, strategy):
        return self.lib.bridct_batch_set_strategy(self.ptr, strategy)

    @property
    def strategy(self):
        return self.lib.bridct_batch_get_strategy(self.ptr)

    @property
    def stream(self):
        return self.lib.bridct_batch_stream(self.ptr)

    def split(self, flat):
        """Views of the arrays of a flat batch buffer."""
        return [flat[o:o + h * w].reshape(h, w) for o, (h, w) in zip(self.offsets, self.shapes)]

    def close(self):
        if self.ptr:
            self.lib.bridct_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Codec8:
    """Owning wrapper of a bridct_codec8 with quantizer steps `quant` (64 values, row-major
    u * 8 + v).  encode: u8 image (H, W) -> int16 coefficients (H/8 * W/8, 64); decode: the
    reverse."""

    def __init__(self, quant):
        import numpy as np
        self.lib = load()
        self.quant = np.ascontiguousarray(quant, dtype=np.float32).reshape(64)
        self.ptr = self.lib.bridct_codec8_create(ptr(self.quant))
        if not self.ptr:
            raise RuntimeError('bridct_codec8_create failed')

    def encode_device(self, d_image, height, width, stride, d_coeffs):
        return self.lib.bridct_codec8_encode_device(self.ptr, d_image, height, width, stride, d_coeffs)

    def decode_device(self, d_coeffs, height, width, stride, d_image):
        return self.lib.bridct_codec8_decode_device(self.ptr, d_coeffs, height, width, stride, d_image)

    def encode(self, image):
        """Host call: image is a 2D uint8 array whose rows may be strided."""
        import numpy as np
        assert image.dtype == np.uint8 and image.ndim == 2 and image.strides[1] == 1
        h, w = image.shape
        out = np.empty(((h // 8) * (w // 8), 64), np.int16)
        rc = self.lib.bridct_codec8_encode(self.ptr, ptr(image), h, w, image.strides[0], ptr(out))
        if rc:
            raise RuntimeError('bridct_codec8_encode: status %d' % rc)
        return out

    def decode(self, coeffs, height, width):
        import numpy as np
        coeffs = np.ascontiguousarray(coeffs, np.int16)
        out = np.empty((height, width), np.uint8)
        rc = self.lib.bridct_codec8_decode(self.ptr, ptr(coeffs), height, width, width, ptr(out))
        if rc:
            raise RuntimeError('bridct_codec8_decode: status %d' % rc)
        return out

    def synchronize(self):
        return self.lib.bridct_codec8_synchronize(self.ptr)

    @property
    def stream(self):
        return self.lib.bridct_codec8_stream(self.ptr)

    def close(self):
        if self.ptr:
            self.lib.bridct_codec8_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
