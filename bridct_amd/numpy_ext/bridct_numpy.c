/* bridct_numpy -- the NumPy face of the MI355X BRiDCT library (no PyTorch, no CPU path).
 *
 * Drop-in contract (the reference module of the same name, described by its
 * python/README.md and checked by python/test_api.py and python/test_hardening.py):
 *
 *   Plan(height, width)                 both axes powers of two in 8..1024, else ValueError
 *     .forward(a, *, out=None)          orthonormal 2D DCT-II over the last two axes
 *     .inverse(a, *, out=None)          orthonormal 2D DCT-III
 *     .roundtrip(a, *, out=None)        forward, FP32 spectrum materialized, then inverse;
 *                                       equal bitwise to inverse(forward(a))
 *     .close()                          idempotent
 *     .shape / .route / .closed
 *   core_version                        the C library's BRIDCT_VERSION
 *
 *   a: an exact ndarray (no subclass, no conversion) of native float32, C order, float
 *      aligned, shape (height, width) or (B, height, width) with B >= 1; read-only is fine
 *      and it is never modified.  out: a writable ndarray of exactly a's shape under the
 *      same rules, sharing no byte with a; returned as is.  Without out every call
 *      returns a new owning array.
 *   Exceptions: TypeError (argument kinds, dtype, byte order, keywords), ValueError
 *   (shape, layout, alignment, writability, overlap), RuntimeError (object not open, in
 *   use by a nested call or another thread, re-initialization, a failed GPU call),
 *   MemoryError (allocation).
 *   Calls of 16384 elements or more run without the GIL.  An object is used by one call
 *   at a time: a second caller -- another thread, or Python code run from inside the call
 *   (an __index__ of a constructor argument, a NumPy allocator hook) -- gets RuntimeError
 *   and the first call is unaffected.
 *   The module refuses to load in any interpreter but the main one (ImportError), and
 *   refuses to build for a free-threaded CPython.
 *
 * Beyond the reference:
 *   RaggedPlan(shapes)                  a list of arrays of mixed supported shapes, one
 *                                       batch launch per direction (bridct_batch_*);
 *                                       methods take and return lists of 2D arrays.
 *   Codec8(quant)                       the 8x8 block codec step (bridct_codec8_*):
 *                                       encode uint8 (H, W) -> int16 (H/8, W/8, 8, 8)
 *                                       quantized coefficients, decode the reverse.
 *
 * Every object goes through one lifecycle (struct life below): NEW (allocated, not
 * initialized) -> SETUP (inside __init__) -> OPEN -> BUSY (inside a call) -> OPEN ...
 * -> SHUT (closed; never opens again).  A failed __init__ returns to NEW, so it may be
 * retried; close() of a NEW object is a no-op.
 */
#define PY_SSIZE_T_CLEAN
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <Python.h>
#include <numpy/arrayobject.h>

#include "bridct.h"
#include "bridct_gpu.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef Py_GIL_DISABLED
#error "bridct_numpy relies on the GIL to guard each object's state; build it for a CPython with the GIL."
#endif

/* Element count from which a call releases the GIL while the GPU works. */
#define NOGIL_ELEMENTS 16384

/* ---- lifecycle ---------------------------------------------------------------------------- */

enum { LIFE_NEW, LIFE_SETUP, LIFE_OPEN, LIFE_BUSY, LIFE_SHUT };

/* Every state change happens with the GIL held, so test-and-set is atomic across threads. */
typedef struct {
    int state;
    const char *kind;   /* type name for messages */
} life;

static int life_in_use(const life *l)
{
    PyErr_Format(PyExc_RuntimeError,
                 "This %s is in use by another call (another thread, or Python code running inside "
                 "the call); use one %s per thread.",
                 l->kind, l->kind);
    return -1;
}

/* Begins __init__: only a NEW object may be initialized. */
static int life_begin_setup(life *l)
{
    switch (l->state) {
    case LIFE_NEW:
        l->state = LIFE_SETUP;
        return 0;
    case LIFE_SETUP:
    case LIFE_BUSY:
        return life_in_use(l);
    default:
        PyErr_Format(PyExc_RuntimeError, "A %s is initialized once; create a new one instead.", l->kind);
        return -1;
    }
}

static void life_end_setup(life *l, int ok)
{
    l->state = ok ? LIFE_OPEN : LIFE_NEW;
}

/* Begins a transform call: only an OPEN object may run one. */
static int life_begin_call(life *l)
{
    switch (l->state) {
    case LIFE_OPEN:
        l->state = LIFE_BUSY;
        return 0;
    case LIFE_SETUP:
    case LIFE_BUSY:
        return life_in_use(l);
    case LIFE_NEW:
        PyErr_Format(PyExc_RuntimeError, "This %s was never initialized.", l->kind);
        return -1;
    default:
        PyErr_Format(PyExc_RuntimeError, "This %s is closed.", l->kind);
        return -1;
    }
}

static void life_end_call(life *l)
{
    l->state = LIFE_OPEN;
}

/* close(): 1 when the caller must release the native object now, 0 when there is nothing
 * to release, -1 (error set) while a call or an initialization is running. */
static int life_close(life *l)
{
    switch (l->state) {
    case LIFE_SETUP:
    case LIFE_BUSY:
        return life_in_use(l);
    case LIFE_OPEN:
        l->state = LIFE_SHUT;
        return 1;
    default:
        return 0;
    }
}

static int life_open(const life *l)
{
    return l->state == LIFE_OPEN || l->state == LIFE_BUSY;
}

/* ---- argument helpers ------------------------------------------------------------------------ */

/* method(input, *, out=None) in vectorcall form. */
static int parse_input_out(PyObject *const *args, Py_ssize_t nargs, PyObject *kwnames, PyObject **input,
                           PyObject **out)
{
    const Py_ssize_t nkw = kwnames ? PyTuple_GET_SIZE(kwnames) : 0;
    if (nargs != 1) {
        PyErr_SetString(PyExc_TypeError, "Expected exactly one positional argument; pass out= by keyword.");
        return -1;
    }
    *input = args[0];
    *out = Py_None;
    for (Py_ssize_t k = 0; k < nkw; k++) {
        if (PyUnicode_CompareWithASCIIString(PyTuple_GET_ITEM(kwnames, k), "out") != 0 || nkw > 1) {
            PyErr_SetString(PyExc_TypeError, "out is the only keyword argument.");
            return -1;
        }
        *out = args[nargs + k];
    }
    return 0;
}

/* Byte range [lo, hi) spanned by an array (strides may be of either sign). */
static void byte_span(PyArrayObject *a, uintptr_t *lo, uintptr_t *hi)
{
    intptr_t first = 0, last = 0;
    for (int d = 0; d < PyArray_NDIM(a); d++) {
        const intptr_t reach = (intptr_t)(PyArray_DIM(a, d) - 1) * (intptr_t)PyArray_STRIDE(a, d);
        if (reach < 0)
            first += reach;
        else
            last += reach;
    }
    *lo = (uintptr_t)PyArray_DATA(a) + first;
    *hi = (uintptr_t)PyArray_DATA(a) + last + (uintptr_t)PyArray_ITEMSIZE(a);
}

static int spans_meet(PyArrayObject *a, PyArrayObject *b)
{
    if (PyArray_SIZE(a) == 0 || PyArray_SIZE(b) == 0)
        return 0;
    uintptr_t a0, a1, b0, b1;
    byte_span(a, &a0, &a1);
    byte_span(b, &b0, &b1);
    return a0 < b1 && b0 < a1;
}

/* The kind checks shared by every array argument: exact ndarray of `type` in native byte
 * order.  Returns the array (borrowed) or NULL with TypeError. */
static PyArrayObject *exact_array(PyObject *o, int type, const char *what)
{
    if (!PyArray_CheckExact(o)) {
        PyErr_Format(PyExc_TypeError, "%s must be an exact numpy.ndarray (no subclass, no conversion).", what);
        return NULL;
    }
    PyArrayObject *a = (PyArrayObject *)o;
    if (PyArray_TYPE(a) != type || !PyArray_ISNOTSWAPPED(a)) {
        PyErr_Format(PyExc_TypeError, "%s must have native-endian dtype %s.", what,
                     type == NPY_FLOAT32 ? "float32" : type == NPY_INT16 ? "int16" : "uint8");
        return NULL;
    }
    return a;
}

/* Layout rules shared by float32 inputs and outputs: C order, aligned, writable if asked. */
static int dense_layout(PyArrayObject *a, int writable, const char *what)
{
    if (!PyArray_IS_C_CONTIGUOUS(a) || !PyArray_ISALIGNED(a)) {
        PyErr_Format(PyExc_ValueError, "%s must be C-contiguous and aligned.", what);
        return -1;
    }
    if (writable && !PyArray_ISWRITEABLE(a)) {
        PyErr_Format(PyExc_ValueError, "%s must be writable.", what);
        return -1;
    }
    return 0;
}

/* Validates an array argument of one object (`ctx`); `is_out` selects the output rules. */
typedef PyArrayObject *(*array_check)(PyObject *o, int is_out, const void *ctx);

/* The output of a call: a new array of shape `dims` and `type`, or the caller's `out` after
 * `check` and the shape and overlap rules.  Returns a new reference or NULL. */
static PyArrayObject *output_for(PyObject *out, PyArrayObject *in, int nd, const npy_intp *dims, int type,
                                 array_check check, const void *ctx)
{
    if (out == Py_None)
        return (PyArrayObject *)PyArray_SimpleNew(nd, (npy_intp *)dims, type);
    PyArrayObject *o = check(out, 1, ctx);
    if (!o)
        return NULL;
    if (PyArray_NDIM(o) != nd || memcmp(PyArray_DIMS(o), dims, (size_t)nd * sizeof(npy_intp)) != 0) {
        PyErr_SetString(PyExc_ValueError, "out must have exactly the shape of the result.");
        return NULL;
    }
    if (spans_meet(in, o)) {
        PyErr_SetString(PyExc_ValueError, "out must not share memory with the input.");
        return NULL;
    }
    Py_INCREF(o);
    return o;
}

static PyObject *gpu_status_error(const char *what, int status)
{
    PyErr_Format(PyExc_RuntimeError, "%s failed on the GPU with status %d.", what, status);
    return NULL;
}

/* ---- Plan ------------------------------------------------------------------------------------ */

typedef struct {
    PyObject_HEAD
    life life;
    bridct_plan *plan;
    int height, width;
} PlanObject;

static PyObject *Plan_new(PyTypeObject *type, PyObject *args, PyObject *kwargs)
{
    (void)args;
    (void)kwargs;
    PlanObject *self = (PlanObject *)type->tp_alloc(type, 0);
    if (self)
        self->life = (life){LIFE_NEW, "Plan"};
    return (PyObject *)self;
}

static int Plan_init(PlanObject *self, PyObject *args, PyObject *kwargs)
{
    static char *names[] = {"height", "width", NULL};
    int h = 0, w = 0;
    if (life_begin_setup(&self->life))
        return -1;
    /* Parsing can run Python code (__index__); the object is already in SETUP then. */
    int ok = PyArg_ParseTupleAndKeywords(args, kwargs, "ii:Plan", names, &h, &w);
    if (ok && !bridct_supported(h, w)) {
        PyErr_Format(PyExc_ValueError, "Unsupported shape (%d, %d): each axis must be a power of two "
                     "from 8 through 1024.", h, w);
        ok = 0;
    }
    if (ok && !bridct_gpu_available()) {
        PyErr_SetString(PyExc_RuntimeError, "No usable gfx950 GPU; this library has no CPU path.");
        ok = 0;
    }
    bridct_plan *plan = NULL;
    if (ok && !(plan = bridct_plan_create(h, w))) {
        PyErr_SetString(PyExc_MemoryError, "Could not allocate the GPU plan.");
        ok = 0;
    }
    if (ok) {
        self->plan = plan;
        self->height = h;
        self->width = w;
    }
    life_end_setup(&self->life, ok);
    return ok ? 0 : -1;
}

static void Plan_dealloc(PlanObject *self)
{
    bridct_plan_destroy(self->plan);
    Py_TYPE(self)->tp_free((PyObject *)self);
}

static PyObject *Plan_close(PlanObject *self, PyObject *unused)
{
    (void)unused;
    const int r = life_close(&self->life);
    if (r < 0)
        return NULL;
    if (r) {
        bridct_plan_destroy(self->plan);
        self->plan = NULL;
    }
    Py_RETURN_NONE;
}

/* A Plan's input or output: (h, w) or (B, h, w) float32, B >= 1. */
static PyArrayObject *plan_array(PyObject *o, int is_out, const void *ctx)
{
    const PlanObject *p = ctx;
    const char *what = is_out ? "out" : "The input";
    PyArrayObject *a = exact_array(o, NPY_FLOAT32, what);
    if (!a)
        return NULL;
    const int nd = PyArray_NDIM(a);
    if (nd < 2 || nd > 3 || PyArray_DIM(a, nd - 2) != p->height || PyArray_DIM(a, nd - 1) != p->width ||
        (nd == 3 && PyArray_DIM(a, 0) < 1)) {
        PyErr_Format(PyExc_ValueError, "%s must have shape (%d, %d) or (batch, %d, %d) with batch >= 1.", what,
                     p->height, p->width, p->height, p->width);
        return NULL;
    }
    return dense_layout(a, is_out, what) ? NULL : a;
}

static PyObject *plan_run(PlanObject *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kwnames, int mode)
{
    if (life_begin_call(&self->life))
        return NULL;
    PyObject *input, *out_arg, *result = NULL;
    PyArrayObject *in, *out = NULL;
    if (parse_input_out(args, nargs, kwnames, &input, &out_arg))
        goto done;
    if (!(in = plan_array(input, 0, self)))
        goto done;
    /* May allocate through NumPy's data allocator, which can run Python code. */
    out = output_for(out_arg, in, PyArray_NDIM(in), PyArray_DIMS(in), NPY_FLOAT32, plan_array, self);
    if (!out)
        goto done;
    const size_t batch = PyArray_NDIM(in) == 3 ? (size_t)PyArray_DIM(in, 0) : 1;
    const float *src = PyArray_DATA(in);
    float *dst = PyArray_DATA(out);
    int status;
    if (PyArray_SIZE(in) >= NOGIL_ELEMENTS) {
        Py_BEGIN_ALLOW_THREADS
        status = bridct_plan_apply_batch(self->plan, mode, src, dst, batch);
        Py_END_ALLOW_THREADS
    } else {
        status = bridct_plan_apply_batch(self->plan, mode, src, dst, batch);
    }
    if (status) {
        gpu_status_error("The transform", status);
        goto done;
    }
    result = (PyObject *)out;
    out = NULL;
done:
    Py_XDECREF(out);
    life_end_call(&self->life);
    return result;
}

static PyObject *Plan_forward(PlanObject *self, PyObject *const *a, Py_ssize_t n, PyObject *k)
{
    return plan_run(self, a, n, k, BRIDCT_FORWARD);
}

static PyObject *Plan_inverse(PlanObject *self, PyObject *const *a, Py_ssize_t n, PyObject *k)
{
    return plan_run(self, a, n, k, BRIDCT_INVERSE);
}

static PyObject *Plan_roundtrip(PlanObject *self, PyObject *const *a, Py_ssize_t n, PyObject *k)
{
    return plan_run(self, a, n, k, BRIDCT_GPU_FORWARD_THEN_INVERSE);
}

static PyObject *Plan_get_shape(PlanObject *self, void *unused)
{
    (void)unused;
    if (self->life.state == LIFE_NEW || self->life.state == LIFE_SETUP)
        Py_RETURN_NONE;
    return Py_BuildValue("(ii)", self->height, self->width);
}

static PyObject *Plan_get_route(PlanObject *self, void *unused)
{
    (void)unused;
    if (!life_open(&self->life))
        Py_RETURN_NONE;
    return PyUnicode_FromString(bridct_plan_route(self->plan));
}

static PyObject *Plan_get_closed(PlanObject *self, void *unused)
{
    (void)unused;
    return PyBool_FromLong(!life_open(&self->life));
}

#define VECTORCALL_METHOD(fn) (PyCFunction)(void (*)(void))(fn)

static PyMethodDef Plan_methods[] = {
    {"forward", VECTORCALL_METHOD(Plan_forward), METH_FASTCALL | METH_KEYWORDS,
     "forward(input, *, out=None) -> orthonormal 2D DCT-II of the last two axes (GPU)."},
    {"inverse", VECTORCALL_METHOD(Plan_inverse), METH_FASTCALL | METH_KEYWORDS,
     "inverse(input, *, out=None) -> orthonormal 2D DCT-III of the last two axes (GPU)."},
    {"roundtrip", VECTORCALL_METHOD(Plan_roundtrip), METH_FASTCALL | METH_KEYWORDS,
     "roundtrip(input, *, out=None) -> inverse(forward(input)), spectrum kept in float32."},
    {"close", VECTORCALL_METHOD(Plan_close), METH_NOARGS, "close() -> None; frees the GPU plan (idempotent)."},
    {NULL, NULL, 0, NULL}};

static PyGetSetDef Plan_getset[] = {
    {"shape", (getter)Plan_get_shape, NULL, "(height, width) of the plan; None before initialization.", NULL},
    {"route", (getter)Plan_get_route, NULL, "Shape category string of the plan; None unless open.", NULL},
    {"closed", (getter)Plan_get_closed, NULL, "True unless the plan is initialized and not yet closed.", NULL},
    {NULL, NULL, NULL, NULL, NULL}};

static PyTypeObject Plan_Type = {
    PyVarObject_HEAD_INIT(NULL, 0)
    .tp_name = "bridct_numpy.Plan",
    .tp_doc = "Plan(height, width): reusable orthonormal float32 2D DCT on the GPU; one per thread.",
    .tp_basicsize = sizeof(PlanObject),
    .tp_flags = Py_TPFLAGS_DEFAULT,
    .tp_new = Plan_new,
    .tp_init = (initproc)Plan_init,
    .tp_dealloc = (destructor)Plan_dealloc,
    .tp_methods = Plan_methods,
    .tp_getset = Plan_getset};

/* ---- RaggedPlan ------------------------------------------------------------------------------ */

typedef struct {
    PyObject_HEAD
    life life;
    bridct_batch *batch;
    PyObject *shapes;             /* tuple of (h, w) tuples, kept after close */
    int *heights, *widths;
    size_t count, elements;
    float *stage_in, *stage_out;  /* host staging buffers, made on the first call */
} RaggedObject;

static PyObject *Ragged_new(PyTypeObject *type, PyObject *args, PyObject *kwargs)
{
    (void)args;
    (void)kwargs;
    RaggedObject *self = (RaggedObject *)type->tp_alloc(type, 0);
    if (self)
        self->life = (life){LIFE_NEW, "RaggedPlan"};
    return (PyObject *)self;
}

static void ragged_free_native(RaggedObject *self)
{
    bridct_batch_destroy(self->batch);
    self->batch = NULL;
    free(self->stage_in);
    free(self->stage_out);
    self->stage_in = self->stage_out = NULL;
}

/* Reads the shape list into self->heights / widths / shapes (new objects on success). */
static int ragged_read_shapes(RaggedObject *self, PyObject *spec)
{
    PyObject *seq = PySequence_Fast(spec, "shapes must be a sequence of (height, width) pairs.");
    if (!seq)
        return -1;
    const Py_ssize_t n = PySequence_Fast_GET_SIZE(seq);
    PyObject *tuple = NULL;
    int *hs = NULL, *ws = NULL;
    if (n < 1) {
        PyErr_SetString(PyExc_ValueError, "A RaggedPlan needs at least one shape.");
        goto fail;
    }
    hs = PyMem_Malloc((size_t)n * sizeof(int));
    ws = PyMem_Malloc((size_t)n * sizeof(int));
    if (!hs || !ws || !(tuple = PyTuple_New(n))) {
        if (!PyErr_Occurred())
            PyErr_NoMemory();
        goto fail;
    }
    for (Py_ssize_t i = 0; i < n; i++) {
        PyObject *item = PySequence_Fast_GET_ITEM(seq, i);
        if (!PyTuple_Check(item) && !PyList_Check(item)) {
            PyErr_Format(PyExc_TypeError, "shapes[%zd] must be a (height, width) tuple or list.", i);
            goto fail;
        }
        if (PySequence_Size(item) != 2) {
            PyErr_Format(PyExc_TypeError, "shapes[%zd] must hold exactly two integers.", i);
            goto fail;
        }
        int dims[2];
        for (int d = 0; d < 2; d++) {
            PyObject *v = PySequence_GetItem(item, d);
            if (!v)
                goto fail;
            const long x = PyLong_AsLong(v);
            Py_DECREF(v);
            if (x == -1 && PyErr_Occurred())
                goto fail;
            dims[d] = x < INT32_MIN || x > INT32_MAX ? 0 : (int)x;
        }
        if (!bridct_supported(dims[0], dims[1])) {
            PyErr_Format(PyExc_ValueError, "shapes[%zd]: each axis must be a power of two from 8 through 1024.", i);
            goto fail;
        }
        hs[i] = dims[0];
        ws[i] = dims[1];
        PyObject *pair = Py_BuildValue("(ii)", dims[0], dims[1]);
        if (!pair)
            goto fail;
        PyTuple_SET_ITEM(tuple, i, pair);
    }
    Py_DECREF(seq);
    self->heights = hs;
    self->widths = ws;
    self->shapes = tuple;
    self->count = (size_t)n;
    return 0;
fail:
    PyMem_Free(hs);
    PyMem_Free(ws);
    Py_XDECREF(tuple);
    Py_DECREF(seq);
    return -1;
}

static int Ragged_init(RaggedObject *self, PyObject *args, PyObject *kwargs)
{
    static char *names[] = {"shapes", NULL};
    PyObject *spec = NULL;
    if (life_begin_setup(&self->life))
        return -1;
    int ok = PyArg_ParseTupleAndKeywords(args, kwargs, "O:RaggedPlan", names, &spec) &&
             ragged_read_shapes(self, spec) == 0;
    if (ok && !bridct_gpu_available()) {
        PyErr_SetString(PyExc_RuntimeError, "No usable gfx950 GPU; this library has no CPU path.");
        ok = 0;
    }
    if (ok && !(self->batch = bridct_batch_create(self->count, self->heights, self->widths))) {
        PyErr_SetString(PyExc_MemoryError, "Could not allocate the GPU batch.");
        ok = 0;
    }
    if (ok) {
        self->elements = bridct_batch_elements(self->batch);
    } else {
        PyMem_Free(self->heights);
        PyMem_Free(self->widths);
        self->heights = self->widths = NULL;
        Py_CLEAR(self->shapes);
        self->count = 0;
    }
    life_end_setup(&self->life, ok);
    return ok ? 0 : -1;
}

static void Ragged_dealloc(RaggedObject *self)
{
    ragged_free_native(self);
    PyMem_Free(self->heights);
    PyMem_Free(self->widths);
    Py_XDECREF(self->shapes);
    Py_TYPE(self)->tp_free((PyObject *)self);
}

static PyObject *Ragged_close(RaggedObject *self, PyObject *unused)
{
    (void)unused;
    const int r = life_close(&self->life);
    if (r < 0)
        return NULL;
    if (r)
        ragged_free_native(self);
    Py_RETURN_NONE;
}

/* Array i of a list argument: a 2D float32 array of planned shape i. */
static PyArrayObject *ragged_member(PyObject *o, int h, int w, int is_out)
{
    const char *what = is_out ? "Each out array" : "Each input array";
    PyArrayObject *a = exact_array(o, NPY_FLOAT32, what);
    if (!a)
        return NULL;
    if (PyArray_NDIM(a) != 2 || PyArray_DIM(a, 0) != h || PyArray_DIM(a, 1) != w) {
        PyErr_Format(PyExc_ValueError, "%s must have its planned shape (%d, %d).", what, h, w);
        return NULL;
    }
    return dense_layout(a, is_out, what) ? NULL : a;
}

static PyObject *list_of(PyObject *o, const char *what)
{
    if (!PyList_Check(o) && !PyTuple_Check(o)) {
        PyErr_Format(PyExc_TypeError, "%s must be a list or tuple of arrays.", what);
        return NULL;
    }
    return PySequence_Fast(o, what);
}

static PyObject *ragged_run(RaggedObject *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kwnames,
                            int mode)
{
    if (life_begin_call(&self->life))
        return NULL;
    PyObject *input, *out_arg, *ins = NULL, *outs = NULL, *result = NULL;
    if (parse_input_out(args, nargs, kwnames, &input, &out_arg) || !(ins = list_of(input, "The input")))
        goto done;
    if ((size_t)PySequence_Fast_GET_SIZE(ins) != self->count) {
        PyErr_Format(PyExc_ValueError, "Expected %zu arrays, one per planned shape.", self->count);
        goto done;
    }
    if (out_arg != Py_None) {
        if (!(outs = list_of(out_arg, "out")))
            goto done;
        if ((size_t)PySequence_Fast_GET_SIZE(outs) != self->count) {
            PyErr_Format(PyExc_ValueError, "out must hold %zu arrays, one per planned shape.", self->count);
            goto done;
        }
    }
    for (size_t i = 0; i < self->count; i++) {
        PyArrayObject *a = ragged_member(PySequence_Fast_GET_ITEM(ins, (Py_ssize_t)i), self->heights[i],
                                         self->widths[i], 0);
        if (!a)
            goto done;
        if (outs) {
            PyArrayObject *o = ragged_member(PySequence_Fast_GET_ITEM(outs, (Py_ssize_t)i), self->heights[i],
                                             self->widths[i], 1);
            if (!o)
                goto done;
            if (spans_meet(a, o)) {
                PyErr_Format(PyExc_ValueError, "out[%zu] shares memory with its input.", i);
                goto done;
            }
        }
    }
    if (!self->stage_in) {
        self->stage_in = malloc(self->elements * sizeof(float));
        self->stage_out = malloc(self->elements * sizeof(float));
        if (!self->stage_in || !self->stage_out) {
            free(self->stage_in);
            free(self->stage_out);
            self->stage_in = self->stage_out = NULL;
            PyErr_NoMemory();
            goto done;
        }
    }
    for (size_t i = 0; i < self->count; i++) {
        PyArrayObject *a = (PyArrayObject *)PySequence_Fast_GET_ITEM(ins, (Py_ssize_t)i);
        memcpy(self->stage_in + bridct_batch_offset(self->batch, i), PyArray_DATA(a), (size_t)PyArray_NBYTES(a));
    }
    int status;
    if (self->elements >= NOGIL_ELEMENTS) {
        Py_BEGIN_ALLOW_THREADS
        status = bridct_batch_apply(self->batch, mode, self->stage_in, self->stage_out);
        Py_END_ALLOW_THREADS
    } else {
        status = bridct_batch_apply(self->batch, mode, self->stage_in, self->stage_out);
    }
    if (status) {
        gpu_status_error("The batch transform", status);
        goto done;
    }
    PyObject *list = PyList_New((Py_ssize_t)self->count);
    if (!list)
        goto done;
    for (size_t i = 0; i < self->count; i++) {
        PyArrayObject *o;
        if (outs) {
            o = (PyArrayObject *)PySequence_Fast_GET_ITEM(outs, (Py_ssize_t)i);
            Py_INCREF(o);
        } else {
            const npy_intp dims[2] = {self->heights[i], self->widths[i]};
            if (!(o = (PyArrayObject *)PyArray_SimpleNew(2, (npy_intp *)dims, NPY_FLOAT32))) {
                Py_DECREF(list);
                goto done;
            }
        }
        memcpy(PyArray_DATA(o), self->stage_out + bridct_batch_offset(self->batch, i), (size_t)PyArray_NBYTES(o));
        PyList_SET_ITEM(list, (Py_ssize_t)i, (PyObject *)o);
    }
    result = list;
done:
    Py_XDECREF(ins);
    Py_XDECREF(outs);
    life_end_call(&self->life);
    return result;
}

static PyObject *Ragged_forward(RaggedObject *self, PyObject *const *a, Py_ssize_t n, PyObject *k)
{
    return ragged_run(self, a, n, k, BRIDCT_FORWARD);
}

static PyObject *Ragged_inverse(RaggedObject *self, PyObject *const *a, Py_ssize_t n, PyObject *k)
{
    return ragged_run(self, a, n, k, BRIDCT_INVERSE);
}

static PyObject *Ragged_roundtrip(RaggedObject *self, PyObject *const *a, Py_ssize_t n, PyObject *k)
{
    return ragged_run(self, a, n, k, BRIDCT_GPU_FORWARD_THEN_INVERSE);
}

static PyObject *Ragged_get_shapes(RaggedObject *self, void *unused)
{
    (void)unused;
    if (!self->shapes)
        Py_RETURN_NONE;
    Py_INCREF(self->shapes);
    return self->shapes;
}

static PyObject *Ragged_get_closed(RaggedObject *self, void *unused)
{
    (void)unused;
    return PyBool_FromLong(!life_open(&self->life));
}

static PyMethodDef Ragged_methods[] = {
    {"forward", VECTORCALL_METHOD(Ragged_forward), METH_FASTCALL | METH_KEYWORDS,
     "forward(arrays, *, out=None) -> list: orthonormal 2D DCT-II of every array in one batch call."},
    {"inverse", VECTORCALL_METHOD(Ragged_inverse), METH_FASTCALL | METH_KEYWORDS,
     "inverse(arrays, *, out=None) -> list: orthonormal 2D DCT-III of every array in one batch call."},
    {"roundtrip", VECTORCALL_METHOD(Ragged_roundtrip), METH_FASTCALL | METH_KEYWORDS,
     "roundtrip(arrays, *, out=None) -> list: inverse(forward(a)) per array, spectrum in float32."},
    {"close", VECTORCALL_METHOD(Ragged_close), METH_NOARGS, "close() -> None; frees the GPU batch (idempotent)."},
    {NULL, NULL, 0, NULL}};

static PyGetSetDef Ragged_getset[] = {
    {"shapes", (getter)Ragged_get_shapes, NULL, "((height, width), ...) of the plan; None before initialization.",
     NULL},
    {"closed", (getter)Ragged_get_closed, NULL, "True unless the plan is initialized and not yet closed.", NULL},
    {NULL, NULL, NULL, NULL, NULL}};

static PyTypeObject Ragged_Type = {
    PyVarObject_HEAD_INIT(NULL, 0)
    .tp_name = "bridct_numpy.RaggedPlan",
    .tp_doc = "RaggedPlan(shapes): orthonormal float32 2D DCT of a list of arrays of mixed shapes on the GPU.",
    .tp_basicsize = sizeof(RaggedObject),
    .tp_flags = Py_TPFLAGS_DEFAULT,
    .tp_new = Ragged_new,
    .tp_init = (initproc)Ragged_init,
    .tp_dealloc = (destructor)Ragged_dealloc,
    .tp_methods = Ragged_methods,
    .tp_getset = Ragged_getset};

/* ---- Codec8 ---------------------------------------------------------------------------------- */

typedef struct {
    PyObject_HEAD
    life life;
    bridct_codec8 *codec;
    float quant[64];
} CodecObject;

static PyObject *Codec_new(PyTypeObject *type, PyObject *args, PyObject *kwargs)
{
    (void)args;
    (void)kwargs;
    CodecObject *self = (CodecObject *)type->tp_alloc(type, 0);
    if (self)
        self->life = (life){LIFE_NEW, "Codec8"};
    return (PyObject *)self;
}

/* 64 finite positive float32 steps, shape (64,) or (8, 8), C order. */
static int codec_read_steps(PyObject *o, float steps[64])
{
    PyArrayObject *q = exact_array(o, NPY_FLOAT32, "quant");
    if (!q)
        return -1;
    if (PyArray_SIZE(q) != 64 || !(PyArray_NDIM(q) == 1 || (PyArray_NDIM(q) == 2 && PyArray_DIM(q, 0) == 8)) ||
        dense_layout(q, 0, "quant")) {
        if (!PyErr_Occurred())
            PyErr_SetString(PyExc_ValueError, "quant must hold 64 steps, shape (64,) or (8, 8).");
        return -1;
    }
    memcpy(steps, PyArray_DATA(q), 64 * sizeof(float));
    for (int i = 0; i < 64; i++)
        if (!isfinite(steps[i]) || !(steps[i] > 0.0f)) {
            PyErr_SetString(PyExc_ValueError, "Every quantizer step must be finite and positive.");
            return -1;
        }
    return 0;
}

static int Codec_init(CodecObject *self, PyObject *args, PyObject *kwargs)
{
    static char *names[] = {"quant", NULL};
    PyObject *q = NULL;
    float steps[64];
    if (life_begin_setup(&self->life))
        return -1;
    int ok = PyArg_ParseTupleAndKeywords(args, kwargs, "O:Codec8", names, &q) && codec_read_steps(q, steps) == 0;
    if (ok && !bridct_gpu_available()) {
        PyErr_SetString(PyExc_RuntimeError, "No usable gfx950 GPU; this library has no CPU path.");
        ok = 0;
    }
    if (ok && !(self->codec = bridct_codec8_create(steps))) {
        PyErr_SetString(PyExc_MemoryError, "Could not allocate the GPU codec.");
        ok = 0;
    }
    if (ok)
        memcpy(self->quant, steps, sizeof(steps));
    life_end_setup(&self->life, ok);
    return ok ? 0 : -1;
}

static void Codec_dealloc(CodecObject *self)
{
    bridct_codec8_destroy(self->codec);
    Py_TYPE(self)->tp_free((PyObject *)self);
}

static PyObject *Codec_close(CodecObject *self, PyObject *unused)
{
    (void)unused;
    const int r = life_close(&self->life);
    if (r < 0)
        return NULL;
    if (r) {
        bridct_codec8_destroy(self->codec);
        self->codec = NULL;
    }
    Py_RETURN_NONE;
}

/* A uint8 (H, W) raster, H and W positive multiples of 8, unit column stride and rows at
 * least W bytes apart (the row stride is passed to the GPU). */
static PyArrayObject *codec_image(PyObject *o, int is_out, const void *ctx)
{
    (void)ctx;
    const char *what = is_out ? "out" : "The image";
    PyArrayObject *a = exact_array(o, NPY_UINT8, what);
    if (!a)
        return NULL;
    if (PyArray_NDIM(a) != 2 || PyArray_DIM(a, 0) < 8 || PyArray_DIM(a, 1) < 8 || PyArray_DIM(a, 0) % 8 ||
        PyArray_DIM(a, 1) % 8) {
        PyErr_Format(PyExc_ValueError, "%s must be (height, width) with both multiples of 8.", what);
        return NULL;
    }
    if (PyArray_STRIDE(a, 1) != 1 || PyArray_STRIDE(a, 0) < PyArray_DIM(a, 1)) {
        PyErr_Format(PyExc_ValueError, "%s must have contiguous rows in increasing order.", what);
        return NULL;
    }
    if (is_out && !PyArray_ISWRITEABLE(a)) {
        PyErr_Format(PyExc_ValueError, "%s must be writable.", what);
        return NULL;
    }
    return a;
}

/* int16 (H/8, W/8, 8, 8) coefficients, C order. */
static PyArrayObject *codec_coeffs(PyObject *o, int is_out, const void *ctx)
{
    (void)ctx;
    const char *what = is_out ? "out" : "The coefficients";
    PyArrayObject *a = exact_array(o, NPY_INT16, what);
    if (!a)
        return NULL;
    if (PyArray_NDIM(a) != 4 || PyArray_DIM(a, 0) < 1 || PyArray_DIM(a, 1) < 1 || PyArray_DIM(a, 2) != 8 ||
        PyArray_DIM(a, 3) != 8) {
        PyErr_Format(PyExc_ValueError, "%s must have shape (height/8, width/8, 8, 8).", what);
        return NULL;
    }
    return dense_layout(a, is_out, what) ? NULL : a;
}

static PyObject *codec_run(CodecObject *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kwnames,
                           int decode)
{
    if (life_begin_call(&self->life))
        return NULL;
    PyObject *input, *out_arg, *result = NULL;
    PyArrayObject *in, *out = NULL;
    if (parse_input_out(args, nargs, kwnames, &input, &out_arg))
        goto done;
    if (!(in = decode ? codec_coeffs(input, 0, NULL) : codec_image(input, 0, NULL)))
        goto done;
    const npy_intp h = decode ? 8 * PyArray_DIM(in, 0) : PyArray_DIM(in, 0);
    const npy_intp w = decode ? 8 * PyArray_DIM(in, 1) : PyArray_DIM(in, 1);
    const npy_intp img_dims[2] = {h, w}, coeff_dims[4] = {h / 8, w / 8, 8, 8};
    out = decode ? output_for(out_arg, in, 2, img_dims, NPY_UINT8, codec_image, NULL)
                 : output_for(out_arg, in, 4, coeff_dims, NPY_INT16, codec_coeffs, NULL);
    if (!out)
        goto done;
    const size_t stride = (size_t)PyArray_STRIDE(decode ? out : in, 0);
    int status;
    Py_BEGIN_ALLOW_THREADS
    status = decode ? bridct_codec8_decode(self->codec, PyArray_DATA(in), (size_t)h, (size_t)w, stride,
                                           PyArray_DATA(out))
                    : bridct_codec8_encode(self->codec, PyArray_DATA(in), (size_t)h, (size_t)w, stride,
                                           PyArray_DATA(out));
    Py_END_ALLOW_THREADS
    if (status) {
        gpu_status_error(decode ? "Decoding" : "Encoding", status);
        goto done;
    }
    result = (PyObject *)out;
    out = NULL;
done:
    Py_XDECREF(out);
    life_end_call(&self->life);
    return result;
}

static PyObject *Codec_encode(CodecObject *self, PyObject *const *a, Py_ssize_t n, PyObject *k)
{
    return codec_run(self, a, n, k, 0);
}

static PyObject *Codec_decode(CodecObject *self, PyObject *const *a, Py_ssize_t n, PyObject *k)
{
    return codec_run(self, a, n, k, 1);
}

static PyObject *Codec_get_quant(CodecObject *self, void *unused)
{
    (void)unused;
    if (self->life.state == LIFE_NEW || self->life.state == LIFE_SETUP)
        Py_RETURN_NONE;
    const npy_intp dims[2] = {8, 8};
    PyObject *q = PyArray_SimpleNew(2, (npy_intp *)dims, NPY_FLOAT32);
    if (q)
        memcpy(PyArray_DATA((PyArrayObject *)q), self->quant, sizeof(self->quant));
    return q;
}

static PyObject *Codec_get_closed(CodecObject *self, void *unused)
{
    (void)unused;
    return PyBool_FromLong(!life_open(&self->life));
}

static PyMethodDef Codec_methods[] = {
    {"encode", VECTORCALL_METHOD(Codec_encode), METH_FASTCALL | METH_KEYWORDS,
     "encode(image, *, out=None) -> int16 (H/8, W/8, 8, 8) quantized 8x8 DCT coefficients of a uint8 image."},
    {"decode", VECTORCALL_METHOD(Codec_decode), METH_FASTCALL | METH_KEYWORDS,
     "decode(coeffs, *, out=None) -> uint8 (H, W) image from quantized coefficients."},
    {"close", VECTORCALL_METHOD(Codec_close), METH_NOARGS, "close() -> None; frees the GPU codec (idempotent)."},
    {NULL, NULL, 0, NULL}};

static PyGetSetDef Codec_getset[] = {
    {"quant", (getter)Codec_get_quant, NULL, "The 64 quantizer steps as (8, 8) float32; None before initialization.",
     NULL},
    {"closed", (getter)Codec_get_closed, NULL, "True unless the codec is initialized and not yet closed.", NULL},
    {NULL, NULL, NULL, NULL, NULL}};

static PyTypeObject Codec_Type = {
    PyVarObject_HEAD_INIT(NULL, 0)
    .tp_name = "bridct_numpy.Codec8",
    .tp_doc = "Codec8(quant): 8x8 block codec step on the GPU (level shift, 8x8 DCT, quantization, and back).",
    .tp_basicsize = sizeof(CodecObject),
    .tp_flags = Py_TPFLAGS_DEFAULT,
    .tp_new = Codec_new,
    .tp_init = (initproc)Codec_init,
    .tp_dealloc = (destructor)Codec_dealloc,
    .tp_methods = Codec_methods,
    .tp_getset = Codec_getset};

/* ---- module ---------------------------------------------------------------------------------- */

static int bridct_numpy_exec(PyObject *module)
{
    /* The types are static and each object's state is guarded by one GIL: another
     * interpreter would share the types without sharing that guard. */
    if (PyInterpreterState_Get() != PyInterpreterState_Main()) {
        PyErr_SetString(PyExc_ImportError,
                        "bridct_numpy loads in the main interpreter only; subinterpreters are not supported.");
        return -1;
    }
    if (_import_array() < 0)
        return -1;
    PyTypeObject *types[] = {&Plan_Type, &Ragged_Type, &Codec_Type};
    for (size_t i = 0; i < sizeof(types) / sizeof(types[0]); i++)
        if (PyModule_AddType(module, types[i]) < 0)
            return -1;
    return PyModule_AddStringConstant(module, "core_version", BRIDCT_VERSION);
}

static PyModuleDef_Slot bridct_numpy_slots[] = {{Py_mod_exec, (void *)bridct_numpy_exec}, {0, NULL}};

static struct PyModuleDef bridct_numpy_module = {
    PyModuleDef_HEAD_INIT,
    .m_name = "bridct_numpy",
    .m_doc = "NumPy interface to the MI355X BRiDCT library: Plan, RaggedPlan, Codec8, core_version.",
    .m_size = 0,
    .m_slots = bridct_numpy_slots};

PyMODINIT_FUNC PyInit_bridct_numpy(void)
{
    return PyModuleDef_Init(&bridct_numpy_module);
}
