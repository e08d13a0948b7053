/* bridct_numpy: CPython/NumPy interface to the MI355X BRiDCT plan library.
 *
 * Same module, type and contract as the reference extension (python/bridct_numpy.c):
 *   Plan(height, width); forward / inverse / roundtrip(input, *, out=None); close();
 *   getters shape, route, closed; module constant core_version.
 * Inputs are exact float32 ndarrays of shape (h, w) or (B, h, w), native byte order,
 * C-contiguous and aligned; `out` must also be writable, of the same ndim and size, and
 * must not overlap the input.  TypeError for type/keyword problems, ValueError for
 * shape/layout/overlap, RuntimeError for closed/busy/re-initialized plans and failed
 * calls, MemoryError for allocation.  roundtrip materializes the normalized FP32
 * spectrum (python/bridct_numpy.c:114-133), so it equals inverse(forward(x)) bitwise.
 * The GIL is released for inputs of 16384 elements or more, and a per-plan busy flag
 * rejects concurrent use (python/bridct_numpy.c:142-147, 191-198).
 *
 * The whole batch goes to the GPU in one call (bridct_plan_apply_batch), not one call
 * per image as in the reference loop (python/bridct_numpy.c:114-133).
 *
 * RaggedPlan(shapes) has no counterpart in the reference: a list of arrays of mixed
 * supported shapes goes through one persistent GPU launch per direction (bridct_batch_*).
 * forward / inverse / roundtrip(arrays, *, out=None) take a list or tuple of float32
 * (h_i, w_i) arrays, validated like Plan inputs, and return a list (or fill `out`, a list
 * of writable arrays of the same shapes, each disjoint from its input).
 *
 * Codec8(quant) has no counterpart in the reference either: the 8x8 block codec step
 * (bridct_codec8_*).  quant: float32 ndarray of 64 finite positive steps (shape (64,) or
 * (8, 8), row-major u * 8 + v).  encode(image, *, out=None): uint8 (H, W) image, H and W
 * multiples of 8, rows may be strided -> int16 (H/8, W/8, 8, 8) quantized coefficients;
 * decode(coeffs, *, out=None): the reverse (out may have strided rows).  Same exception
 * split, busy flag and GIL rule as Plan.
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
#error "bridct_numpy needs a CPython build with the GIL."
#endif

typedef struct {
    PyObject_HEAD
    bridct_plan *plan;
    size_t elements;
    int height, width, initialized, busy;
} Plan;

static int Plan_init(Plan *self, PyObject *args, PyObject *kwargs)
{
    static char *keywords[] = {"height", "width", NULL};
    int height, width;
    if (self->initialized) {
        PyErr_SetString(PyExc_RuntimeError, "A Plan cannot be reinitialized.");
        return -1;
    }
    if (self->busy) {
        PyErr_SetString(PyExc_RuntimeError, "This Plan is already initializing or executing.");
        return -1;
    }
    self->busy = 1;
    if (!PyArg_ParseTupleAndKeywords(args, kwargs, "ii:Plan", keywords, &height, &width))
        goto fail;
    if (!bridct_supported(height, width)) {
        PyErr_SetString(PyExc_ValueError, "Both dimensions must be powers of two from 8 through 1024.");
        goto fail;
    }
    if (!bridct_gpu_available()) {
        PyErr_SetString(PyExc_RuntimeError, "No usable gfx950 GPU: the BRiDCT GPU plan cannot be created.");
        goto fail;
    }
    bridct_plan *plan = bridct_plan_create(height, width);
    if (!plan) {
        PyErr_NoMemory();
        goto fail;
    }
    self->plan = plan;
    self->elements = (size_t)height * (size_t)width;
    self->height = height;
    self->width = width;
    self->initialized = 1;
    self->busy = 0;
    return 0;
fail:
    self->busy = 0;
    return -1;
}

static void Plan_dealloc(Plan *self)
{
    bridct_plan_destroy(self->plan);
    Py_TYPE(self)->tp_free((PyObject *)self);
}

static PyObject *Plan_close(Plan *self, PyObject *unused)
{
    (void)unused;
    if (self->busy) {
        PyErr_SetString(PyExc_RuntimeError,
                        "This Plan is already initializing or executing, possibly in another thread.");
        return NULL;
    }
    bridct_plan_destroy(self->plan);
    self->plan = NULL;
    Py_RETURN_NONE;
}

static PyArrayObject *validated(PyObject *object, const Plan *self, int writable)
{
    if (!PyArray_CheckExact(object)) {
        PyErr_SetString(PyExc_TypeError, "Expected an exact NumPy ndarray; no implicit conversion is performed.");
        return NULL;
    }
    PyArrayObject *a = (PyArrayObject *)object;
    if (PyArray_TYPE(a) != NPY_FLOAT32 || !PyArray_ISNOTSWAPPED(a)) {
        PyErr_SetString(PyExc_TypeError, "Expected native-endian float32.");
        return NULL;
    }
    const int nd = PyArray_NDIM(a);
    if ((nd != 2 && nd != 3) || PyArray_DIM(a, nd - 2) != self->height ||
        PyArray_DIM(a, nd - 1) != self->width || (nd == 3 && PyArray_DIM(a, 0) < 1)) {
        PyErr_SetString(PyExc_ValueError, "Expected (height, width) or (batch, height, width), with batch >= 1.");
        return NULL;
    }
    if (!PyArray_IS_C_CONTIGUOUS(a) || !PyArray_ISALIGNED(a)) {
        PyErr_SetString(PyExc_ValueError, "Expected C-contiguous, float-aligned storage.");
        return NULL;
    }
    if (writable && !PyArray_ISWRITEABLE(a)) {
        PyErr_SetString(PyExc_ValueError, "The output array must be writable.");
        return NULL;
    }
    return a;
}

static PyObject *execute(Plan *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kwnames,
                         int mode)
{
    if (!self->plan) {
        PyErr_SetString(PyExc_RuntimeError, "This Plan is closed or uninitialized.");
        return NULL;
    }
    if (self->busy) {
        PyErr_SetString(PyExc_RuntimeError, "This Plan is executing in another thread; use one Plan per thread.");
        return NULL;
    }
    /* Array allocation can run Python code; the flag covers setup as well as compute. */
    self->busy = 1;
    PyArrayObject *out = NULL;
    const Py_ssize_t nkw = kwnames ? PyTuple_GET_SIZE(kwnames) : 0;
    if (nargs != 1 || nkw > 1) {
        PyErr_SetString(PyExc_TypeError, "Use method(input, *, out=None).");
        goto fail;
    }
    PyObject *out_obj = Py_None;
    if (nkw) {
        if (PyUnicode_CompareWithASCIIString(PyTuple_GET_ITEM(kwnames, 0), "out") != 0) {
            PyErr_SetString(PyExc_TypeError, "The only accepted keyword is out.");
            goto fail;
        }
        out_obj = args[1];
    }
    PyArrayObject *in = validated(args[0], self, 0);
    if (!in)
        goto fail;
    if (out_obj == Py_None) {
        out = (PyArrayObject *)PyArray_SimpleNew(PyArray_NDIM(in), PyArray_DIMS(in), NPY_FLOAT32);
        if (!out)
            goto fail;
    } else {
        PyArrayObject *given = validated(out_obj, self, 1);
        if (!given)
            goto fail;
        if (PyArray_NDIM(in) != PyArray_NDIM(given) || PyArray_SIZE(in) != PyArray_SIZE(given)) {
            PyErr_SetString(PyExc_ValueError, "Input and output shapes must match exactly.");
            goto fail;
        }
        const uintptr_t a = (uintptr_t)PyArray_DATA(in), b = (uintptr_t)PyArray_DATA(given);
        if ((a <= b ? b - a : a - b) < (size_t)PyArray_NBYTES(in)) {
            PyErr_SetString(PyExc_ValueError, "Input and output must not overlap.");
            goto fail;
        }
        out = given;
        Py_INCREF(out);
    }
    const float *x = PyArray_DATA(in);
    float *y = PyArray_DATA(out);
    const size_t batch = PyArray_NDIM(in) == 3 ? (size_t)PyArray_DIM(in, 0) : 1;
    int status;
    if (PyArray_SIZE(in) >= 16384) {
        Py_BEGIN_ALLOW_THREADS
        status = bridct_plan_apply_batch(self->plan, mode, x, y, batch);
        Py_END_ALLOW_THREADS
    } else {
        status = bridct_plan_apply_batch(self->plan, mode, x, y, batch);
    }
    if (status) {
        PyErr_Format(PyExc_RuntimeError, "BRiDCT failed with status %d.", status);
        goto fail;
    }
    self->busy = 0;
    return (PyObject *)out;
fail:
    self->busy = 0;
    Py_XDECREF(out);
    return NULL;
}

static PyObject *Plan_forward(Plan *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kw)
{
    return execute(self, args, nargs, kw, BRIDCT_FORWARD);
}

static PyObject *Plan_inverse(Plan *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kw)
{
    return execute(self, args, nargs, kw, BRIDCT_INVERSE);
}

static PyObject *Plan_roundtrip(Plan *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kw)
{
    return execute(self, args, nargs, kw, BRIDCT_GPU_FORWARD_THEN_INVERSE);
}

static PyObject *Plan_shape(Plan *self, void *unused)
{
    (void)unused;
    return Py_BuildValue("(ii)", self->height, self->width);
}

static PyObject *Plan_route(Plan *self, void *unused)
{
    (void)unused;
    if (!self->plan)
        Py_RETURN_NONE;
    return PyUnicode_FromString(bridct_plan_route(self->plan));
}

static PyObject *Plan_closed(Plan *self, void *unused)
{
    (void)unused;
    return PyBool_FromLong(self->plan == NULL);
}

#define AS_METHOD(fn) (PyCFunction)(void (*)(void))(fn)
static PyMethodDef Plan_methods[] = {
    {"forward", AS_METHOD(Plan_forward), METH_FASTCALL | METH_KEYWORDS,
     "forward(input, *, out=None): orthonormal 2D DCT-II over the last two axes, on the GPU."},
    {"inverse", AS_METHOD(Plan_inverse), METH_FASTCALL | METH_KEYWORDS,
     "inverse(input, *, out=None): orthonormal 2D DCT-III over the last two axes, on the GPU."},
    {"roundtrip", AS_METHOD(Plan_roundtrip), METH_FASTCALL | METH_KEYWORDS,
     "roundtrip(input, *, out=None): forward then inverse with a materialized spectrum."},
    {"close", AS_METHOD(Plan_close), METH_NOARGS, "Release the GPU plan and its buffers."},
    {NULL, NULL, 0, NULL}};

static PyGetSetDef Plan_getset[] = {
    {"shape", (getter)Plan_shape, NULL, "Planned (height, width).", NULL},
    {"route", (getter)Plan_route, NULL, "Route category, or None when closed.", NULL},
    {"closed", (getter)Plan_closed, NULL, "Whether the plan has been closed.", NULL},
    {NULL, NULL, NULL, NULL, NULL}};

static PyTypeObject PlanType = {
    PyVarObject_HEAD_INIT(NULL, 0)
    .tp_name = "bridct_numpy.Plan",
    .tp_basicsize = sizeof(Plan),
    .tp_dealloc = (destructor)Plan_dealloc,
    .tp_flags = Py_TPFLAGS_DEFAULT,
    .tp_doc = "Reusable orthonormal float32 2D DCT plan on the GPU; one plan per concurrent worker.",
    .tp_methods = Plan_methods,
    .tp_getset = Plan_getset,
    .tp_init = (initproc)Plan_init,
    .tp_new = PyType_GenericNew};

/* ---- RaggedPlan ------------------------------------------------------------------ */

typedef struct {
    PyObject_HEAD
    bridct_batch *batch;
    PyObject *shapes;           /* tuple of (h, w) tuples */
    size_t count, elements;
    float *stage_in, *stage_out;   /* host staging, allocated on first use */
    int initialized, busy;
} RaggedPlan;

static void ragged_release(RaggedPlan *self)
{
    bridct_batch_destroy(self->batch);
    self->batch = NULL;
    free(self->stage_in);
    free(self->stage_out);
    self->stage_in = self->stage_out = NULL;
}

static int RaggedPlan_init(RaggedPlan *self, PyObject *args, PyObject *kwargs)
{
    static char *keywords[] = {"shapes", NULL};
    PyObject *seq = NULL, *shapes = NULL;
    int *hs = NULL, *ws = NULL;
    if (self->initialized) {
        PyErr_SetString(PyExc_RuntimeError, "A RaggedPlan cannot be reinitialized.");
        return -1;
    }
    if (self->busy) {
        PyErr_SetString(PyExc_RuntimeError, "This RaggedPlan is already initializing or executing.");
        return -1;
    }
    self->busy = 1;
    if (!PyArg_ParseTupleAndKeywords(args, kwargs, "O:RaggedPlan", keywords, &seq))
        goto fail;
    seq = PySequence_Fast(seq, "shapes must be a sequence of (height, width) pairs.");
    if (!seq)
        goto fail;
    const Py_ssize_t n = PySequence_Fast_GET_SIZE(seq);
    if (n < 1) {
        PyErr_SetString(PyExc_ValueError, "A RaggedPlan needs at least one shape.");
        goto fail;
    }
    hs = malloc((size_t)n * sizeof(int));
    ws = malloc((size_t)n * sizeof(int));
    shapes = PyTuple_New(n);
    if (!hs || !ws || !shapes) {
        PyErr_NoMemory();
        goto fail;
    }
    for (Py_ssize_t i = 0; i < n; i++) {
        int h, w;
        PyObject *item = PySequence_Fast_GET_ITEM(seq, i);
        if (!PyTuple_Check(item) && !PyList_Check(item)) {
            PyErr_SetString(PyExc_TypeError, "Each shape must be a (height, width) pair.");
            goto fail;
        }
        PyObject *pair = PySequence_Tuple(item);
        if (!pair)
            goto fail;
        const int ok = PyArg_ParseTuple(pair, "ii", &h, &w);
        Py_DECREF(pair);
        if (!ok)
            goto fail;
        if (!bridct_supported(h, w)) {
            PyErr_SetString(PyExc_ValueError, "Both dimensions must be powers of two from 8 through 1024.");
            goto fail;
        }
        hs[i] = h;
        ws[i] = w;
        PyObject *t = Py_BuildValue("(ii)", h, w);
        if (!t)
            goto fail;
        PyTuple_SET_ITEM(shapes, i, t);
    }
    if (!bridct_gpu_available()) {
        PyErr_SetString(PyExc_RuntimeError, "No usable gfx950 GPU: the BRiDCT GPU batch cannot be created.");
        goto fail;
    }
    self->batch = bridct_batch_create((size_t)n, hs, ws);
    if (!self->batch) {
        PyErr_NoMemory();
        goto fail;
    }
    self->count = (size_t)n;
    self->elements = bridct_batch_elements(self->batch);
    self->shapes = shapes;
    self->initialized = 1;
    self->busy = 0;
    free(hs);
    free(ws);
    Py_DECREF(seq);
    return 0;
fail:
    self->busy = 0;
    free(hs);
    free(ws);
    Py_XDECREF(shapes);
    Py_XDECREF(seq);
    return -1;
}

static void RaggedPlan_dealloc(RaggedPlan *self)
{
    ragged_release(self);
    Py_XDECREF(self->shapes);
    Py_TYPE(self)->tp_free((PyObject *)self);
}

static PyObject *RaggedPlan_close(RaggedPlan *self, PyObject *unused)
{
    (void)unused;
    if (self->busy) {
        PyErr_SetString(PyExc_RuntimeError,
                        "This RaggedPlan is already initializing or executing, possibly in another thread.");
        return NULL;
    }
    ragged_release(self);
    Py_RETURN_NONE;
}

/* Validates array i of a list against shape i (the Plan rules, for a 2D array). */
static PyArrayObject *ragged_item(PyObject *object, int h, int w, int writable)
{
    if (!PyArray_CheckExact(object)) {
        PyErr_SetString(PyExc_TypeError, "Expected exact NumPy ndarrays; no implicit conversion is performed.");
        return NULL;
    }
    PyArrayObject *a = (PyArrayObject *)object;
    if (PyArray_TYPE(a) != NPY_FLOAT32 || !PyArray_ISNOTSWAPPED(a)) {
        PyErr_SetString(PyExc_TypeError, "Expected native-endian float32.");
        return NULL;
    }
    if (PyArray_NDIM(a) != 2 || PyArray_DIM(a, 0) != h || PyArray_DIM(a, 1) != w) {
        PyErr_SetString(PyExc_ValueError, "Each array must have the planned (height, width).");
        return NULL;
    }
    if (!PyArray_IS_C_CONTIGUOUS(a) || !PyArray_ISALIGNED(a)) {
        PyErr_SetString(PyExc_ValueError, "Expected C-contiguous, float-aligned storage.");
        return NULL;
    }
    if (writable && !PyArray_ISWRITEABLE(a)) {
        PyErr_SetString(PyExc_ValueError, "The output arrays must be writable.");
        return NULL;
    }
    return a;
}

static PyObject *ragged_execute(RaggedPlan *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kwnames,
                                int mode)
{
    if (!self->batch) {
        PyErr_SetString(PyExc_RuntimeError, "This RaggedPlan is closed or uninitialized.");
        return NULL;
    }
    if (self->busy) {
        PyErr_SetString(PyExc_RuntimeError, "This RaggedPlan is executing in another thread; use one per thread.");
        return NULL;
    }
    self->busy = 1;
    PyObject *in_seq = NULL, *out_seq = NULL, *result = NULL;
    const Py_ssize_t nkw = kwnames ? PyTuple_GET_SIZE(kwnames) : 0;
    if (nargs != 1 || nkw > 1) {
        PyErr_SetString(PyExc_TypeError, "Use method(arrays, *, out=None).");
        goto fail;
    }
    PyObject *out_obj = Py_None;
    if (nkw) {
        if (PyUnicode_CompareWithASCIIString(PyTuple_GET_ITEM(kwnames, 0), "out") != 0) {
            PyErr_SetString(PyExc_TypeError, "The only accepted keyword is out.");
            goto fail;
        }
        out_obj = args[1];
    }
    if (!PyList_Check(args[0]) && !PyTuple_Check(args[0])) {
        PyErr_SetString(PyExc_TypeError, "Expected a list or tuple of arrays.");
        goto fail;
    }
    in_seq = PySequence_Fast(args[0], "Expected a list or tuple of arrays.");
    if (!in_seq)
        goto fail;
    if ((size_t)PySequence_Fast_GET_SIZE(in_seq) != self->count) {
        PyErr_SetString(PyExc_ValueError, "The list length must equal the number of planned shapes.");
        goto fail;
    }
    if (out_obj != Py_None) {
        if (!PyList_Check(out_obj) && !PyTuple_Check(out_obj)) {
            PyErr_SetString(PyExc_TypeError, "out must be a list or tuple of arrays.");
            goto fail;
        }
        out_seq = PySequence_Fast(out_obj, "out must be a list or tuple of arrays.");
        if (!out_seq)
            goto fail;
        if ((size_t)PySequence_Fast_GET_SIZE(out_seq) != self->count) {
            PyErr_SetString(PyExc_ValueError, "out must hold one array per planned shape.");
            goto fail;
        }
    }
    for (size_t i = 0; i < self->count; i++) {
        PyObject *shape = PyTuple_GET_ITEM(self->shapes, (Py_ssize_t)i);
        const int h = (int)PyLong_AsLong(PyTuple_GET_ITEM(shape, 0)), w = (int)PyLong_AsLong(PyTuple_GET_ITEM(shape, 1));
        PyArrayObject *a = ragged_item(PySequence_Fast_GET_ITEM(in_seq, (Py_ssize_t)i), h, w, 0);
        if (!a)
            goto fail;
        if (out_seq) {
            PyArrayObject *o = ragged_item(PySequence_Fast_GET_ITEM(out_seq, (Py_ssize_t)i), h, w, 1);
            if (!o)
                goto fail;
            const uintptr_t x = (uintptr_t)PyArray_DATA(a), y = (uintptr_t)PyArray_DATA(o);
            if ((x <= y ? y - x : x - y) < (size_t)PyArray_NBYTES(a)) {
                PyErr_SetString(PyExc_ValueError, "Input and output arrays must not overlap.");
                goto fail;
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
            goto fail;
        }
    }
    for (size_t i = 0; i < self->count; i++) {
        PyArrayObject *a = (PyArrayObject *)PySequence_Fast_GET_ITEM(in_seq, (Py_ssize_t)i);
        memcpy(self->stage_in + bridct_batch_offset(self->batch, i), PyArray_DATA(a), (size_t)PyArray_NBYTES(a));
    }
    int status;
    if (self->elements >= 16384) {
        Py_BEGIN_ALLOW_THREADS
        status = bridct_batch_apply(self->batch, mode, self->stage_in, self->stage_out);
        Py_END_ALLOW_THREADS
    } else {
        status = bridct_batch_apply(self->batch, mode, self->stage_in, self->stage_out);
    }
    if (status) {
        PyErr_Format(PyExc_RuntimeError, "BRiDCT failed with status %d.", status);
        goto fail;
    }
    result = PyList_New((Py_ssize_t)self->count);
    if (!result)
        goto fail;
    for (size_t i = 0; i < self->count; i++) {
        PyArrayObject *o;
        if (out_seq) {
            o = (PyArrayObject *)PySequence_Fast_GET_ITEM(out_seq, (Py_ssize_t)i);
            Py_INCREF(o);
        } else {
            PyArrayObject *a = (PyArrayObject *)PySequence_Fast_GET_ITEM(in_seq, (Py_ssize_t)i);
            o = (PyArrayObject *)PyArray_SimpleNew(2, PyArray_DIMS(a), NPY_FLOAT32);
            if (!o)
                goto fail;
        }
        memcpy(PyArray_DATA(o), self->stage_out + bridct_batch_offset(self->batch, i), (size_t)PyArray_NBYTES(o));
        PyList_SET_ITEM(result, (Py_ssize_t)i, (PyObject *)o);
    }
    Py_DECREF(in_seq);
    Py_XDECREF(out_seq);
    self->busy = 0;
    return result;
fail:
    self->busy = 0;
    Py_XDECREF(in_seq);
    Py_XDECREF(out_seq);
    Py_XDECREF(result);
    return NULL;
}

static PyObject *RaggedPlan_forward(RaggedPlan *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kw)
{
    return ragged_execute(self, args, nargs, kw, BRIDCT_FORWARD);
}

static PyObject *RaggedPlan_inverse(RaggedPlan *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kw)
{
    return ragged_execute(self, args, nargs, kw, BRIDCT_INVERSE);
}

static PyObject *RaggedPlan_roundtrip(RaggedPlan *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kw)
{
    return ragged_execute(self, args, nargs, kw, BRIDCT_GPU_FORWARD_THEN_INVERSE);
}

static PyObject *RaggedPlan_shapes(RaggedPlan *self, void *unused)
{
    (void)unused;
    if (!self->shapes)
        Py_RETURN_NONE;
    Py_INCREF(self->shapes);
    return self->shapes;
}

static PyObject *RaggedPlan_closed(RaggedPlan *self, void *unused)
{
    (void)unused;
    return PyBool_FromLong(self->batch == NULL);
}

static PyMethodDef RaggedPlan_methods[] = {
    {"forward", AS_METHOD(RaggedPlan_forward), METH_FASTCALL | METH_KEYWORDS,
     "forward(arrays, *, out=None): orthonormal 2D DCT-II of every array, one GPU launch."},
    {"inverse", AS_METHOD(RaggedPlan_inverse), METH_FASTCALL | METH_KEYWORDS,
     "inverse(arrays, *, out=None): orthonormal 2D DCT-III of every array, one GPU launch."},
    {"roundtrip", AS_METHOD(RaggedPlan_roundtrip), METH_FASTCALL | METH_KEYWORDS,
     "roundtrip(arrays, *, out=None): forward then inverse with a materialized spectrum."},
    {"close", AS_METHOD(RaggedPlan_close), METH_NOARGS, "Release the GPU batch and its buffers."},
    {NULL, NULL, 0, NULL}};

static PyGetSetDef RaggedPlan_getset[] = {
    {"shapes", (getter)RaggedPlan_shapes, NULL, "Planned ((height, width), ...).", NULL},
    {"closed", (getter)RaggedPlan_closed, NULL, "Whether the plan has been closed.", NULL},
    {NULL, NULL, NULL, NULL, NULL}};

static PyTypeObject RaggedPlanType = {
    PyVarObject_HEAD_INIT(NULL, 0)
    .tp_name = "bridct_numpy.RaggedPlan",
    .tp_basicsize = sizeof(RaggedPlan),
    .tp_dealloc = (destructor)RaggedPlan_dealloc,
    .tp_flags = Py_TPFLAGS_DEFAULT,
    .tp_doc = "Orthonormal float32 2D DCT of a list of arrays of mixed shapes, one GPU launch per direction.",
    .tp_methods = RaggedPlan_methods,
    .tp_getset = RaggedPlan_getset,
    .tp_init = (initproc)RaggedPlan_init,
    .tp_new = PyType_GenericNew};

/* ---- Codec8 ---------------------------------------------------------------------------- */

typedef struct {
    PyObject_HEAD
    bridct_codec8 *codec;
    float quant[64];
    int initialized, busy;
} Codec8;

static int Codec8_init(Codec8 *self, PyObject *args, PyObject *kwargs)
{
    static char *keywords[] = {"quant", NULL};
    PyObject *q_obj;
    if (self->initialized) {
        PyErr_SetString(PyExc_RuntimeError, "A Codec8 cannot be reinitialized.");
        return -1;
    }
    if (self->busy) {
        PyErr_SetString(PyExc_RuntimeError, "This Codec8 is already initializing or executing.");
        return -1;
    }
    if (!PyArg_ParseTupleAndKeywords(args, kwargs, "O:Codec8", keywords, &q_obj))
        return -1;
    if (!PyArray_CheckExact(q_obj)) {
        PyErr_SetString(PyExc_TypeError, "Expected the 64 quantizer steps as an exact NumPy ndarray.");
        return -1;
    }
    PyArrayObject *q = (PyArrayObject *)q_obj;
    if (PyArray_TYPE(q) != NPY_FLOAT32 || !PyArray_ISNOTSWAPPED(q)) {
        PyErr_SetString(PyExc_TypeError, "Expected native-endian float32 quantizer steps.");
        return -1;
    }
    if (PyArray_SIZE(q) != 64 || !(PyArray_NDIM(q) == 1 || (PyArray_NDIM(q) == 2 && PyArray_DIM(q, 0) == 8)) ||
        !PyArray_IS_C_CONTIGUOUS(q) || !PyArray_ISALIGNED(q)) {
        PyErr_SetString(PyExc_ValueError, "Expected 64 C-contiguous steps, shape (64,) or (8, 8).");
        return -1;
    }
    const float *steps = PyArray_DATA(q);
    for (int i = 0; i < 64; i++)
        if (!isfinite(steps[i]) || !(steps[i] > 0.0f)) {
            PyErr_SetString(PyExc_ValueError, "Every quantizer step must be finite and positive.");
            return -1;
        }
    if (!bridct_gpu_available()) {
        PyErr_SetString(PyExc_RuntimeError, "No usable gfx950 GPU: the BRiDCT codec cannot be created.");
        return -1;
    }
    self->busy = 1;
    bridct_codec8 *codec = bridct_codec8_create(steps);
    self->busy = 0;
    if (!codec) {
        PyErr_NoMemory();
        return -1;
    }
    memcpy(self->quant, steps, sizeof(self->quant));
    self->codec = codec;
    self->initialized = 1;
    return 0;
}

static void Codec8_dealloc(Codec8 *self)
{
    bridct_codec8_destroy(self->codec);
    Py_TYPE(self)->tp_free((PyObject *)self);
}

static PyObject *Codec8_close(Codec8 *self, PyObject *unused)
{
    (void)unused;
    if (self->busy) {
        PyErr_SetString(PyExc_RuntimeError,
                        "This Codec8 is already initializing or executing, possibly in another thread.");
        return NULL;
    }
    bridct_codec8_destroy(self->codec);
    self->codec = NULL;
    Py_RETURN_NONE;
}

/* A uint8 (H, W) raster with unit column stride and rows at least W bytes apart. */
static PyArrayObject *codec_image(PyObject *object, int writable, npy_intp h, npy_intp w)
{
    if (!PyArray_CheckExact(object)) {
        PyErr_SetString(PyExc_TypeError, "Expected an exact NumPy ndarray; no implicit conversion is performed.");
        return NULL;
    }
    PyArrayObject *a = (PyArrayObject *)object;
    if (PyArray_TYPE(a) != NPY_UINT8) {
        PyErr_SetString(PyExc_TypeError, "Expected a uint8 image.");
        return NULL;
    }
    if (PyArray_NDIM(a) != 2 || PyArray_DIM(a, 0) < 8 || PyArray_DIM(a, 1) < 8 || PyArray_DIM(a, 0) % 8 ||
        PyArray_DIM(a, 1) % 8 || (h >= 0 && (PyArray_DIM(a, 0) != h || PyArray_DIM(a, 1) != w))) {
        PyErr_SetString(PyExc_ValueError, "Expected a (height, width) image, both multiples of 8 (and matching).");
        return NULL;
    }
    if (PyArray_STRIDE(a, 1) != 1 || PyArray_STRIDE(a, 0) < PyArray_DIM(a, 1)) {
        PyErr_SetString(PyExc_ValueError, "Expected contiguous rows (unit column stride, forward row stride).");
        return NULL;
    }
    if (writable && !PyArray_ISWRITEABLE(a)) {
        PyErr_SetString(PyExc_ValueError, "The output array must be writable.");
        return NULL;
    }
    return a;
}

/* int16 (H/8, W/8, 8, 8) coefficients, C-contiguous. */
static PyArrayObject *codec_coeffs(PyObject *object, int writable, npy_intp bh, npy_intp bw)
{
    if (!PyArray_CheckExact(object)) {
        PyErr_SetString(PyExc_TypeError, "Expected an exact NumPy ndarray; no implicit conversion is performed.");
        return NULL;
    }
    PyArrayObject *a = (PyArrayObject *)object;
    if (PyArray_TYPE(a) != NPY_INT16 || !PyArray_ISNOTSWAPPED(a)) {
        PyErr_SetString(PyExc_TypeError, "Expected native-endian int16 coefficients.");
        return NULL;
    }
    if (PyArray_NDIM(a) != 4 || PyArray_DIM(a, 0) < 1 || PyArray_DIM(a, 1) < 1 || PyArray_DIM(a, 2) != 8 ||
        PyArray_DIM(a, 3) != 8 || (bh >= 0 && (PyArray_DIM(a, 0) != bh || PyArray_DIM(a, 1) != bw))) {
        PyErr_SetString(PyExc_ValueError, "Expected (height/8, width/8, 8, 8) coefficients (and matching).");
        return NULL;
    }
    if (!PyArray_IS_C_CONTIGUOUS(a) || !PyArray_ISALIGNED(a)) {
        PyErr_SetString(PyExc_ValueError, "Expected C-contiguous, aligned coefficients.");
        return NULL;
    }
    if (writable && !PyArray_ISWRITEABLE(a)) {
        PyErr_SetString(PyExc_ValueError, "The output array must be writable.");
        return NULL;
    }
    return a;
}

/* First and one-past-last byte of an array's storage (non-negative strides). */
static void extent(PyArrayObject *a, uintptr_t *lo, uintptr_t *hi)
{
    uintptr_t end = (uintptr_t)PyArray_DATA(a);
    for (int d = 0; d < PyArray_NDIM(a); d++)
        end += (uintptr_t)(PyArray_DIM(a, d) - 1) * (uintptr_t)PyArray_STRIDE(a, d);
    *lo = (uintptr_t)PyArray_DATA(a);
    *hi = end + (uintptr_t)PyArray_ITEMSIZE(a);
}

static PyObject *codec_execute(Codec8 *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kwnames,
                               int decode)
{
    if (!self->codec) {
        PyErr_SetString(PyExc_RuntimeError, "This Codec8 is closed or uninitialized.");
        return NULL;
    }
    if (self->busy) {
        PyErr_SetString(PyExc_RuntimeError, "This Codec8 is executing in another thread; use one per thread.");
        return NULL;
    }
    self->busy = 1;
    PyArrayObject *out = NULL;
    const Py_ssize_t nkw = kwnames ? PyTuple_GET_SIZE(kwnames) : 0;
    if (nargs != 1 || nkw > 1) {
        PyErr_SetString(PyExc_TypeError, "Use method(input, *, out=None).");
        goto fail;
    }
    PyObject *out_obj = Py_None;
    if (nkw) {
        if (PyUnicode_CompareWithASCIIString(PyTuple_GET_ITEM(kwnames, 0), "out") != 0) {
            PyErr_SetString(PyExc_TypeError, "The only accepted keyword is out.");
            goto fail;
        }
        out_obj = args[1];
    }
    PyArrayObject *in = decode ? codec_coeffs(args[0], 0, -1, -1) : codec_image(args[0], 0, -1, -1);
    if (!in)
        goto fail;
    const npy_intp h = decode ? 8 * PyArray_DIM(in, 0) : PyArray_DIM(in, 0);
    const npy_intp w = decode ? 8 * PyArray_DIM(in, 1) : PyArray_DIM(in, 1);
    if (out_obj == Py_None) {
        npy_intp img_dims[2] = {h, w}, co_dims[4] = {h / 8, w / 8, 8, 8};
        out = (PyArrayObject *)(decode ? PyArray_SimpleNew(2, img_dims, NPY_UINT8)
                                       : PyArray_SimpleNew(4, co_dims, NPY_INT16));
        if (!out)
            goto fail;
    } else {
        PyArrayObject *given = decode ? codec_image(out_obj, 1, h, w) : codec_coeffs(out_obj, 1, h / 8, w / 8);
        if (!given)
            goto fail;
        uintptr_t a0, a1, b0, b1;
        extent(in, &a0, &a1);
        extent(given, &b0, &b1);
        if (a0 < b1 && b0 < a1) {
            PyErr_SetString(PyExc_ValueError, "Input and output must not overlap.");
            goto fail;
        }
        out = given;
        Py_INCREF(out);
    }
    PyArrayObject *image = decode ? out : in;
    const size_t stride = (size_t)PyArray_STRIDE(image, 0);
    int status;
    Py_BEGIN_ALLOW_THREADS
    if (decode)
        status = bridct_codec8_decode(self->codec, PyArray_DATA(in), (size_t)h, (size_t)w, stride,
                                      PyArray_DATA(out));
    else
        status = bridct_codec8_encode(self->codec, PyArray_DATA(in), (size_t)h, (size_t)w, stride,
                                      PyArray_DATA(out));
    Py_END_ALLOW_THREADS
    if (status) {
        PyErr_Format(PyExc_RuntimeError, "BRiDCT codec failed with status %d.", status);
        goto fail;
    }
    self->busy = 0;
    return (PyObject *)out;
fail:
    self->busy = 0;
    Py_XDECREF(out);
    return NULL;
}

static PyObject *Codec8_encode(Codec8 *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kw)
{
    return codec_execute(self, args, nargs, kw, 0);
}

static PyObject *Codec8_decode(Codec8 *self, PyObject *const *args, Py_ssize_t nargs, PyObject *kw)
{
    return codec_execute(self, args, nargs, kw, 1);
}

static PyObject *Codec8_quant(Codec8 *self, void *unused)
{
    (void)unused;
    if (!self->initialized)
        Py_RETURN_NONE;
    npy_intp dims[2] = {8, 8};
    PyObject *q = PyArray_SimpleNew(2, dims, NPY_FLOAT32);
    if (q)
        memcpy(PyArray_DATA((PyArrayObject *)q), self->quant, sizeof(self->quant));
    return q;
}

static PyObject *Codec8_closed(Codec8 *self, void *unused)
{
    (void)unused;
    return PyBool_FromLong(self->codec == NULL);
}

static PyMethodDef Codec8_methods[] = {
    {"encode", AS_METHOD(Codec8_encode), METH_FASTCALL | METH_KEYWORDS,
     "encode(image, *, out=None): uint8 (H, W) -> int16 (H/8, W/8, 8, 8) quantized 8x8 DCT coefficients."},
    {"decode", AS_METHOD(Codec8_decode), METH_FASTCALL | METH_KEYWORDS,
     "decode(coeffs, *, out=None): int16 (H/8, W/8, 8, 8) -> uint8 (H, W) image."},
    {"close", AS_METHOD(Codec8_close), METH_NOARGS, "Release the GPU codec and its buffers."},
    {NULL, NULL, 0, NULL}};

static PyGetSetDef Codec8_getset[] = {
    {"quant", (getter)Codec8_quant, NULL, "The quantizer steps, (8, 8) float32.", NULL},
    {"closed", (getter)Codec8_closed, NULL, "Whether the codec has been closed.", NULL},
    {NULL, NULL, NULL, NULL, NULL}};

static PyTypeObject Codec8Type = {
    PyVarObject_HEAD_INIT(NULL, 0)
    .tp_name = "bridct_numpy.Codec8",
    .tp_basicsize = sizeof(Codec8),
    .tp_dealloc = (destructor)Codec8_dealloc,
    .tp_flags = Py_TPFLAGS_DEFAULT,
    .tp_doc = "8x8 block codec step on the GPU: level shift, 8x8 DCT and quantization, and back.",
    .tp_methods = Codec8_methods,
    .tp_getset = Codec8_getset,
    .tp_init = (initproc)Codec8_init,
    .tp_new = PyType_GenericNew};

static int module_exec(PyObject *m)
{
    if (PyInterpreterState_Get() != PyInterpreterState_Main()) {
        PyErr_SetString(PyExc_ImportError, "bridct_numpy supports the main interpreter only.");
        return -1;
    }
    if (_import_array() < 0)
        return -1;
    if (PyType_Ready(&PlanType) < 0)
        return -1;
    Py_INCREF(&PlanType);
    if (PyModule_AddObject(m, "Plan", (PyObject *)&PlanType) < 0) {
        Py_DECREF(&PlanType);
        return -1;
    }
    if (PyType_Ready(&RaggedPlanType) < 0)
        return -1;
    Py_INCREF(&RaggedPlanType);
    if (PyModule_AddObject(m, "RaggedPlan", (PyObject *)&RaggedPlanType) < 0) {
        Py_DECREF(&RaggedPlanType);
        return -1;
    }
    if (PyType_Ready(&Codec8Type) < 0)
        return -1;
    Py_INCREF(&Codec8Type);
    if (PyModule_AddObject(m, "Codec8", (PyObject *)&Codec8Type) < 0) {
        Py_DECREF(&Codec8Type);
        return -1;
    }
    if (PyModule_AddStringConstant(m, "core_version", BRIDCT_VERSION) < 0)
        return -1;
    return 0;
}

#if defined(__GNUC__) && !defined(__clang__)
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wpedantic"
#endif
static PyModuleDef_Slot module_slots[] = {{Py_mod_exec, (void *)module_exec}, {0, NULL}};
#if defined(__GNUC__) && !defined(__clang__)
#pragma GCC diagnostic pop
#endif

static PyModuleDef module_def = {
    PyModuleDef_HEAD_INIT,
    .m_name = "bridct_numpy",
    .m_doc = "NumPy interface to the MI355X BRiDCT 2D DCT plan library.",
    .m_size = 0,
    .m_slots = module_slots};

PyMODINIT_FUNC PyInit_bridct_numpy(void)
{
    return PyModuleDef_Init(&module_def);
}
