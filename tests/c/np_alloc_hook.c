/* Test-only CPython module `np_alloc_hook`: runs a callable with a NumPy data-memory handler
 * installed that calls back into Python on the first array-data allocation and can make
 * that allocation fail.  It lets the tests drive two paths of the bridct_numpy objects
 * that ordinary inputs never reach:
 *   - Python code running inside a transform call (the output allocation) that tries to
 *     use or close the same object, which must be refused without harming the call;
 *   - a failed output allocation, which must raise MemoryError and leave the object usable.
 *
 *   np_alloc_hook.intercept(call, on_alloc, fail=False) -> call()'s result
 *
 * Uses NumPy's public handler API (PyDataMem_SetHandler, numpy >= 1.22).  Test
 * infrastructure only; built by the tests with the host C compiler.
 */
#define PY_SSIZE_T_CLEAN
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#define NPY_TARGET_VERSION NPY_1_22_API_VERSION   /* the handler API */
#include <Python.h>
#include <numpy/arrayobject.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    PyDataMem_Handler handler;   /* first member: NumPy reads the capsule as this */
    PyObject *on_alloc;
    int fail, fired;
} trap;

/* 0: go ahead and allocate; -1: refuse (with or without a Python error set). */
static int spring(trap *t)
{
    if (t->fired)
        return 0;
    t->fired = 1;
    PyObject *r = PyObject_CallObject(t->on_alloc, NULL);
    if (!r)
        return -1;
    Py_DECREF(r);
    return t->fail ? -1 : 0;
}

static void *trap_malloc(void *ctx, size_t n)
{
    return spring(ctx) ? NULL : malloc(n);
}

static void *trap_calloc(void *ctx, size_t k, size_t n)
{
    return spring(ctx) ? NULL : calloc(k, n);
}

static void *trap_realloc(void *ctx, void *p, size_t n)
{
    return spring(ctx) ? NULL : realloc(p, n);
}

static void trap_free(void *ctx, void *p, size_t n)
{
    (void)ctx;
    (void)n;
    free(p);
}

/* Arrays keep a reference to their handler's capsule, so the trap lives as long as the
 * last array allocated through it. */
static void trap_release(PyObject *capsule)
{
    trap *t = PyCapsule_GetPointer(capsule, "mem_handler");
    if (!t)
        return;
    Py_XDECREF(t->on_alloc);
    PyMem_RawFree(t);
}

static PyObject *intercept(PyObject *self, PyObject *args, PyObject *kwargs)
{
    (void)self;
    static char *names[] = {"call", "on_alloc", "fail", NULL};
    PyObject *call, *on_alloc;
    int fail = 0;
    if (!PyArg_ParseTupleAndKeywords(args, kwargs, "OO|p:intercept", names, &call, &on_alloc, &fail))
        return NULL;
    trap *t = PyMem_RawCalloc(1, sizeof(*t));
    if (!t)
        return PyErr_NoMemory();
    snprintf(t->handler.name, sizeof(t->handler.name), "np_alloc_hook");
    t->handler.version = 1;
    t->handler.allocator = (PyDataMemAllocator){t, trap_malloc, trap_calloc, trap_realloc, trap_free};
    Py_INCREF(on_alloc);
    t->on_alloc = on_alloc;
    t->fail = fail;
    PyObject *capsule = PyCapsule_New(t, "mem_handler", trap_release);
    if (!capsule) {
        Py_DECREF(on_alloc);
        PyMem_RawFree(t);
        return NULL;
    }
    PyObject *saved = PyDataMem_SetHandler(capsule);
    Py_DECREF(capsule);
    if (!saved)
        return NULL;
    PyObject *result = PyObject_CallObject(call, NULL);
    /* restore the previous handler whatever call() did, keeping its exception */
    PyObject *et, *ev, *tb;
    PyErr_Fetch(&et, &ev, &tb);
    PyObject *mine = PyDataMem_SetHandler(saved);
    Py_DECREF(saved);
    Py_XDECREF(mine);
    if (!mine) {
        Py_XDECREF(et);
        Py_XDECREF(ev);
        Py_XDECREF(tb);
        Py_XDECREF(result);
        return NULL;
    }
    PyErr_Restore(et, ev, tb);
    return result;
}

static PyMethodDef hook_methods[] = {
    {"intercept", (PyCFunction)(void (*)(void))intercept, METH_VARARGS | METH_KEYWORDS,
     "intercept(call, on_alloc, fail=False): call() with a one-shot NumPy allocation trap."},
    {NULL, NULL, 0, NULL}};

static struct PyModuleDef hook_module = {PyModuleDef_HEAD_INIT, "np_alloc_hook", NULL, -1, hook_methods,
                                         NULL, NULL, NULL, NULL};

PyMODINIT_FUNC PyInit_np_alloc_hook(void)
{
    import_array();
    return PyModule_Create(&hook_module);
}
