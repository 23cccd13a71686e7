/*
 * edpy.c -- the `bitlev._edpy` CPython extension: host glue of the HIP backend.
 *
 * Reads Python str (1-byte kind only, i.e. every character <= U+00FF) and bytes
 * objects in place and packs them into the buffer + int64 offsets layout the
 * C ABI takes. Same contract as the reference's _edpy.pack and its view()
 * (csrc/edpy.c:14-35, :200-224): ValueError for wider str, TypeError for
 * other items, a never-empty buffer. It does no distance arithmetic: every
 * distance of this package is computed on the GPU.
 *
 *   pack(seq) -> (buf: bytes, offsets: bytes of int64[n + 1])
 *   pack_into(seq, buf, offsets) -> total bytes   (writable buffers, no copy out)
 *   sizes(seq) -> (n, total_bytes, max_len)
 *   pack_range(seq, lo, hi, buf, offsets) -> (total_bytes, max_len)
 *       items lo..hi-1 of a list or tuple, entirely without the GIL, so that
 *       several threads pack chunks of one batch at once
 *
 * The same module carries the C ABI of libbitlev_hip.so (edpy_hip.c):
 * hip_open(path) binds the library, then bl_*(...) call its entry points with
 * the GIL released.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <stdint.h>
#include <string.h>

static int item_bytes(PyObject *o, const char **p, Py_ssize_t *len)
{
    if (PyUnicode_Check(o)) {
#if PY_VERSION_HEX < 0x030C0000
        if (PyUnicode_READY(o) < 0) return -1;
#endif
        if (PyUnicode_KIND(o) != PyUnicode_1BYTE_KIND) {
            PyErr_SetString(PyExc_ValueError,
                            "string has characters above U+00FF; pass bytes (e.g. UTF-8 encoded) instead");
            return -1;
        }
        *p = (const char *)PyUnicode_1BYTE_DATA(o);
        *len = PyUnicode_GET_LENGTH(o);
        return 0;
    }
    if (PyBytes_Check(o)) {
        *p = PyBytes_AS_STRING(o);
        *len = PyBytes_GET_SIZE(o);
        return 0;
    }
    PyErr_SetString(PyExc_TypeError, "expected str or bytes");
    return -1;
}

/* Walk a sequence once: total bytes and longest item; validates every item. */
static PyObject *fast_seq(PyObject *obj, Py_ssize_t *n, int64_t *total, int64_t *maxlen)
{
    PyObject *fast = PySequence_Fast(obj, "expected a sequence of str/bytes");
    if (!fast) return NULL;
    *n = PySequence_Fast_GET_SIZE(fast);
    PyObject **items = PySequence_Fast_ITEMS(fast);
    int64_t t = 0, mx = 0;
    for (Py_ssize_t i = 0; i < *n; i++) {
        const char *p;
        Py_ssize_t len;
        if (item_bytes(items[i], &p, &len) < 0) { Py_DECREF(fast); return NULL; }
        t += len;
        if (len > mx) mx = len;
    }
    *total = t;
    *maxlen = mx;
    return fast;
}

static void fill(PyObject *fast, Py_ssize_t n, uint8_t *buf, int64_t *off)
{
    PyObject **items = PySequence_Fast_ITEMS(fast);
    off[0] = 0;
    for (Py_ssize_t i = 0; i < n; i++) {
        const char *p;
        Py_ssize_t len;
        item_bytes(items[i], &p, &len);          /* validated by fast_seq */
        memcpy(buf + off[i], p, (size_t)len);
        off[i + 1] = off[i] + len;
    }
}

static PyObject *py_pack(PyObject *self, PyObject *args)
{
    (void)self;
    PyObject *obj;
    if (!PyArg_ParseTuple(args, "O", &obj)) return NULL;
    Py_ssize_t n;
    int64_t total, mx;
    PyObject *fast = fast_seq(obj, &n, &total, &mx);
    if (!fast) return NULL;
    PyObject *buf = PyBytes_FromStringAndSize(NULL, total ? (Py_ssize_t)total : 1);
    PyObject *off = PyBytes_FromStringAndSize(NULL, (n + 1) * (Py_ssize_t)sizeof(int64_t));
    PyObject *ret = NULL;
    if (buf && off) {
        uint8_t *b = (uint8_t *)PyBytes_AS_STRING(buf);
        if (!total) b[0] = 0;
        Py_BEGIN_ALLOW_THREADS
        fill(fast, n, b, (int64_t *)PyBytes_AS_STRING(off));
        Py_END_ALLOW_THREADS
        ret = PyTuple_Pack(2, buf, off);
    }
    Py_XDECREF(buf);
    Py_XDECREF(off);
    Py_DECREF(fast);
    return ret;
}

static PyObject *py_pack_into(PyObject *self, PyObject *args)
{
    (void)self;
    PyObject *obj;
    Py_buffer bv, ov;
    if (!PyArg_ParseTuple(args, "Ow*w*", &obj, &bv, &ov)) return NULL;
    Py_ssize_t n;
    int64_t total, mx;
    PyObject *ret = NULL;
    PyObject *fast = fast_seq(obj, &n, &total, &mx);
    if (fast) {
        if (bv.len < total || ov.len < (Py_ssize_t)((n + 1) * sizeof(int64_t))) {
            PyErr_SetString(PyExc_ValueError, "output buffer too small");
        } else {
            Py_BEGIN_ALLOW_THREADS
            fill(fast, n, (uint8_t *)bv.buf, (int64_t *)ov.buf);
            Py_END_ALLOW_THREADS
            ret = PyLong_FromLongLong(total);
        }
        Py_DECREF(fast);
    }
    PyBuffer_Release(&bv);
    PyBuffer_Release(&ov);
    return ret;
}

/* One item read without the GIL (type flags, kind and length are fixed for
 * the object's life): 0 ok, 1 neither str nor bytes, 2 str wider than one
 * byte per character, 3 str still in the legacy non-canonical form (made
 * canonical under the GIL by the caller, then read again). */
static int item_bytes_nogil(PyObject *o, const char **p, Py_ssize_t *len)
{
    if (PyUnicode_Check(o)) {
#if PY_VERSION_HEX < 0x030C0000
        if (!PyUnicode_IS_READY(o)) return 3;
#endif
        if (PyUnicode_KIND(o) != PyUnicode_1BYTE_KIND) return 2;
        *p = (const char *)PyUnicode_1BYTE_DATA(o);
        *len = PyUnicode_GET_LENGTH(o);
        return 0;
    }
    if (PyBytes_Check(o)) {
        *p = PyBytes_AS_STRING(o);
        *len = PyBytes_GET_SIZE(o);
        return 0;
    }
    return 1;
}

/* Offsets (relative to the first item) and bytes of n items in one pass. The
 * objects are scattered over the heap, so their headers are prefetched a few
 * items ahead; bytes past `cap` are counted but not written (the caller
 * grows its buffer to the returned total and packs again). */
static int fill_range(PyObject *const *items, Py_ssize_t n, uint8_t *buf, int64_t cap,
                      int64_t *off, int64_t *total, int64_t *maxlen, Py_ssize_t *bad)
{
    enum { AHEAD = 12 };
    int64_t t = 0, mx = 0;
    off[0] = 0;
    for (Py_ssize_t i = 0; i < n; i++) {
        if (i + AHEAD < n) {
            const char *q = (const char *)items[i + AHEAD];
            __builtin_prefetch(q);
            __builtin_prefetch(q + 64);
        }
        const char *p;
        Py_ssize_t len;
        int e = item_bytes_nogil(items[i], &p, &len);
        if (e) {
            *bad = i;
            return e;
        }
        if (t + len <= cap) memcpy(buf + t, p, (size_t)len);
        t += len;
        off[i + 1] = t;
        if (len > mx) mx = len;
    }
    *total = t;
    *maxlen = mx;
    return 0;
}

static PyObject *py_pack_range(PyObject *self, PyObject *args)
{
    (void)self;
    PyObject *seq;
    Py_ssize_t lo, hi;
    Py_buffer bv, ov;
    if (!PyArg_ParseTuple(args, "Onnw*w*", &seq, &lo, &hi, &bv, &ov)) return NULL;
    PyObject *ret = NULL;
    if (!PyList_Check(seq) && !PyTuple_Check(seq)) {
        PyErr_SetString(PyExc_TypeError, "pack_range needs a list or tuple");
    } else if (lo < 0 || hi < lo || hi > PySequence_Fast_GET_SIZE(seq)) {
        PyErr_SetString(PyExc_IndexError, "pack_range: range outside the sequence");
    } else if (ov.len < (Py_ssize_t)((hi - lo + 1) * sizeof(int64_t))) {
        PyErr_SetString(PyExc_ValueError, "offsets buffer too small");
    } else {
        PyObject *const *items = PySequence_Fast_ITEMS(seq) + lo;
        for (int attempt = 0;; attempt++) {
            int64_t total = 0, mx = 0;
            Py_ssize_t bad = 0;
            int err;
            Py_BEGIN_ALLOW_THREADS
            err = fill_range(items, hi - lo, (uint8_t *)bv.buf, (int64_t)bv.len,
                             (int64_t *)ov.buf, &total, &mx, &bad);
            Py_END_ALLOW_THREADS
            if (err == 3 && attempt == 0) {
                Py_ssize_t i;
                for (i = 0; i < hi - lo; i++) {
                    const char *p;
                    Py_ssize_t len;
                    if (PyUnicode_Check(items[i]) && item_bytes(items[i], &p, &len) < 0) break;
                }
                if (i < hi - lo) break;            /* item_bytes set the error */
                continue;
            }
            if (err == 1) {
                PyErr_SetString(PyExc_TypeError, "expected str or bytes");
            } else if (err == 2) {
                PyErr_SetString(PyExc_ValueError,
                                "string has characters above U+00FF; pass bytes (e.g. UTF-8 encoded) instead");
            } else if (err == 3) {
                PyErr_SetString(PyExc_RuntimeError, "pack_range: str could not be made canonical");
            } else {
                ret = Py_BuildValue("(LL)", (long long)total, (long long)mx);
            }
            break;
        }
    }
    PyBuffer_Release(&bv);
    PyBuffer_Release(&ov);
    return ret;
}

static PyObject *py_sizes(PyObject *self, PyObject *args)
{
    (void)self;
    PyObject *obj;
    if (!PyArg_ParseTuple(args, "O", &obj)) return NULL;
    Py_ssize_t n;
    int64_t total, mx;
    PyObject *fast = fast_seq(obj, &n, &total, &mx);
    if (!fast) return NULL;
    Py_DECREF(fast);
    return Py_BuildValue("(nLL)", n, (long long)total, (long long)mx);
}

static PyMethodDef methods[] = {
    {"pack", py_pack, METH_VARARGS, "pack(seq) -> (buf bytes, int64 offsets bytes)"},
    {"pack_into", py_pack_into, METH_VARARGS, "pack_into(seq, buf, offsets) -> total bytes"},
    {"pack_range", py_pack_range, METH_VARARGS,
     "pack_range(seq, lo, hi, buf, offsets) -> (total bytes, longest item)"},
    {"sizes", py_sizes, METH_VARARGS, "sizes(seq) -> (n, total_bytes, max_len)"},
    {NULL, NULL, 0, NULL},
};

static struct PyModuleDef module = {PyModuleDef_HEAD_INIT, "_edpy", NULL, -1, methods,
                                    NULL, NULL, NULL, NULL};

/* The C ABI of libbitlev_hip.so, one function per entry point (edpy_hip.c). */
extern PyMethodDef edpy_hip_methods[];

PyMODINIT_FUNC PyInit__edpy(void)
{
    PyObject *m = PyModule_Create(&module);
    if (m && PyModule_AddFunctions(m, edpy_hip_methods) < 0) {
        Py_DECREF(m);
        return NULL;
    }
    return m;
}
