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
    {"sizes", py_sizes, METH_VARARGS, "sizes(seq) -> (n, total_bytes, max_len)"},
    {NULL, NULL, 0, NULL},
};

static struct PyModuleDef module = {PyModuleDef_HEAD_INIT, "_edpy", NULL, -1, methods,
                                    NULL, NULL, NULL, NULL};

PyMODINIT_FUNC PyInit__edpy(void) { return PyModule_Create(&module); }
