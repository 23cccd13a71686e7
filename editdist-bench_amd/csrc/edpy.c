/*
 * edpy.c -- the `bitlev._edpy` CPython extension: host glue of the HIP backend.
 *
 * Reads Python str (1-byte kind only, i.e. every character <= U+00FF) and bytes
 * objects in place and packs them into the buffer + int64 offsets layout the
 * C ABI takes. Same contract as the reference's _edpy.pack and its view()
 * (csrc/edpy.c:14-35, :200-224): ValueError for wider str, TypeError for
 * other items, a never-empty buffer.
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
 * the GIL released; and the CPU engine behind the reference's CPU methods
 * (cpu_engine.c): pairs / cross (scalar), pairs_simd / cross_simd (AVX2),
 * simd_available, and pairs_packed / cross_packed on packed buffers.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cpu_engine.h"

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
        const char *p = NULL;
        Py_ssize_t len = 0;
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

/* ------------------------------------------------ CPU engine (cpu_engine.c) */
/* Lists -> a malloc'd packed copy (the reference's cross_simd packs too,
 * csrc/edpy.c:152-164 there). */
static int pack_list(PyObject *obj, Py_ssize_t *n, uint8_t **buf, int64_t **off)
{
    int64_t total, mx;
    PyObject *fast = fast_seq(obj, n, &total, &mx);
    if (!fast) return -1;
    *buf = malloc((size_t)(total ? total : 1));
    *off = malloc(sizeof(int64_t) * (size_t)(*n + 1));
    if (!*buf || !*off) {
        free(*buf);
        free(*off);
        Py_DECREF(fast);
        PyErr_NoMemory();
        return -1;
    }
    fill(fast, *n, *buf, *off);
    Py_DECREF(fast);
    return 0;
}

static PyObject *cpu_status(int rc)
{
    if (rc == -1) return PyErr_NoMemory();
    if (rc == -2) {
        PyErr_SetString(PyExc_ValueError, "invalid argument");
        return NULL;
    }
    Py_RETURN_NONE;
}

/* pairs(a, b, out, trim=True, k=-1) / pairs_simd(a, b, out, trim=True, groups=4):
 * the reference's _edpy.pairs / pairs_simd (csrc/edpy.c:59-88, :130-149 there) */
static PyObject *pairs_common(PyObject *args, int simd)
{
    PyObject *oa, *ob;
    Py_buffer out;
    int trim = 1, x = simd ? 4 : -1;
    if (!PyArg_ParseTuple(args, "OOw*|pi", &oa, &ob, &out, &trim, &x)) return NULL;
    if (simd && x != 1 && x != 2 && x != 4) {
        PyBuffer_Release(&out);
        PyErr_SetString(PyExc_ValueError, "groups must be 1, 2 or 4");
        return NULL;
    }
    Py_ssize_t na, nb;
    uint8_t *ba = NULL, *bb = NULL;
    int64_t *oa_ = NULL, *ob_ = NULL;
    PyObject *ret = NULL;
    if (pack_list(oa, &na, &ba, &oa_) < 0) goto done;
    if (pack_list(ob, &nb, &bb, &ob_) < 0) goto done;
    if (na != nb) {
        PyErr_SetString(PyExc_ValueError, "length mismatch");
        goto done;
    }
    if (out.len < (Py_ssize_t)(na * (Py_ssize_t)sizeof(int32_t))) {
        PyErr_SetString(PyExc_ValueError, "output buffer too small");
        goto done;
    }
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = simd ? cpu_pairs_simd(ba, oa_, bb, ob_, na, (int32_t *)out.buf, trim)
              : cpu_pairs(ba, oa_, bb, ob_, na, (int32_t *)out.buf, trim, x);
    Py_END_ALLOW_THREADS
    ret = cpu_status(rc);
done:
    free(ba); free(bb); free(oa_); free(ob_);
    PyBuffer_Release(&out);
    return ret;
}

static PyObject *py_cpu_pairs(PyObject *self, PyObject *args) { (void)self; return pairs_common(args, 0); }
static PyObject *py_cpu_pairs_simd(PyObject *self, PyObject *args) { (void)self; return pairs_common(args, 1); }

/* cross(q, c, out, k=-1) / cross_simd(q, c, out, groups=4, k=-1): the
 * reference's _edpy.cross / cross_simd (csrc/edpy.c:91-127, :169-194 there);
 * the SIMD form here takes the threshold exits too */
static PyObject *cross_common(PyObject *args, int simd)
{
    PyObject *oq, *oc;
    Py_buffer out;
    int groups = 4, k = -1;
    if (simd) {
        if (!PyArg_ParseTuple(args, "OOw*|ii", &oq, &oc, &out, &groups, &k)) return NULL;
    } else if (!PyArg_ParseTuple(args, "OOw*|i", &oq, &oc, &out, &k)) {
        return NULL;
    }
    if (simd && groups != 1 && groups != 2 && groups != 4) {
        PyBuffer_Release(&out);
        PyErr_SetString(PyExc_ValueError, "groups must be 1, 2 or 4");
        return NULL;
    }
    Py_ssize_t nq, nc;
    uint8_t *bq = NULL, *bc = NULL;
    int64_t *oq_ = NULL, *oc_ = NULL;
    PyObject *ret = NULL;
    if (pack_list(oq, &nq, &bq, &oq_) < 0) goto done;
    if (pack_list(oc, &nc, &bc, &oc_) < 0) goto done;
    if (out.len < (Py_ssize_t)(nq * nc * (Py_ssize_t)sizeof(int32_t))) {
        PyErr_SetString(PyExc_ValueError, "output buffer too small");
        goto done;
    }
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = simd ? cpu_cross_simd(bq, oq_, nq, bc, oc_, nc, (int32_t *)out.buf, k)
              : cpu_cross(bq, oq_, nq, bc, oc_, nc, (int32_t *)out.buf, k);
    Py_END_ALLOW_THREADS
    ret = cpu_status(rc);
done:
    free(bq); free(bc); free(oq_); free(oc_);
    PyBuffer_Release(&out);
    return ret;
}

static PyObject *py_cpu_cross(PyObject *self, PyObject *args) { (void)self; return cross_common(args, 0); }
static PyObject *py_cpu_cross_simd(PyObject *self, PyObject *args) { (void)self; return cross_common(args, 1); }

static PyObject *py_simd_available(PyObject *self, PyObject *args)
{
    (void)self;
    (void)args;
    return PyBool_FromLong(cpu_simd_available());
}

/* The same on packed inputs (buffer objects, e.g. numpy arrays):
 * pairs_packed(buf_a, off_a, buf_b, off_b, out, trim, k, simd) and
 * cross_packed(buf_q, off_q, buf_c, off_c, out, k, simd); simd pairs are
 * exact, then capped at k + 1 here. */
static PyObject *py_pairs_packed(PyObject *self, PyObject *args)
{
    (void)self;
    Py_buffer a, oa, b, ob, out;
    int trim, k, simd;
    if (!PyArg_ParseTuple(args, "y*y*y*y*w*pip", &a, &oa, &b, &ob, &out, &trim, &k, &simd)) return NULL;
    const int64_t n = oa.len / (Py_ssize_t)sizeof(int64_t) - 1;
    PyObject *ret = NULL;
    if (n < 0 || ob.len != oa.len || out.len < (Py_ssize_t)(n * (int64_t)sizeof(int32_t))) {
        PyErr_SetString(PyExc_ValueError, "offsets or output of the wrong size");
    } else {
        int rc;
        int32_t *o = (int32_t *)out.buf;
        Py_BEGIN_ALLOW_THREADS
        rc = simd ? cpu_pairs_simd(a.buf, oa.buf, b.buf, ob.buf, n, o, trim)
                  : cpu_pairs(a.buf, oa.buf, b.buf, ob.buf, n, o, trim, k);
        if (rc == 0 && simd && k >= 0)
            for (int64_t i = 0; i < n; i++) if (o[i] > k) o[i] = k + 1;
        Py_END_ALLOW_THREADS
        ret = cpu_status(rc);
    }
    PyBuffer_Release(&a); PyBuffer_Release(&oa); PyBuffer_Release(&b); PyBuffer_Release(&ob);
    PyBuffer_Release(&out);
    return ret;
}

static PyObject *py_cross_packed(PyObject *self, PyObject *args)
{
    (void)self;
    Py_buffer q, oq, c, oc, out;
    int k, simd;
    if (!PyArg_ParseTuple(args, "y*y*y*y*w*ip", &q, &oq, &c, &oc, &out, &k, &simd)) return NULL;
    const int64_t nq = oq.len / (Py_ssize_t)sizeof(int64_t) - 1, nc = oc.len / (Py_ssize_t)sizeof(int64_t) - 1;
    PyObject *ret = NULL;
    if (nq < 0 || nc < 0 || out.len < (Py_ssize_t)(nq * nc * (int64_t)sizeof(int32_t))) {
        PyErr_SetString(PyExc_ValueError, "offsets or output of the wrong size");
    } else {
        int rc;
        Py_BEGIN_ALLOW_THREADS
        rc = simd ? cpu_cross_simd(q.buf, oq.buf, nq, c.buf, oc.buf, nc, (int32_t *)out.buf, k)
                  : cpu_cross(q.buf, oq.buf, nq, c.buf, oc.buf, nc, (int32_t *)out.buf, k);
        Py_END_ALLOW_THREADS
        ret = cpu_status(rc);
    }
    PyBuffer_Release(&q); PyBuffer_Release(&oq); PyBuffer_Release(&c); PyBuffer_Release(&oc);
    PyBuffer_Release(&out);
    return ret;
}

static PyMethodDef methods[] = {
    {"pairs", py_cpu_pairs, METH_VARARGS, "pairs(a, b, out, trim=True, k=-1): CPU engine, scalar"},
    {"cross", py_cpu_cross, METH_VARARGS, "cross(q, c, out, k=-1): CPU engine, scalar"},
    {"pairs_simd", py_cpu_pairs_simd, METH_VARARGS,
     "pairs_simd(a, b, out, trim=True, groups=4): CPU engine, AVX2 (exact distances)"},
    {"cross_simd", py_cpu_cross_simd, METH_VARARGS,
     "cross_simd(q, c, out, groups=4, k=-1): CPU engine, AVX2"},
    {"simd_available", py_simd_available, METH_NOARGS, "simd_available() -> AVX2 kernels usable"},
    {"pairs_packed", py_pairs_packed, METH_VARARGS,
     "pairs_packed(buf_a, off_a, buf_b, off_b, out, trim, k, simd)"},
    {"cross_packed", py_cross_packed, METH_VARARGS, "cross_packed(buf_q, off_q, buf_c, off_c, out, k, simd)"},
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
