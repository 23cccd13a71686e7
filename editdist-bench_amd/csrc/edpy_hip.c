/*
 * edpy_hip.c -- the C ABI of libbitlev_hip.so (include/bitlev_hip.h) as
 * functions of the `bitlev._edpy` extension, so Python reaches the kernels
 * through the package's own C extension, as the reference reaches its C core
 * through _edpy (csrc/edpy.c:12, :231-239).
 *
 *   hip_open(path) -> None      dlopen the library (once) and bind every entry
 *                               point; OSError if it cannot be loaded
 *   hip_symbols() -> tuple      the entry points bound
 *   bl_*(...)                   one function per C entry point, same arguments
 *                               in the same order; returns the entry point's
 *                               status or value as a Python int (bl_last_error:
 *                               bytes or None)
 *
 * Argument conversion: int and size arguments take Python ints; pointer
 * arguments take None (NULL), an int address (device pointers, numpy
 * `.ctypes.data`), or any object exporting a buffer (a ctypes instance such as
 * c_void_p or a Structure, a numpy array, bytearray), whose first byte is
 * passed. Buffers must stay alive and unresized during the call. Every call
 * runs with the GIL released: no entry point calls back into Python.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <dlfcn.h>
#include <stddef.h>
#include <stdint.h>

#include "bitlev_hip.h"

#define BL_FNS(X)                                                                              \
    X(bl_version) X(bl_last_error) X(bl_device_count) X(bl_set_device) X(bl_device_info)       \
    X(bl_malloc) X(bl_free) X(bl_host_malloc) X(bl_host_free) X(bl_memcpy_htod)               \
    X(bl_memcpy_dtoh) X(bl_memcpy_dtod) X(bl_memset) X(bl_stream_create) X(bl_stream_destroy) \
    X(bl_stream_sync) X(bl_device_sync) X(bl_event_create) X(bl_event_destroy)                \
    X(bl_event_record) X(bl_stream_wait_event) X(bl_event_sync) X(bl_event_query)             \
    X(bl_event_elapsed_ms) X(bl_graph_begin) X(bl_graph_end) X(bl_graph_launch)               \
    X(bl_graph_destroy) X(bl_pairs_class) X(bl_stripes_layout) X(bl_stripes_build)            \
    X(bl_stripes_layout_sorted) X(bl_stripes_build_order) X(bl_stripes_layout_workspace_size)       \
    X(bl_stripes_layout_device) X(bl_gather_cols) X(bl_cdist_class) X(bl_cdist_class_stripes) X(bl_long) \
    X(bl_cdist_workspace_size) X(bl_cdist) X(bl_cdist_tuned) X(bl_cdist_matches)              \
    X(bl_pairs_workspace_size) X(bl_pairs) X(bl_pairs_tuned) X(bl_cdist_noalloc)

#define BL_DECL(n) __typeof__(&n) n;
static struct { BL_FNS(BL_DECL) } H;
static void *handle;

#define BL_ENTRY(n) {#n, (void **)&H.n},
static const struct { const char *name; void **slot; } TABLE[] = { BL_FNS(BL_ENTRY) };
#define NFNS (sizeof(TABLE) / sizeof(TABLE[0]))

static PyObject *py_hip_open(PyObject *self, PyObject *args)
{
    (void)self;
    const char *path;
    if (!PyArg_ParseTuple(args, "s", &path)) return NULL;
    if (handle) Py_RETURN_NONE;
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        PyErr_Format(PyExc_OSError, "cannot load %s: %s", path, dlerror());
        return NULL;
    }
    for (size_t i = 0; i < NFNS; i++) {
        void *f = dlsym(h, TABLE[i].name);
        if (!f) {
            PyErr_Format(PyExc_OSError, "%s does not export %s", path, TABLE[i].name);
            for (size_t j = 0; j < i; j++) *TABLE[j].slot = NULL;
            dlclose(h);
            return NULL;
        }
        *TABLE[i].slot = f;
    }
    handle = h;
    Py_RETURN_NONE;
}

static PyObject *py_hip_symbols(PyObject *self, PyObject *args)
{
    (void)self;
    (void)args;
    PyObject *t = PyTuple_New((Py_ssize_t)NFNS);
    if (!t) return NULL;
    for (size_t i = 0; i < NFNS; i++) {
        PyObject *s = PyUnicode_FromString(TABLE[i].name);
        if (!s) {
            Py_DECREF(t);
            return NULL;
        }
        PyTuple_SET_ITEM(t, (Py_ssize_t)i, s);
    }
    return t;
}

/* "O&" converter: None -> NULL, int -> address, buffer -> its first byte. */
static int ptr_arg(PyObject *o, void *out)
{
    void **p = (void **)out;
    if (o == Py_None) {
        *p = NULL;
        return 1;
    }
    if (PyLong_Check(o)) {
        *p = PyLong_AsVoidPtr(o);
        return !PyErr_Occurred();
    }
    if (PyObject_CheckBuffer(o)) {
        Py_buffer v;
        if (PyObject_GetBuffer(o, &v, PyBUF_SIMPLE) < 0) return 0;
        *p = v.buf;
        PyBuffer_Release(&v);
        return 1;
    }
    PyErr_Format(PyExc_TypeError, "expected an address (int), None or a buffer, got %s",
                 Py_TYPE(o)->tp_name);
    return 0;
}

/* "O&" converter for size_t: a non-negative Python int. */
static int size_arg(PyObject *o, void *out)
{
    size_t v = PyLong_AsSize_t(o);
    if (v == (size_t)-1 && PyErr_Occurred()) return 0;
    *(size_t *)out = v;
    return 1;
}

static int bound(void)
{
    if (handle) return 1;
    PyErr_SetString(PyExc_RuntimeError, "libbitlev_hip.so is not loaded (call hip_open first)");
    return 0;
}

/* Status-returning call with the GIL released. */
#define RUN(expr)                                                                              \
    do {                                                                                       \
        int rc_;                                                                               \
        Py_BEGIN_ALLOW_THREADS rc_ = (expr);                                                   \
        Py_END_ALLOW_THREADS return PyLong_FromLong(rc_);                                      \
    } while (0)

#define P(x) ptr_arg, &x
#define Z(x) size_arg, &x
#define FN(name) static PyObject *py_##name(PyObject *self, PyObject *args)
#define PARSE(...)                                                                             \
    (void)self;                                                                                \
    if (!bound() || !PyArg_ParseTuple(args, __VA_ARGS__)) return NULL

FN(bl_version) { (void)args; PARSE(""); RUN(H.bl_version()); }
FN(bl_device_count) { (void)args; PARSE(""); RUN(H.bl_device_count()); }
FN(bl_device_sync) { (void)args; PARSE(""); RUN(H.bl_device_sync()); }

FN(bl_last_error)
{
    PARSE("");
    const char *m = H.bl_last_error();
    if (!m) Py_RETURN_NONE;
    return PyBytes_FromString(m);
}

FN(bl_set_device) { int d; PARSE("i", &d); RUN(H.bl_set_device(d)); }

FN(bl_device_info)
{
    void *name, *cu, *ma, *mi;
    size_t len;
    PARSE("O&O&O&O&O&", P(name), Z(len), P(cu), P(ma), P(mi));
    RUN(H.bl_device_info((char *)name, len, (int *)cu, (int *)ma, (int *)mi));
}

FN(bl_malloc) { void *p; size_t n; PARSE("O&O&", P(p), Z(n)); RUN(H.bl_malloc((void **)p, n)); }
FN(bl_free) { void *p; PARSE("O&", P(p)); RUN(H.bl_free(p)); }
FN(bl_host_malloc) { void *p; size_t n; PARSE("O&O&", P(p), Z(n)); RUN(H.bl_host_malloc((void **)p, n)); }
FN(bl_host_free) { void *p; PARSE("O&", P(p)); RUN(H.bl_host_free(p)); }

FN(bl_memcpy_htod)
{
    void *d, *h, *s;
    size_t n;
    PARSE("O&O&O&O&", P(d), P(h), Z(n), P(s));
    RUN(H.bl_memcpy_htod(d, h, n, s));
}

FN(bl_memcpy_dtoh)
{
    void *h, *d, *s;
    size_t n;
    PARSE("O&O&O&O&", P(h), P(d), Z(n), P(s));
    RUN(H.bl_memcpy_dtoh(h, d, n, s));
}

FN(bl_memcpy_dtod)
{
    void *d, *src, *s;
    size_t n;
    PARSE("O&O&O&O&", P(d), P(src), Z(n), P(s));
    RUN(H.bl_memcpy_dtod(d, src, n, s));
}

FN(bl_memset)
{
    void *d, *s;
    int v;
    size_t n;
    PARSE("O&iO&O&", P(d), &v, Z(n), P(s));
    RUN(H.bl_memset(d, v, n, s));
}

FN(bl_stream_create) { void *p; PARSE("O&", P(p)); RUN(H.bl_stream_create((void **)p)); }
FN(bl_stream_destroy) { void *s; PARSE("O&", P(s)); RUN(H.bl_stream_destroy(s)); }
FN(bl_stream_sync) { void *s; PARSE("O&", P(s)); RUN(H.bl_stream_sync(s)); }
FN(bl_event_create) { void *p; PARSE("O&", P(p)); RUN(H.bl_event_create((void **)p)); }
FN(bl_event_destroy) { void *e; PARSE("O&", P(e)); RUN(H.bl_event_destroy(e)); }
FN(bl_event_record) { void *e, *s; PARSE("O&O&", P(e), P(s)); RUN(H.bl_event_record(e, s)); }
FN(bl_stream_wait_event) { void *s, *e; PARSE("O&O&", P(s), P(e)); RUN(H.bl_stream_wait_event(s, e)); }
FN(bl_event_sync) { void *e; PARSE("O&", P(e)); RUN(H.bl_event_sync(e)); }
FN(bl_event_query) { void *e; PARSE("O&", P(e)); RUN(H.bl_event_query(e)); }

FN(bl_event_elapsed_ms)
{
    void *a, *b, *ms;
    PARSE("O&O&O&", P(a), P(b), P(ms));
    RUN(H.bl_event_elapsed_ms(a, b, (float *)ms));
}

FN(bl_graph_begin) { void *s; PARSE("O&", P(s)); RUN(H.bl_graph_begin(s)); }
FN(bl_graph_end) { void *s, *g; PARSE("O&O&", P(s), P(g)); RUN(H.bl_graph_end(s, (void **)g)); }
FN(bl_graph_launch) { void *g, *s; PARSE("O&O&", P(g), P(s)); RUN(H.bl_graph_launch(g, s)); }
FN(bl_graph_destroy) { void *g; PARSE("O&", P(g)); RUN(H.bl_graph_destroy(g)); }

FN(bl_pairs_class)
{
    int cap, dtype, trim, k;
    void *ba, *oa, *bb, *ob, *idx, *out, *s;
    long long cnt;
    PARSE("iO&O&O&O&O&LO&iiiO&", &cap, P(ba), P(oa), P(bb), P(ob), P(idx), &cnt, P(out), &dtype,
          &trim, &k, P(s));
    RUN(H.bl_pairs_class(cap, ba, oa, bb, ob, idx, cnt, out, dtype, trim, k, s));
}

FN(bl_stripes_layout)
{
    void *off, *gbase;
    long long n;
    PARSE("O&LO&", P(off), &n, P(gbase));
    long long r;
    Py_BEGIN_ALLOW_THREADS r = H.bl_stripes_layout(off, n, gbase);
    Py_END_ALLOW_THREADS return PyLong_FromLongLong(r);
}

FN(bl_stripes_build)
{
    void *buf, *off, *gbase, *words, *len, *counts, *counts2, *s;
    long long n;
    PARSE("O&O&LO&O&O&O&O&O&", P(buf), P(off), &n, P(gbase), P(words), P(len), P(counts),
          P(counts2), P(s));
    RUN(H.bl_stripes_build(buf, off, n, gbase, words, len, counts, counts2, s));
}

FN(bl_stripes_layout_sorted)
{
    void *off, *order, *gbase;
    long long n, block;
    PARSE("O&LLO&O&", P(off), &n, &block, P(order), P(gbase));
    long long r;
    Py_BEGIN_ALLOW_THREADS r = H.bl_stripes_layout_sorted(off, n, block, order, gbase);
    Py_END_ALLOW_THREADS return PyLong_FromLongLong(r);
}

FN(bl_stripes_layout_workspace_size)
{
    long long n;
    PARSE("L", &n);
    return PyLong_FromSize_t(H.bl_stripes_layout_workspace_size(n));
}

FN(bl_stripes_layout_device)
{
    void *off, *gbase, *work, *s;
    long long n;
    unsigned long long work_bytes;
    PARSE("O&LO&O&KO&", P(off), &n, P(gbase), P(work), &work_bytes, P(s));
    long long r;
    Py_BEGIN_ALLOW_THREADS r = H.bl_stripes_layout_device(off, n, gbase, work, (size_t)work_bytes, s);
    Py_END_ALLOW_THREADS return PyLong_FromLongLong(r);
}

FN(bl_stripes_build_order)
{
    void *buf, *off, *order, *gbase, *words, *len, *counts, *counts2, *counts3, *s;
    long long n;
    PARSE("O&O&LO&O&O&O&O&O&O&O&", P(buf), P(off), &n, P(order), P(gbase), P(words), P(len),
          P(counts), P(counts2), P(counts3), P(s));
    RUN(H.bl_stripes_build_order(buf, off, n, order, gbase, words, len, counts, counts2, counts3, s));
}

FN(bl_gather_cols)
{
    void *src, *inv, *dst, *s;
    long long rows, cols;
    int dtype;
    PARSE("O&O&O&LLiO&", P(src), P(inv), P(dst), &rows, &cols, &dtype, P(s));
    RUN(H.bl_gather_cols(src, inv, dst, rows, cols, dtype, s));
}

FN(bl_cdist_class)
{
    int nw, dtype, k;
    void *bq, *oq, *qidx, *bc, *oc, *out, *s;
    long long nqc, nc;
    PARSE("iO&O&O&LO&O&LO&iiO&", &nw, P(bq), P(oq), P(qidx), &nqc, P(bc), P(oc), &nc, P(out),
          &dtype, &k, P(s));
    RUN(H.bl_cdist_class(nw, bq, oq, qidx, nqc, bc, oc, nc, out, dtype, k, s));
}

FN(bl_cdist_class_stripes)
{
    int nw, dtype, k;
    void *bq, *oq, *qidx, *st, *out, *s;
    long long nqc;
    PARSE("iO&O&O&LO&O&iiO&", &nw, P(bq), P(oq), P(qidx), &nqc, P(st), P(out), &dtype, &k, P(s));
    RUN(H.bl_cdist_class_stripes(nw, bq, oq, qidx, nqc, st, out, dtype, k, s));
}

FN(bl_long)
{
    int is_pair, dtype, trim, k;
    void *ba, *oa, *bb, *ob, *idx, *out, *s;
    long long cnt, nc, max_len;
    PARSE("iO&O&O&O&O&LLO&iiiLO&", &is_pair, P(ba), P(oa), P(bb), P(ob), P(idx), &cnt, &nc, P(out),
          &dtype, &trim, &k, &max_len, P(s));
    RUN(H.bl_long(is_pair, ba, oa, bb, ob, idx, cnt, nc, out, dtype, trim, k, max_len, s));
}

FN(bl_cdist_workspace_size)
{
    long long nq;
    int mq;
    PARSE("Li", &nq, &mq);
    return PyLong_FromSize_t(H.bl_cdist_workspace_size(nq, mq));
}

FN(bl_pairs_workspace_size)
{
    long long n, max_len;
    PARSE("LL", &n, &max_len);
    return PyLong_FromSize_t(H.bl_pairs_workspace_size(n, max_len));
}

FN(bl_cdist)
{
    void *bq, *oq, *st, *out, *work, *s;
    long long nq;
    int mq, dtype, k;
    size_t wb;
    PARSE("O&O&LO&iO&iiO&O&O&", P(bq), P(oq), &nq, P(st), &mq, P(out), &dtype, &k, P(work), Z(wb),
          P(s));
    RUN(H.bl_cdist(bq, oq, nq, st, mq, out, dtype, k, work, wb, s));
}

FN(bl_cdist_tuned)
{
    void *bq, *oq, *st, *out, *kvec, *work, *prm, *s;
    long long nq;
    int mq, dtype, k;
    size_t wb;
    PARSE("O&O&LO&iO&iiO&O&O&O&O&", P(bq), P(oq), &nq, P(st), &mq, P(out), &dtype, &k, P(kvec),
          P(work), Z(wb), P(prm), P(s));
    RUN(H.bl_cdist_tuned(bq, oq, nq, st, mq, out, dtype, k, kvec, work, wb, prm, s));
}

FN(bl_cdist_noalloc)
{
    void *bq, *oq, *st, *out, *kvec, *prm, *s;
    long long nq;
    int mq, dtype, k;
    PARSE("O&O&LO&iO&iiO&O&O&", P(bq), P(oq), &nq, P(st), &mq, P(out), &dtype, &k, P(kvec), P(prm),
          P(s));
    RUN(H.bl_cdist_noalloc(bq, oq, nq, st, mq, out, dtype, k, kvec, prm, s));
}

FN(bl_cdist_matches)
{
    void *bq, *oq, *st, *kvec, *dq, *dc, *dd, *count, *work, *prm, *s;
    long long nq, cap;
    int mq, k;
    size_t wb;
    PARSE("O&O&LO&iiO&O&O&O&LO&O&O&O&O&", P(bq), P(oq), &nq, P(st), &mq, &k, P(kvec), P(dq), P(dc),
          P(dd), &cap, P(count), P(work), Z(wb), P(prm), P(s));
    RUN(H.bl_cdist_matches(bq, oq, nq, st, mq, k, kvec, dq, dc, dd, cap, count, work, wb, prm, s));
}

FN(bl_pairs)
{
    void *ba, *oa, *bb, *ob, *out, *work, *s;
    long long n, max_len;
    int dtype, trim, k;
    size_t wb;
    PARSE("O&O&O&O&LLO&iiiO&O&O&", P(ba), P(oa), P(bb), P(ob), &n, &max_len, P(out), &dtype, &trim,
          &k, P(work), Z(wb), P(s));
    RUN(H.bl_pairs(ba, oa, bb, ob, n, max_len, out, dtype, trim, k, work, wb, s));
}

FN(bl_pairs_tuned)
{
    void *ba, *oa, *bb, *ob, *out, *kvec, *work, *prm, *s;
    long long n, max_len;
    int dtype, trim, k;
    size_t wb;
    PARSE("O&O&O&O&LLO&iiiO&O&O&O&O&", P(ba), P(oa), P(bb), P(ob), &n, &max_len, P(out), &dtype,
          &trim, &k, P(kvec), P(work), Z(wb), P(prm), P(s));
    RUN(H.bl_pairs_tuned(ba, oa, bb, ob, n, max_len, out, dtype, trim, k, kvec, work, wb, prm, s));
}

#define BL_METHOD(n) {#n, py_##n, METH_VARARGS, "C ABI entry point " #n " (include/bitlev_hip.h)"},
PyMethodDef edpy_hip_methods[] = {
    {"hip_open", py_hip_open, METH_VARARGS, "hip_open(path): load libbitlev_hip.so and bind it"},
    {"hip_symbols", py_hip_symbols, METH_NOARGS, "hip_symbols() -> the bound entry points"},
    BL_FNS(BL_METHOD)
    {NULL, NULL, 0, NULL},
};
