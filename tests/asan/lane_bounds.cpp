/*
 * lane_bounds.cpp -- the kernels' per-lane code (bl_device.h, through the
 * host build bl_hostcheck.cpp) under AddressSanitizer and UBSan.
 *
 * The device code reads strings in whole aligned units: dwords (load4,
 * ByteStream, the trims, the bit-plane build) or 16-byte blocks (BlockStream,
 * the band lanes), and only units that hold a byte of the string. Such a unit
 * never crosses a page, so on the device it stays inside the caller's
 * allocation. Here every string sits in a heap block of its own that begins
 * at the start of the first such unit and ends exactly at the end of the last
 * (an empty string gets a block with nothing readable past it), at every
 * offset within a unit. Any read before or past those units, or any undefined
 * shift or overflow, aborts the run. Results are checked against
 * Wagner-Fischer on the way.
 *
 *   lane_bounds            run every check; prints "lane bounds: ok"
 *   lane_bounds overread   read the dword after a string's last one (the
 *                          check the run above relies on: it must abort)
 */
#include "bl_hostcheck.cpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

int fails = 0;

#define CHECK(cond, ...)                                                                           \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            if (fails++ < 20) {                                                                    \
                std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                          \
                std::fprintf(stderr, __VA_ARGS__);                                                 \
                std::fprintf(stderr, "\n");                                                        \
            }                                                                                      \
        }                                                                                          \
    } while (0)

/* A string at byte `off` (0 <= off < unit) of its own heap block of
 * round_up(off + n, unit) bytes (off bytes when n == 0). */
struct Placed {
    u8 *base = nullptr;
    const u8 *s = nullptr;
    int n = 0;
    Placed(const std::string &src, int off, int unit)
    {
        n = (int)src.size();
        const size_t size = n ? (size_t)((off + n + unit - 1) / unit * unit) : (size_t)off;
        void *p = nullptr;
        if (posix_memalign(&p, 64, size)) std::abort();
        base = (u8 *)p;
        std::memset(base, 0xA5, size);
        std::memcpy(base + off, src.data(), (size_t)n);
        s = base + off;
    }
    ~Placed() { std::free(base); }
    Placed(const Placed &) = delete;
    Placed &operator=(const Placed &) = delete;
};

/* A packed set (buffer + offsets) in one block of round_up(total, unit) bytes. */
struct PackedSet {
    u8 *buf = nullptr;
    std::vector<int64_t> off;
    PackedSet(const std::vector<std::string> &v, int unit) : off(v.size() + 1, 0)
    {
        for (size_t i = 0; i < v.size(); i++) off[i + 1] = off[i] + (int64_t)v[i].size();
        const size_t total = (size_t)off.back();
        const size_t size = (total + (size_t)unit - 1) / (size_t)unit * (size_t)unit;
        void *p = nullptr;
        if (posix_memalign(&p, 64, size)) std::abort();
        buf = (u8 *)p;
        for (size_t i = 0; i < v.size(); i++) std::memcpy(buf + off[i], v[i].data(), v[i].size());
    }
    ~PackedSet() { std::free(buf); }
    PackedSet(const PackedSet &) = delete;
    PackedSet &operator=(const PackedSet &) = delete;
};

int wf(const std::string &a, const std::string &b)
{
    std::vector<int> row(b.size() + 1);
    for (size_t j = 0; j <= b.size(); j++) row[j] = (int)j;
    for (size_t i = 1; i <= a.size(); i++) {
        int diag = row[0];
        row[0] = (int)i;
        for (size_t j = 1; j <= b.size(); j++) {
            const int up = row[j];
            const int sub = diag + (a[i - 1] != b[j - 1]);
            row[j] = std::min(std::min(row[j - 1] + 1, up + 1), sub);
            diag = up;
        }
    }
    return row[b.size()];
}

int capped(int d, int k) { return (k >= 0 && d > k) ? k + 1 : d; }

std::mt19937 rng(20261204);

std::string random_string(int n, int alpha)
{
    std::string s((size_t)n, '\0');
    for (auto &c : s) c = (char)(alpha == 256 ? rng() & 255 : 'a' + (int)(rng() % (unsigned)alpha));
    return s;
}

/* b from a: a few substitutions, insertions and deletions, then a tail. */
std::string mutate(const std::string &a, int edits, int alpha, int tail)
{
    std::string b = a;
    for (int e = 0; e < edits; e++) {
        const unsigned op = rng() % 3;
        const size_t at = b.empty() ? 0 : rng() % b.size();
        const char c = random_string(1, alpha)[0];
        if (op == 0 && !b.empty()) b[at] = c;
        else if (op == 1) b.insert(b.begin() + (long)at, c);
        else if (!b.empty()) b.erase(b.begin() + (long)at);
    }
    return b + random_string(tail, alpha);
}

u32 bytes_le(const std::string &s, int i, int cnt)
{
    u32 v = 0;
    for (int b = 0; b < cnt; b++) v |= (u32)(u8)s[(size_t)(i + b)] << (8 * b);
    return v;
}

u32 low_mask(int cnt) { return cnt >= 4 ? 0xFFFFFFFFu : (1u << (8 * cnt)) - 1u; }

/* ---------------------------------------------------------------- readers */
void check_readers()
{
    for (int n = 1; n <= 300; n++) {
        const std::string src = random_string(n, 256);
        for (int off = 0; off < 16; off++) {
            {
                Placed p(src, off % 4, 4);
                for (int i = 0; i < n; i++) {
                    const int cnt = std::min(4, n - i);
                    CHECK((load4(p.s, i, n) & low_mask(cnt)) == bytes_le(src, i, cnt),
                          "load4 n=%d off=%d i=%d", n, off, i);
                }
                ByteStream bs;
                bs.open(p.s, n);
                for (int j = 0; j < n / 4 + 3; j++) {
                    const u32 v = bs.next4();
                    const int cnt = std::min(4, n - 4 * j);
                    if (cnt > 0)
                        CHECK((v & low_mask(cnt)) == bytes_le(src, 4 * j, cnt),
                              "ByteStream n=%d off=%d j=%d", n, off, j);
                }
                for (int64_t j = 0; j < n / 4 + 3; j++) {
                    const int cnt = std::max(0, std::min(4, n - 4 * (int)j));
                    const u32 want = cnt ? bytes_le(src, 4 * (int)j, cnt) : 0u;
                    CHECK(stripe_word(p.s, n, j) == want, "stripe_word n=%d off=%d j=%d", n, off,
                          (int)j);
                }
                if (n <= 64) {
                    u64 pl[8];
                    build_planes(p.s, n, pl);
                    for (int i = 0; i < n; i++)
                        for (int b = 0; b < 8; b++)
                            CHECK(((pl[b] >> i) & 1) == (u64)(((u8)src[(size_t)i] >> b) & 1),
                                  "build_planes n=%d off=%d row=%d plane=%d", n, off, i, b);
                }
            }
            {
                Placed p(src, off, 16);
                BlockStream bs;
                bs.open(p.s, n);
                for (int j = 0; j < n / 4 + 3; j++) {
                    const u32 v = bs.next4();
                    const int cnt = std::min(4, n - 4 * j);
                    if (cnt > 0)
                        CHECK((v & low_mask(cnt)) == bytes_le(src, 4 * j, cnt),
                              "BlockStream n=%d off=%d j=%d", n, off, j);
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ trims */
void check_trim()
{
    for (int la = 0; la <= 300; la += (la < 40 ? 1 : 7)) {
        for (int rep = 0; rep < 6; rep++) {
            const int alpha = rep % 3 == 0 ? 2 : rep % 3 == 1 ? 4 : 256;
            const std::string core = random_string(la, alpha);
            const std::string pre = random_string((int)(rng() % 9), alpha);
            const std::string suf = random_string((int)(rng() % 9), alpha);
            const std::string a = pre + core + suf;
            const std::string b = pre + mutate(core, (int)(rng() % 3), alpha, 0) + suf;
            Placed pa(a, (int)(rng() % 4), 4), pb(b, (int)(rng() % 4), 4);
            const u8 *x = pa.s, *y = pb.s;
            int lx = (int)a.size(), ly = (int)b.size();
            trim_pair(x, lx, y, ly);
            size_t p = 0;
            while (p < a.size() && p < b.size() && a[p] == b[p]) p++;
            size_t s = 0;
            while (s < a.size() - p && s < b.size() - p && a[a.size() - 1 - s] == b[b.size() - 1 - s])
                s++;
            CHECK(x == pa.s + p && y == pb.s + p && lx == (int)(a.size() - p - s) &&
                      ly == (int)(b.size() - p - s),
                  "trim_pair la=%d lb=%d", (int)a.size(), (int)b.size());
        }
    }
}

/* ------------------------------------------------------------------ lanes */
void check_pair(const std::string &a, const std::string &b)
{
    const int d = wf(a, b);
    const int la = (int)a.size(), lb = (int)b.size();
    const int offa = (int)(rng() % 16), offb = (int)(rng() % 16);
    Placed pa(a, offa % 4, 4), pb(b, offb % 4, 4);
    Placed qa(a, offa, 16), qb(b, offb, 16);
    const int ks[] = {-1, 0, 2, 7, 40};
    for (int k : ks) {
        for (int trim = 0; trim < 2; trim++) {
            CHECK(blh_pair(pa.s, la, pb.s, lb, trim, k) == capped(d, k), "pair la=%d lb=%d k=%d",
                  la, lb, k);
            const int m = std::min(la, lb);
            if (m <= 256) {
                const int w32 = std::max(1, (m + 31) / 32);
                CHECK(blh_pair_w32(pa.s, la, pb.s, lb, trim, k, w32) == capped(d, k),
                      "pair_w32 la=%d lb=%d k=%d w32=%d", la, lb, k, w32);
            }
            if (k >= 0 && k <= 30) {
                for (int win : {32, 64}) {
                    const int d0 = blh_band_offset(la - lb, la - lb, win);
                    int reach = 0;
                    const int r = blh_band(qa.s, la, qb.s, lb, trim, k, d0, win, &reach);
                    CHECK(r == -1 || r == capped(d, k), "band la=%d lb=%d k=%d win=%d", la, lb, k,
                          win);
                }
            }
        }
        CHECK(blh_peq(pa.s, la, pb.s, lb, k) == capped(d, k), "peq m=%d n=%d k=%d", la, lb, k);
    }
    if (la + lb <= 400)
        CHECK(blh_long(pa.s, la, pb.s, lb, 1, 1, -1) == d, "long la=%d lb=%d", la, lb);
}

void check_lanes()
{
    for (int la = 0; la <= 300; la += (la < 70 ? 1 : la < 140 ? 3 : 9)) {
        for (int rep = 0; rep < 3; rep++) {
            const int alpha = rep == 0 ? 2 : rep == 1 ? 4 : 256;
            const std::string a = random_string(la, alpha);
            check_pair(a, mutate(a, (int)(rng() % 6), alpha, (int)(rng() % 5)));
            check_pair(a, random_string((int)(rng() % 301), alpha));
        }
    }
    check_pair("", "");
    check_pair("", "abc");
    check_pair("abc", "");
}

/* ------------------------------------------------------------ packed sets */
void check_packed()
{
    std::vector<std::string> q, c;
    for (int i = 0; i < 40; i++) q.push_back(random_string((int)(rng() % 131), i % 2 ? 4 : 256));
    for (int i = 0; i < 150; i++) c.push_back(random_string((int)(rng() % 131), i % 2 ? 4 : 256));
    q.push_back("");
    c.push_back("");
    for (int unit : {4, 16}) {
        PackedSet Q(q, unit), C(c, unit);
        std::vector<int32_t> out(q.size() * c.size());
        for (int k : {-1, 3}) {
            blh_cross(Q.buf, Q.off.data(), (int64_t)q.size(), C.buf, C.off.data(), (int64_t)c.size(),
                      out.data(), k, 0);
            for (size_t i = 0; i < q.size(); i++)
                for (size_t j = 0; j < c.size(); j++)
                    CHECK(out[i * c.size() + j] == capped(wf(q[i], c[j]), k), "cross q=%d c=%d k=%d",
                          (int)i, (int)j, k);
        }
        std::vector<int32_t> pout(c.size());
        std::vector<std::string> a(c.rbegin(), c.rend());
        PackedSet A(a, unit);
        blh_pairs(A.buf, A.off.data(), C.buf, C.off.data(), (int64_t)c.size(), pout.data(), 1, -1, 0);
        for (size_t i = 0; i < c.size(); i++)
            CHECK(pout[i] == wf(a[i], c[i]), "pairs i=%d", (int)i);
    }
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc > 1 && std::string(argv[1]) == "overread") {
        Placed p(random_string(5, 256), 0, 4);
        volatile u32 v = ld32(p.s + 8);          /* the dword after the last one */
        std::printf("read %u\n", (unsigned)v);
        return 0;
    }
    check_readers();
    check_trim();
    check_lanes();
    check_packed();
    if (fails) {
        std::fprintf(stderr, "%d check(s) failed\n", fails);
        return 1;
    }
    std::printf("lane bounds: ok\n");
    return 0;
}
