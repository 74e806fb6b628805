// Host-side code of the SpTRSV library: Matrix Market input, lower triangle
// extraction, right-hand side, level analysis, the mix chunk table, CSR->CSC,
// layout validation, the seeded benchmark generators, and the serial CPU
// solves (the reference's CPU path, exported as helpers of their own).
//
// The preprocessing runs once per matrix on the CPU (as in the reference).
// Every solve entry of the GPU library runs on the GPU: none of them calls
// the serial solves.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/capellini_sptrsv.h"
#include "cap_internal.h"

extern "C" void cap_free(void *p) { std::free(p); }

// ---------------------------------------------------------------- .mtx input
namespace {

struct MMType {
    bool complex_ = false, real = false, pattern = false, integer = false, symmetric = false;
};

// Banner rules of mmio.h:249-332: five whitespace-separated tokens; the first
// must start with "%%MatrixMarket" (case-sensitive), the rest are compared
// lower-cased.  Skew-symmetric is accepted but not mirrored (read_mtx.h:49
// only tests symmetric / hermitian).
int parse_banner(const std::string &line, MMType *t)
{
    char tok[5][64];
    if (std::sscanf(line.c_str(), "%63s %63s %63s %63s %63s", tok[0], tok[1], tok[2], tok[3], tok[4]) != 5)
        return 12;
    for (int k = 1; k < 5; ++k)
        for (char *p = tok[k]; *p; ++p) *p = (char)std::tolower((unsigned char)*p);
    if (std::strncmp(tok[0], "%%MatrixMarket", 14) != 0) return 14;
    if (std::strcmp(tok[1], "matrix") != 0) return 15;
    if (std::strcmp(tok[2], "coordinate") != 0 && std::strcmp(tok[2], "array") != 0) return 15;
    if (!std::strcmp(tok[3], "real")) t->real = true;
    else if (!std::strcmp(tok[3], "complex")) t->complex_ = true;
    else if (!std::strcmp(tok[3], "pattern")) t->pattern = true;
    else if (!std::strcmp(tok[3], "integer")) t->integer = true;
    else return 15;
    if (!std::strcmp(tok[4], "symmetric") || !std::strcmp(tok[4], "hermitian")) t->symmetric = true;
    else if (std::strcmp(tok[4], "general") != 0 && std::strcmp(tok[4], "skew-symmetric") != 0) return 15;
    return 0;
}

// Whitespace tokenizer over the whole file buffer: fscanf's "%d %d %lg\n"
// treats every whitespace run the same, so entries may span lines.
struct Cursor {
    const char *p, *e;
    bool skip() {
        while (p < e && std::isspace((unsigned char)*p)) ++p;
        return p < e;
    }
    bool next_int(int *out) {
        if (!skip()) return false;
        char *end = nullptr;
        long v = std::strtol(p, &end, 10);
        if (end == p || v < INT_MIN || v > INT_MAX) return false;
        p = end;
        *out = (int)v;
        return true;
    }
    bool next_double(double *out) {
        if (!skip()) return false;
        char *end = nullptr;
        double v = std::strtod(p, &end);  // the conversion fscanf's %lg performs
        if (end == p) return false;
        p = end;
        *out = v;
        return true;
    }
};

template <typename T>
int read_mtx_body(const char *path, int *m_out, int *n_out, int *nnz_out, int **rp_out, int **ci_out, T **v_out)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return -1;
    std::fseek(f, 0, SEEK_END);
    long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (size < 0) {
        std::fclose(f);
        return -1;
    }
    std::vector<char> buf((size_t)size + 1, '\0');
    size_t got = size > 0 ? std::fread(buf.data(), 1, (size_t)size, f) : 0;
    std::fclose(f);
    const char *p = buf.data(), *e = buf.data() + got;

    auto next_line = [&](std::string *line) -> bool {
        if (p >= e) return false;
        const char *q = (const char *)std::memchr(p, '\n', (size_t)(e - p));
        const char *stop = q ? q : e;
        line->assign(p, stop);
        p = q ? q + 1 : e;
        return true;
    };
    std::string line;
    MMType t;
    if (!next_line(&line) || parse_banner(line, &t) != 0) return -2;
    if (t.complex_) return -3;
    // CUDA/ourWrtFst/mmio.h:343-356: skip '%' lines, the next line must be "M N nz"
    do {
        if (!next_line(&line)) return -4;
    } while (!line.empty() && line[0] == '%');
    int m, n, nr;
    if (std::sscanf(line.c_str(), "%d %d %d", &m, &n, &nr) != 3 || m < 0 || n < 0 || nr < 0) return -4;
    // every entry takes at least four bytes ("i j" and a separator): a count
    // the rest of the file cannot hold is malformed, not an allocation size
    if ((long long)nr > (long long)(e - p) / 4 + 1) return CAP_ERR_IO;

    std::vector<int> ri((size_t)nr), cj((size_t)nr);
    std::vector<T> vv((size_t)nr);
    std::vector<long long> cnt((size_t)m + 1, 0);
    Cursor cur{p, e};
    for (int k = 0; k < nr; ++k) {
        int a, b;
        double fv = 1.0;  // pattern entries are 1.0 (read_mtx.h:93-94)
        if (!cur.next_int(&a) || !cur.next_int(&b)) return CAP_ERR_IO;
        if (t.real) {
            if (!cur.next_double(&fv)) return CAP_ERR_IO;
        } else if (t.integer) {
            int iv;
            if (!cur.next_int(&iv)) return CAP_ERR_IO;
            fv = iv;
        }
        if (a < 1 || a > m || b < 1 || b > n) return CAP_ERR_IO;
        --a;
        --b;  // 1-based -> 0-based (read_mtx.h:98-99)
        if (t.symmetric && a != b && b >= m) return CAP_ERR_IO;
        ri[k] = a;
        cj[k] = b;
        vv[k] = (T)fv;
        cnt[a]++;
        if (t.symmetric && a != b) cnt[b]++;  // mirrored entry (read_mtx.h:110-117)
    }
    // mirrored symmetric entries can take the total past int32 even when the
    // entry count fits: refuse it before any index is formed
    long long total = 0;
    for (int i = 0; i < m; ++i) total += cnt[i];
    if (total > (long long)INT_MAX) return CAP_ERR_IO;
    int *rp = (int *)std::malloc(sizeof(int) * ((size_t)m + 1));
    if (!rp) return CAP_ERR_ALLOC;
    rp[0] = 0;
    for (int i = 0; i < m; ++i) rp[i + 1] = rp[i] + cnt[i];
    const int nnz = rp[m];
    int *ci = (int *)std::malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    T *v = (T *)std::malloc(sizeof(T) * (size_t)(nnz > 0 ? nnz : 1));
    if (!ci || !v) {
        std::free(rp); std::free(ci); std::free(v);
        return CAP_ERR_ALLOC;
    }
    std::fill(cnt.begin(), cnt.end(), 0);
    // entries land in their row in file order, mirrored copies interleaved in
    // file order too (read_mtx.h:139-173); duplicates are kept
    for (int k = 0; k < nr; ++k) {
        int a = ri[k], b = cj[k];
        int o = rp[a] + cnt[a]++;
        ci[o] = b;
        v[o] = vv[k];
        if (t.symmetric && a != b) {
            o = rp[b] + cnt[b]++;
            ci[o] = a;
            v[o] = vv[k];
        }
    }
    *m_out = m; *n_out = n; *nnz_out = nnz;
    *rp_out = rp; *ci_out = ci; *v_out = v;
    return 0;
}

// The reader's working vectors are sized from the file header: an
// allocation failure is reported as CAP_ERR_ALLOC, never thrown across the
// C ABI.
template <typename T>
int read_mtx_impl(const char *path, int *m_out, int *n_out, int *nnz_out, int **rp_out, int **ci_out, T **v_out)
{
    try {
        return read_mtx_body<T>(path, m_out, n_out, nnz_out, rp_out, ci_out, v_out);
    } catch (const std::bad_alloc &) {
        return CAP_ERR_ALLOC;
    }
}

// change2tran (read_mtx.h:194-257): for row i < m keep the entries with
// col < i in ascending column order (duplicates kept), values 1, then a unit
// diagonal.  The reference sorts the whole row with an exchange sort and
// stops at the first col >= i; selecting col < i first and sorting only those
// gives the same columns in the same order.  With keep_values the entries keep
// A's values (stable order among duplicate columns) and the diagonal becomes
// 1 + sum|offdiag|, which makes L diagonally dominant.
template <typename T>
int change2tran_impl(int m, int nnzA, const int *rpA, const int *ciA, const T *vA, int keep,
                     int *nnzL_out, int **rpL_out, int **ciL_out, T **vL_out)
{
    if (m < 0 || nnzA < 0 || !rpA || (nnzA > 0 && (!ciA || !vA))) return CAP_ERR_ARG;
    long long total = 0;
    for (int i = 0; i < m; ++i) {
        int c = 0;
        for (int j = rpA[i]; j < rpA[i + 1]; ++j) c += ciA[j] < i;
        total += c + 1;
    }
    if (total > 0x7fffffffLL) return CAP_ERR_ARG;
    int *rp = (int *)std::malloc(sizeof(int) * ((size_t)m + 1));
    int *ci = (int *)std::malloc(sizeof(int) * (size_t)(total ? total : 1));
    T *v = (T *)std::malloc(sizeof(T) * (size_t)(total ? total : 1));
    if (!rp || !ci || !v) {
        std::free(rp); std::free(ci); std::free(v);
        return CAP_ERR_ALLOC;
    }
    std::vector<std::pair<int, T>> row;
    int p = 0;
    rp[0] = 0;
    for (int i = 0; i < m; ++i) {
        row.clear();
        for (int j = rpA[i]; j < rpA[i + 1]; ++j)
            if (ciA[j] < i) row.emplace_back(ciA[j], vA[j]);
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<int, T> &a, const std::pair<int, T> &b) { return a.first < b.first; });
        double absum = 0.0;
        for (auto &e : row) {
            ci[p] = e.first;
            v[p] = keep ? e.second : (T)1;
            absum += std::fabs((double)v[p]);
            ++p;
        }
        ci[p] = i;
        v[p] = keep ? (T)(1.0 + absum) : (T)1.0;
        ++p;
        rp[i + 1] = p;
    }
    *nnzL_out = p; *rpL_out = rp; *ciL_out = ci; *vL_out = v;
    return CAP_OK;
}

template <typename T>
void get_x_b_impl(int m, int n, const int *rp, const int *ci, const T *v, T *x_ref, T *b)
{
    for (int i = 0; i < n; ++i) x_ref[i] = (T)1;
    for (int i = 0; i < m; ++i) {
        T s = (T)0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) s += v[j] * x_ref[ci[j]];
        b[i] = s;
    }
}

template <typename T>
void transpose_impl(int m, int n, int nnz, const int *rp, const int *ci, const T *v, int *crow, int *cp, T *cv)
{
    std::fill(cp, cp + n + 1, 0);
    for (int k = 0; k < nnz; ++k) cp[ci[k] + 1]++;
    for (int c = 0; c < n; ++c) cp[c + 1] += cp[c];
    std::vector<int> fill(cp, cp + n);
    for (int r = 0; r < m; ++r)
        for (int j = rp[r]; j < rp[r + 1]; ++j) {
            int o = fill[ci[j]]++;
            crow[o] = r;
            cv[o] = v[j];
        }
}

template <typename T>
int validate_impl(int n, int nnz, const int *rp, const int *ci, const T *v, int *bad_row)
{
    if (bad_row) *bad_row = -1;
    if (n < 0 || nnz < 0 || !rp) return CAP_ERR_ARG;
    if (n > 0 && (!ci || !v)) return CAP_ERR_ARG;
    if (rp[0] != 0 || rp[n] != nnz) return CAP_ERR_ROWPTR;
    for (int i = 0; i < n; ++i) {
        const int a = rp[i], e = rp[i + 1];
        if (e < a || e > nnz) { if (bad_row) *bad_row = i; return CAP_ERR_ROWPTR; }
        if (e == a) { if (bad_row) *bad_row = i; return CAP_ERR_NO_DIAG; }
        for (int j = a; j < e - 1; ++j)
            if (ci[j] < 0 || ci[j] >= i) { if (bad_row) *bad_row = i; return CAP_ERR_UPPER; }
        if (ci[e - 1] != i) { if (bad_row) *bad_row = i; return CAP_ERR_NO_DIAG; }
        if (v[e - 1] == (T)0) { if (bad_row) *bad_row = i; return CAP_ERR_ZERO_DIAG; }
    }
    return CAP_OK;
}

// ---------------------------------------------------------------- generators
inline unsigned long long splitmix64(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline double unit53(unsigned long long r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }
// uniform integer in [0, bound) (bound >= 1), multiply-shift on the top 32 bits
inline int below(unsigned long long *s, int bound)
{
    return (int)(((splitmix64(s) >> 32) * (unsigned long long)bound) >> 32);
}
// state of row i: rows are independent streams, so the two passes (count,
// fill) and any host thread count give the same matrix
inline unsigned long long row_state(unsigned long long seed, long long i)
{
    unsigned long long s = seed ^ (0xD1B54A32D192ED03ULL * (unsigned long long)(i + 1));
    splitmix64(&s);
    return s;
}

// k distinct values uniform in [lo, lo+w) by Floyd's algorithm, sorted
void floyd_sample(unsigned long long *s, int lo, int w, int k, std::vector<int> *out)
{
    out->clear();
    for (int j = w - k; j < w; ++j) {
        int t = below(s, j + 1);
        bool seen = false;
        for (int q : *out) if (q == lo + t) { seen = true; break; }
        out->push_back(seen ? lo + j : lo + t);
    }
    std::sort(out->begin(), out->end());
}

// columns of row i (off-diagonals only, ascending)
void gen_row(const cap_gen_params &g, int i, unsigned long long *s, std::vector<int> *cols)
{
    cols->clear();
    switch (g.kind) {
    case 0: {  // band, fixed count
        int lo = std::max(0, i - g.band);
        int k = std::min(g.offdiag, i - lo);
        floyd_sample(s, lo, i - lo, k, cols);
        break;
    }
    case 1: {  // level structure
        int lvl = i / g.width;
        if (lvl == 0) break;
        int lo = std::max(0, lvl - g.depth) * g.width, hi = lvl * g.width;
        int k = 1 + below(s, g.offdiag);
        k = std::min(k, hi - lo);
        floyd_sample(s, lo, hi - lo, k, cols);
        break;
    }
    case 2: {  // banded Bernoulli
        int lo = std::max(0, i - g.band);
        for (int c = lo; c < i; ++c)
            if (unit53(splitmix64(s)) < g.p) cols->push_back(c);
        break;
    }
    case 3: {  // band, per-row count U{0..offdiag-1}
        int lo = std::max(0, i - g.band);
        int k = std::min(below(s, g.offdiag), i - lo);
        floyd_sample(s, lo, i - lo, k, cols);
        break;
    }
    default: break;
    }
}

template <typename T>
int generate_impl(const cap_gen_params *g, int *nnz_out, int *rp, int *ci, T *v)
{
    if (!g || !nnz_out || g->n < 0 || g->kind < 0 || g->kind > 3) return CAP_ERR_ARG;
    if ((g->kind == 0 || g->kind == 3) && (g->offdiag < 0 || g->band < 0)) return CAP_ERR_ARG;
    if (g->kind == 3 && g->offdiag < 1) return CAP_ERR_ARG;
    if (g->kind == 1 && (g->width < 1 || g->depth < 1 || g->offdiag < 1 || (long long)g->levels * g->width != g->n))
        return CAP_ERR_ARG;
    if (g->kind == 2 && (g->band < 0 || !(g->p >= 0.0 && g->p <= 1.0))) return CAP_ERR_ARG;
    std::vector<int> cols;
    long long p = 0;
    if (rp) rp[0] = 0;
    for (int i = 0; i < g->n; ++i) {
        unsigned long long s = row_state(g->seed, i);
        gen_row(*g, i, &s, &cols);
        if (ci && v) {
            double absum = 0.0;
            for (int c : cols) {
                double val = 2.0 * unit53(splitmix64(&s)) - 1.0;
                ci[p] = c;
                v[p] = (T)val;
                absum += std::fabs(val);
                ++p;
            }
            ci[p] = i;
            v[p] = (T)(1.0 + absum);
            ++p;
        } else {
            p += (long long)cols.size() + 1;
        }
        if (p > 0x7fffffffLL) return CAP_ERR_ARG;
        if (rp) rp[i + 1] = (int)p;
    }
    *nnz_out = (int)p;
    return CAP_OK;
}

// ------------------------------------------------------------ serial solves
// Structure the column-driven executor relies on
// (openCL/SynFree_csc/sptrsv_syncfree_serialref.h:25-108): col_ptr runs from
// 0 to nnz without decreasing; FORWARD takes each column's first entry as its
// diagonal and needs every other entry below it, BACKWARD takes the last entry
// and needs every other entry above it.  The reference checks none of this: a
// column whose producers come later makes its wait loop (:51-55, :80-84) spin
// forever, and a row index out of range writes outside left_sum.
int check_csc_factor(int m, int n, int nnz, const int *cp, const int *ri, int substitution)
{
    if (m < 0 || n < 0 || nnz < 0 || !cp || (nnz > 0 && !ri)) return CAP_ERR_ARG;
    if (substitution != CAP_SUBSTITUTION_FORWARD && substitution != CAP_SUBSTITUTION_BACKWARD) return CAP_ERR_ARG;
    if (cp[0] != 0 || cp[n] != nnz) return CAP_ERR_ROWPTR;
    const bool fwd = substitution == CAP_SUBSTITUTION_FORWARD;
    for (int c = 0; c < n; ++c) {
        const int lo = cp[c], hi = cp[c + 1];
        if (hi < lo) return CAP_ERR_ROWPTR;
        if (hi == lo) return CAP_ERR_NO_DIAG;
        const int dg = fwd ? lo : hi - 1;
        if (ri[dg] != c) return CAP_ERR_NO_DIAG;
        for (int j = lo; j < hi; ++j) {
            if (j == dg) continue;
            const int r = ri[j];
            if (r < 0 || r >= m) return CAP_ERR_ARG;
            if (fwd ? r <= c : r >= c) return CAP_ERR_UPPER;
        }
    }
    return CAP_OK;
}

// Column-driven substitution: visit the columns in solve order; a column's
// x is final once all of its row's contributions have arrived (counted
// against the in-degree table), then its entries are scattered into the
// left sums of later rows.  Arithmetic: left += x * v (a rounded product,
// then a rounded sum: this file is compiled without FP contraction) and
// x = (b - left) / diag, as serialref.h:57-60, :67-68.
template <typename T>
int serial_executor_impl(const int *cp, const int *ri, const T *v, const int *in_degree, int m, int n,
                         int substitution, int rhs, const T *b, T *x)
{
    if (m != n) return CAP_ERR_NOT_SQUARE;
    if (!cp || !v || !in_degree || !b || !x || rhs < 1 || n < 0) return CAP_ERR_ARG;
    CAP_TRY(check_csc_factor(m, n, cp[n], cp, ri, substitution));
    const size_t k = (size_t)rhs;
    std::vector<T> left;
    std::vector<int> arrived;
    try {
        left.assign((size_t)m * k, (T)0);
        arrived.assign((size_t)m, 0);
    } catch (const std::bad_alloc &) {
        return CAP_ERR_ALLOC;
    }
    const bool fwd = substitution == CAP_SUBSTITUTION_FORWARD;
    for (int t = 0; t < n; ++t) {
        const int c = fwd ? t : n - 1 - t;
        // a table that disagrees with the matrix would leave the reference waiting
        if (arrived[c] != in_degree[c] - 1) return CAP_ERR_ARG;
        const int lo = cp[c], hi = cp[c + 1];
        const int dg = fwd ? lo : hi - 1;
        const T d = v[dg];
        T *xc = x + (size_t)c * k;
        const T *bc = b + (size_t)c * k;
        const T *lc = left.data() + (size_t)c * k;
        for (size_t q = 0; q < k; ++q) xc[q] = (bc[q] - lc[q]) / d;
        for (int j = lo; j < hi; ++j) {
            if (j == dg) continue;
            const int r = ri[j];
            T *lr = left.data() + (size_t)r * k;
            const T w = v[j];
            for (size_t q = 0; q < k; ++q) {
                const T p = xc[q] * w;
                lr[q] = lr[q] + p;
            }
            ++arrived[r];
        }
    }
    return CAP_OK;
}

// Row-ordered forward substitution over the solve contract's CSR (lower
// triangle, diagonal last, entries in stored order): per right-hand side,
// s accumulates the row's off-diagonal terms in stored order and
// x = (b - s) / diag.  arith picks fma(v, x, s) or s + v*x per term.
template <typename T>
int serial_csr_impl(int n, int nnz, const int *rp, const int *ci, const T *v, int rhs, const T *b, T *x, int arith)
{
    if (!b || !x || rhs < 1 || (arith != CAP_ARITH_FUSED && arith != CAP_ARITH_REFERENCE)) return CAP_ERR_ARG;
    CAP_TRY(validate_impl<T>(n, nnz, rp, ci, v, nullptr));
    const size_t k = (size_t)rhs;
    std::vector<T> s;
    try {
        s.assign(k, (T)0);
    } catch (const std::bad_alloc &) {
        return CAP_ERR_ALLOC;
    }
    const bool fused = arith == CAP_ARITH_FUSED;
    for (int i = 0; i < n; ++i) {
        std::fill(s.begin(), s.end(), (T)0);
        const int dg = rp[i + 1] - 1;
        for (int j = rp[i]; j < dg; ++j) {
            const T w = v[j];
            const T *xc = x + (size_t)ci[j] * k;
            if (fused) {
                for (size_t q = 0; q < k; ++q) s[q] = std::fma(w, xc[q], s[q]);
            } else {
                for (size_t q = 0; q < k; ++q) {
                    const T p = w * xc[q];
                    s[q] = s[q] + p;
                }
            }
        }
        const T d = v[dg];
        for (size_t q = 0; q < k; ++q) x[(size_t)i * k + q] = (b[(size_t)i * k + q] - s[q]) / d;
    }
    return CAP_OK;
}

double ms_since(const std::chrono::steady_clock::time_point &t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// The timed, validated wrapper (serialref.h:110-177): analysis, execution,
// relative L1 error against x_ref with the reference's 1e-4 bar, and its
// report lines on stdout.  Returns -1 for a non-square matrix, as the
// reference does, and the executor's code for a malformed one.
template <typename T>
int serialref_impl(const int *cp, const int *ri, const T *v, int m, int n, int nnz, int substitution, int rhs, T *x,
                   const T *b, const T *x_ref)
{
    if (m != n) {
        std::printf("This is not a square matrix, return.\n");
        return CAP_ERR_NOT_SQUARE;
    }
    if (!cp || !x_ref || rhs < 1 || n < 0 || nnz < 0 || (nnz > 0 && !ri)) return CAP_ERR_ARG;
    if (cp[n] != nnz) return CAP_ERR_ROWPTR;
    std::vector<int> indeg;
    try {
        indeg.assign((size_t)(m > 0 ? m : 1), 0);
    } catch (const std::bad_alloc &) {
        return CAP_ERR_ALLOC;
    }
    std::printf(" - SpTRSV Serial analyser start!\n");
    auto t0 = std::chrono::steady_clock::now();
    CAP_TRY(cap_serial_analyser(ri, m, n, nnz, indeg.data()));
    std::printf("SpTRSV Serial analyser on L used %4.2f ms\n", ms_since(t0));
    std::printf(" - SpTRSV Serial executor start!\n");
    t0 = std::chrono::steady_clock::now();
    CAP_TRY(serial_executor_impl<T>(cp, ri, v, indeg.data(), m, n, substitution, rhs, b, x));
    const double ms = ms_since(t0);
    std::printf("SpTRSV Serial executor used %4.2f ms, throughput is %4.2f gflops\n", ms,
                2.0 * rhs * (double)nnz / (1e6 * ms));
    double ref = 0, res = 0;
    for (size_t i = 0; i < (size_t)n * rhs; ++i) {
        ref += std::fabs((double)x_ref[i]);
        res += std::fabs((double)x[i] - (double)x_ref[i]);
    }
    if (ref != 0) res /= ref;
    if (res < 1e-4)
        std::printf("SpTRSV Serial executor passed! |x-xref|/|xref| = %8.2e\n", res);
    else
        std::printf("SpTRSV Serial executor _NOT_ passed! |x-xref|/|xref| = %8.2e\n", res);
    return CAP_OK;
}

}  // namespace

extern "C" {

int cap_read_mtx_d(const char *path, int *m, int *n, int *nnz, int **rp, int **ci, double **v)
{
    return read_mtx_impl<double>(path, m, n, nnz, rp, ci, v);
}
int cap_read_mtx_s(const char *path, int *m, int *n, int *nnz, int **rp, int **ci, float **v)
{
    return read_mtx_impl<float>(path, m, n, nnz, rp, ci, v);
}
int cap_change2tran_d(int m, int nnzA, const int *rpA, const int *ciA, const double *vA, int keep, int *nnzL,
                      int **rpL, int **ciL, double **vL)
{
    return change2tran_impl<double>(m, nnzA, rpA, ciA, vA, keep, nnzL, rpL, ciL, vL);
}
int cap_change2tran_s(int m, int nnzA, const int *rpA, const int *ciA, const float *vA, int keep, int *nnzL,
                      int **rpL, int **ciL, float **vL)
{
    return change2tran_impl<float>(m, nnzA, rpA, ciA, vA, keep, nnzL, rpL, ciL, vL);
}
void cap_get_x_b_d(int m, int n, const int *rp, const int *ci, const double *v, double *x, double *b)
{
    get_x_b_impl<double>(m, n, rp, ci, v, x, b);
}
void cap_get_x_b_s(int m, int n, const int *rp, const int *ci, const float *v, float *x, float *b)
{
    get_x_b_impl<float>(m, n, rp, ci, v, x, b);
}

// matrix_layer (tranpose.h:78-142): level[r] = max over the row's entries of
// level[col] + 1, the diagonal included while level[r] is still 0.
int cap_matrix_layer(int m, const int *rp, const int *ci, int *levels, double *parallelism, int *row_level)
{
    std::vector<int> own;
    int *lv = row_level;
    if (!lv) {
        own.assign((size_t)(m > 0 ? m : 1), 0);
        lv = own.data();
    } else {
        std::fill(lv, lv + m, 0);
    }
    int mx = 0;
    for (int r = 0; r < m; ++r) {
        int l = 0;
        for (int j = rp[r]; j < rp[r + 1]; ++j) l = std::max(l, lv[ci[j]] + 1);
        lv[r] = l;
        mx = std::max(mx, l);
    }
    if (levels) *levels = mx;
    if (parallelism) *parallelism = (double)m / mx;
    return mx;
}

// matrix_warp4 (CUDA/mix/tranpose.h:1423-1477)
void cap_matrix_warp4(int m, int nnz, const int *rp, int border, int chunk, int *len, int *warp_num,
                      double *warp_occupy, double *element_occupy)
{
    int k = 1, row = 0;
    double wave_rows = 0, wave_nnz = 0;
    warp_num[0] = 0;
    for (int i = 0; i < m; i += chunk) {
        const int end = std::min(i + chunk, m);
        const int e = rp[end] - rp[i];
        if ((double)e / (end - i) >= border) {
            wave_nnz += e;
            wave_rows += end - i;
            for (int j = i; j < end; ++j) warp_num[k++] = ++row;
        } else {
            row += end - i;
            warp_num[k++] = row;
        }
    }
    *len = k;
    *warp_occupy = wave_rows / m;
    *element_occupy = wave_nnz / nnz;
}

void cap_transpose_d(int m, int n, int nnz, const int *rp, const int *ci, const double *v, int *cr, int *cp,
                     double *cv)
{
    transpose_impl<double>(m, n, nnz, rp, ci, v, cr, cp, cv);
}
void cap_transpose_s(int m, int n, int nnz, const int *rp, const int *ci, const float *v, int *cr, int *cp,
                     float *cv)
{
    transpose_impl<float>(m, n, nnz, rp, ci, v, cr, cp, cv);
}
int cap_validate_csr_d(int n, int nnz, const int *rp, const int *ci, const double *v, int *bad_row)
{
    return validate_impl<double>(n, nnz, rp, ci, v, bad_row);
}
int cap_validate_csr_s(int n, int nnz, const int *rp, const int *ci, const float *v, int *bad_row)
{
    return validate_impl<float>(n, nnz, rp, ci, v, bad_row);
}
int cap_generate_d(const cap_gen_params *g, int *nnz, int *rp, int *ci, double *v)
{
    return generate_impl<double>(g, nnz, rp, ci, v);
}
int cap_generate_s(const cap_gen_params *g, int *nnz, int *rp, int *ci, float *v)
{
    return generate_impl<float>(g, nnz, rp, ci, v);
}
void cap_uniform_d(unsigned long long seed, long long count, double lo, double hi, double *out)
{
    for (long long i = 0; i < count; ++i) {
        unsigned long long s = row_state(seed ^ 0x5851F42D4C957F2DULL, i);
        out[i] = lo + (hi - lo) * unit53(splitmix64(&s));
    }
}
int cap_serial_analyser(const int *csc_row_idx, int m, int n, int nnz, int *in_degree)
{
    (void)n;
    if (m < 0 || nnz < 0 || !in_degree || (nnz > 0 && !csc_row_idx)) return CAP_ERR_ARG;
    std::fill(in_degree, in_degree + m, 0);
    for (int j = 0; j < nnz; ++j) {
        const int r = csc_row_idx[j];
        if (r < 0 || r >= m) return CAP_ERR_ARG;
        ++in_degree[r];
    }
    return CAP_OK;
}
int cap_serial_executor_d(const int *cp, const int *ri, const double *v, const int *in_degree, int m, int n,
                          int substitution, int rhs, const double *b, double *x)
{
    return serial_executor_impl<double>(cp, ri, v, in_degree, m, n, substitution, rhs, b, x);
}
int cap_serial_executor_s(const int *cp, const int *ri, const float *v, const int *in_degree, int m, int n,
                          int substitution, int rhs, const float *b, float *x)
{
    return serial_executor_impl<float>(cp, ri, v, in_degree, m, n, substitution, rhs, b, x);
}
int sptrsv_syncfree_serialref_d(const int *cp, const int *ri, const double *v, int m, int n, int nnz,
                                int substitution, int rhs, double *x, const double *b, const double *x_ref)
{
    return serialref_impl<double>(cp, ri, v, m, n, nnz, substitution, rhs, x, b, x_ref);
}
int sptrsv_syncfree_serialref_s(const int *cp, const int *ri, const float *v, int m, int n, int nnz,
                                int substitution, int rhs, float *x, const float *b, const float *x_ref)
{
    return serialref_impl<float>(cp, ri, v, m, n, nnz, substitution, rhs, x, b, x_ref);
}
int cap_serial_csr_d(int n, int nnz, const int *rp, const int *ci, const double *v, int rhs, const double *b,
                     double *x, int arith)
{
    return serial_csr_impl<double>(n, nnz, rp, ci, v, rhs, b, x, arith);
}
int cap_serial_csr_s(int n, int nnz, const int *rp, const int *ci, const float *v, int rhs, const float *b,
                     float *x, int arith)
{
    return serial_csr_impl<float>(n, nnz, rp, ci, v, rhs, b, x, arith);
}

double cap_datasize(int m, int n, int nnzL, int vb, int len)
{
    return (double)(n + 1) * 4 + (double)(nnzL + m) * 4 + (double)nnzL * vb + 2.0 * n * vb + (double)len * 4;
}

const char *cap_strerror(int code)
{
    switch (code) {
    case CAP_OK: return "ok";
    case CAP_ERR_NOT_SQUARE: return "matrix is not square";
    case CAP_ERR_ARG: return "invalid argument";
    case CAP_ERR_ROWPTR: return "row_ptr is not monotone or does not end at nnz";
    case CAP_ERR_UPPER: return "off-diagonal entry on or above the diagonal";
    case CAP_ERR_NO_DIAG: return "last entry of a row is not its diagonal";
    case CAP_ERR_ZERO_DIAG: return "zero diagonal";
    case CAP_ERR_HIP: return "HIP runtime error";
    case CAP_ERR_NO_DEVICE: return "no gfx950 device";
    case CAP_ERR_ALLOC: return "allocation failed";
    case CAP_ERR_UNSUPPORTED: return "unsupported option combination";
    case CAP_ERR_IO: return "malformed Matrix Market file";
    case CAP_ERR_NOT_CONVERGED: return "refinement did not converge";
    case CAP_ERR_WATCHDOG: return "a dependency wait exceeded the spin limit";
    default: return "unknown error";
    }
}
const char *cap_version(void) { return "capellini-sptrsv-amd 0.1 (gfx950)"; }

}  // extern "C"
