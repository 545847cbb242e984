/*
 * The oracle (oracle/rpp_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: every entry point on small shapes, with each
 * operand in its own exactly sized allocation ((n - 1) * stride + dim
 * elements, strides at and above dim), over random degree windows including
 * empty ones, random sparse matrices, and ragged signature batches with
 * empty paths.  Any out-of-bounds access, use of freed memory, leak or
 * undefined arithmetic aborts the run (-fno-sanitize-recover); the last line
 * is "ORACLE SAN OK" when none occurred.  Host-only test infrastructure.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define DECL(T, S)                                                                                                  \
    int orc_batch_ft_mul_##S(int, int, long long, T*, long long, int, int, const T*, long long, int, int, const T*,  \
                             long long, int, int, T);                                                               \
    int orc_batch_ft_fma_##S(int, int, long long, T*, long long, int, int, const T*, long long, int, int, const T*,  \
                             long long, int, int, const T*, long long, int, int, T, T);                            \
    int orc_batch_ft_inplace_fma_##S(int, int, int, long long, T*, long long, int, int, const T*, long long, int,   \
                                     int, const T*, long long, int, int, T, T);                                    \
    int orc_batch_ft_inplace_mul_##S(int, int, long long, T*, long long, int, int, const T*, long long, int, int,   \
                                     T);                                                                            \
    int orc_batch_ft_exp_##S(int, int, long long, T*, long long, int, int, const T*, long long, int, int);          \
    int orc_batch_ft_log_##S(int, int, long long, T*, long long, int, int, const T*, long long, int, int);          \
    int orc_batch_ft_fmexp_##S(int, int, long long, T*, long long, int, int, const T*, long long, int, int,        \
                               const T*, long long, int, int);                                                      \
    int orc_batch_tensor_antipode_##S(int, int, int, long long, T*, long long, int, int, const T*, long long, int, \
                                      int);                                                                         \
    int orc_batch_vector_assign_##S(const long long*, int, long long, T*, long long, int, int, const T*, long long, \
                                    int, int);                                                                      \
    int orc_batch_vector_add_##S(const long long*, int, long long, T*, long long, int, int, const T*, long long,    \
                                 int, int, const T*, long long, int, int, T, T);                                   \
    int orc_batch_vector_inplace_add_##S(const long long*, int, long long, T*, long long, int, int, const T*,       \
                                         long long, int, int, T);                                                   \
    int orc_batch_vector_scalar_##S(int, const long long*, int, long long, T*, long long, int, int, T);            \
    int orc_batch_st_fma_##S(int, int, int, long long, T*, long long, int, int, const T*, long long, int, int,      \
                             const T*, long long, int, int, const T*, long long, int, int, T, T);                  \
    int orc_batch_shuffle2_##S(int, int, int, long long, T*, long long, int, int, const T*, long long, int, int,    \
                               const T*, long long, int, int, T);                                                   \
    int orc_batch_tensor_pairing_##S(int, int, long long, T*, long long, const T*, long long, int, int, const T*,   \
                                     long long, int, int);                                                          \
    int orc_batch_ft_adj_lmul_##S(int, int, long long, T*, long long, int, int, const T*, long long, int, int,      \
                                  const T*, long long, int, int);                                                   \
    int orc_batch_ft_adj_rmul_##S(int, int, long long, T*, long long, int, int, const T*, long long, int, int,      \
                                  const T*, long long, int, int);                                                   \
    int orc_batch_spmv_##S(int, long long, T*, long long, long long, long long, const T*, long long, long long,    \
                           const T*, const long long*, const long long*, long long, T);                            \
    int orc_signature_##S(int, int, long long, const long long*, const T*, T*, long long);

DECL(double, f64)
DECL(float, f32)

long long orc_tensor_dimension(int width, int depth);
int orc_fmexp_vjp_f64(int, int, long long, const double*, const double*, const double*, double*, double*);
int orc_signature_vjp_f64(int, int, long long, const long long*, const double*, const double*, double*);
int orc_log_vjp_f64(int, int, long long, const double*, const double*, double*);

static uint64_t g_state = 0x9e3779b97f4a7c15ull;
static long g_calls = 0;
static int g_failures = 0;

static uint64_t next(void) {
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return g_state;
}
static int rnd(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }
static double unif(void) { return (double)(next() >> 11) * (1.0 / 9007199254740992.0) - 0.5; }

static void expect(int ok, const char* what) {
    if (!ok && ++g_failures <= 20) printf("FAIL %s\n", what);
}

/* An operand: n objects at stride s (>= dim), exactly (n - 1) s + dim elements. */
typedef struct {
    void* p;
    long long s;
} buf;

static buf make(long long n, long long dim, size_t elem, int is_float) {
    buf b;
    b.s = dim + (rnd(0, 2) == 0 ? rnd(1, 5) : 0);
    const long long count = n > 0 ? (n - 1) * b.s + dim : 0;
    b.p = malloc((size_t)(count > 0 ? count : 1) * elem);
    for (long long i = 0; i < count; ++i) {
        if (is_float)
            ((float*)b.p)[i] = (float)unif();
        else
            ((double*)b.p)[i] = unif();
    }
    return b;
}

/* a random window of [0, depth], empty one time in eight */
static void window(int depth, int* lo, int* hi) {
    if (rnd(0, 7) == 0) {
        *lo = rnd(1, depth + 1);
        *hi = *lo - 1;
        return;
    }
    *lo = rnd(0, depth);
    *hi = rnd(*lo, depth);
}

#define CHECK(call) (++g_calls, expect((call) == 0, #call))

#define PRODUCTS(T, S, IS_FLOAT)                                                                                   \
    static void products_##S(int w, int d, int rounds) {                                                           \
        const long long dim = orc_tensor_dimension(w, d);                                                          \
        const size_t e = sizeof(T);                                                                                \
        for (int r = 0; r < rounds; ++r) {                                                                         \
            const long long n = rnd(0, 3);                                                                         \
            buf o = make(n, dim, e, IS_FLOAT), a = make(n, dim, e, IS_FLOAT), b = make(n, dim, e, IS_FLOAT),       \
                c = make(n, dim, e, IS_FLOAT);                                                                     \
            T* O = (T*)o.p;                                                                                        \
            const T *A = (const T*)a.p, *Bp = (const T*)b.p, *C = (const T*)c.p;                                   \
            int ol, oh, al, ah, bl, bh, cl, ch;                                                                    \
            window(d, &ol, &oh);                                                                                   \
            window(d, &al, &ah);                                                                                   \
            window(d, &bl, &bh);                                                                                   \
            window(d, &cl, &ch);                                                                                   \
            CHECK(orc_batch_ft_mul_##S(w, d, n, O, o.s, ol, oh, A, a.s, al, ah, Bp, b.s, bl, bh, (T)0.5));         \
            CHECK(orc_batch_ft_fma_##S(w, d, n, O, o.s, ol, oh, A, a.s, al, ah, Bp, b.s, bl, bh, C, c.s, cl, ch,   \
                                       (T)1, (T)2));                                                               \
            for (int v = 0; v < 3; ++v)                                                                            \
                CHECK(orc_batch_ft_inplace_fma_##S(v == 0 ? 231 : v == 1 ? 123 : 213, w, d, n, O, o.s, ol, oh, A,  \
                                                   a.s, al, ah, Bp, b.s, bl, bh, (T)1, (T)1));                     \
            CHECK(orc_batch_ft_inplace_mul_##S(w, d, n, O, o.s, ol, oh, A, a.s, al, ah, (T)1));                    \
            CHECK(orc_batch_ft_exp_##S(w, d, n, O, o.s, ol, oh, A, a.s, al, ah));                                  \
            CHECK(orc_batch_ft_log_##S(w, d, n, O, o.s, ol, oh, A, a.s, al, ah));                                  \
            CHECK(orc_batch_ft_fmexp_##S(w, d, n, O, o.s, ol, oh, A, a.s, al, ah, Bp, b.s, bl, bh));               \
            CHECK(orc_batch_tensor_antipode_##S(r & 1, w, d, n, O, o.s, ol, oh, A, a.s, al, ah));                  \
            CHECK(orc_batch_ft_adj_lmul_##S(w, d, n, O, o.s, ol, oh, A, a.s, al, ah, Bp, b.s, bl, bh));            \
            CHECK(orc_batch_ft_adj_rmul_##S(w, d, n, O, o.s, ol, oh, A, a.s, al, ah, Bp, b.s, bl, bh));            \
            CHECK(orc_batch_st_fma_##S(0, w, d, n, O, o.s, ol, oh, NULL, 0, 0, -1, Bp, b.s, bl, bh, C, c.s, cl,    \
                                       ch, (T)1, (T)1));                                                           \
            CHECK(orc_batch_st_fma_##S(1, w, d, n, O, o.s, ol, oh, A, a.s, al, ah, Bp, b.s, bl, bh, C, c.s, cl,    \
                                       ch, (T)1, (T)1));                                                           \
            CHECK(orc_batch_st_fma_##S(2, w, d, n, O, o.s, ol, oh, O, o.s, ol, oh, Bp, b.s, bl, bh, C, c.s, cl,    \
                                       ch, (T)1, (T)1));                                                           \
            for (int op = 0; op < 4; ++op)                                                                         \
                CHECK(orc_batch_shuffle2_##S(op, w, d, n, O, o.s, ol, oh, A, a.s, al, ah, Bp, b.s, bl, bh, (T)1)); \
            {                                                                                                      \
                T* sc = (T*)malloc((size_t)(n > 0 ? n : 1) * e);                                                   \
                CHECK(orc_batch_tensor_pairing_##S(w, d, n, sc, 1, A, a.s, al, ah, Bp, b.s, bl, bh));              \
                free(sc);                                                                                          \
            }                                                                                                      \
            {                                                                                                      \
                long long start[40];                                                                               \
                long long p = 1;                                                                                   \
                start[0] = 0;                                                                                      \
                for (int k = 0; k <= d; ++k, p *= w) start[k + 1] = start[k] + p;                                  \
                CHECK(orc_batch_vector_assign_##S(start, d, n, O, o.s, ol, oh, A, a.s, al, ah));                   \
                CHECK(orc_batch_vector_add_##S(start, d, n, O, o.s, ol, oh, A, a.s, al, ah, Bp, b.s, bl, bh,       \
                                               (T)2, (T)-1));                                                      \
                CHECK(orc_batch_vector_inplace_add_##S(start, d, n, O, o.s, ol, oh, A, a.s, al, ah, (T)3));        \
                for (int op = 0; op < 4; ++op)                                                                     \
                    CHECK(orc_batch_vector_scalar_##S(op, start, d, n, O, o.s, ol, oh, (T)0.25));                  \
            }                                                                                                      \
            free(o.p);                                                                                             \
            free(a.p);                                                                                             \
            free(b.p);                                                                                             \
            free(c.p);                                                                                             \
        }                                                                                                          \
    }

PRODUCTS(double, f64, 0)
PRODUCTS(float, f32, 1)

/* A random compressed matrix, outer x inner, in exactly sized arrays. */
static void spmv(int rounds) {
    for (int r = 0; r < rounds; ++r) {
        const long long outer = rnd(0, 12), inner = rnd(0, 12), n = rnd(0, 3);
        long long* off = (long long*)malloc((size_t)(outer + 1) * sizeof(long long));
        off[0] = 0;
        for (long long i = 0; i < outer; ++i) off[i + 1] = off[i] + (inner > 0 ? rnd(0, 3) : 0);
        const long long nnz = off[outer];
        long long* idx = (long long*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(long long));
        double* val = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
        for (long long i = 0; i < nnz; ++i) {
            idx[i] = rnd(0, (int)inner - 1);
            val[i] = unif();
        }
        const int csc = r & 1;
        /* CSR: rows = outer (out), columns = inner (arg); CSC the other way */
        const long long rows = csc ? inner : outer, cols = csc ? outer : inner;
        const long long oo = rnd(0, 2), ao = rnd(0, 2);
        buf o = make(n, oo + rows, sizeof(double), 0), a = make(n, ao + cols, sizeof(double), 0);
        CHECK(orc_batch_spmv_f64(csc, n, (double*)o.p, o.s, oo, rows, (const double*)a.p, a.s, ao, val, idx, off,
                                 outer, 1.5));
        free(o.p);
        free(a.p);
        free(off);
        free(idx);
        free(val);
    }
}

static void signatures(int w, int d) {
    const long long dim = orc_tensor_dimension(w, d);
    const long long n = rnd(0, 4);
    long long* off = (long long*)malloc((size_t)(n + 1) * sizeof(long long));
    off[0] = rnd(0, 2);
    for (long long p = 0; p < n; ++p) off[p + 1] = off[p] + (rnd(0, 3) == 0 ? 0 : rnd(1, 6));
    const long long rows = off[n];
    double* inc = (double*)malloc((size_t)(rows > 0 ? rows : 1) * (size_t)w * sizeof(double));
    for (long long i = 0; i < rows * w; ++i) inc[i] = unif() * 0.5;
    double* out = (double*)malloc((size_t)(n > 0 ? n : 1) * (size_t)dim * sizeof(double));
    CHECK(orc_signature_f64(w, d, n, off, inc, out, dim));
    double* ell = (double*)malloc((size_t)dim * sizeof(double));
    for (long long i = 0; i < dim; ++i) ell[i] = unif();
    double* grad = (double*)malloc((size_t)(rows > 0 ? rows : 1) * (size_t)w * sizeof(double));
    CHECK(orc_signature_vjp_f64(w, d, n, off, inc, ell, grad));
    {
        double* mult = (double*)malloc((size_t)(n > 0 ? n : 1) * (size_t)dim * sizeof(double));
        double* x = (double*)malloc((size_t)(n > 0 ? n : 1) * (size_t)w * sizeof(double));
        double* cot = (double*)malloc((size_t)(n > 0 ? n : 1) * (size_t)dim * sizeof(double));
        double* gm = (double*)malloc((size_t)(n > 0 ? n : 1) * (size_t)dim * sizeof(double));
        double* gx = (double*)malloc((size_t)(n > 0 ? n : 1) * (size_t)w * sizeof(double));
        for (long long i = 0; i < n * dim; ++i) mult[i] = cot[i] = unif();
        for (long long i = 0; i < n * w; ++i) x[i] = unif();
        CHECK(orc_fmexp_vjp_f64(w, d, n, mult, x, cot, gm, gx));
        CHECK(orc_log_vjp_f64(w, d, n, mult, cot, gm));
        free(mult);
        free(x);
        free(cot);
        free(gm);
        free(gx);
    }
    {
        float* incf = (float*)malloc((size_t)(rows > 0 ? rows : 1) * (size_t)w * sizeof(float));
        float* outf = (float*)malloc((size_t)(n > 0 ? n : 1) * (size_t)dim * sizeof(float));
        for (long long i = 0; i < rows * w; ++i) incf[i] = (float)inc[i];
        CHECK(orc_signature_f32(w, d, n, off, incf, outf, dim));
        free(incf);
        free(outf);
    }
    free(off);
    free(inc);
    free(out);
    free(ell);
    free(grad);
}

int main(int argc, char** argv) {
    const int rounds = argc > 1 ? atoi(argv[1]) : 60;
    /* tile geometries of the reference's product: widths 1-8 (F, R per
     * width), 32 (F = 3) and above 32 (F = 2); depth 0 and 1 edge cases */
    static const int shapes[][2] = {{1, 5}, {2, 5}, {3, 4}, {4, 3}, {5, 3}, {6, 2}, {7, 2}, {8, 2},
                                    {32, 2}, {33, 2}, {3, 0}, {2, 1}, {3, 7}, {2, 11}};
    for (size_t i = 0; i < sizeof(shapes) / sizeof(shapes[0]); ++i) {
        const int w = shapes[i][0], d = shapes[i][1];
        const int rr = orc_tensor_dimension(w, d) > 2000 ? 2 : rounds;
        products_f64(w, d, rr);
        products_f32(w, d, rr / 2 + 1);
        if (orc_tensor_dimension(w, d) <= 2000)
            for (int r = 0; r < 4; ++r) signatures(w, d);
    }
    spmv(rounds * 20);
    printf("%s: %ld calls, %d failures\n", g_failures ? "ORACLE SAN FAIL" : "ORACLE SAN OK", g_calls, g_failures);
    return g_failures ? 1 : 0;
}
