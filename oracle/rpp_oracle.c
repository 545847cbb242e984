/*
 * rpp_oracle.c -- CPU oracle for the truncated tensor algebra hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the semantics of
 * the reference strategy cpu::single_thread of RoughPathPrimitives v1.0.2 for
 * ft_mul, ft_fma, ft_inplace_fma, ft_inplace_mul, ft_exp, ft_fmexp, ft_log,
 * tensor_pairing, sparse_matrix_vector_product (CSR and CSC), ft_adj_lmul and
 * ft_adj_rmul, plus the caller-side Chen recursion (signature) and its
 * gradient composed from the reference adjoints.  Only tests/, the smoke test
 * and the CPU leg of bench.py load it, and only as the checker.
 *
 * Pinning (each function is checked by tests/test_oracle_*.py):
 *   - against the real reference compiled from its own headers (oracle/_ref,
 *     built by oracle/Makefile in the container that holds the reference) on
 *     seeded inputs and on every degree-window case of
 *     tests/rpp/ft_degree_range_cases.hpp;
 *   - against committed golden fixtures written from oracle/_ref
 *     (tests/golden/);
 *   - against mathematics: exact dyadic brute force of the concatenation
 *     product, unit identity, associativity, exp(x)_w = prod x_{w_i}/k! for a
 *     degree-one x, log(exp(x)) = x, fmexp(exp x, y) = exp(x) exp(y),
 *     the adjoint pairing identity <L_a^* g, t> = <g, a t>, and central finite
 *     differences of the signature for the gradient.
 */
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_DEPTH 32

typedef struct {
    int width;
    int depth;
    long long pow[ORC_MAX_DEPTH + 2];   /* width^k                     */
    long long start[ORC_MAX_DEPTH + 2]; /* start_of_degree(k), k<=N+1  */
} orc_basis;

/* Degree offsets of TensorBasis (basis/tensor_basis.hpp:122-130):
 * start(k) = 1 + d + ... + d^(k-1). */
static int orc_basis_init(orc_basis* B, int width, int depth)
{
    if (width < 1 || depth < 0 || depth > ORC_MAX_DEPTH) return 1;
    B->width = width;
    B->depth = depth;
    B->pow[0] = 1;
    B->start[0] = 0;
    for (int k = 1; k <= depth + 1; ++k) {
        B->pow[k] = B->pow[k - 1] * width;
        B->start[k] = B->start[k - 1] + B->pow[k - 1];
    }
    return 0;
}

/* Any graded basis given by its degree offsets start[0..depth+1] (a Lie
 * basis, for the linear operations, which read nothing else of it). */
static int orc_graded_init(orc_basis* G, const long long* start, int depth)
{
    if (depth < 0 || depth > ORC_MAX_DEPTH || start == 0) return 1;
    G->width = 0;
    G->depth = depth;
    for (int k = 0; k <= depth + 1; ++k) {
        G->start[k] = start[k];
        G->pow[k] = 0;
    }
    return 0;
}

long long orc_tensor_dimension(int width, int depth)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return -1;
    return B.start[depth + 1];
}

/* Order of the terms of one output level in cpu::single_thread's product.
 *
 * The reference multiplies in tiles (cpu/single_thread/operations/basic/
 * detail/ft_multiply.hpp) with tile power 5: 32-wide subtiles, 1024-entry
 * tiles (:25-26).  Its Operation constructor (:114-145) fixes, per width,
 * the tile letters T and the first tiled output degree F; multiply_oneside
 * (:1554-1598) then forms output degree k
 *   - for k < F by contiguous_terms / low_degree_terms
 *     (contiguous_level_terms, :985-1027): rhs degree j ascending;
 *   - for k >= F by right_tiled_terms (:1405-1502) with right blocks of
 *     R = (T == 0 ? 1 : 2T) letters: j = k (lhs[0] * rhs_k, skipped when
 *     lhs[0] == 0), j = 0 (lhs_k * rhs[0], skipped when rhs[0] == 0), then
 *     right_tiled_normal_terms j = R .. k-1 ascending (:1297-1336), then
 *     right_tiled_small_terms j = R-1 .. 1 descending (:1338-1403);
 *   - at width 1 by width_one_terms (:1126-1173): lhs degree ascending, so
 *     j descending.
 * Every term is rounded on its own (x86-64 without FMA: a product, then a
 * sum), so the order fixes the result bit for bit. */
static void orc_tile_geometry(int width, int* first_tiled, int* right_letters)
{
    const long long tile_width = 32, tile_extent = 1024;
    int letters = width > tile_width ? 0 : 1;
    long long size = letters == 0 ? 1 : width;
    int first = 2 + (width <= tile_width ? 1 : 0);
    if (width > 1 && width < tile_width) {
        for (; size * width <= tile_width; ++letters, size *= width) {
        }
        first = 2 * letters + 1;
        long long level = size * size * width;
        for (; level <= tile_extent; ++first, level *= width) {
        }
    }
    *first_tiled = first;
    *right_letters = letters == 0 ? 1 : 2 * letters;
}

/* The rhs degrees j of output degree k in summation order, restricted to
 * [jlo, jhi] (the degrees both operand windows allow); lhs0_zero / rhs0_zero
 * say whether lhs[0] / rhs[0] is zero (the tiled path skips those scalar
 * terms).  Returns the count written to seq (at most k + 1). */
static int orc_term_order(int width, int k, int jlo, int jhi, int lhs0_zero, int rhs0_zero, int* seq)
{
    int n = 0;
    if (jlo > jhi) return 0;
    if (width == 1) {
        for (int j = jhi; j >= jlo; --j) seq[n++] = j;
        return n;
    }
    int first, R;
    orc_tile_geometry(width, &first, &R);
    if (k < first) {
        for (int j = jlo; j <= jhi; ++j) seq[n++] = j;
        return n;
    }
#define ORC_IN(j) ((j) >= jlo && (j) <= jhi)
    if (ORC_IN(k) && !lhs0_zero) seq[n++] = k;
    if (k > 0 && ORC_IN(0) && !rhs0_zero) seq[n++] = 0;
    for (int j = R; j <= k - 1; ++j)
        if (ORC_IN(j)) seq[n++] = j;
    for (int j = R - 1; j >= 1; --j)
        if (ORC_IN(j)) seq[n++] = j;
#undef ORC_IN
    return n;
}

/* Writer kinds of the reference (ft_multiply.hpp:176-435):
 *   W_SCALED:  out = beta * tile                       (ScaledTileWriter)
 *   W_ADD:     out = alpha * add + beta * tile, where the
 *              addend is read only inside its window   (ScaledTileAddWriter)
 */
enum { W_SCALED = 0, W_ADD = 1 };

/* View::truncate(a, b) (views/dense_tensor_view.hpp:29-35). */
static void trunc_window(int mn, int mx, int a, int b, int* omn, int* omx)
{
    *omn = mn > a ? mn : a;
    *omx = mx < b ? mx : b;
}

#define T double
#define SFX f64
#include "rpp_oracle_impl.h"
#undef T
#undef SFX

#define T float
#define SFX f32
#include "rpp_oracle_impl.h"
#undef T
#undef SFX

/* Reverse mode of one Chen step S' = S (x) exp(x) for a degree-one x, composed
 * from the reference adjoints (operations/basic/ft_adj_lmul.hpp,
 * ft_adj_rmul.hpp) as the reference leaves it to the caller:
 *   E     = exp(x)
 *   Ebar  = ft_adj_lmul(op = S, arg = G)
 *   xbar  = J_exp(x)^T Ebar,   exp(x)_w = prod_i x_{w_i} / |w|!
 *   Gprev = ft_adj_rmul(op = E, arg = G)
 * Work arrays: scratch (width^depth + 1), e (dim, zero outside degree 1), E,
 * Eb (dim), fact (depth + 1), letters (depth + 1). */
static void orc_step_vjp_f64(const orc_basis* B, double* scratch, double* e, double* E, double* Eb,
                             const double* fact, int* letters, const double* S, const double* x,
                             const double* G, double* xb, double* Gprev)
{
    const int width = B->width, depth = B->depth;
    for (int l = 0; l < width; ++l) e[1 + l] = x[l];
    orc_ft_exp_f64(B, scratch, E, 0, depth, e, 1, 1);
    orc_ft_adj_lmul_f64(B, Eb, 0, depth, S, 0, depth, G, 0, depth);
    for (int l = 0; l < width; ++l) xb[l] = 0.0;
    for (int k = 1; k <= depth; ++k) {
        for (long long w = 0; w < B->pow[k]; ++w) {
            const double g = Eb[B->start[k] + w] / fact[k];
            if (g == 0.0) continue;
            long long r = w;
            for (int i = k - 1; i >= 0; --i) { letters[i] = (int)(r % width); r /= width; }
            for (int i = 0; i < k; ++i) {
                double prod = 1.0;
                for (int j = 0; j < k; ++j)
                    if (j != i) prod *= x[letters[j]];
                xb[letters[i]] += g * prod;
            }
        }
    }
    orc_ft_adj_rmul_f64(B, Gprev, 0, depth, E, 0, depth, G, 0, depth);
}

typedef struct {
    double *scratch, *e, *E, *Eb, *fact;
    int* letters;
} orc_vjp_work;

static void orc_vjp_work_init(orc_vjp_work* w, const orc_basis* B)
{
    const long long dim = B->start[B->depth + 1];
    w->scratch = (double*)malloc(sizeof(double) * (size_t)(B->pow[B->depth] + 1));
    /* degree-one slots 1..width exist even at depth 0 (the exponent view
     * (1,1) then lies outside the basis and is never read) */
    w->e = (double*)calloc((size_t)(dim + B->width), sizeof(double));
    w->E = (double*)malloc(sizeof(double) * (size_t)dim);
    w->Eb = (double*)malloc(sizeof(double) * (size_t)dim);
    w->fact = (double*)malloc(sizeof(double) * (size_t)(B->depth + 1));
    w->letters = (int*)malloc(sizeof(int) * (size_t)(B->depth + 1));
    w->fact[0] = 1.0;
    for (int k = 1; k <= B->depth; ++k) w->fact[k] = w->fact[k - 1] * k;
}

static void orc_vjp_work_free(orc_vjp_work* w)
{
    free(w->scratch); free(w->e); free(w->E); free(w->Eb); free(w->fact); free(w->letters);
}

/* Adjoint of one ft_fmexp with a degree-one exponent over a batch of n
 * objects: mult, cot, grad_mult are [n, dim], x and grad_x [n, width].
 * FP64 only. */
int orc_fmexp_vjp_f64(int width, int depth, long long n, const double* mult, const double* x,
                      const double* cot, double* grad_mult, double* grad_x)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    const long long dim = B.start[depth + 1];
    orc_vjp_work w;
    orc_vjp_work_init(&w, &B);
    for (long long b = 0; b < n; ++b)
        orc_step_vjp_f64(&B, w.scratch, w.e, w.E, w.Eb, w.fact, w.letters, mult + b * dim, x + b * width,
                         cot + b * dim, grad_x + b * width, grad_mult + b * dim);
    orc_vjp_work_free(&w);
    return 0;
}

/* Gradient of loss_p = <ell, S_p> with respect to the increments of each path:
 * the forward chain of ft_fmexp, then orc_step_vjp_f64 for t = L..1 with
 * G_L = ell.  FP64 only.  ell is one functional (dim) shared by every path. */
int orc_signature_vjp_f64(int width, int depth, long long n_paths,
                          const long long* offsets, const double* increments,
                          const double* ell, double* grad)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    const long long dim = B.start[depth + 1];
    orc_vjp_work w;
    orc_vjp_work_init(&w, &B);
    double* G = (double*)malloc(sizeof(double) * (size_t)dim);
    double* G2 = (double*)malloc(sizeof(double) * (size_t)dim);
    for (long long p = 0; p < n_paths; ++p) {
        const long long L = offsets[p + 1] - offsets[p];
        double* S = (double*)malloc(sizeof(double) * (size_t)(dim * (L + 1)));
        for (long long i = 0; i < dim; ++i) S[i] = 0.0;
        S[0] = 1.0;
        for (long long t = 0; t < L; ++t) {
            const double* x = increments + (offsets[p] + t) * width;
            for (int l = 0; l < width; ++l) w.e[1 + l] = x[l];
            orc_ft_fmexp_f64(&B, w.scratch, S + (t + 1) * dim, 0, depth,
                             S + t * dim, 0, depth, w.e, 1, 1);
        }
        for (long long i = 0; i < dim; ++i) G[i] = ell[i];
        for (long long t = L - 1; t >= 0; --t) {
            const double* x = increments + (offsets[p] + t) * width;
            orc_step_vjp_f64(&B, w.scratch, w.e, w.E, w.Eb, w.fact, w.letters, S + t * dim, x, G,
                             grad + (offsets[p] + t) * width, G2);
            double* tmp = G; G = G2; G2 = tmp;
        }
        free(S);
    }
    orc_vjp_work_free(&w);
    free(G); free(G2);
    return 0;
}

/* Gradient of ft_log (output over the full window [0, depth]) with respect to
 * its argument, from the series log(1 + Y) = sum_k c_k Y^k, c_k = (-1)^(k+1)/k,
 * Y = arg without its degree-0 term (ft_log reads degrees >= 1,
 * operations/intermediate/ft_log.hpp:66-85):
 *   d(Y^k)[delta] = sum_{j<k} Y^j delta Y^(k-1-j), so
 *   Ybar = sum_k c_k sum_{j<k} R^*_{Y^(k-1-j)} L^*_{Y^j} G
 * with L^*, R^* the reference adjoints (ft_adj_lmul.hpp, ft_adj_rmul.hpp) and
 * every product truncated at depth.  This sums the derivative term by term
 * instead of reversing the Horner scheme.  grad[0] = 0.  FP64 only;
 * arg, cot, grad are [n, dim]. */
int orc_log_vjp_f64(int width, int depth, long long n, const double* arg, const double* cot, double* grad)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    const long long dim = B.start[depth + 1];
    double* scratch = (double*)malloc(sizeof(double) * (size_t)(B.pow[depth] + 1));
    double* pw = (double*)calloc((size_t)(dim * (depth + 1)), sizeof(double)); /* Y^0 .. Y^depth */
    double* t1 = (double*)malloc(sizeof(double) * (size_t)dim);
    double* t2 = (double*)malloc(sizeof(double) * (size_t)dim);
    for (long long b = 0; b < n; ++b) {
        const double* a = arg + b * dim;
        const double* G = cot + b * dim;
        double* g = grad + b * dim;
        for (long long i = 0; i < dim * (depth + 1); ++i) pw[i] = 0.0;
        pw[0] = 1.0;
        for (long long i = 1; i < dim; ++i) pw[dim + i] = a[i];
        for (int k = 2; k <= depth; ++k)
            orc_ft_mul_f64(&B, scratch, pw + k * dim, 0, depth, pw + (k - 1) * dim, 0, depth, pw + dim, 0, depth,
                           1.0);
        for (long long i = 0; i < dim; ++i) g[i] = 0.0;
        for (int k = 1; k <= depth; ++k) {
            const double ck = (k % 2 == 0 ? -1.0 : 1.0) / k;
            for (int j = 0; j < k; ++j) {
                orc_ft_adj_lmul_f64(&B, t1, 0, depth, pw + j * dim, 0, depth, G, 0, depth);
                orc_ft_adj_rmul_f64(&B, t2, 0, depth, pw + (k - 1 - j) * dim, 0, depth, t1, 0, depth);
                for (long long i = 1; i < dim; ++i) g[i] += ck * t2[i];
            }
        }
        g[0] = 0.0;
    }
    free(scratch); free(pw); free(t1); free(t2);
    return 0;
}
