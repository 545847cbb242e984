/*
 * Type-generic body of the oracle, included twice by rpp_oracle.c with
 *   T   = the scalar and accumulator type (float or double),
 *   SFX = the name suffix (f32 or f64).
 *
 * TEST INFRASTRUCTURE ONLY: this is a plain-C restatement of the reference
 * strategy cpu::single_thread (RoughPathPrimitives v1.0.2).  It is the parity
 * checker for the HIP library and is never linked into it.
 *
 * Every object is addressed as (base pointer, min_degree, max_degree), where
 * the base pointer is the address of coefficient 0 (the empty word), exactly
 * like DenseGradedVectorView (views/dense_graded_vector_view.hpp:17-128).
 * Batches are (pointer, stride in elements) like layouts::StrideLayout
 * (views/batch.hpp:49-68); a stride of 0 broadcasts one object.
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* ---- product core ---------------------------------------------------------
 * Level k of lhs (x) rhs restricted to the operand windows, accumulated in T:
 *   tile[w] = sum_{j} lhs_{k-j}[w / d^j] * rhs_j[w % d^j]
 * over j in [max(rhs.min, k - lhs.max), min(rhs.max, k - lhs.min)], in the
 * order of the reference's tiled product (orc_term_order: ascending below the
 * first tiled degree, cpu/single_thread/operations/basic/detail/
 * ft_multiply.hpp:985-1027; the right-tiled order at and above it,
 * :1297-1502), each product and each sum rounded on its own.
 */
static void FN(level_product)(const orc_basis* B, T* tile, int k,
                              const T* lhs, int lmin, int lmax,
                              const T* rhs, int rmin, int rmax)
{
    const long long n = B->pow[k];
    for (long long w = 0; w < n; ++w) tile[w] = (T)0;
    int jlo = rmin > k - lmax ? rmin : k - lmax;
    int jhi = rmax < k - lmin ? rmax : k - lmin;
    if (jlo < 0) jlo = 0;
    if (jhi > k) jhi = k;
    int seq[ORC_MAX_DEPTH + 2];
    const int lhs0_zero = lmin == 0 && lhs[0] == (T)0;
    const int rhs0_zero = rmin == 0 && rhs[0] == (T)0;
    const int count = orc_term_order(B->width, k, jlo, jhi, lhs0_zero, rhs0_zero, seq);
    for (int t = 0; t < count; ++t) {
        /* word w = p * d^j + s of level k: prefix p of degree k - j, suffix s
         * of degree j; each word receives one term per j, in seq's order */
        const int j = seq[t];
        const long long pj = B->pow[j], np = B->pow[k - j];
        const T* L = lhs + B->start[k - j];
        const T* R = rhs + B->start[j];
        for (long long p = 0; p < np; ++p) {
            const T lv = L[p];
            T* row = tile + p * pj;
            for (long long sfx = 0; sfx < pj; ++sfx) {
                const T prod = lv * R[sfx];
                row[sfx] = row[sfx] + prod;
            }
        }
    }
}

/* Writer kinds W_SCALED / W_ADD are defined in rpp_oracle.c. */

/* multiply(): output degrees are visited in descending order and each level
 * is fully formed in scratch before it is written, so out may alias lhs, rhs
 * or the addend (ft_multiply.hpp:1596-1640). */
static void FN(multiply)(const orc_basis* B, T* scratch,
                         T* out, int omin, int omax,
                         const T* lhs, int lmin, int lmax,
                         const T* rhs, int rmin, int rmax,
                         int writer, const T* add, int amin, int amax,
                         T alpha, T beta)
{
    for (int k = omax; k >= omin; --k) {
        FN(level_product)(B, scratch, k, lhs, lmin, lmax, rhs, rmin, rmax);
        const long long n = B->pow[k];
        T* o = out + B->start[k];
        if (writer == W_ADD && k >= amin && k <= amax) {
            const T* a = add + B->start[k];
            for (long long w = 0; w < n; ++w) o[w] = alpha * a[w] + beta * scratch[w];
        } else {
            for (long long w = 0; w < n; ++w) o[w] = beta * scratch[w];
        }
    }
}

/* --- sub-operations of the composites ------------------------------------ */

/* TensorSetIdentity (cpu/.../basic/tensor_set_identity.hpp): zero the window,
 * then put `s` in degree 0 if the window holds it. */
static void FN(set_identity)(const orc_basis* B, T* t, int mn, int mx, T s)
{
    if (mn > mx) return;
    for (long long i = B->start[mn]; i < B->start[mx + 1]; ++i) t[i] = (T)0;
    if (mn == 0) t[0] = s;
}

/* TensorAddIdentity (cpu/.../basic/tensor_add_identity.hpp). */
static void FN(add_identity)(T* t, int mn, int mx, T s)
{
    if (mn == 0 && mx >= 0) t[0] = t[0] + s;
}

/* VectorSetConstant (cpu/.../linalg/vector_set_constant.hpp). */
static void FN(set_constant)(const orc_basis* B, T* t, int mn, int mx, T v)
{
    if (mn > mx) return;
    for (long long i = B->start[mn]; i < B->start[mx + 1]; ++i) t[i] = v;
}

/* VectorAssign (cpu/.../linalg/vector_assign.hpp:32-61): copy on the overlap
 * of the windows, zero the rest of out's window. */
static void FN(assign)(const orc_basis* B, T* out, int omin, int omax,
                       const T* arg, int amin, int amax)
{
    if (omin > omax) return;
    int lo = omin > amin ? omin : amin;
    int hi = omax < amax ? omax : amax;
    const long long ob = B->start[omin], oe = B->start[omax + 1];
    if (hi < lo) {
        for (long long i = ob; i < oe; ++i) out[i] = (T)0;
        return;
    }
    const long long b = B->start[lo], e = B->start[hi + 1];
    for (long long i = ob; i < b; ++i) out[i] = (T)0;
    for (long long i = b; i < e; ++i) out[i] = arg[i];
    for (long long i = e; i < oe; ++i) out[i] = (T)0;
}

/* --- public single-object operations ------------------------------------- */

/* FTMul (operations/basic/ft_mul.hpp; cpu/.../basic/ft_mul.hpp:42-56). */
void FN(orc_ft_mul)(const orc_basis* B, T* scratch, T* out, int omin, int omax,
                    const T* lhs, int lmin, int lmax, const T* rhs, int rmin,
                    int rmax, T beta)
{
    FN(multiply)(B, scratch, out, omin, omax, lhs, lmin, lmax, rhs, rmin, rmax,
                 W_SCALED, 0, 1, 0, (T)0, beta);
}

/* FTFma (cpu/.../basic/ft_fma.hpp): out = alpha*a + beta*b(x)c. */
void FN(orc_ft_fma)(const orc_basis* B, T* scratch, T* out, int omin, int omax,
                    const T* a, int amin, int amax, const T* b, int bmin,
                    int bmax, const T* c, int cmin, int cmax, T alpha, T beta)
{
    FN(multiply)(B, scratch, out, omin, omax, b, bmin, bmax, c, cmin, cmax,
                 W_ADD, a, amin, amax, alpha, beta);
}

/* FTInplaceFma (cpu/.../basic/ft_inplace_fma.hpp:54-67):
 *   231: a <- alpha*a + beta*b(x)c
 *   123: a <- alpha*c + beta*a(x)b
 *   213: a <- alpha*c + beta*b(x)a */
void FN(orc_ft_inplace_fma)(const orc_basis* B, T* scratch, int variant,
                            T* a, int amin, int amax, const T* b, int bmin,
                            int bmax, const T* c, int cmin, int cmax, T alpha,
                            T beta)
{
    if (variant == 231) {
        FN(multiply)(B, scratch, a, amin, amax, b, bmin, bmax, c, cmin, cmax,
                     W_ADD, a, amin, amax, alpha, beta);
    } else if (variant == 123) {
        FN(multiply)(B, scratch, a, amin, amax, a, amin, amax, b, bmin, bmax,
                     W_ADD, c, cmin, cmax, alpha, beta);
    } else {
        FN(multiply)(B, scratch, a, amin, amax, b, bmin, bmax, a, amin, amax,
                     W_ADD, c, cmin, cmax, alpha, beta);
    }
}

/* FTInplaceMul (cpu/.../basic/ft_inplace_mul.hpp): lhs <- beta*lhs(x)rhs. */
void FN(orc_ft_inplace_mul)(const orc_basis* B, T* scratch, T* lhs, int lmin,
                            int lmax, const T* rhs, int rmin, int rmax, T beta)
{
    FN(multiply)(B, scratch, lhs, lmin, lmax, lhs, lmin, lmax, rhs, rmin, rmax,
                 W_SCALED, 0, 1, 0, (T)0, beta);
}

/* FTExp (operations/intermediate/ft_exp.hpp:70-81), literal Horner order. */
void FN(orc_ft_exp)(const orc_basis* B, T* scratch, T* out, int omin, int omax,
                    const T* arg, int amin, int amax)
{
    const int N = B->depth;
    int tmin, tmax;
    FN(set_identity)(B, out, omin, omax, (T)1);
    trunc_window(amin, amax, 1, N, &tmin, &tmax);
    for (int d = N; d > 0; --d) {
        const T divisor = (T)1 / (T)d;
        FN(multiply)(B, scratch, out, omin, omax, out, omin, omax, arg, tmin,
                     tmax, W_SCALED, 0, 1, 0, (T)0, divisor);
        FN(add_identity)(out, omin, omax, (T)1);
    }
}

/* FTFMExp (operations/intermediate/ft_fmexp.hpp:66-88), literal Horner order:
 * out = mult; for d=N..1: out = mult + (1/d) out (x) exponent_[1, N-d+1]. */
void FN(orc_ft_fmexp)(const orc_basis* B, T* scratch, T* out, int omin,
                      int omax, const T* mult, int mmin, int mmax,
                      const T* ex, int emin, int emax)
{
    const int N = B->depth;
    FN(assign)(B, out, omin, omax, mult, mmin, mmax);
    for (int d = N; d > 0; --d) {
        int tmin, tmax;
        trunc_window(emin, emax, 1, N - d + 1, &tmin, &tmax);
        const T divisor = (T)1 / (T)d;
        FN(multiply)(B, scratch, out, omin, omax, out, omin, omax, ex, tmin,
                     tmax, W_ADD, mult, mmin, mmax, (T)1, divisor);
    }
}

/* FTLog (operations/intermediate/ft_log.hpp:70-84), literal Horner order. */
void FN(orc_ft_log)(const orc_basis* B, T* scratch, T* out, int omin, int omax,
                    const T* arg, int amin, int amax)
{
    const int N = B->depth;
    FN(set_constant)(B, out, omin, omax, (T)0);
    for (int d = N; d > 0; --d) {
        const T val = (d % 2 == 0 ? (T)-1 : (T)1) / (T)d;
        FN(add_identity)(out, omin, omax, val);
        int o_lo, o_hi, a_lo, a_hi;
        trunc_window(omin, omax, 0, N - d + 1, &o_lo, &o_hi);
        trunc_window(amin, amax, 1, N - d + 1, &a_lo, &a_hi);
        FN(multiply)(B, scratch, out, o_lo, o_hi, out, o_lo, o_hi, arg, a_lo,
                     a_hi, W_SCALED, 0, 1, 0, (T)0, (T)1);
    }
}

/* TensorPairing (cpu/.../basic/tensor_pairing.hpp:43-65): sum over the
 * overlap of the two windows; 0 when it is empty. */
T FN(orc_tensor_pairing)(const orc_basis* B, const T* f, int fmin, int fmax,
                         const T* x, int xmin, int xmax)
{
    int lo = fmin > xmin ? fmin : xmin;
    int hi = fmax < xmax ? fmax : xmax;
    T acc = (T)0;
    if (hi < lo) return acc;
    for (long long i = B->start[lo]; i < B->start[hi + 1]; ++i) acc += f[i] * x[i];
    return acc;
}

/* FTAdjLMul (cpu/.../basic/ft_adj_lmul.hpp:40-115): out[w] = sum_u op[u] arg[uw]
 * with argument-degree-major traversal and per-step rounding to T. */
void FN(orc_ft_adj_lmul)(const orc_basis* B, T* out, int omin, int omax,
                         const T* op, int pmin, int pmax, const T* arg,
                         int gmin, int gmax)
{
    FN(set_constant)(B, out, omin, omax, (T)0);
    int gmx = gmax < omax + pmax ? gmax : omax + pmax;
    int gmn = gmin > omin + pmin ? gmin : omin + pmin;
    if (gmn > gmx) return;
    for (int ad = gmx; ad >= gmn; --ad) {
        int olo = ad - pmax > omin ? ad - pmax : omin;
        int ohi = ad - pmin < omax ? ad - pmin : omax;
        const T* A = arg + B->start[ad];
        for (int od = ohi; od >= olo; --od) {
            const int pd = ad - od;
            const T* P = op + B->start[pd];
            T* O = out + B->start[od];
            const long long osz = B->pow[od], psz = B->pow[pd];
            for (long long u = 0; u < psz; ++u) {
                const T pv = P[u];
                for (long long i = 0; i < osz; ++i) O[i] = O[i] + pv * A[i + u * osz];
            }
        }
    }
}

static void FN(antipode)(const orc_basis* B, T* out, int omin, int omax, const T* arg, int amin, int amax,
                         int sign_by_degree);

/* FTAdjRMul (cpu/.../basic/ft_adj_rmul.hpp:102-134), by the reference's own
 * route: zeroed scratch, antipode of op and of arg, adj_lmul of those into the
 * scratch, antipode of the result into out, each over the windows of out, op
 * and arg.  The antipodes only flip signs and reverse words, which rounding
 * commutes with, so the sums are those of adj_lmul on the reversed words:
 * out[w] = sum_u arg[wu] op[u], term for term as the reference forms it. */
void FN(orc_ft_adj_rmul)(const orc_basis* B, T* out, int omin, int omax,
                         const T* op, int pmin, int pmax, const T* arg,
                         int gmin, int gmax)
{
    const long long dim = B->start[B->depth + 1];
    T* ws = (T*)calloc((size_t)(3 * dim), sizeof(T));
    if (!ws) return;
    T* ow = ws;
    T* pw = ws + dim;
    T* aw = ws + 2 * dim;
    FN(antipode)(B, pw, pmin, pmax, op, pmin, pmax, 1);
    FN(antipode)(B, aw, gmin, gmax, arg, gmin, gmax, 1);
    FN(orc_ft_adj_lmul)(B, ow, omin, omax, pw, pmin, pmax, aw, gmin, gmax);
    FN(antipode)(B, out, omin, omax, ow, omin, omax, 1);
    free(ws);
}

/* SparseMatrixVectorProduct, CSR rows (cpu/.../linalg/sparse_matrix_vector.hpp
 * :45-73).  out_begin/arg_begin are view.begin(): row 0 is out_begin. */
void FN(orc_spmv_csr)(T* out_begin, const T* arg_begin, const T* values,
                      const long long* inner, const long long* offsets,
                      long long rows, T alpha)
{
    for (long long r = 0; r < rows; ++r) {
        T acc = (T)0;
        for (long long e = offsets[r]; e < offsets[r + 1]; ++e)
            acc += values[e] * arg_begin[inner[e]];
        out_begin[r] = alpha * acc;
    }
}

/* CSC scatter (sparse_matrix_vector.hpp:133-163): zero the out window first. */
void FN(orc_spmv_csc)(T* out_begin, long long out_size, const T* arg_begin,
                      const T* values, const long long* inner,
                      const long long* offsets, long long cols, T alpha)
{
    for (long long i = 0; i < out_size; ++i) out_begin[i] = (T)0;
    for (long long c = 0; c < cols; ++c) {
        const T av = alpha * arg_begin[c];
        for (long long e = offsets[c]; e < offsets[c + 1]; ++e)
            out_begin[inner[e]] = out_begin[inner[e]] + values[e] * av;
    }
}

/* --- batched wrappers (the strategy's serial batch loop,
 * cpu/single_thread/strategy.hpp:115-124) ---------------------------------- */

#define OBJ(p, s, i) ((p) + (long long)(s) * (long long)(i))

int FN(orc_batch_ft_mul)(int width, int depth, long long n, T* out,
                         long long os, int omin, int omax, const T* lhs,
                         long long ls, int lmin, int lmax, const T* rhs,
                         long long rs, int rmin, int rmax, T beta)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    T* scratch = (T*)malloc(sizeof(T) * (size_t)(B.pow[depth] + 1));
    for (long long i = 0; i < n; ++i)
        FN(orc_ft_mul)(&B, scratch, OBJ(out, os, i), omin, omax, OBJ(lhs, ls, i),
                       lmin, lmax, OBJ(rhs, rs, i), rmin, rmax, beta);
    free(scratch);
    return 0;
}

int FN(orc_batch_ft_fma)(int width, int depth, long long n, T* out,
                         long long os, int omin, int omax, const T* a,
                         long long as, int amin, int amax, const T* b,
                         long long bs, int bmin, int bmax, const T* c,
                         long long cs, int cmin, int cmax, T alpha, T beta)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    T* scratch = (T*)malloc(sizeof(T) * (size_t)(B.pow[depth] + 1));
    for (long long i = 0; i < n; ++i)
        FN(orc_ft_fma)(&B, scratch, OBJ(out, os, i), omin, omax, OBJ(a, as, i),
                       amin, amax, OBJ(b, bs, i), bmin, bmax, OBJ(c, cs, i),
                       cmin, cmax, alpha, beta);
    free(scratch);
    return 0;
}

int FN(orc_batch_ft_inplace_fma)(int variant, int width, int depth,
                                 long long n, T* a, long long as, int amin,
                                 int amax, const T* b, long long bs, int bmin,
                                 int bmax, const T* c, long long cs, int cmin,
                                 int cmax, T alpha, T beta)
{
    orc_basis B;
    if (variant != 231 && variant != 123 && variant != 213) return 3;
    if (orc_basis_init(&B, width, depth)) return 3;
    T* scratch = (T*)malloc(sizeof(T) * (size_t)(B.pow[depth] + 1));
    for (long long i = 0; i < n; ++i)
        FN(orc_ft_inplace_fma)(&B, scratch, variant, OBJ(a, as, i), amin, amax,
                               OBJ(b, bs, i), bmin, bmax, OBJ(c, cs, i), cmin,
                               cmax, alpha, beta);
    free(scratch);
    return 0;
}

int FN(orc_batch_ft_inplace_mul)(int width, int depth, long long n, T* lhs,
                                 long long ls, int lmin, int lmax,
                                 const T* rhs, long long rs, int rmin,
                                 int rmax, T beta)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    T* scratch = (T*)malloc(sizeof(T) * (size_t)(B.pow[depth] + 1));
    for (long long i = 0; i < n; ++i)
        FN(orc_ft_inplace_mul)(&B, scratch, OBJ(lhs, ls, i), lmin, lmax,
                               OBJ(rhs, rs, i), rmin, rmax, beta);
    free(scratch);
    return 0;
}

int FN(orc_batch_ft_exp)(int width, int depth, long long n, T* out,
                         long long os, int omin, int omax, const T* arg,
                         long long as, int amin, int amax)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    T* scratch = (T*)malloc(sizeof(T) * (size_t)(B.pow[depth] + 1));
    for (long long i = 0; i < n; ++i)
        FN(orc_ft_exp)(&B, scratch, OBJ(out, os, i), omin, omax,
                       OBJ(arg, as, i), amin, amax);
    free(scratch);
    return 0;
}

int FN(orc_batch_ft_fmexp)(int width, int depth, long long n, T* out,
                           long long os, int omin, int omax, const T* mult,
                           long long ms, int mmin, int mmax, const T* ex,
                           long long es, int emin, int emax)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    T* scratch = (T*)malloc(sizeof(T) * (size_t)(B.pow[depth] + 1));
    for (long long i = 0; i < n; ++i)
        FN(orc_ft_fmexp)(&B, scratch, OBJ(out, os, i), omin, omax,
                         OBJ(mult, ms, i), mmin, mmax, OBJ(ex, es, i), emin,
                         emax);
    free(scratch);
    return 0;
}

int FN(orc_batch_ft_log)(int width, int depth, long long n, T* out,
                         long long os, int omin, int omax, const T* arg,
                         long long as, int amin, int amax)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    T* scratch = (T*)malloc(sizeof(T) * (size_t)(B.pow[depth] + 1));
    for (long long i = 0; i < n; ++i)
        FN(orc_ft_log)(&B, scratch, OBJ(out, os, i), omin, omax,
                       OBJ(arg, as, i), amin, amax);
    free(scratch);
    return 0;
}

/* TensorGeneralisedAntipode (cpu/.../basic/tensor_generalised_antipode.hpp:40-70):
 * on each degree k of out's window, out[reverse(w)] = s_k * arg[w] with
 * s_k = (-1)^k (antipode) or 1 (reflect); 0 where arg lacks degree k. */
static void FN(antipode)(const orc_basis* B, T* out, int omin, int omax, const T* arg, int amin, int amax,
                         int sign_by_degree)
{
    for (int k = omin; k <= omax; ++k) {
        T* o = out + B->start[k];
        const long long nk = B->pow[k];
        if (k < amin || k > amax) {
            for (long long i = 0; i < nk; ++i) o[i] = (T)0;
            continue;
        }
        const T* a = arg + B->start[k];
        const T s = (sign_by_degree && (k % 2)) ? (T)-1 : (T)1;
        for (long long i = 0; i < nk; ++i) {
            long long r = 0, v = i;
            for (int d = 0; d < k; ++d) {
                r = r * B->width + v % B->width;
                v /= B->width;
            }
            o[r] = a[i] * s;
        }
    }
}

int FN(orc_batch_tensor_antipode)(int sign_by_degree, int width, int depth, long long n, T* out, long long os,
                                  int omin, int omax, const T* arg, long long as, int amin, int amax)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    for (long long i = 0; i < n; ++i)
        FN(antipode)(&B, OBJ(out, os, i), omin, omax, OBJ(arg, as, i), amin, amax, sign_by_degree);
    return 0;
}

/* --- linear operations over any graded basis ------------------------------
 * `start` holds the degree offsets start_of_degree(k), k = 0..depth+1, of a
 * tensor or Lie basis. */

/* VectorAdd (cpu/.../linalg/vector_add.hpp:36-66): on every degree of out's
 * window out = alpha*lhs + beta*rhs, each operand contributing only on its own
 * window; the sum starts from 0 and adds the two products in that order. */
static void FN(vector_add)(const orc_basis* G, T* out, int omin, int omax, const T* lhs, int lmin, int lmax,
                           const T* rhs, int rmin, int rmax, T alpha, T beta)
{
    for (int k = omin; k <= omax; ++k) {
        const int hl = lmin <= k && k <= lmax, hr = rmin <= k && k <= rmax;
        for (long long i = G->start[k]; i < G->start[k + 1]; ++i) {
            T v = (T)0;
            if (hl) v = v + alpha * lhs[i];
            if (hr) v = v + beta * rhs[i];
            out[i] = v;
        }
    }
}

/* VectorInplaceAdd (cpu/.../linalg/vector_inplace_add.hpp:36-63): lhs +=
 * alpha*rhs on the overlap of the two windows. */
static void FN(vector_inplace_add)(const orc_basis* G, T* lhs, int lmin, int lmax, const T* rhs, int rmin,
                                   int rmax, T alpha)
{
    const int lo = lmin > rmin ? lmin : rmin, hi = lmax < rmax ? lmax : rmax;
    if (hi < lo) return;
    for (long long i = G->start[lo]; i < G->start[hi + 1]; ++i) lhs[i] = lhs[i] + alpha * rhs[i];
}

int FN(orc_batch_vector_assign)(const long long* start, int depth, long long n, T* out, long long os, int omin,
                                int omax, const T* arg, long long as, int amin, int amax)
{
    orc_basis G;
    if (orc_graded_init(&G, start, depth)) return 3;
    for (long long i = 0; i < n; ++i)
        FN(assign)(&G, OBJ(out, os, i), omin, omax, OBJ(arg, as, i), amin, amax);
    return 0;
}

int FN(orc_batch_vector_add)(const long long* start, int depth, long long n, T* out, long long os, int omin,
                             int omax, const T* lhs, long long ls, int lmin, int lmax, const T* rhs,
                             long long rs, int rmin, int rmax, T alpha, T beta)
{
    orc_basis G;
    if (orc_graded_init(&G, start, depth)) return 3;
    for (long long i = 0; i < n; ++i)
        FN(vector_add)(&G, OBJ(out, os, i), omin, omax, OBJ(lhs, ls, i), lmin, lmax, OBJ(rhs, rs, i), rmin,
                       rmax, alpha, beta);
    return 0;
}

int FN(orc_batch_vector_inplace_add)(const long long* start, int depth, long long n, T* lhs, long long ls,
                                     int lmin, int lmax, const T* rhs, long long rs, int rmin, int rmax,
                                     T alpha)
{
    orc_basis G;
    if (orc_graded_init(&G, start, depth)) return 3;
    for (long long i = 0; i < n; ++i)
        FN(vector_inplace_add)(&G, OBJ(lhs, ls, i), lmin, lmax, OBJ(rhs, rs, i), rmin, rmax, alpha);
    return 0;
}

/* op 0: VectorScalarMultiply (cpu/.../linalg/vector_scalar_multiply.hpp:37-49),
 * 1: VectorSetConstant, 2: TensorSetIdentity, 3: TensorAddIdentity. */
int FN(orc_batch_vector_scalar)(int op, const long long* start, int depth, long long n, T* vec, long long vs,
                                int vmin, int vmax, T scalar)
{
    orc_basis G;
    if (orc_graded_init(&G, start, depth)) return 3;
    for (long long b = 0; b < n; ++b) {
        T* v = OBJ(vec, vs, b);
        switch (op) {
            case 0:
                if (vmin <= vmax)
                    for (long long i = G.start[vmin]; i < G.start[vmax + 1]; ++i) v[i] = v[i] * scalar;
                break;
            case 1: FN(set_constant)(&G, v, vmin, vmax, scalar); break;
            case 2: FN(set_identity)(&G, v, vmin, vmax, scalar); break;
            case 3: FN(add_identity)(v, vmin, vmax, scalar); break;
            default: return 3;
        }
    }
    return 0;
}

/* --- shuffle and half-shuffle products -------------------------------------
 * A word of degree k is held as its letters lt[0..k-1], first letter first.
 * A subset S of its positions is a mask whose bit (k-1-p) marks position p,
 * and masks run in ascending order, as in
 * operations/implementation/word_shuffle_product.hpp:49-93 with
 * basis/tensor_basis.hpp:58-91.  `pre` >= 0 prepends that letter to the
 * first factor's word (the left half-shuffle,
 * word_half_shuffle_product.hpp:20-66). */

static void FN(letters_of)(const orc_basis* B, int k, long long idx, int* lt)
{
    for (int p = k - 1; p >= 0; --p) {
        lt[p] = (int)(idx % B->width);
        idx /= B->width;
    }
}

/* sum over S of l[(pre) w|S] r[w|S^c], each factor only on its window */
static T FN(shuffle_word)(const orc_basis* B, const int* lt, int k, int pre, const T* l, int lmin, int lmax,
                          const T* r, int rmin, int rmax)
{
    T acc = (T)0;
    const int extra = pre >= 0 ? 1 : 0;
    for (unsigned long long mask = 0; mask < (1ULL << k); ++mask) {
        int ld = 0, rd = 0;
        long long li = 0, ri = 0;
        for (int p = 0; p < k; ++p) {
            if ((mask >> (k - 1 - p)) & 1ULL) {
                ++ld;
                li = li * B->width + lt[p];
            } else {
                ++rd;
                ri = ri * B->width + lt[p];
            }
        }
        if (ld + extra < lmin || ld + extra > lmax || rd < rmin || rd > rmax) continue;
        const long long lidx = pre >= 0 ? pre * B->pow[ld] + li : li;
        acc = acc + l[B->start[ld + extra] + lidx] * r[B->start[rd] + ri];
    }
    return acc;
}

/* The adjoint sum of out word u (letters ul, degree m): over op degrees
 * j in [jlo, jhi] and op words v of degree j, the sum over the masks of weight
 * j on m+j positions (bit set: the next letter of v, else the next letter of
 * u) of op[(pre) v] * arg[(pre) merge], each (j, v) summed on its own first
 * (operations/implementation/word_shuffle_adjoint_multiply.hpp:22-95,
 * shuffle_adjoint_op_loop.hpp:20-60). */
static T FN(adjoint_word)(const orc_basis* B, const int* ul, int m, int pre, const T* op, const T* arg, int jlo,
                          int jhi)
{
    T acc = (T)0;
    int vl[ORC_MAX_DEPTH + 1];
    const int extra = pre >= 0 ? 1 : 0;
    for (int j = jlo; j <= jhi; ++j) {
        const int L = m + j;
        for (long long v = 0; v < B->pow[j]; ++v) {
            FN(letters_of)(B, j, v, vl);
            const T opv = op[B->start[j + extra] + (pre >= 0 ? pre * B->pow[j] : 0) + v];
            T inner = (T)0;
            for (unsigned long long mask = 0; mask < (1ULL << L); ++mask) {
                if (__builtin_popcountll(mask) != j) continue;
                int iu = 0, iv = 0;
                long long idx = 0;
                for (int p = 0; p < L; ++p) {
                    const int letter = ((mask >> (L - 1 - p)) & 1ULL) ? vl[iv++] : ul[iu++];
                    idx = idx * B->width + letter;
                }
                inner = inner + opv * arg[B->start[L + extra] + (pre >= 0 ? pre * B->pow[L] : 0) + idx];
            }
            acc = acc + inner;
        }
    }
    return acc;
}

/* STMul / STFma / STInplaceFma (cpu/.../basic/st_mul.hpp:37-66,
 * st_fma.hpp:41-79, st_inplace_fma.hpp:38-68).  kind 0: out = beta b.c;
 * 1: out = alpha a + beta b.c; 2: a = alpha a + beta b.c (out == a). */
static void FN(st_fma_obj)(const orc_basis* B, int kind, T* out, int omin, int omax, const T* a, int amin, int amax,
                           const T* b, int bmin, int bmax, const T* c, int cmin, int cmax, T alpha, T beta)
{
    int lt[ORC_MAX_DEPTH + 1];
    const int b0 = bmin <= 0 && 0 <= bmax, c0 = cmin <= 0 && 0 <= cmax, a0 = amin <= 0 && 0 <= amax;
    if (omin == 0 && omax >= 0) {
        T v = (T)0;
        if (kind == 0) {
            if (b0 && c0) v = beta * b[0] * c[0];
        } else if (kind == 1) {
            if (a0) v = v + alpha * a[0];
            if (b0 && c0) v = v + beta * b[0] * c[0];
        } else {
            v = alpha * out[0];
            if (b0 && c0) v = v + beta * b[0] * c[0];
        }
        out[0] = v;
    }
    for (int k = omin > 1 ? omin : 1; k <= omax; ++k) {
        for (long long i = 0; i < B->pow[k]; ++i) {
            FN(letters_of)(B, k, i, lt);
            const T ws = FN(shuffle_word)(B, lt, k, -1, b, bmin, bmax, c, cmin, cmax);
            T* o = out + B->start[k] + i;
            if (kind == 0) {
                *o = beta * ws;
            } else if (kind == 1) {
                T v = (T)0;
                if (amin <= k && k <= amax) v = v + alpha * a[B->start[k] + i];
                *o = v + beta * ws;
            } else {
                *o = alpha * *o + beta * ws;
            }
        }
    }
}

/* STAdjMul (cpu/.../basic/st_adj_mul.hpp:43-82): <out, x> = <arg, op . x>. */
static void FN(st_adj_mul_obj)(const orc_basis* B, T* out, int omin, int omax, const T* op, int pmin, int pmax,
                               const T* arg, int gmin, int gmax)
{
    int ul[ORC_MAX_DEPTH + 1];
    for (int m = omin; m <= omax; ++m) {
        const int jlo = pmin > gmin - m ? pmin : gmin - m;
        const int jhi = pmax < gmax - m ? pmax : gmax - m;
        for (long long i = 0; i < B->pow[m]; ++i) {
            FN(letters_of)(B, m, i, ul);
            out[B->start[m] + i] = FN(adjoint_word)(B, ul, m, -1, op, arg, jlo, jhi);
        }
    }
}

/* LeftHSMul (cpu/.../basic/left_hs_mul.hpp:37-84): out[a w] = beta sum over
 * S of x[a w|S] y[w|S^c]; degree 0 of out is 0. */
static void FN(left_hs_mul_obj)(const orc_basis* B, T* out, int omin, int omax, const T* x, int xmin, int xmax,
                                const T* y, int ymin, int ymax, T beta)
{
    int lt[ORC_MAX_DEPTH + 1];
    if (omin == 0 && omax >= 0) out[0] = (T)0;
    for (int k = omin > 1 ? omin : 1; k <= omax; ++k) {
        for (long long i = 0; i < B->pow[k]; ++i) {
            FN(letters_of)(B, k, i, lt);
            out[B->start[k] + i] = beta * FN(shuffle_word)(B, lt + 1, k - 1, lt[0], x, xmin, xmax, y, ymin, ymax);
        }
    }
}

/* LeftHSAdjLMul (cpu/.../basic/left_hs_adj_lmul.hpp:43-118):
 * <out, z> = <y, x < z>.  out[u] = sum over prefixes a of the adjoint sum
 * with op x[a v] and arg y[a merge], scaled by beta as the reference's GPU
 * operator does (gpu/block/operations/basic/left_hs_adj_lmul.hpp:114-177;
 * the CPU operator has no beta, i.e. beta = 1). */
static void FN(left_hs_adj_lmul_obj)(const orc_basis* B, T* out, int omin, int omax, const T* x, int xmin, int xmax,
                                     const T* y, int ymin, int ymax, T beta)
{
    int ul[ORC_MAX_DEPTH + 1];
    for (int m = omin; m <= omax; ++m) {
        int jlo = xmin - 1 > ymin - 1 - m ? xmin - 1 : ymin - 1 - m;
        if (jlo < 0) jlo = 0;
        const int jhi = xmax - 1 < ymax - 1 - m ? xmax - 1 : ymax - 1 - m;
        for (long long i = 0; i < B->pow[m]; ++i) {
            FN(letters_of)(B, m, i, ul);
            T acc = (T)0;
            for (int a = 0; a < B->width; ++a) acc = acc + FN(adjoint_word)(B, ul, m, a, x, y, jlo, jhi);
            out[B->start[m] + i] = beta * acc;
        }
    }
}

/* LeftHSAdjRMul (cpu/.../basic/left_hs_adj_rmul.hpp:37-103):
 * <out, z> = <x, z < y>.  out[a u] = beta times the adjoint sum of u with op
 * y[v] and arg x[a merge] (beta as for the left adjoint); degree 0 of out is
 * 0. */
static void FN(left_hs_adj_rmul_obj)(const orc_basis* B, T* out, int omin, int omax, const T* x, int xmin, int xmax,
                                     const T* y, int ymin, int ymax, T beta)
{
    int ul[ORC_MAX_DEPTH + 1];
    if (omin == 0 && omax >= 0) out[0] = (T)0;
    for (int m = omin > 1 ? omin : 1; m <= omax; ++m) {
        const int jlo = ymin > xmin - m ? ymin : xmin - m;
        const int jhi = ymax < xmax - m ? ymax : xmax - m;
        for (long long i = 0; i < B->pow[m]; ++i) {
            FN(letters_of)(B, m, i, ul);
            /* the prefix letter goes to the arg side; op = y carries none */
            T acc = (T)0;
            for (int j = jlo; j <= jhi; ++j) {
                const int L = m - 1 + j;
                int vl[ORC_MAX_DEPTH + 1];
                for (long long v = 0; v < B->pow[j]; ++v) {
                    FN(letters_of)(B, j, v, vl);
                    const T opv = y[B->start[j] + v];
                    T inner = (T)0;
                    for (unsigned long long mask = 0; mask < (1ULL << L); ++mask) {
                        if (__builtin_popcountll(mask) != j) continue;
                        int iu = 1, iv = 0;
                        long long idx = 0;
                        for (int p = 0; p < L; ++p) {
                            const int letter = ((mask >> (L - 1 - p)) & 1ULL) ? vl[iv++] : ul[iu++];
                            idx = idx * B->width + letter;
                        }
                        inner = inner + opv * x[B->start[L + 1] + ul[0] * B->pow[L] + idx];
                    }
                    acc = acc + inner;
                }
            }
            out[B->start[m] + i] = beta * acc;
        }
    }
}

/* batched forms; st ops: kind as st_fma_obj, out == a for kind 2 */
int FN(orc_batch_st_fma)(int kind, int width, int depth, long long n, T* out, long long os, int omin, int omax,
                         const T* a, long long as, int amin, int amax, const T* b, long long bs, int bmin, int bmax,
                         const T* c, long long cs, int cmin, int cmax, T alpha, T beta)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    for (long long i = 0; i < n; ++i)
        FN(st_fma_obj)(&B, kind, OBJ(out, os, i), omin, omax, a ? OBJ(a, as, i) : 0, amin, amax, OBJ(b, bs, i), bmin,
                       bmax, OBJ(c, cs, i), cmin, cmax, alpha, beta);
    return 0;
}

/* op 0: st_adj_mul(out, op=p, arg=q); 1: left_hs_mul(out, x=p, y=q, beta);
 * 2: left_hs_adj_lmul(out, x=p, y=q, beta); 3: left_hs_adj_rmul(out, x=p, y=q, beta) */
int FN(orc_batch_shuffle2)(int opcode, int width, int depth, long long n, T* out, long long os, int omin, int omax,
                           const T* p, long long ps, int pmin, int pmax, const T* q, long long qs, int qmin, int qmax,
                           T beta)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    for (long long i = 0; i < n; ++i) {
        T* o = OBJ(out, os, i);
        const T* P = OBJ(p, ps, i);
        const T* Q = OBJ(q, qs, i);
        switch (opcode) {
            case 0: FN(st_adj_mul_obj)(&B, o, omin, omax, P, pmin, pmax, Q, qmin, qmax); break;
            case 1: FN(left_hs_mul_obj)(&B, o, omin, omax, P, pmin, pmax, Q, qmin, qmax, beta); break;
            case 2: FN(left_hs_adj_lmul_obj)(&B, o, omin, omax, P, pmin, pmax, Q, qmin, qmax, beta); break;
            case 3: FN(left_hs_adj_rmul_obj)(&B, o, omin, omax, P, pmin, pmax, Q, qmin, qmax, beta); break;
            default: return 3;
        }
    }
    return 0;
}

int FN(orc_batch_tensor_pairing)(int width, int depth, long long n, T* out,
                                 long long ostride, const T* f, long long fs,
                                 int fmin, int fmax, const T* x, long long xs,
                                 int xmin, int xmax)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    for (long long i = 0; i < n; ++i)
        out[ostride * i] = FN(orc_tensor_pairing)(&B, OBJ(f, fs, i), fmin, fmax,
                                                  OBJ(x, xs, i), xmin, xmax);
    return 0;
}

int FN(orc_batch_ft_adj_lmul)(int width, int depth, long long n, T* out,
                              long long os, int omin, int omax, const T* op,
                              long long ps, int pmin, int pmax, const T* arg,
                              long long gs, int gmin, int gmax)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    for (long long i = 0; i < n; ++i)
        FN(orc_ft_adj_lmul)(&B, OBJ(out, os, i), omin, omax, OBJ(op, ps, i),
                            pmin, pmax, OBJ(arg, gs, i), gmin, gmax);
    return 0;
}

int FN(orc_batch_ft_adj_rmul)(int width, int depth, long long n, T* out,
                              long long os, int omin, int omax, const T* op,
                              long long ps, int pmin, int pmax, const T* arg,
                              long long gs, int gmin, int gmax)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    for (long long i = 0; i < n; ++i)
        FN(orc_ft_adj_rmul)(&B, OBJ(out, os, i), omin, omax, OBJ(op, ps, i),
                            pmin, pmax, OBJ(arg, gs, i), gmin, gmax);
    return 0;
}

/* Batched SpMV over graded views: out_begin = out + os*i + out_offset, etc. */
int FN(orc_batch_spmv)(int csc, long long n, T* out, long long os,
                       long long out_offset, long long out_size, const T* arg,
                       long long as, long long arg_offset, const T* values,
                       const long long* inner, const long long* offsets,
                       long long outer, T alpha)
{
    for (long long i = 0; i < n; ++i) {
        T* o = OBJ(out, os, i) + out_offset;
        const T* a = OBJ(arg, as, i) + arg_offset;
        if (csc)
            FN(orc_spmv_csc)(o, out_size, a, values, inner, offsets, outer, alpha);
        else
            FN(orc_spmv_csr)(o, a, values, inner, offsets, outer, alpha);
    }
    return 0;
}

/* Signature of piecewise-linear paths by the caller-side Chen recursion the
 * reference leaves to its users: S_0 = 1, S_t = ft_fmexp(S_{t-1}, dX_t) with
 * the increment held in a full tensor whose only non-zero degree is 1 and the
 * exponent view (1,1).  Increments are packed [sum(len), width] row-major,
 * path p owning rows offsets[p]..offsets[p+1]. */
int FN(orc_signature)(int width, int depth, long long n_paths,
                      const long long* offsets, const T* increments, T* out,
                      long long os)
{
    orc_basis B;
    if (orc_basis_init(&B, width, depth)) return 3;
    const long long dim = B.start[depth + 1];
    T* scratch = (T*)malloc(sizeof(T) * (size_t)(B.pow[depth] + 1));
    T* a = (T*)malloc(sizeof(T) * (size_t)dim);
    T* b = (T*)malloc(sizeof(T) * (size_t)dim);
    T* e = (T*)calloc((size_t)(dim + width), sizeof(T)); /* slots 1..width even at depth 0 */
    for (long long p = 0; p < n_paths; ++p) {
        for (long long i = 0; i < dim; ++i) a[i] = (T)0;
        a[0] = (T)1;
        T* cur = a;
        T* nxt = b;
        for (long long t = offsets[p]; t < offsets[p + 1]; ++t) {
            for (int l = 0; l < width; ++l) e[1 + l] = increments[t * width + l];
            FN(orc_ft_fmexp)(&B, scratch, nxt, 0, depth, cur, 0, depth, e, 1, 1);
            T* tmp = cur; cur = nxt; nxt = tmp;
        }
        T* o = out + os * p;
        for (long long i = 0; i < dim; ++i) o[i] = cur[i];
    }
    free(scratch); free(a); free(b); free(e);
    return 0;
}

#undef OBJ
#undef FN
#undef CAT
#undef CAT2
