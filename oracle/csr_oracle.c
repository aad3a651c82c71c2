/*
 * csr_oracle.c -- CPU restatement of FlagSparse's CSR SpMV / SpMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker
 * (or the timed CPU baseline).  The product path never calls it.
 *
 * Everything here works on widened host arrays: values as double (fp32 -> fp64
 * is exact), complex values as interleaved (re, im) doubles, and every index
 * array as int64.  The Python wrapper (oracle/oracle.py) does the widening.
 *
 * Reference behaviour restated (paths relative to the reference repository):
 *   - SpMV y = A x over stored entries, duplicates and unsorted columns summed
 *     as stored:   src/flagsparse/sparse_operations/spmv_csr.py:158-184
 *     (the default path runs fp32 inputs in fp64: spmv_csr.py:116-122, 664-689)
 *   - complex SpMV: spmv_csr.py:187-219
 *   - SpMM C = A B, row-major B and C: spmm_csr.py:92-134, 525-599
 *   - transpose for op=trans/conj (stable argsort by column):
 *     spmv_csr.py:456-494
 *   - SpMV opt row buckets: spmv_csr.py:128-153, 345-391
 *   - SpMM opt buckets and 256-nnz split parts: spmm_csr.py:395-477
 *   - SpMM alg2 buckets: spmm_csr_opt_alg2.py:60-166, 307-335
 *
 * Accumulation order is the plain sequential CSR order (entry ptr[r] first),
 * which is also SciPy's csr_matvec / csr_matvecs order; the tests pin the fp64
 * results bit-for-bit against SciPy and against exactly rounded sums.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

/* ------------------------------------------------------------------ SpMV */

/* y[r] = sum_j data[j] * x[indices[j]], j in [indptr[r], indptr[r+1]). */
void orc_spmv_f64(int64_t n_rows, const int64_t *indptr, const int64_t *indices,
                  const double *data, const double *x, double *y)
{
    for (int64_t r = 0; r < n_rows; ++r) {
        double acc = 0.0;
        for (int64_t j = indptr[r]; j < indptr[r + 1]; ++j)
            acc += data[j] * x[indices[j]];
        y[r] = acc;
    }
}

/* Condition denominator: s[r] = sum_j |data[j] * x[indices[j]]|. */
void orc_spmv_abs_f64(int64_t n_rows, const int64_t *indptr, const int64_t *indices,
                      const double *data, const double *x, double *s)
{
    for (int64_t r = 0; r < n_rows; ++r) {
        double acc = 0.0;
        for (int64_t j = indptr[r]; j < indptr[r + 1]; ++j)
            acc += fabs(data[j] * x[indices[j]]);
        s[r] = acc;
    }
}

/* Complex SpMV, interleaved (re, im) pairs; same product formula as
 * spmv_csr.py:210-213: re = ar*xr - ai*xi, im = ar*xi + ai*xr. */
void orc_spmv_c128(int64_t n_rows, const int64_t *indptr, const int64_t *indices,
                   const double *data_ri, const double *x_ri, double *y_ri)
{
    for (int64_t r = 0; r < n_rows; ++r) {
        double re = 0.0, im = 0.0;
        for (int64_t j = indptr[r]; j < indptr[r + 1]; ++j) {
            double ar = data_ri[2 * j], ai = data_ri[2 * j + 1];
            int64_t c = indices[j];
            double xr = x_ri[2 * c], xi = x_ri[2 * c + 1];
            re += ar * xr - ai * xi;
            im += ar * xi + ai * xr;
        }
        y_ri[2 * r] = re;
        y_ri[2 * r + 1] = im;
    }
}

/* ------------------------------------------------------------------ SpMM */

/* C[r, :] = sum_j data[j] * B[indices[j], :]; B is (n_cols x n_dense) with
 * leading dimension ldb, C is (n_rows x n_dense) with leading dimension ldc. */
void orc_spmm_f64(int64_t n_rows, int64_t n_dense, const int64_t *indptr,
                  const int64_t *indices, const double *data, const double *B,
                  int64_t ldb, double *C, int64_t ldc)
{
    for (int64_t r = 0; r < n_rows; ++r) {
        double *c = C + r * ldc;
        for (int64_t k = 0; k < n_dense; ++k) c[k] = 0.0;
        for (int64_t j = indptr[r]; j < indptr[r + 1]; ++j) {
            const double a = data[j];
            const double *b = B + indices[j] * ldb;
            for (int64_t k = 0; k < n_dense; ++k) c[k] += a * b[k];
        }
    }
}

/* ----------------------------------------------------------------- SDDMM */

/* out[p] = alpha * sum_k x[row(p), k] * y[indices[p], k] + beta * data[p]
 * (data may be NULL: then no beta term), fp64, k in increasing order;
 * follows sddmm_csr.py:108-151 (per entry: dot of an x row and a y row) and
 * its torch reference sddmm_csr.py:423-437.  Also the condition
 * abs_out[p] = |alpha| * sum_k |x[row(p), k] * y[indices[p], k]| + |beta * data[p]|. */
void orc_sddmm_f64(int64_t n_rows, int64_t k_dim, const int64_t *indptr, const int64_t *indices,
                   const double *x, const double *y, const double *data, double alpha,
                   double beta, double *out, double *abs_out)
{
    for (int64_t r = 0; r < n_rows; ++r) {
        const double *xr = x + r * k_dim;
        for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
            const double *yc = y + indices[p] * k_dim;
            double s = 0.0, a = 0.0;
            for (int64_t k = 0; k < k_dim; ++k) {
                s += xr[k] * yc[k];
                a += fabs(xr[k] * yc[k]);
            }
            double v = alpha * s, c = fabs(alpha) * a;
            if (data) {
                v += beta * data[p];
                c += fabs(beta * data[p]);
            }
            out[p] = v;
            if (abs_out) abs_out[p] = c;
        }
    }
}

void orc_spmm_abs_f64(int64_t n_rows, int64_t n_dense, const int64_t *indptr,
                      const int64_t *indices, const double *data, const double *B,
                      int64_t ldb, double *S, int64_t lds)
{
    for (int64_t r = 0; r < n_rows; ++r) {
        double *s = S + r * lds;
        for (int64_t k = 0; k < n_dense; ++k) s[k] = 0.0;
        for (int64_t j = indptr[r]; j < indptr[r + 1]; ++j) {
            const double a = data[j];
            const double *b = B + indices[j] * ldb;
            for (int64_t k = 0; k < n_dense; ++k) s[k] += fabs(a * b[k]);
        }
    }
}

void orc_spmm_c128(int64_t n_rows, int64_t n_dense, const int64_t *indptr,
                   const int64_t *indices, const double *data_ri, const double *B_ri,
                   int64_t ldb, double *C_ri, int64_t ldc)
{
    for (int64_t r = 0; r < n_rows; ++r) {
        double *c = C_ri + 2 * r * ldc;
        for (int64_t k = 0; k < 2 * n_dense; ++k) c[k] = 0.0;
        for (int64_t j = indptr[r]; j < indptr[r + 1]; ++j) {
            const double ar = data_ri[2 * j], ai = data_ri[2 * j + 1];
            const double *b = B_ri + 2 * indices[j] * ldb;
            for (int64_t k = 0; k < n_dense; ++k) {
                const double br = b[2 * k], bi = b[2 * k + 1];
                c[2 * k] += ar * br - ai * bi;
                c[2 * k + 1] += ar * bi + ai * br;
            }
        }
    }
}

/* ------------------------------------------------------------ transpose */

/* CSR of A^T (n_cols x n_rows).  Entries with equal column keep their CSR
 * order, matching the stable argsort of spmv_csr.py:476-479.  perm[k] is the
 * source entry of transposed entry k (data_T = data[perm]). */
void orc_csr_transpose(int64_t n_rows, int64_t n_cols, const int64_t *indptr,
                       const int64_t *indices, int64_t *t_indptr, int64_t *t_indices,
                       int64_t *perm)
{
    const int64_t nnz = indptr[n_rows];
    for (int64_t c = 0; c <= n_cols; ++c) t_indptr[c] = 0;
    for (int64_t j = 0; j < nnz; ++j) t_indptr[indices[j] + 1] += 1;
    for (int64_t c = 0; c < n_cols; ++c) t_indptr[c + 1] += t_indptr[c];
    int64_t *cursor = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_cols > 0 ? n_cols : 1));
    for (int64_t c = 0; c < n_cols; ++c) cursor[c] = t_indptr[c];
    for (int64_t r = 0; r < n_rows; ++r) {
        for (int64_t j = indptr[r]; j < indptr[r + 1]; ++j) {
            const int64_t dst = cursor[indices[j]]++;
            t_indices[dst] = r;
            perm[dst] = j;
        }
    }
    free(cursor);
}

/* -------------------------------------------------------------- buckets */

/* SpMV opt bucket of a row of length len (spmv_csr.py:128-153, 345-391):
 * fp32 upper bounds {64, 512, 4096, inf}; fp64 {64, 256, 2048, 8192, inf}.
 * Empty rows go to the first bucket. */
int orc_spmv_opt_bucket(int64_t len, int fp64)
{
    static const int64_t f32_ub[] = {64, 512, 4096};
    static const int64_t f64_ub[] = {64, 256, 2048, 8192};
    const int64_t *ub = fp64 ? f64_ub : f32_ub;
    const int nb = fp64 ? 4 : 3;
    for (int b = 0; b < nb; ++b)
        if (len <= ub[b]) return b;
    return nb;
}

/* SpMM opt bucket (spmm_csr.py:395-401): upper bounds {32, 128, 512, 2048, inf}. */
int orc_spmm_opt_bucket(int64_t len)
{
    static const int64_t ub[] = {32, 128, 512, 2048};
    for (int b = 0; b < 4; ++b)
        if (len <= ub[b]) return b;
    return 4;
}

/* SpMM alg2 bucket (spmm_csr_opt_alg2.py:60-166, 307-335): labels short_16,
 * short_64, row_256, seg_1024, seg_long with inclusive ranges
 * [0,16], [17,64], [65,256], [257,1024], [1025, inf). */
int orc_spmm_alg2_bucket(int64_t len)
{
    if (len <= 16) return 0;
    if (len <= 64) return 1;
    if (len <= 256) return 2;
    if (len <= 1024) return 3;
    return 4;
}

/* Split the listed rows into parts of at most part_nnz entries
 * (spmm_csr.py:417-441).  Returns the number of parts; writes at most
 * max_parts of them.  row_part_ptr has n_long + 1 entries. */
int64_t orc_split_parts(int64_t n_long, const int64_t *long_rows, const int64_t *indptr,
                        int64_t part_nnz, int64_t max_parts, int64_t *part_rows,
                        int64_t *part_starts, int64_t *part_ends, int64_t *row_part_ptr)
{
    int64_t np = 0;
    row_part_ptr[0] = 0;
    for (int64_t i = 0; i < n_long; ++i) {
        const int64_t r = long_rows[i];
        int64_t cur = indptr[r];
        const int64_t end = indptr[r + 1];
        while (cur < end) {
            const int64_t nxt = (cur + part_nnz < end) ? cur + part_nnz : end;
            if (np < max_parts) {
                part_rows[np] = r;
                part_starts[np] = cur;
                part_ends[np] = nxt;
            }
            ++np;
            cur = nxt;
        }
        row_part_ptr[i + 1] = np;
    }
    return np;
}

/* ------------------------------------------------------------- features */

/* Sixteen row-length statistics used by the kernel selector.  Layout:
 *  0 log2(n_rows+1)        1 log2(n_cols+1)        2 log2(nnz+1)
 *  3 mean length           4 std of length         5 std/mean (0 if mean 0)
 *  6 log2(max length + 1)  7 min length            8 fraction of empty rows
 *  9..14 fractions of rows with length in (0,4], (4,16], (16,64], (64,256],
 *        (256,4096], (4096, inf)
 * 15 fraction of nnz held by rows longer than 256
 * Sums are taken in exact integer arithmetic (length, length^2, histogram) so
 * the device version is bit-reproducible. */
void orc_row_features(int64_t n_rows, int64_t n_cols, const int64_t *indptr, double *f)
{
    int64_t nnz = n_rows > 0 ? indptr[n_rows] - indptr[0] : 0;
    int64_t mx = 0, mn = n_rows > 0 ? INT64_MAX : 0, empty = 0, long_nnz = 0;
    int64_t hist[6] = {0, 0, 0, 0, 0, 0};
    unsigned __int128 sq = 0;
    for (int64_t r = 0; r < n_rows; ++r) {
        const int64_t len = indptr[r + 1] - indptr[r];
        if (len > mx) mx = len;
        if (len < mn) mn = len;
        sq += (unsigned __int128)len * (unsigned __int128)len;
        if (len == 0) ++empty;
        else if (len <= 4) ++hist[0];
        else if (len <= 16) ++hist[1];
        else if (len <= 64) ++hist[2];
        else if (len <= 256) ++hist[3];
        else if (len <= 4096) ++hist[4];
        else ++hist[5];
        if (len > 256) long_nnz += len;
    }
    const double n = (double)n_rows;
    const double mean = n_rows > 0 ? (double)nnz / n : 0.0;
    double var = 0.0;
    if (n_rows > 0) {
        var = (double)sq / n - mean * mean;
        if (var < 0.0) var = 0.0;
    }
    const double sd = sqrt(var);
    f[0] = log2((double)n_rows + 1.0);
    f[1] = log2((double)n_cols + 1.0);
    f[2] = log2((double)nnz + 1.0);
    f[3] = mean;
    f[4] = sd;
    f[5] = mean > 0.0 ? sd / mean : 0.0;
    f[6] = log2((double)mx + 1.0);
    f[7] = (double)mn;
    f[8] = n_rows > 0 ? (double)empty / n : 0.0;
    for (int b = 0; b < 6; ++b) f[9 + b] = n_rows > 0 ? (double)hist[b] / n : 0.0;
    f[15] = nnz > 0 ? (double)long_nnz / (double)nnz : 0.0;
}

/* ------------------------------------------------------------------ MLP */

/* Three-layer perceptron 16 -> H -> H -> K with ReLU hidden activations,
 * row-major weights W1[16][H], W2[H][H], W3[H][K].  Forward keeps the hidden
 * activations for the backward pass.  Loss = mean over the batch and outputs
 * of 0.5 * mask * (out - target)^2, mask in {0, 1} marks valid labels. */
void orc_mlp_forward(int64_t batch, int64_t in_dim, int64_t hid, int64_t out_dim,
                     const double *X, const double *W1, const double *b1,
                     const double *W2, const double *b2, const double *W3,
                     const double *b3, double *H1, double *H2, double *Y)
{
    for (int64_t s = 0; s < batch; ++s) {
        const double *x = X + s * in_dim;
        double *h1 = H1 + s * hid, *h2 = H2 + s * hid, *y = Y + s * out_dim;
        for (int64_t j = 0; j < hid; ++j) {
            double a = b1[j];
            for (int64_t i = 0; i < in_dim; ++i) a += x[i] * W1[i * hid + j];
            h1[j] = a > 0.0 ? a : 0.0;
        }
        for (int64_t j = 0; j < hid; ++j) {
            double a = b2[j];
            for (int64_t i = 0; i < hid; ++i) a += h1[i] * W2[i * hid + j];
            h2[j] = a > 0.0 ? a : 0.0;
        }
        for (int64_t k = 0; k < out_dim; ++k) {
            double a = b3[k];
            for (int64_t i = 0; i < hid; ++i) a += h2[i] * W3[i * out_dim + k];
            y[k] = a;
        }
    }
}

/* Gradients of the masked MSE loss defined above; returns the loss. */
double orc_mlp_backward(int64_t batch, int64_t in_dim, int64_t hid, int64_t out_dim,
                        const double *X, const double *W2, const double *W3,
                        const double *H1, const double *H2, const double *Y,
                        const double *T, const double *M, double *gW1, double *gb1,
                        double *gW2, double *gb2, double *gW3, double *gb3)
{
    memset(gW1, 0, sizeof(double) * (size_t)(in_dim * hid));
    memset(gb1, 0, sizeof(double) * (size_t)hid);
    memset(gW2, 0, sizeof(double) * (size_t)(hid * hid));
    memset(gb2, 0, sizeof(double) * (size_t)hid);
    memset(gW3, 0, sizeof(double) * (size_t)(hid * out_dim));
    memset(gb3, 0, sizeof(double) * (size_t)out_dim);
    const double scale = 1.0 / (double)(batch * out_dim);
    double loss = 0.0;
    double *d3 = (double *)malloc(sizeof(double) * (size_t)out_dim);
    double *d2 = (double *)malloc(sizeof(double) * (size_t)hid);
    double *d1 = (double *)malloc(sizeof(double) * (size_t)hid);
    for (int64_t s = 0; s < batch; ++s) {
        const double *x = X + s * in_dim, *h1 = H1 + s * hid, *h2 = H2 + s * hid;
        const double *y = Y + s * out_dim, *t = T + s * out_dim, *m = M + s * out_dim;
        for (int64_t k = 0; k < out_dim; ++k) {
            const double e = m[k] * (y[k] - t[k]);
            loss += 0.5 * e * e;
            d3[k] = e * scale;
        }
        for (int64_t i = 0; i < hid; ++i) {
            double a = 0.0;
            for (int64_t k = 0; k < out_dim; ++k) {
                gW3[i * out_dim + k] += h2[i] * d3[k];
                a += W3[i * out_dim + k] * d3[k];
            }
            d2[i] = h2[i] > 0.0 ? a : 0.0;
        }
        for (int64_t k = 0; k < out_dim; ++k) gb3[k] += d3[k];
        for (int64_t i = 0; i < hid; ++i) {
            double a = 0.0;
            for (int64_t j = 0; j < hid; ++j) {
                gW2[i * hid + j] += h1[i] * d2[j];
                a += W2[i * hid + j] * d2[j];
            }
            d1[i] = h1[i] > 0.0 ? a : 0.0;
        }
        for (int64_t j = 0; j < hid; ++j) gb2[j] += d2[j];
        for (int64_t i = 0; i < in_dim; ++i)
            for (int64_t j = 0; j < hid; ++j) gW1[i * hid + j] += x[i] * d1[j];
        for (int64_t j = 0; j < hid; ++j) gb1[j] += d1[j];
    }
    free(d3);
    free(d2);
    free(d1);
    return loss * scale;
}

/* One Adam step (Kingma & Ba 2015, bias-corrected), in place. */
void orc_adam_step(int64_t n, double *p, const double *g, double *m, double *v,
                   double lr, double beta1, double beta2, double eps, int64_t step)
{
    const double c1 = 1.0 - pow(beta1, (double)step);
    const double c2 = 1.0 - pow(beta2, (double)step);
    for (int64_t i = 0; i < n; ++i) {
        m[i] = beta1 * m[i] + (1.0 - beta1) * g[i];
        v[i] = beta2 * v[i] + (1.0 - beta2) * g[i] * g[i];
        const double mh = m[i] / c1, vh = v[i] / c2;
        p[i] -= lr * mh / (sqrt(vh) + eps);
    }
}

/* ---------------------------------------------------- split value storage */

/* Round-to-nearest-even fp32 -> bf16 (upper 16 bits), NaN kept quiet. */
static uint16_t f32_to_bf16_rne(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
        return (uint16_t)((u >> 16) | 0x0040u);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

static float bf16_to_f32(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* hi = bf16(v), lo = bf16(v - hi); v - hi is exact in fp32. */
void orc_bf16_split(int64_t n, const float *v, uint16_t *hi, uint16_t *lo)
{
    for (int64_t i = 0; i < n; ++i) {
        const uint16_t h = f32_to_bf16_rne(v[i]);
        const float r = v[i] - bf16_to_f32(h);
        hi[i] = h;
        lo[i] = f32_to_bf16_rne(r);
    }
}

/* Decoded value of a split pair, exactly, in fp64. */
void orc_bf16_join(int64_t n, const uint16_t *hi, const uint16_t *lo, double *out)
{
    for (int64_t i = 0; i < n; ++i)
        out[i] = (double)bf16_to_f32(hi[i]) + (lo ? (double)bf16_to_f32(lo[i]) : 0.0);
}
