/*
 * fft_oracle.c -- CPU restatement of the reference's transform, used ONLY as
 * the expected-result source for the GPU library (tests, smoke check, CPU baseline leg of the
 * bench).  Nothing in the product library links or calls this code.
 *
 * Everything is computed in double (long double for the naive DFT).  Each
 * function names the reference file:line it restates.  Conventions
 * (api/aoclfftz.h:86-311, docs/rst/data_layout_conventions.rst:9-60):
 *   - forward  X[k] = sum_j x[j] exp(-2 pi i j k / N), backward uses +i,
 *     both unnormalised (test/bench/dft_reference.c:556-636);
 *   - complex data interleaved, strides in complex elements (real elements on
 *     the real side of R2C/C2R), dims[0] innermost.
 *
 * Pinning (see tests/test_oracle.py): naive long-double DFT, closed forms
 * (impulse, single tones, Parseval, linearity, time shift), numpy.fft as an
 * independent third-party cross-check, and the reference library itself built
 * from its sources under oracle/_ref where that build is present.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cd;

static inline cd cmul(cd a, cd b) { cd r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cd cconj(cd a) { cd r = {a.re, -a.im}; return r; }
static inline cd cadd(cd a, cd b) { cd r = {a.re + b.re, a.im + b.im}; return r; }

/* exp(sign * 2 pi i p / n) with exact integer reduction of p modulo n and an
 * octant fold, so the angle handed to sin/cos is at most pi/4 (the reference
 * computes the same quantity as cos/sin of 2*pi*k*j/(r*m), core/common/
 * twiddle.c:124-171). */
static cd unit_root(int64_t p, int64_t n, int sign)
{
    p %= n;
    if (p < 0) p += n;
    /* reduce to quadrant: p/n = q/4 + rest */
    int64_t p4 = 4 * p;            /* angle = 2 pi p / n = (pi/2) * (4p/n) */
    int64_t q = p4 / n;            /* quadrant 0..3 */
    int64_t r = p4 - q * n;        /* remainder in [0, n): angle_in_q = (pi/2) r/n */
    double s, c;
    if (2 * r <= n) {
        double a = (M_PI / 2.0) * ((double)r / (double)n);
        s = sin(a); c = cos(a);
    } else {
        double a = (M_PI / 2.0) * ((double)(n - r) / (double)n);
        c = sin(a); s = cos(a);
    }
    double re, im;
    switch (q & 3) {
        case 0: re = c; im = s; break;
        case 1: re = -s; im = c; break;
        case 2: re = -c; im = -s; break;
        default: re = s; im = -c; break;
    }
    cd w = {re, sign * im};
    return w;
}

/* --------------------------------------------------------------------------
 * Naive O(N^2) DFT in long double with exact (j*k mod N) reduction.
 * Restates test/bench/dft_reference.c:556-636 (dft_ref_d) for one transform.
 * ------------------------------------------------------------------------ */
void orc_dft_naive(int64_t n, const double *in, int64_t is, double *out,
                   int64_t os, int sign)
{
    const long double two_pi = 6.283185307179586476925286766559005768L;
    long double *wr = (long double *)malloc(sizeof(long double) * (size_t)n);
    long double *wi = (long double *)malloc(sizeof(long double) * (size_t)n);
    for (int64_t p = 0; p < n; p++) {
        long double a = two_pi * (long double)p / (long double)n;
        wr[p] = cosl(a);
        wi[p] = (long double)sign * sinl(a);
    }
    for (int64_t k = 0; k < n; k++) {
        long double sr = 0.0L, si = 0.0L;
        int64_t p = 0;
        for (int64_t j = 0; j < n; j++) {
            long double xr = in[2 * j * is], xi = in[2 * j * is + 1];
            sr += xr * wr[p] - xi * wi[p];
            si += xr * wi[p] + xi * wr[p];
            p += k;
            if (p >= n) p -= n;
        }
        out[2 * k * os] = (double)sr;
        out[2 * k * os + 1] = (double)si;
    }
    free(wr);
    free(wi);
}

/* --------------------------------------------------------------------------
 * Bluestein extended length: smallest m >= 2n-1 whose prime factors are all
 * in {2,3,5,7,11,13}.  Restates core/common/bluestein_utils.c:39-70.
 * ------------------------------------------------------------------------ */
int64_t orc_bluestein_length(int64_t n)
{
    static const int64_t primes[] = {2, 3, 5, 7, 11, 13};
    for (int64_t m = 2 * n - 1;; m++) {
        int64_t q = m;
        for (int i = 0; i < 6; i++)
            while (q % primes[i] == 0) q /= primes[i];
        if (q == 1) return m;
    }
}

/* Radices the reference has straight-line codelets for (core/kernels/
 * kernel_list.h:72-92: 2..16, 20, 48).  The recursive split below uses them in
 * descending order. */
static const int64_t k_radices[] = {16, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2};

static int64_t pick_radix(int64_t n)
{
    for (size_t i = 0; i < sizeof(k_radices) / sizeof(k_radices[0]); i++)
        if (n % k_radices[i] == 0 && n != k_radices[i]) return k_radices[i];
    return 0;
}

static void fft_rec(int64_t n, const cd *in, int64_t is, cd *out, int64_t os,
                    int sign);

/* Cached tables W[p] = exp(sign 2 pi i p / n), p in [0, n), built with
 * unit_root (octant-folded, ~1 ulp).  A table handed out stays valid until
 * the exported call that asked for it returns: nothing is evicted inside a
 * transform (a recursive transform holds its table across the sub-transforms
 * and leaf DFTs that fetch theirs).  roots_trim(), at the entry of each
 * exported transform, drops the whole cache once it exceeds a byte budget.
 * The oracle runs single-threaded. */
#define ROOT_CACHE_BYTES ((size_t)1 << 29)
static struct root_entry { int64_t n; int sign; cd *w; } *g_roots = NULL;
static size_t g_roots_len = 0, g_roots_cap = 0, g_roots_bytes = 0;

static const cd *roots(int64_t n, int sign)
{
    for (size_t i = 0; i < g_roots_len; i++)
        if (g_roots[i].n == n && g_roots[i].sign == sign) return g_roots[i].w;
    cd *w = (cd *)malloc(sizeof(cd) * (size_t)n);
    for (int64_t p = 0; p < n; p++) w[p] = unit_root(p, n, sign);
    if (g_roots_len == g_roots_cap) {
        g_roots_cap = g_roots_cap ? 2 * g_roots_cap : 64;
        g_roots = (struct root_entry *)realloc(g_roots, sizeof(*g_roots) * g_roots_cap);
    }
    g_roots[g_roots_len].n = n;
    g_roots[g_roots_len].sign = sign;
    g_roots[g_roots_len].w = w;
    g_roots_len++;
    g_roots_bytes += sizeof(cd) * (size_t)n;
    return w;
}

void orc_clear_cache(void)
{
    for (size_t i = 0; i < g_roots_len; i++) free(g_roots[i].w);
    g_roots_len = 0;
    g_roots_bytes = 0;
}

/* Only called where no table is in use (entry of an exported transform). */
static void roots_trim(void)
{
    if (g_roots_bytes > ROOT_CACHE_BYTES) orc_clear_cache();
}

/* Small DFT by definition, in double, for leaf sizes (<= 16 or 20/48 in the
 * reference's codelet list). */
static void dft_small(int64_t n, const cd *in, int64_t is, cd *out, int64_t os,
                      int sign)
{
    cd tmp[64];
    const cd *w = roots(n, sign);
    for (int64_t k = 0; k < n; k++) {
        cd s = {0.0, 0.0};
        for (int64_t j = 0; j < n; j++)
            s = cadd(s, cmul(in[j * is], w[(j * k) % n]));
        tmp[k] = s;
    }
    for (int64_t k = 0; k < n; k++) out[k * os] = tmp[k];
}

/* Bluestein (chirp-z) for lengths with a prime factor > 16.  Restates
 * core/solvers/complex/bluestein_solver_dft.c:137-234 with the chirp of
 * core/common/bluestein_utils.c:88-158:
 *   B_k = exp(+i pi (k^2 mod 2n) / n), mirrored at m-k, zero elsewhere;
 *   forward: a = x * conj(B); A = FFT_m(a) * FFT_m(B); y = IFFT_m(A);
 *            X = (1/m) * conj(B) * y.   Backward conjugates the chirp. */
static void bluestein(int64_t n, const cd *in, int64_t is, cd *out, int64_t os,
                      int sign)
{
    int64_t m = orc_bluestein_length(n);
    cd *b = (cd *)calloc((size_t)m, sizeof(cd));
    cd *fb = (cd *)malloc(sizeof(cd) * (size_t)m);
    cd *a = (cd *)calloc((size_t)m, sizeof(cd));
    cd *fa = (cd *)malloc(sizeof(cd) * (size_t)m);
    for (int64_t k = 0; k < n; k++) {
        int64_t sq = (int64_t)(((__int128)k * k) % (2 * n));
        cd w = unit_root(sq, 2 * n, +1); /* exp(+i pi sq / n) */
        b[k] = w;
        if (k > 0) b[m - k] = w;
    }
    fft_rec(m, b, 1, fb, 1, -1);
    for (int64_t k = 0; k < n; k++) {
        cd bk = b[k];
        a[k] = cmul(in[k * is], sign < 0 ? cconj(bk) : bk);
    }
    fft_rec(m, a, 1, fa, 1, -1);
    for (int64_t k = 0; k < m; k++) fa[k] = cmul(fa[k], sign < 0 ? fb[k] : cconj(fb[k]));
    /* inverse FFT_m (unnormalised) */
    fft_rec(m, fa, 1, a, 1, +1);
    double inv_m = 1.0 / (double)m;
    for (int64_t k = 0; k < n; k++) {
        cd bk = b[k];
        cd v = cmul(a[k], sign < 0 ? cconj(bk) : bk);
        v.re *= inv_m;
        v.im *= inv_m;
        out[k * os] = v;
    }
    free(b); free(fb); free(a); free(fa);
}

/* Recursive decimation-in-time Cooley-Tukey, n = r * m.  Restates the CT
 * solver (core/solvers/complex/ct_solver_dft.c:89-116; twiddles
 * selector/selector.c:1938-1953): radix-m sub-FFTs over the r interleaved
 * sequences x[q + r j] (input stride r*s), then a radix-r twiddle pass
 * X[k1 + m k2] = sum_q (w_n^{q k1} Y_q[k1]) w_r^{q k2}. */
static void fft_rec(int64_t n, const cd *in, int64_t is, cd *out, int64_t os,
                    int sign)
{
    if (n == 1) { out[0] = in[0]; return; }
    int64_t r = pick_radix(n);
    if (r == 0) {
        if (n <= 16 || n == 20 || n == 48) { dft_small(n, in, is, out, os, sign); return; }
        /* n has no supported radix: smallest prime factor > 16 */
        int64_t f = 0;
        for (int64_t p = 17; p * p <= n; p += 2)
            if (n % p == 0) { f = p; break; }
        if (f == 0 || f == n) { bluestein(n, in, is, out, os, sign); return; }
        r = f; /* split off a large prime factor, handled by Bluestein leaves */
    }
    int64_t m = n / r;
    cd *y = (cd *)malloc(sizeof(cd) * (size_t)n);
    for (int64_t q = 0; q < r; q++)
        fft_rec(m, in + q * is, is * r, y + q * m, 1, sign);
    cd *t = (cd *)malloc(sizeof(cd) * (size_t)r);
    cd *o = (cd *)malloc(sizeof(cd) * (size_t)r);
    const cd *w = roots(n, sign);
    for (int64_t k1 = 0; k1 < m; k1++) {
        for (int64_t q = 0; q < r; q++)
            t[q] = cmul(y[q * m + k1], w[q * k1]);
        if (r <= 16) {
            dft_small(r, t, 1, o, 1, sign);
        } else {
            fft_rec(r, t, 1, o, 1, sign);
        }
        for (int64_t k2 = 0; k2 < r; k2++) out[(k1 + m * k2) * os] = o[k2];
    }
    free(t); free(o); free(y);
}

/* One 1-D complex transform, arbitrary strides (complex elements). */
void orc_fft1d(int64_t n, const double *in, int64_t is, double *out, int64_t os,
               int sign)
{
    roots_trim();
    cd *src = (cd *)malloc(sizeof(cd) * (size_t)n);
    cd *dst = (cd *)malloc(sizeof(cd) * (size_t)n);
    for (int64_t j = 0; j < n; j++) { src[j].re = in[2 * j * is]; src[j].im = in[2 * j * is + 1]; }
    fft_rec(n, src, 1, dst, 1, sign);
    for (int64_t k = 0; k < n; k++) { out[2 * k * os] = dst[k].re; out[2 * k * os + 1] = dst[k].im; }
    free(src); free(dst);
}

/* --------------------------------------------------------------------------
 * Batched multi-dimensional C2C with the descriptor's dims/vecs tensor.
 * Row-column over dims (core/solvers/complex/ndim_solver_dft.c:96-121), batch
 * loop over vecs (core/solvers/complex/batched_solver_dft.c:33-110).
 * All input is gathered before any output is written, so in-place layouts with
 * equal strides behave as in the reference.
 *   dims: dim_rank triples (n, in_stride, out_stride); vecs: vec_rank triples.
 * ------------------------------------------------------------------------ */
static int64_t prod_n(const int64_t *t, int rank)
{
    int64_t p = 1;
    for (int i = 0; i < rank; i++) p *= t[3 * i];
    return p;
}

static int64_t offset_of(int64_t lin, const int64_t *t, int rank, int which)
{
    /* lin enumerates the tensor with entry 0 fastest */
    int64_t off = 0;
    for (int i = 0; i < rank; i++) {
        int64_t n = t[3 * i];
        int64_t idx = lin % n;
        lin /= n;
        off += idx * t[3 * i + 1 + which];
    }
    return off;
}

void orc_c2c(const double *in, double *out, int dim_rank, const int64_t *dims,
             int vec_rank, const int64_t *vecs, int sign)
{
    roots_trim();
    int64_t n = prod_n(dims, dim_rank);
    int64_t nb = prod_n(vecs, vec_rank);
    if (n == 0 || nb == 0) return;
    cd *buf = (cd *)malloc(sizeof(cd) * (size_t)(n * nb));
    for (int64_t b = 0; b < nb; b++) {
        int64_t vo = offset_of(b, vecs, vec_rank, 0);
        for (int64_t e = 0; e < n; e++) {
            int64_t o = vo + offset_of(e, dims, dim_rank, 0);
            buf[b * n + e].re = in[2 * o];
            buf[b * n + e].im = in[2 * o + 1];
        }
    }
    /* transform along each dim d of the dense [batch][dims reversed] buffer */
    int64_t inner = 1;
    cd *tmp = NULL;
    for (int d = 0; d < dim_rank; d++) {
        int64_t nd = dims[3 * d];
        int64_t outer = n / (inner * nd);
        tmp = (cd *)realloc(tmp, sizeof(cd) * (size_t)nd);
        for (int64_t b = 0; b < nb; b++)
            for (int64_t o = 0; o < outer; o++)
                for (int64_t i = 0; i < inner; i++) {
                    cd *base = buf + b * n + o * inner * nd + i;
                    fft_rec(nd, base, inner, tmp, 1, sign);
                    for (int64_t k = 0; k < nd; k++) base[k * inner] = tmp[k];
                }
        inner *= nd;
    }
    free(tmp);
    for (int64_t b = 0; b < nb; b++) {
        int64_t vo = offset_of(b, vecs, vec_rank, 1);
        for (int64_t e = 0; e < n; e++) {
            int64_t o = vo + offset_of(e, dims, dim_rank, 1);
            out[2 * o] = buf[b * n + e].re;
            out[2 * o + 1] = buf[b * n + e].im;
        }
    }
    free(buf);
}

/* --------------------------------------------------------------------------
 * Real-packed R2C for even n: treat the n reals as M = n/2 complex values,
 * FFT_M, then recombine bins k and M-k.  Restates
 * core/solvers/real/packed_solver_rdft.c:154-186 and the pair formula of
 * core/kernels/non_dft/real_packed/packed_rdft_common.h:19-60 with the
 * half-scaled twiddle (0.5 cos, -0.5 sin) of core/common/twiddle.c:438-489.
 * Output: M+1 complex bins.
 * ------------------------------------------------------------------------ */
void orc_r2c_packed_1d(int64_t n, const double *in, double *out)
{
    roots_trim();
    int64_t M = n / 2;
    cd *z = (cd *)malloc(sizeof(cd) * (size_t)M);
    cd *x = (cd *)malloc(sizeof(cd) * (size_t)M);
    for (int64_t j = 0; j < M; j++) { x[j].re = in[2 * j]; x[j].im = in[2 * j + 1]; }
    fft_rec(M, x, 1, z, 1, -1);
    /* DC and Nyquist */
    out[0] = z[0].re + z[0].im; out[1] = 0.0;
    out[2 * M] = z[0].re - z[0].im; out[2 * M + 1] = 0.0;
    for (int64_t k = 1; k <= M / 2; k++) {
        cd zk = z[k], zmk = z[M - k];
        cd w = unit_root(k, n, -1);            /* exp(-2 pi i k / n) */
        double twr = 0.5 * w.re, twi = 0.5 * w.im;
        double sum_re = zk.re + zmk.re;
        double diff_im = zk.im - zmk.im;
        double rot_arg_re = zk.im + zmk.im;
        double rot_arg_im = zmk.re - zk.re;
        double rot_re = twr * rot_arg_re - twi * rot_arg_im;
        double rot_im = twr * rot_arg_im + twi * rot_arg_re;
        double half_re = 0.5 * sum_re, half_im = 0.5 * diff_im;
        out[2 * k] = half_re + rot_re;
        out[2 * k + 1] = half_im + rot_im;
        out[2 * (M - k)] = half_re - rot_re;
        out[2 * (M - k) + 1] = rot_im - half_im;
    }
    free(z); free(x);
}

/* R2C by definition (real input promoted to complex), any n, n/2+1 bins. */
void orc_r2c_1d(int64_t n, const double *in, int64_t is, double *out, int64_t os)
{
    roots_trim();
    cd *x = (cd *)malloc(sizeof(cd) * (size_t)n);
    cd *z = (cd *)malloc(sizeof(cd) * (size_t)n);
    for (int64_t j = 0; j < n; j++) { x[j].re = in[j * is]; x[j].im = 0.0; }
    fft_rec(n, x, 1, z, 1, -1);
    for (int64_t k = 0; k <= n / 2; k++) { out[2 * k * os] = z[k].re; out[2 * k * os + 1] = z[k].im; }
    free(x); free(z);
}

/* C2R (backward, unnormalised) from n/2+1 Hermitian bins; the imaginary parts
 * of DC and (even n) Nyquist are ignored, as the layout convention says they
 * are zero (docs/rst/data_layout_conventions.rst:49-60). */
void orc_c2r_1d(int64_t n, const double *in, int64_t is, double *out, int64_t os)
{
    roots_trim();
    cd *x = (cd *)malloc(sizeof(cd) * (size_t)n);
    cd *z = (cd *)malloc(sizeof(cd) * (size_t)n);
    for (int64_t k = 0; k <= n / 2; k++) { x[k].re = in[2 * k * is]; x[k].im = in[2 * k * is + 1]; }
    x[0].im = 0.0;
    if (n % 2 == 0) x[n / 2].im = 0.0;
    for (int64_t k = n / 2 + 1; k < n; k++) x[k] = cconj(x[n - k]);
    fft_rec(n, x, 1, z, 1, +1);
    for (int64_t j = 0; j < n; j++) out[j * os] = z[j].re;
    free(x); free(z);
}

/* --------------------------------------------------------------------------
 * Error metrics.
 * ------------------------------------------------------------------------ */
/* relative L2: ||a - b|| / ||b|| over len scalars */
double orc_rel_l2(const double *a, const double *b, int64_t len)
{
    long double num = 0.0L, den = 0.0L;
    for (int64_t i = 0; i < len; i++) {
        long double d = (long double)a[i] - (long double)b[i];
        num += d * d;
        den += (long double)b[i] * (long double)b[i];
    }
    if (den == 0.0L) return (double)sqrtl(num);
    return (double)sqrtl(num / den);
}

/* The reference's pass/fail metric: max abs error over scalars divided by the
 * largest min(|a_i|,|b_i|); test/bench/utils/compare.c:88-147. */
double orc_ref_rel_err(const double *a, const double *b, int64_t len)
{
    double max_abs = 0.0, max_mag = 0.0;
    for (int64_t i = 0; i < len; i++) {
        double e = fabs(a[i] - b[i]);
        double mag = fmin(fabs(a[i]), fabs(b[i]));
        if (e > max_abs) max_abs = e;
        if (mag > max_mag) max_mag = mag;
    }
    if (max_abs == 0.0 && max_mag == 0.0) return 0.0;
    if (max_mag == 0.0) return max_abs;
    return max_abs / max_mag;
}
