/* An FFTW3-style caller: fftwf_malloc buffers, fftwf_plan_many_dft with
 * embedded strides, fftwf_execute and fftwf_execute_dft, 2-D r2c.  Built
 * against include/fftw_wrapper.h and linked with -laocl_fftz only.
 * Prints "PASS" lines and exits 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "fftw_wrapper.h"

static double dft_err(const fftwf_complex *x, int xs, const fftwf_complex *y, int ys, int n)
{
    double num = 0, den = 0;
    for (int k = 0; k < n; k++) {
        double sr = 0, si = 0;
        for (int j = 0; j < n; j++) {
            double a = -2.0 * M_PI * (double)((long)j * k % n) / n;
            sr += x[j * xs][0] * cos(a) - x[j * xs][1] * sin(a);
            si += x[j * xs][0] * sin(a) + x[j * xs][1] * cos(a);
        }
        double dr = y[k * ys][0] - sr, di = y[k * ys][1] - si;
        num += dr * dr + di * di;
        den += sr * sr + si * si;
    }
    return sqrt(num / den);
}

int main(void)
{
    int ok = 1;
    /* 6 transforms of n = 256, interleaved input (istride 2, idist 1),
     * contiguous output (ostride 1, odist 256) */
    const int n = 256, howmany = 6;
    fftwf_complex *in = fftwf_alloc_complex((size_t)2 * n * howmany);
    fftwf_complex *out = fftwf_alloc_complex((size_t)n * howmany);
    srand(3);
    for (int i = 0; i < 2 * n * howmany; i++) { in[i][0] = rand() / (float)RAND_MAX - 0.5f; in[i][1] = rand() / (float)RAND_MAX - 0.5f; }
    int nn[1] = {n};
    fftwf_init_threads();
    fftwf_plan_with_nthreads(4);
    fftwf_plan p = fftwf_plan_many_dft(1, nn, howmany, in, NULL, 2 * howmany, 1, out, NULL, 1, n, FFTW_FORWARD, FFTW_ESTIMATE);
    if (!p) { printf("FAIL plan_many_dft\n"); return 1; }
    fftwf_execute(p);
    double worst = 0;
    for (int b = 0; b < howmany; b++) {
        double e = dft_err(in + b, 2 * howmany, out + (size_t)b * n, 1, n);
        if (e > worst) worst = e;
    }
    printf("%s plan_many_dft strided rel %.3e\n", worst < 2e-6 ? "PASS" : "FAIL", worst);
    ok &= worst < 2e-6;
    /* new arrays through execute_dft */
    fftwf_complex *in2 = fftwf_alloc_complex((size_t)2 * n * howmany);
    for (int i = 0; i < 2 * n * howmany; i++) { in2[i][0] = (float)sin(i * 0.37); in2[i][1] = (float)cos(i * 0.11); }
    fftwf_execute_dft(p, in2, out);
    double e2 = dft_err(in2 + 1, 2 * howmany, out + n, 1, n);
    printf("%s execute_dft rel %.3e\n", e2 < 2e-6 ? "PASS" : "FAIL", e2);
    ok &= e2 < 2e-6;
    fftwf_destroy_plan(p);
    /* 2-D r2c 16 x 20 -> 16 x 11 */
    const int r0 = 16, r1 = 20;
    float *rin = fftwf_alloc_real((size_t)r0 * r1);
    fftwf_complex *rout = fftwf_alloc_complex((size_t)r0 * (r1 / 2 + 1));
    for (int i = 0; i < r0 * r1; i++) rin[i] = (float)cos(0.7 * i) + 0.25f;
    fftwf_plan pr = fftwf_plan_dft_r2c_2d(r0, r1, rin, rout, FFTW_MEASURE);
    if (!pr) { printf("FAIL plan_dft_r2c_2d\n"); return 1; }
    fftwf_execute(pr);
    double num = 0, den = 0;
    for (int k0 = 0; k0 < r0; k0++)
        for (int k1 = 0; k1 <= r1 / 2; k1++) {
            double sr = 0, si = 0;
            for (int j0 = 0; j0 < r0; j0++)
                for (int j1 = 0; j1 < r1; j1++) {
                    double a = -2.0 * M_PI * ((double)(j0 * k0 % r0) / r0 + (double)(j1 * k1 % r1) / r1);
                    sr += rin[j0 * r1 + j1] * cos(a);
                    si += rin[j0 * r1 + j1] * sin(a);
                }
            double dr = rout[k0 * (r1 / 2 + 1) + k1][0] - sr, di = rout[k0 * (r1 / 2 + 1) + k1][1] - si;
            num += dr * dr + di * di;
            den += sr * sr + si * si;
        }
    double er = sqrt(num / den);
    printf("%s plan_dft_r2c_2d rel %.3e\n", er < 2e-6 ? "PASS" : "FAIL", er);
    ok &= er < 2e-6;
    fftwf_destroy_plan(pr);
    fftwf_free(in); fftwf_free(out); fftwf_free(in2); fftwf_free(rin); fftwf_free(rout);
    fftwf_cleanup_threads();
    printf("%s %s\n", ok ? "PASS" : "FAIL", fftwf_version);
    return ok ? 0 : 1;
}
