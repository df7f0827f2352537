/* A caller of the aoclfftz C API, written the way the reference's
 * examples/example_one_dim_complex.c uses it: host buffers, a batched 1-D
 * complex transform, forward then backward, checked against a direct DFT.
 * Built against include/aoclfftz.h and linked with -laocl_fftz only.
 * Prints "PASS <rel_err>" and exits 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "aoclfftz.h"

int main(void)
{
    const int n = 360, batch = 3;
    double *in = malloc(sizeof(double) * 2 * n * batch), *out = malloc(sizeof(double) * 2 * n * batch);
    double *back = malloc(sizeof(double) * 2 * n * batch);
    srand(7);
    for (int i = 0; i < 2 * n * batch; i++) in[i] = 2.0 * rand() / RAND_MAX - 1.0;
    aoclfftz_dim_t dims[1] = {{n, 1, 1}};
    aoclfftz_dim_t vecs[1] = {{batch, n, n}};
    aoclfftz_prob_desc_d pd = {0};
    pd.in = in; pd.out = out; pd.dim_rank = 1; pd.vec_rank = 1; pd.dims = dims; pd.vecs = vecs;
    pd.flags.fft_placement = 1; pd.pthr_fft.num_threads = 1; pd.cntrl_params.opt_level = 3;
    void *h = aoclfftz_setup_d(&pd);
    if (!h) { printf("FAIL setup\n"); return 1; }
    if (aoclfftz_execute(h) != AOCLFFTZ_SUCCESS) { printf("FAIL execute\n"); return 1; }
    pd.flags.fft_direction = 1; pd.in = out; pd.out = back;
    void *hb = aoclfftz_setup_d(&pd);
    if (!hb || aoclfftz_execute(hb) != AOCLFFTZ_SUCCESS) { printf("FAIL backward\n"); return 1; }
    double num = 0, den = 0, rt = 0;
    for (int b = 0; b < batch; b++)
        for (int k = 0; k < n; k++) {
            long double sr = 0, si = 0;
            for (int j = 0; j < n; j++) {
                long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)((long)j * k % n) / n;
                long double xr = in[2 * (b * n + j)], xi = in[2 * (b * n + j) + 1];
                sr += xr * cosl(a) - xi * sinl(a);
                si += xr * sinl(a) + xi * cosl(a);
            }
            double dr = out[2 * (b * n + k)] - (double)sr, di = out[2 * (b * n + k) + 1] - (double)si;
            num += dr * dr + di * di;
            den += (double)(sr * sr + si * si);
            double er = back[2 * (b * n + k)] / n - in[2 * (b * n + k)], ei = back[2 * (b * n + k) + 1] / n - in[2 * (b * n + k) + 1];
            rt += er * er + ei * ei;
        }
    double rel = sqrt(num / den);
    aoclfftz_destroy(h);
    aoclfftz_destroy(hb);
    printf("%s %.3e roundtrip %.3e %s\n", rel < 1e-13 ? "PASS" : "FAIL", rel, sqrt(rt), aoclfftz_version());
    return rel < 1e-13 ? 0 : 1;
}
