/* fftw_fortran.h -- legacy Fortran interface of the FFTW wrapper.
 *
 * C prototypes of the subroutines a Fortran 77 / 90 program calls (every
 * argument by reference, the plan returned through the first argument as an
 * INTEGER*8): dfftw_* in double precision and dfftwf_* in single precision,
 * the entry points of the reference's wrappers/fftw/src/fortran/
 * fftw_plan_wrapper.c, fftw_execute_wrapper.c, fftw_destroy_wrapper.c and
 * fftw_misc_wrapper.c; the single-precision set is also exported under
 * FFTW's own legacy names (sfftw_*).  Fortran callers include
 * fftw_wrapper.f for the constants (FFTW_FORWARD, FFTW_ESTIMATE, ...).
 *
 * Dimension sizes are given in Fortran (column-major) order: the planner
 * arrays are reversed before the C planners (and the 2-D / 3-D planners swap
 * their sizes); the guru planners take axis i as (n[i], is[i], os[i]),
 * innermost first, with a zero stride meaning 1.
 */
#ifndef FFTW_FORTRAN_H
#define FFTW_FORTRAN_H

#include "fftw_wrapper.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FFTZ_FORTRAN_DECLARE(F, X, R, C)                                                                     \
    void F##plan_dft_(X##plan *p, int *rank, const int *n, C *in, C *out, int *sign, int *flags);           \
    void F##plan_dft_1d_(X##plan *p, int *n, C *in, C *out, int *sign, int *flags);                         \
    void F##plan_dft_2d_(X##plan *p, int *nx, int *ny, C *in, C *out, int *sign, int *flags);               \
    void F##plan_dft_3d_(X##plan *p, int *nx, int *ny, int *nz, C *in, C *out, int *sign, int *flags);      \
    void F##plan_many_dft_(X##plan *p, int *rank, const int *n, int *howmany, C *in, const int *inembed,   \
                           int *istride, int *idist, C *out, const int *onembed, int *ostride, int *odist,  \
                           int *sign, int *flags);                                                          \
    void F##plan_guru_dft_(X##plan *p, int *rank, const int *n, const int *is, const int *os, int *hrank,   \
                           const int *hn, const int *his, const int *hos, C *in, C *out, int *sign,         \
                           int *flags);                                                                     \
    void F##plan_dft_r2c_(X##plan *p, int *rank, const int *n, R *in, C *out, int *flags);                  \
    void F##plan_dft_r2c_1d_(X##plan *p, int *n, R *in, C *out, int *flags);                                \
    void F##plan_dft_r2c_2d_(X##plan *p, int *nx, int *ny, R *in, C *out, int *flags);                      \
    void F##plan_dft_r2c_3d_(X##plan *p, int *nx, int *ny, int *nz, R *in, C *out, int *flags);             \
    void F##plan_many_dft_r2c_(X##plan *p, int *rank, const int *n, int *howmany, R *in,                    \
                               const int *inembed, int *istride, int *idist, C *out, const int *onembed,    \
                               int *ostride, int *odist, int *flags);                                       \
    void F##plan_guru_dft_r2c_(X##plan *p, int *rank, const int *n, const int *is, const int *os,           \
                               int *hrank, const int *hn, const int *his, const int *hos, R *in, C *out,    \
                               int *flags);                                                                 \
    void F##plan_dft_c2r_(X##plan *p, int *rank, const int *n, C *in, R *out, int *flags);                  \
    void F##plan_dft_c2r_1d_(X##plan *p, int *n, C *in, R *out, int *flags);                                \
    void F##plan_dft_c2r_2d_(X##plan *p, int *nx, int *ny, C *in, R *out, int *flags);                      \
    void F##plan_dft_c2r_3d_(X##plan *p, int *nx, int *ny, int *nz, C *in, R *out, int *flags);             \
    void F##plan_many_dft_c2r_(X##plan *p, int *rank, const int *n, int *howmany, C *in,                    \
                               const int *inembed, int *istride, int *idist, R *out, const int *onembed,    \
                               int *ostride, int *odist, int *flags);                                       \
    void F##plan_guru_dft_c2r_(X##plan *p, int *rank, const int *n, const int *is, const int *os,           \
                               int *hrank, const int *hn, const int *his, const int *hos, C *in, R *out,    \
                               int *flags);                                                                 \
    void F##execute_(X##plan *p);                                                                           \
    void F##execute_dft_(X##plan *p, C *in, C *out);                                                        \
    void F##execute_dft_r2c_(X##plan *p, R *in, C *out);                                                    \
    void F##execute_dft_c2r_(X##plan *p, C *in, R *out);                                                    \
    void F##destroy_plan_(X##plan *p);                                                                      \
    void F##cleanup_(void);                                                                                 \
    void F##plan_with_nthreads_(int *nthreads);                                                             \
    void F##init_threads_(int *ret);                                                                        \
    void F##cleanup_threads_(void);                                                                         \
    void F##flops_(X##plan *p, double *add, double *mul, double *fmas);                                     \
    void F##estimate_cost_(double *cost, X##plan *p);                                                       \
    void F##cost_(double *cost, X##plan *p);

FFTZ_FORTRAN_DECLARE(dfftw_, fftw_, double, fftw_complex)
FFTZ_FORTRAN_DECLARE(dfftwf_, fftwf_, float, fftwf_complex)
FFTZ_FORTRAN_DECLARE(sfftw_, fftwf_, float, fftwf_complex)

#ifdef __cplusplus
}
#endif

#endif /* FFTW_FORTRAN_H */
