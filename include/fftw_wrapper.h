/* fftw_wrapper.h -- FFTW3-compatible C interface of the MI355X FFT library.
 *
 * Drop-in for the reference's FFTW wrapper (wrappers/fftw/api/fftw_wrapper.h):
 * the plan / execute / destroy entry points an FFTW3 program calls, mapped
 * onto aoclfftz_setup_*_64_ / aoclfftz_execute[_io] / aoclfftz_destroy with
 * the reference's translation (wrappers/fftw/src/translator/
 * fftz_translator.c:379-600: FFTW row-major dims reversed so dim 0 is
 * innermost, embed arrays turned into strides, howmany into vecs[0]).
 *
 * Buffers may be host memory (fftw_malloc) or device memory; plans created
 * here execute synchronously like FFTW.  Planner flags (FFTW_MEASURE, ...)
 * are accepted and ignored, as in the reference.  Real-to-real plans
 * (r2r) are unsupported, as in the reference (wrappers/fftw/src/c/
 * fftw_plan_wrapper.c:457-560): their planners print a warning and return
 * NULL, and execute_r2r does nothing.
 */
#ifndef FFTW_WRAPPER_H
#define FFTW_WRAPPER_H

#include <stddef.h>
#include <stdio.h>

#include "aoclfftz.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FFTW_FORWARD (-1)
#define FFTW_BACKWARD (+1)
#define FFTW_MEASURE (0U)
#define FFTW_DESTROY_INPUT (1U << 0)
#define FFTW_UNALIGNED (1U << 1)
#define FFTW_CONSERVE_MEMORY (1U << 2)
#define FFTW_EXHAUSTIVE (1U << 3)
#define FFTW_PRESERVE_INPUT (1U << 4)
#define FFTW_PATIENT (1U << 5)
#define FFTW_ESTIMATE (1U << 6)
#define FFTW_WISDOM_ONLY (1U << 21)

/* r2r kinds (planners unsupported, as in the reference) */
#define FFTW_R2HC 0
#define FFTW_HC2R 1
#define FFTW_DHT 2
#define FFTW_REDFT00 3
#define FFTW_REDFT01 4
#define FFTW_REDFT10 5
#define FFTW_REDFT11 6
#define FFTW_RODFT00 7
#define FFTW_RODFT01 8
#define FFTW_RODFT10 9
#define FFTW_RODFT11 10

#if defined(_Complex_I) && defined(complex) && defined(I) && !defined(FFTW_NO_Complex)
typedef double _Complex fftw_complex;
typedef float _Complex fftwf_complex;
#else
typedef double fftw_complex[2];
typedef float fftwf_complex[2];
#endif

typedef void *fftw_plan;
typedef void *fftwf_plan;

typedef struct { int n, is, os; } fftw_iodim;
typedef struct { ptrdiff_t n, is, os; } fftw_iodim64;
typedef fftw_iodim fftwf_iodim;
typedef fftw_iodim64 fftwf_iodim64;
typedef int fftw_r2r_kind;
typedef int fftwf_r2r_kind;
typedef void (*fftw_write_char_func)(char c, void *);
typedef int (*fftw_read_char_func)(void *);
typedef fftw_write_char_func fftwf_write_char_func;
typedef fftw_read_char_func fftwf_read_char_func;

#define FFTZ_FFTW_DECLARE(X, R, C)                                                                      \
    X##plan X##plan_dft(int rank, const int *n, C *in, C *out, int sign, unsigned flags);               \
    X##plan X##plan_dft_1d(int n0, C *in, C *out, int sign, unsigned flags);                           \
    X##plan X##plan_dft_2d(int n0, int n1, C *in, C *out, int sign, unsigned flags);                   \
    X##plan X##plan_dft_3d(int n0, int n1, int n2, C *in, C *out, int sign, unsigned flags);           \
    X##plan X##plan_many_dft(int rank, const int *n, int howmany, C *in, const int *inembed,            \
                             int istride, int idist, C *out, const int *onembed, int ostride,           \
                             int odist, int sign, unsigned flags);                                      \
    X##plan X##plan_guru_dft(int rank, const X##iodim *dims, int howmany_rank,                          \
                             const X##iodim *howmany_dims, C *in, C *out, int sign, unsigned flags);    \
    X##plan X##plan_guru64_dft(int rank, const X##iodim64 *dims, int howmany_rank,                      \
                               const X##iodim64 *howmany_dims, C *in, C *out, int sign,                 \
                               unsigned flags);                                                         \
    X##plan X##plan_dft_r2c(int rank, const int *n, R *in, C *out, unsigned flags);                     \
    X##plan X##plan_dft_r2c_1d(int n0, R *in, C *out, unsigned flags);                                 \
    X##plan X##plan_dft_r2c_2d(int n0, int n1, R *in, C *out, unsigned flags);                         \
    X##plan X##plan_dft_r2c_3d(int n0, int n1, int n2, R *in, C *out, unsigned flags);                 \
    X##plan X##plan_many_dft_r2c(int rank, const int *n, int howmany, R *in, const int *inembed,        \
                                 int istride, int idist, C *out, const int *onembed, int ostride,       \
                                 int odist, unsigned flags);                                            \
    X##plan X##plan_guru_dft_r2c(int rank, const X##iodim *dims, int howmany_rank,                      \
                                 const X##iodim *howmany_dims, R *in, C *out, unsigned flags);          \
    X##plan X##plan_guru64_dft_r2c(int rank, const X##iodim64 *dims, int howmany_rank,                  \
                                   const X##iodim64 *howmany_dims, R *in, C *out, unsigned flags);      \
    X##plan X##plan_dft_c2r(int rank, const int *n, C *in, R *out, unsigned flags);                     \
    X##plan X##plan_dft_c2r_1d(int n0, C *in, R *out, unsigned flags);                                 \
    X##plan X##plan_dft_c2r_2d(int n0, int n1, C *in, R *out, unsigned flags);                         \
    X##plan X##plan_dft_c2r_3d(int n0, int n1, int n2, C *in, R *out, unsigned flags);                 \
    X##plan X##plan_many_dft_c2r(int rank, const int *n, int howmany, C *in, const int *inembed,        \
                                 int istride, int idist, R *out, const int *onembed, int ostride,       \
                                 int odist, unsigned flags);                                            \
    X##plan X##plan_guru_dft_c2r(int rank, const X##iodim *dims, int howmany_rank,                      \
                                 const X##iodim *howmany_dims, C *in, R *out, unsigned flags);          \
    X##plan X##plan_guru64_dft_c2r(int rank, const X##iodim64 *dims, int howmany_rank,                  \
                                   const X##iodim64 *howmany_dims, C *in, R *out, unsigned flags);      \
    X##plan X##plan_r2r(int rank, const int *n, R *in, R *out, const X##r2r_kind *kind, unsigned flags); \
    X##plan X##plan_r2r_1d(int n, R *in, R *out, X##r2r_kind kind, unsigned flags);                    \
    X##plan X##plan_r2r_2d(int n0, int n1, R *in, R *out, X##r2r_kind kind0, X##r2r_kind kind1,        \
                           unsigned flags);                                                             \
    X##plan X##plan_r2r_3d(int n0, int n1, int n2, R *in, R *out, X##r2r_kind kind0,                   \
                           X##r2r_kind kind1, X##r2r_kind kind2, unsigned flags);                       \
    X##plan X##plan_guru_r2r(int rank, const X##iodim *dims, int howmany_rank,                          \
                             const X##iodim *howmany_dims, R *in, R *out, const X##r2r_kind *kind,      \
                             unsigned flags);                                                           \
    X##plan X##plan_guru64_r2r(int rank, const X##iodim64 *dims, int howmany_rank,                      \
                               const X##iodim64 *howmany_dims, R *in, R *out,                           \
                               const X##r2r_kind *kind, unsigned flags);                                \
    X##plan X##plan_many_r2r(int rank, const int *n, int howmany, R *in, const int *inembed,           \
                             int istride, int idist, R *out, const int *onembed, int ostride,           \
                             int odist, const X##r2r_kind *kind, unsigned flags);                       \
    void X##execute(const X##plan p);                                                                   \
    void X##execute_dft(const X##plan p, C *in, C *out);                                               \
    void X##execute_dft_r2c(const X##plan p, R *in, C *out);                                           \
    void X##execute_dft_c2r(const X##plan p, C *in, R *out);                                           \
    void X##execute_r2r(const X##plan p, R *in, R *out);                                               \
    void X##destroy_plan(X##plan p);                                                                    \
    void X##cleanup(void);                                                                              \
    void X##set_timelimit(double t);                                                                    \
    void X##plan_with_nthreads(int nthreads);                                                           \
    int X##planner_nthreads(void);                                                                      \
    int X##init_threads(void);                                                                          \
    void X##cleanup_threads(void);                                                                      \
    void X##make_planner_thread_safe(void);                                                             \
    void *X##malloc(size_t n);                                                                          \
    R *X##alloc_real(size_t n);                                                                         \
    C *X##alloc_complex(size_t n);                                                                      \
    void X##free(void *p);                                                                              \
    void X##flops(const X##plan p, double *add, double *mul, double *fmas);                             \
    double X##estimate_cost(const X##plan p);                                                           \
    double X##cost(const X##plan p);                                                                    \
    int X##alignment_of(R *p);                                                                          \
    void X##fprint_plan(const X##plan p, FILE *f);                                                      \
    void X##print_plan(const X##plan p);                                                                \
    char *X##sprint_plan(const X##plan p);                                                              \
    void X##forget_wisdom(void);                                                                        \
    int X##export_wisdom_to_filename(const char *f);                                                    \
    void X##export_wisdom_to_file(FILE *f);                                                             \
    char *X##export_wisdom_to_string(void);                                                             \
    void X##export_wisdom(X##write_char_func write_char, void *data);                                   \
    int X##import_system_wisdom(void);                                                                  \
    int X##import_wisdom_from_filename(const char *f);                                                  \
    int X##import_wisdom_from_file(FILE *f);                                                            \
    int X##import_wisdom_from_string(const char *s);                                                    \
    int X##import_wisdom(X##read_char_func read_char, void *data);                                      \
    extern const char X##version[];

/* Replace wrappers/fftw/api/fftw_wrapper.h:48-260 (double: fftw_, float:
 * fftwf_): plan_dft / _1d / _2d / _3d, plan_many_dft (:91-96), guru and
 * guru64, r2c / c2r planners, execute (src/c/fftw_execute_wrapper.c:13-21),
 * execute_dft (:23-33), execute_dft_r2c / _c2r, destroy_plan, threads,
 * memory and wisdom helpers (src/c/fftw_misc_wrapper.c). */
FFTZ_FFTW_DECLARE(fftw_, double, fftw_complex)
FFTZ_FFTW_DECLARE(fftwf_, float, fftwf_complex)

#ifdef __cplusplus
}
#endif

#endif /* FFTW_WRAPPER_H */
