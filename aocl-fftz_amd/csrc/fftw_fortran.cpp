// Legacy Fortran interface of the FFTW wrapper: the dfftw_* (double) and
// dfftwf_* (single) subroutines a Fortran 77 / 90 program calls with every
// argument by reference and the plan returned through the first argument
// (an INTEGER*8 on the Fortran side).  Same entry points and argument
// meaning as the reference's wrappers/fftw/src/fortran/*.c:
//   * dimension sizes arrive in Fortran (column-major) order, so the planner
//     arrays n / inembed / onembed are reversed before the C planners, and
//     the 2-D / 3-D planners swap their sizes
//     (fortran/fftw_plan_wrapper.c:14-90, 93-150);
//   * the guru planners map n[i] / is[i] / os[i] onto axis i directly
//     (dims[0] innermost), a zero stride meaning 1
//     (translator/fftz_translator.c:950-1002);
//   * execute / destroy / threads / cleanup / flops / cost forward to the C
//     interface (fortran/fftw_execute_wrapper.c, fftw_destroy_wrapper.c,
//     fftw_misc_wrapper.c).
// The single-precision subroutines are also exported under FFTW's own
// legacy names (sfftw_*), which existing Fortran FFTW callers use.
#include <vector>

#include "../../include/fftw_fortran.h"

namespace {

std::vector<int> reversed(const int* v, int rank)
{
    std::vector<int> r;
    if (!v || rank <= 0) return r;
    r.assign(v, v + rank);
    for (int i = 0, j = rank - 1; i < j; i++, j--) std::swap(r[(size_t)i], r[(size_t)j]);
    return r;
}

const int* ptr_or_null(const int* orig, const std::vector<int>& r) { return orig ? r.data() : nullptr; }

// Fortran guru dims: axis i of the transform is (n[i], is[i], os[i]),
// innermost first.  The C guru planners take FFTW's row-major order
// (innermost last), so the arrays are handed over reversed.
template <typename Iodim>
std::vector<Iodim> guru_dims(int rank, const int* n, const int* is, const int* os)
{
    std::vector<Iodim> d;
    for (int i = rank - 1; i >= 0; i--) {
        Iodim x;
        x.n = n[i];
        x.is = is[i] == 0 ? 1 : is[i];
        x.os = os[i] == 0 ? 1 : os[i];
        d.push_back(x);
    }
    return d;
}

bool guru_ok(int rank, const int* n, const int* is, const int* os, int hrank, const int* hn, const int* his,
             const int* hos)
{
    return rank >= 0 && hrank >= 0 && (rank == 0 || (n && is && os)) && (hrank == 0 || (hn && his && hos));
}

}  // namespace

#define FFTZ_FORTRAN_DEFINE(F, X, R, C)                                                                       \
    void F##plan_dft_(X##plan* p, int* rank, const int* n, C* in, C* out, int* sign, int* flags)             \
    {                                                                                                         \
        const std::vector<int> rn = reversed(n, *rank);                                                       \
        *p = X##plan_dft(*rank, ptr_or_null(n, rn), in, out, *sign, (unsigned)*flags);                        \
    }                                                                                                         \
    void F##plan_dft_1d_(X##plan* p, int* n, C* in, C* out, int* sign, int* flags)                           \
    {                                                                                                         \
        *p = X##plan_dft_1d(*n, in, out, *sign, (unsigned)*flags);                                            \
    }                                                                                                         \
    void F##plan_dft_2d_(X##plan* p, int* nx, int* ny, C* in, C* out, int* sign, int* flags)                 \
    {                                                                                                         \
        *p = X##plan_dft_2d(*ny, *nx, in, out, *sign, (unsigned)*flags);                                      \
    }                                                                                                         \
    void F##plan_dft_3d_(X##plan* p, int* nx, int* ny, int* nz, C* in, C* out, int* sign, int* flags)        \
    {                                                                                                         \
        *p = X##plan_dft_3d(*nz, *ny, *nx, in, out, *sign, (unsigned)*flags);                                 \
    }                                                                                                         \
    void F##plan_many_dft_(X##plan* p, int* rank, const int* n, int* howmany, C* in, const int* inembed,     \
                           int* istride, int* idist, C* out, const int* onembed, int* ostride, int* odist,    \
                           int* sign, int* flags)                                                             \
    {                                                                                                         \
        const std::vector<int> rn = reversed(n, *rank), ri = reversed(inembed, *rank),                        \
                               ro = reversed(onembed, *rank);                                                 \
        *p = X##plan_many_dft(*rank, ptr_or_null(n, rn), *howmany, in, ptr_or_null(inembed, ri), *istride,    \
                              *idist, out, ptr_or_null(onembed, ro), *ostride, *odist, *sign,                 \
                              (unsigned)*flags);                                                              \
    }                                                                                                         \
    void F##plan_guru_dft_(X##plan* p, int* rank, const int* n, const int* is, const int* os, int* hrank,     \
                           const int* hn, const int* his, const int* hos, C* in, C* out, int* sign,           \
                           int* flags)                                                                        \
    {                                                                                                         \
        *p = nullptr;                                                                                         \
        if (!guru_ok(*rank, n, is, os, *hrank, hn, his, hos)) return;                                         \
        const auto d = guru_dims<X##iodim>(*rank, n, is, os);                                                 \
        const auto h = guru_dims<X##iodim>(*hrank, hn, his, hos);                                             \
        *p = X##plan_guru_dft(*rank, d.data(), *hrank, h.data(), in, out, *sign, (unsigned)*flags);           \
    }                                                                                                         \
    void F##plan_dft_r2c_(X##plan* p, int* rank, const int* n, R* in, C* out, int* flags)                    \
    {                                                                                                         \
        const std::vector<int> rn = reversed(n, *rank);                                                       \
        *p = X##plan_dft_r2c(*rank, ptr_or_null(n, rn), in, out, (unsigned)*flags);                           \
    }                                                                                                         \
    void F##plan_dft_r2c_1d_(X##plan* p, int* n, R* in, C* out, int* flags)                                  \
    {                                                                                                         \
        *p = X##plan_dft_r2c_1d(*n, in, out, (unsigned)*flags);                                               \
    }                                                                                                         \
    void F##plan_dft_r2c_2d_(X##plan* p, int* nx, int* ny, R* in, C* out, int* flags)                        \
    {                                                                                                         \
        *p = X##plan_dft_r2c_2d(*ny, *nx, in, out, (unsigned)*flags);                                         \
    }                                                                                                         \
    void F##plan_dft_r2c_3d_(X##plan* p, int* nx, int* ny, int* nz, R* in, C* out, int* flags)               \
    {                                                                                                         \
        *p = X##plan_dft_r2c_3d(*nz, *ny, *nx, in, out, (unsigned)*flags);                                    \
    }                                                                                                         \
    void F##plan_many_dft_r2c_(X##plan* p, int* rank, const int* n, int* howmany, R* in, const int* inembed, \
                               int* istride, int* idist, C* out, const int* onembed, int* ostride,            \
                               int* odist, int* flags)                                                        \
    {                                                                                                         \
        const std::vector<int> rn = reversed(n, *rank), ri = reversed(inembed, *rank),                        \
                               ro = reversed(onembed, *rank);                                                 \
        *p = X##plan_many_dft_r2c(*rank, ptr_or_null(n, rn), *howmany, in, ptr_or_null(inembed, ri),          \
                                  *istride, *idist, out, ptr_or_null(onembed, ro), *ostride, *odist,          \
                                  (unsigned)*flags);                                                          \
    }                                                                                                         \
    void F##plan_guru_dft_r2c_(X##plan* p, int* rank, const int* n, const int* is, const int* os, int* hrank, \
                               const int* hn, const int* his, const int* hos, R* in, C* out, int* flags)      \
    {                                                                                                         \
        *p = nullptr;                                                                                         \
        if (!guru_ok(*rank, n, is, os, *hrank, hn, his, hos)) return;                                         \
        const auto d = guru_dims<X##iodim>(*rank, n, is, os);                                                 \
        const auto h = guru_dims<X##iodim>(*hrank, hn, his, hos);                                             \
        *p = X##plan_guru_dft_r2c(*rank, d.data(), *hrank, h.data(), in, out, (unsigned)*flags);              \
    }                                                                                                         \
    void F##plan_dft_c2r_(X##plan* p, int* rank, const int* n, C* in, R* out, int* flags)                    \
    {                                                                                                         \
        const std::vector<int> rn = reversed(n, *rank);                                                       \
        *p = X##plan_dft_c2r(*rank, ptr_or_null(n, rn), in, out, (unsigned)*flags);                           \
    }                                                                                                         \
    void F##plan_dft_c2r_1d_(X##plan* p, int* n, C* in, R* out, int* flags)                                  \
    {                                                                                                         \
        *p = X##plan_dft_c2r_1d(*n, in, out, (unsigned)*flags);                                               \
    }                                                                                                         \
    void F##plan_dft_c2r_2d_(X##plan* p, int* nx, int* ny, C* in, R* out, int* flags)                        \
    {                                                                                                         \
        *p = X##plan_dft_c2r_2d(*ny, *nx, in, out, (unsigned)*flags);                                         \
    }                                                                                                         \
    void F##plan_dft_c2r_3d_(X##plan* p, int* nx, int* ny, int* nz, C* in, R* out, int* flags)               \
    {                                                                                                         \
        *p = X##plan_dft_c2r_3d(*nz, *ny, *nx, in, out, (unsigned)*flags);                                    \
    }                                                                                                         \
    void F##plan_many_dft_c2r_(X##plan* p, int* rank, const int* n, int* howmany, C* in, const int* inembed, \
                               int* istride, int* idist, R* out, const int* onembed, int* ostride,            \
                               int* odist, int* flags)                                                        \
    {                                                                                                         \
        const std::vector<int> rn = reversed(n, *rank), ri = reversed(inembed, *rank),                        \
                               ro = reversed(onembed, *rank);                                                 \
        *p = X##plan_many_dft_c2r(*rank, ptr_or_null(n, rn), *howmany, in, ptr_or_null(inembed, ri),          \
                                  *istride, *idist, out, ptr_or_null(onembed, ro), *ostride, *odist,          \
                                  (unsigned)*flags);                                                          \
    }                                                                                                         \
    void F##plan_guru_dft_c2r_(X##plan* p, int* rank, const int* n, const int* is, const int* os, int* hrank, \
                               const int* hn, const int* his, const int* hos, C* in, R* out, int* flags)      \
    {                                                                                                         \
        *p = nullptr;                                                                                         \
        if (!guru_ok(*rank, n, is, os, *hrank, hn, his, hos)) return;                                         \
        const auto d = guru_dims<X##iodim>(*rank, n, is, os);                                                 \
        const auto h = guru_dims<X##iodim>(*hrank, hn, his, hos);                                             \
        *p = X##plan_guru_dft_c2r(*rank, d.data(), *hrank, h.data(), in, out, (unsigned)*flags);              \
    }                                                                                                         \
    void F##execute_(X##plan* p) { X##execute(*p); }                                                          \
    void F##execute_dft_(X##plan* p, C* in, C* out) { X##execute_dft(*p, in, out); }                          \
    void F##execute_dft_r2c_(X##plan* p, R* in, C* out) { X##execute_dft_r2c(*p, in, out); }                  \
    void F##execute_dft_c2r_(X##plan* p, C* in, R* out) { X##execute_dft_c2r(*p, in, out); }                  \
    void F##destroy_plan_(X##plan* p) { X##destroy_plan(*p); }                                                \
    void F##cleanup_(void) { X##cleanup(); }                                                                  \
    void F##plan_with_nthreads_(int* nthreads) { X##plan_with_nthreads(*nthreads); }                          \
    void F##init_threads_(int* ret) { *ret = X##init_threads(); }                                             \
    void F##cleanup_threads_(void) { X##cleanup_threads(); }                                                  \
    void F##flops_(X##plan* p, double* add, double* mul, double* fmas) { X##flops(*p, add, mul, fmas); }      \
    void F##estimate_cost_(double* cost, X##plan* p) { *cost = X##estimate_cost(*p); }                        \
    void F##cost_(double* cost, X##plan* p) { *cost = X##cost(*p); }

extern "C" {
FFTZ_FORTRAN_DEFINE(dfftw_, fftw_, double, fftw_complex)
FFTZ_FORTRAN_DEFINE(dfftwf_, fftwf_, float, fftwf_complex)
FFTZ_FORTRAN_DEFINE(sfftw_, fftwf_, float, fftwf_complex)
}
