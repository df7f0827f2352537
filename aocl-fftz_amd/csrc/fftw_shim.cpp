// FFTW3-compatible entry points over the aoclfftz C ABI.
//
// Translation rules follow the reference's FFTW wrapper
// (wrappers/fftw/src/translator/fftz_translator.c):
//   plan_many: dims[i].n = n[rank-1-i]; dims[0] strides = istride / ostride
//              (0 -> 1); dims[i] stride = dims[i-1] stride * embed[rank-i]
//              (embed defaults to n); vecs[0] = {howmany, idist, odist}
//              (0 -> 1)                                      (:379-450)
//   r2c / c2r: the complex side's innermost extent is n0/2+1; without an
//              embed array the real side of an in-place plan is padded to
//              2 (n0/2 + 1)                                  (:452-600)
//   guru:      dims and howmany_dims reversed, 0 strides -> 1 (:712-806)
//   sign -1 -> forward; in == out -> in-place                  (:13-28)
//   execute -> aoclfftz_execute, execute_dft -> aoclfftz_execute_io
//              (src/c/fftw_execute_wrapper.c:13-33)
// Plans are built on the ILP64 descriptors so no size overflows 32 bits.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aoclfftz.h"
#include "../../include/aoclfftz_gpu.h"
#include "../../include/fftw_wrapper.h"

namespace {

enum Kind { C2C = 0, R2C = 1, C2R = 2 };

struct DV {
    std::vector<aoclfftz_dim_t_64_> dims, vecs;
};

int g_nthreads = 1;

bool many_desc(int rank, const int* n, int howmany, const int* inembed, int istride, int idist,
               const int* onembed, int ostride, int odist, Kind kind, bool inplace, DV& dv)
{
    if (rank < 0 || (rank > 0 && n == nullptr) || howmany < 0) return false;
    istride = istride == 0 ? 1 : istride;
    ostride = ostride == 0 ? 1 : ostride;
    dv.vecs.push_back({(ptrdiff_t)howmany, (ptrdiff_t)(idist == 0 ? 1 : idist), (ptrdiff_t)(odist == 0 ? 1 : odist)});
    if (rank == 0) {
        dv.dims.push_back({1, 1, 1});
        return true;
    }
    dv.dims.resize((size_t)rank);
    for (int i = 0; i < rank; i++) dv.dims[(size_t)i].n = n[rank - 1 - i];
    const ptrdiff_t n0 = dv.dims[0].n;
    const ptrdiff_t half = n0 / 2 + 1;
    // in strides
    dv.dims[0].in_stride = istride;
    dv.dims[0].out_stride = ostride;
    for (int i = 1; i < rank; i++) {
        ptrdiff_t ein, eout;
        if (inembed) {
            ein = inembed[rank - i];
        } else if (i == 1 && kind == R2C) {
            ein = inplace ? 2 * half : n0;
        } else if (i == 1 && kind == C2R) {
            ein = half;
        } else {
            ein = dv.dims[(size_t)i - 1].n;
        }
        if (onembed) {
            eout = onembed[rank - i];
        } else if (i == 1 && kind == R2C) {
            eout = half;
        } else if (i == 1 && kind == C2R) {
            eout = inplace ? 2 * half : n0;
        } else {
            eout = dv.dims[(size_t)i - 1].n;
        }
        dv.dims[(size_t)i].in_stride = dv.dims[(size_t)i - 1].in_stride * ein;
        dv.dims[(size_t)i].out_stride = dv.dims[(size_t)i - 1].out_stride * eout;
    }
    return true;
}

template <typename IOD>
bool guru_desc(int rank, const IOD* dims, int hrank, const IOD* hdims, DV& dv)
{
    if (rank < 0 || (rank > 0 && dims == nullptr) || (hrank > 0 && hdims == nullptr)) return false;
    if (hrank > 0) {
        for (int i = 0; i < hrank; i++) {
            const IOD& d = hdims[hrank - 1 - i];
            dv.vecs.push_back({(ptrdiff_t)d.n, (ptrdiff_t)(d.is == 0 ? 1 : d.is), (ptrdiff_t)(d.os == 0 ? 1 : d.os)});
        }
    } else {
        dv.vecs.push_back({1, 1, 1});
    }
    if (rank > 0) {
        for (int i = 0; i < rank; i++) {
            const IOD& d = dims[rank - 1 - i];
            dv.dims.push_back({(ptrdiff_t)d.n, (ptrdiff_t)d.is, (ptrdiff_t)d.os});
        }
    } else {
        dv.dims.push_back({1, 1, 1});
    }
    return true;
}

template <typename R>
void* make_plan(DV& dv, void* in, void* out, int sign, Kind kind)
{
    aoclfftz_flags_t fl{};
    fl.fft_type = kind == C2C ? 0 : 1;
    fl.fft_direction = kind == R2C ? 0 : kind == C2R ? 1 : (sign == FFTW_FORWARD ? 0 : 1);
    fl.fft_placement = (in == out) ? 0 : 1;
    if (sizeof(R) == 4) {
        aoclfftz_prob_desc_f_64_ pd{};
        pd.in = (float*)in;
        pd.out = (float*)out;
        pd.dim_rank = (FFTZ_INT32)dv.dims.size();
        pd.vec_rank = (FFTZ_INT32)dv.vecs.size();
        pd.dims = dv.dims.data();
        pd.vecs = dv.vecs.data();
        pd.flags = fl;
        pd.pthr_fft.num_threads = g_nthreads;
        pd.pthr_fft.dynamic_load_model = 1;
        pd.cntrl_params.opt_level = 3;
        return aoclfftz_setup_f_64_(&pd);
    }
    aoclfftz_prob_desc_d_64_ pd{};
    pd.in = (double*)in;
    pd.out = (double*)out;
    pd.dim_rank = (FFTZ_INT32)dv.dims.size();
    pd.vec_rank = (FFTZ_INT32)dv.vecs.size();
    pd.dims = dv.dims.data();
    pd.vecs = dv.vecs.data();
    pd.flags = fl;
    pd.pthr_fft.num_threads = g_nthreads;
    pd.pthr_fft.dynamic_load_model = 1;
    pd.cntrl_params.opt_level = 3;
    return aoclfftz_setup_d_64_(&pd);
}

template <typename R>
void* plan_many(int rank, const int* n, int howmany, void* in, const int* inembed, int istride, int idist,
                void* out, const int* onembed, int ostride, int odist, int sign, Kind kind)
{
    DV dv;
    if (!many_desc(rank, n, howmany, inembed, istride, idist, onembed, ostride, odist, kind, in == out, dv))
        return nullptr;
    return make_plan<R>(dv, in, out, sign, kind);
}

template <typename R, typename IOD>
void* plan_guru(int rank, const IOD* dims, int hrank, const IOD* hdims, void* in, void* out, int sign, Kind kind)
{
    DV dv;
    if (!guru_desc(rank, dims, hrank, hdims, dv)) return nullptr;
    return make_plan<R>(dv, in, out, sign, kind);
}

// Real-to-real planners and execute: unsupported, as in the reference
// (wrappers/fftw/src/c/fftw_plan_wrapper.c:457-560, fftw_execute_wrapper.c:59-73,
// with the warning of translator/fftz_translator.h:54-57); planners return NULL.
void* unsupported_r2r(const char* func)
{
    fprintf(stderr, "[AOCL-FFTZ] WARNING: %s: r2r transforms are not supported\n", func);
    return nullptr;
}

// Wisdom is the library's tuned-tile cache as text (one problem per line);
// the reference's wisdom calls are stubs (src/c/fftw_misc_wrapper.c:300-367):
// exports succeed with nothing written, imports fail.  Here exports write the
// cache and imports return 1 when the text parses (0 otherwise, importing
// nothing).  Both precisions share the cache; its keys carry the precision.
void wisdom_to_file(FILE* f)
{
    char* s = aoclfftz_gpu_tune_export_string();
    if (s && f) fputs(s, f);
    free(s);
}

void wisdom_to_callback(void (*w)(char, void*), void* data)
{
    char* s = aoclfftz_gpu_tune_export_string();
    if (s && w)
        for (const char* c = s; *c; c++) w(*c, data);
    free(s);
}

int wisdom_from_file(FILE* f)
{
    if (!f) return 0;
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    return aoclfftz_gpu_tune_import_string(text.c_str()) >= 0;
}

int wisdom_from_callback(int (*r)(void*), void* data)
{
    if (!r) return 0;
    std::string text;
    for (int c = r(data); c != EOF && c != 0; c = r(data)) text.push_back((char)c);
    return aoclfftz_gpu_tune_import_string(text.c_str()) >= 0;
}

void* host_alloc(size_t bytes)
{
    void* p = nullptr;
    if (posix_memalign(&p, 64, bytes ? bytes : 64) != 0) return nullptr;
    return p;
}

void flops_of(void* p, double* add, double* mul, double* fmas)
{
    // conventional split of the 5 N log2 N count: 3 N log2 N additions and
    // 2 N log2 N multiplications, no fused multiply-adds
    const double f = p ? aoclfftz_gpu_plan_flops(p) : 0.0;
    if (add) *add = 0.6 * f;
    if (mul) *mul = 0.4 * f;
    if (fmas) *fmas = 0.0;
}

void print_to(void* p, FILE* f)
{
    if (p && f) fprintf(f, "(aoclfftz-mi355x %s)\n", aoclfftz_gpu_plan_info(p));
}

char* sprint(void* p)
{
    std::string s = std::string("(aoclfftz-mi355x ") + (p ? aoclfftz_gpu_plan_info(p) : "") + ")";
    char* out = (char*)malloc(s.size() + 1);
    if (out) memcpy(out, s.c_str(), s.size() + 1);
    return out;
}

}  // namespace

#define FFTZ_FFTW_DEFINE(X, R, C)                                                                          \
    X##plan X##plan_many_dft(int rank, const int* n, int howmany, C* in, const int* inembed, int istride,  \
                             int idist, C* out, const int* onembed, int ostride, int odist, int sign,      \
                             unsigned)                                                                     \
    {                                                                                                      \
        return plan_many<R>(rank, n, howmany, in, inembed, istride, idist, out, onembed, ostride, odist,   \
                            sign, C2C);                                                                    \
    }                                                                                                      \
    X##plan X##plan_dft(int rank, const int* n, C* in, C* out, int sign, unsigned f)                       \
    {                                                                                                      \
        return X##plan_many_dft(rank, n, 1, in, nullptr, 1, 0, out, nullptr, 1, 0, sign, f);               \
    }                                                                                                      \
    X##plan X##plan_dft_1d(int n0, C* in, C* out, int sign, unsigned f)                                    \
    {                                                                                                      \
        int n[1] = {n0};                                                                                   \
        return X##plan_dft(1, n, in, out, sign, f);                                                        \
    }                                                                                                      \
    X##plan X##plan_dft_2d(int n0, int n1, C* in, C* out, int sign, unsigned f)                            \
    {                                                                                                      \
        int n[2] = {n0, n1};                                                                               \
        return X##plan_dft(2, n, in, out, sign, f);                                                        \
    }                                                                                                      \
    X##plan X##plan_dft_3d(int n0, int n1, int n2, C* in, C* out, int sign, unsigned f)                    \
    {                                                                                                      \
        int n[3] = {n0, n1, n2};                                                                           \
        return X##plan_dft(3, n, in, out, sign, f);                                                        \
    }                                                                                                      \
    X##plan X##plan_guru_dft(int rank, const X##iodim* dims, int hr, const X##iodim* hd, C* in, C* out,    \
                             int sign, unsigned)                                                           \
    {                                                                                                      \
        return plan_guru<R>(rank, dims, hr, hd, in, out, sign, C2C);                                       \
    }                                                                                                      \
    X##plan X##plan_guru64_dft(int rank, const X##iodim64* dims, int hr, const X##iodim64* hd, C* in,      \
                               C* out, int sign, unsigned)                                                 \
    {                                                                                                      \
        return plan_guru<R>(rank, dims, hr, hd, in, out, sign, C2C);                                       \
    }                                                                                                      \
    X##plan X##plan_many_dft_r2c(int rank, const int* n, int howmany, R* in, const int* inembed,           \
                                 int istride, int idist, C* out, const int* onembed, int ostride,          \
                                 int odist, unsigned)                                                      \
    {                                                                                                      \
        return plan_many<R>(rank, n, howmany, in, inembed, istride, idist, out, onembed, ostride, odist,   \
                            FFTW_FORWARD, R2C);                                                            \
    }                                                                                                      \
    X##plan X##plan_dft_r2c(int rank, const int* n, R* in, C* out, unsigned f)                             \
    {                                                                                                      \
        return X##plan_many_dft_r2c(rank, n, 1, in, nullptr, 1, 0, out, nullptr, 1, 0, f);                 \
    }                                                                                                      \
    X##plan X##plan_dft_r2c_1d(int n0, R* in, C* out, unsigned f)                                          \
    {                                                                                                      \
        int n[1] = {n0};                                                                                   \
        return X##plan_dft_r2c(1, n, in, out, f);                                                          \
    }                                                                                                      \
    X##plan X##plan_dft_r2c_2d(int n0, int n1, R* in, C* out, unsigned f)                                  \
    {                                                                                                      \
        int n[2] = {n0, n1};                                                                               \
        return X##plan_dft_r2c(2, n, in, out, f);                                                          \
    }                                                                                                      \
    X##plan X##plan_dft_r2c_3d(int n0, int n1, int n2, R* in, C* out, unsigned f)                          \
    {                                                                                                      \
        int n[3] = {n0, n1, n2};                                                                           \
        return X##plan_dft_r2c(3, n, in, out, f);                                                          \
    }                                                                                                      \
    X##plan X##plan_guru_dft_r2c(int rank, const X##iodim* dims, int hr, const X##iodim* hd, R* in,        \
                                 C* out, unsigned)                                                         \
    {                                                                                                      \
        return plan_guru<R>(rank, dims, hr, hd, in, out, FFTW_FORWARD, R2C);                               \
    }                                                                                                      \
    X##plan X##plan_guru64_dft_r2c(int rank, const X##iodim64* dims, int hr, const X##iodim64* hd, R* in,  \
                                   C* out, unsigned)                                                       \
    {                                                                                                      \
        return plan_guru<R>(rank, dims, hr, hd, in, out, FFTW_FORWARD, R2C);                               \
    }                                                                                                      \
    X##plan X##plan_many_dft_c2r(int rank, const int* n, int howmany, C* in, const int* inembed,           \
                                 int istride, int idist, R* out, const int* onembed, int ostride,          \
                                 int odist, unsigned)                                                      \
    {                                                                                                      \
        return plan_many<R>(rank, n, howmany, in, inembed, istride, idist, out, onembed, ostride, odist,   \
                            FFTW_BACKWARD, C2R);                                                           \
    }                                                                                                      \
    X##plan X##plan_dft_c2r(int rank, const int* n, C* in, R* out, unsigned f)                             \
    {                                                                                                      \
        return X##plan_many_dft_c2r(rank, n, 1, in, nullptr, 1, 0, out, nullptr, 1, 0, f);                 \
    }                                                                                                      \
    X##plan X##plan_dft_c2r_1d(int n0, C* in, R* out, unsigned f)                                          \
    {                                                                                                      \
        int n[1] = {n0};                                                                                   \
        return X##plan_dft_c2r(1, n, in, out, f);                                                          \
    }                                                                                                      \
    X##plan X##plan_dft_c2r_2d(int n0, int n1, C* in, R* out, unsigned f)                                  \
    {                                                                                                      \
        int n[2] = {n0, n1};                                                                               \
        return X##plan_dft_c2r(2, n, in, out, f);                                                          \
    }                                                                                                      \
    X##plan X##plan_dft_c2r_3d(int n0, int n1, int n2, C* in, R* out, unsigned f)                          \
    {                                                                                                      \
        int n[3] = {n0, n1, n2};                                                                           \
        return X##plan_dft_c2r(3, n, in, out, f);                                                          \
    }                                                                                                      \
    X##plan X##plan_guru_dft_c2r(int rank, const X##iodim* dims, int hr, const X##iodim* hd, C* in,        \
                                 R* out, unsigned)                                                         \
    {                                                                                                      \
        return plan_guru<R>(rank, dims, hr, hd, in, out, FFTW_BACKWARD, C2R);                              \
    }                                                                                                      \
    X##plan X##plan_guru64_dft_c2r(int rank, const X##iodim64* dims, int hr, const X##iodim64* hd, C* in,  \
                                   R* out, unsigned)                                                       \
    {                                                                                                      \
        return plan_guru<R>(rank, dims, hr, hd, in, out, FFTW_BACKWARD, C2R);                              \
    }                                                                                                      \
    X##plan X##plan_r2r(int, const int*, R*, R*, const X##r2r_kind*, unsigned)                               \
    {                                                                                                      \
        return unsupported_r2r(__func__);                                                                  \
    }                                                                                                      \
    X##plan X##plan_r2r_1d(int, R*, R*, X##r2r_kind, unsigned) { return unsupported_r2r(__func__); }       \
    X##plan X##plan_r2r_2d(int, int, R*, R*, X##r2r_kind, X##r2r_kind, unsigned)                           \
    {                                                                                                      \
        return unsupported_r2r(__func__);                                                                  \
    }                                                                                                      \
    X##plan X##plan_r2r_3d(int, int, int, R*, R*, X##r2r_kind, X##r2r_kind, X##r2r_kind, unsigned)         \
    {                                                                                                      \
        return unsupported_r2r(__func__);                                                                  \
    }                                                                                                      \
    X##plan X##plan_many_r2r(int, const int*, int, R*, const int*, int, int, R*, const int*, int, int,      \
                             const X##r2r_kind*, unsigned)                                                 \
    {                                                                                                      \
        return unsupported_r2r(__func__);                                                                  \
    }                                                                                                      \
    X##plan X##plan_guru_r2r(int, const X##iodim*, int, const X##iodim*, R*, R*, const X##r2r_kind*,       \
                             unsigned)                                                                     \
    {                                                                                                      \
        return unsupported_r2r(__func__);                                                                  \
    }                                                                                                      \
    X##plan X##plan_guru64_r2r(int, const X##iodim64*, int, const X##iodim64*, R*, R*,                     \
                               const X##r2r_kind*, unsigned)                                               \
    {                                                                                                      \
        return unsupported_r2r(__func__);                                                                  \
    }                                                                                                      \
    void X##execute(const X##plan p) { (void)aoclfftz_execute(p); }                                        \
    void X##execute_dft(const X##plan p, C* in, C* out) { (void)aoclfftz_execute_io(p, in, out); }         \
    void X##execute_dft_r2c(const X##plan p, R* in, C* out) { (void)aoclfftz_execute_io(p, in, out); }     \
    void X##execute_dft_c2r(const X##plan p, C* in, R* out) { (void)aoclfftz_execute_io(p, in, out); }     \
    void X##execute_r2r(const X##plan, R*, R*) { (void)unsupported_r2r(__func__); }                        \
    void X##destroy_plan(X##plan p) { aoclfftz_destroy(p); }                                               \
    void X##cleanup(void) {}                                                                               \
    void X##set_timelimit(double) {}                                                                       \
    void X##plan_with_nthreads(int nthreads) { g_nthreads = nthreads < 1 ? 1 : nthreads; }                 \
    int X##planner_nthreads(void) { return g_nthreads; }                                                   \
    int X##init_threads(void) { return 1; }                                                                \
    void X##cleanup_threads(void) {}                                                                       \
    void X##make_planner_thread_safe(void) {}                                                              \
    void* X##malloc(size_t n) { return host_alloc(n); }                                                    \
    R* X##alloc_real(size_t n) { return (R*)host_alloc(n * sizeof(R)); }                                   \
    C* X##alloc_complex(size_t n) { return (C*)host_alloc(n * sizeof(C)); }                                \
    void X##free(void* p) { free(p); }                                                                     \
    void X##flops(const X##plan p, double* add, double* mul, double* fmas) { flops_of(p, add, mul, fmas); } \
    double X##estimate_cost(const X##plan p) { return p ? aoclfftz_gpu_plan_flops(p) : 0.0; }              \
    double X##cost(const X##plan p) { return p ? aoclfftz_gpu_plan_flops(p) : 0.0; }                       \
    int X##alignment_of(R* p) { return (int)(((uintptr_t)p) % 64); }                                       \
    void X##fprint_plan(const X##plan p, FILE* f) { print_to(p, f); }                                      \
    void X##print_plan(const X##plan p) { print_to(p, stdout); }                                           \
    char* X##sprint_plan(const X##plan p) { return sprint(p); }                                            \
    void X##forget_wisdom(void) { aoclfftz_gpu_tune_clear(); }                                             \
    int X##export_wisdom_to_filename(const char* f) { return aoclfftz_gpu_tune_export(f) >= 0; }           \
    void X##export_wisdom_to_file(FILE* f) { wisdom_to_file(f); }                                          \
    char* X##export_wisdom_to_string(void) { return aoclfftz_gpu_tune_export_string(); }                   \
    void X##export_wisdom(X##write_char_func w, void* data) { wisdom_to_callback(w, data); }               \
    int X##import_system_wisdom(void) { return 0; }                                                        \
    int X##import_wisdom_from_filename(const char* f) { return aoclfftz_gpu_tune_import(f) >= 0; }         \
    int X##import_wisdom_from_file(FILE* f) { return wisdom_from_file(f); }                                \
    int X##import_wisdom_from_string(const char* s) { return aoclfftz_gpu_tune_import_string(s) >= 0; }    \
    int X##import_wisdom(X##read_char_func r, void* data) { return wisdom_from_callback(r, data); }        \
    const char X##version[] = "fftw-3.3.10-aoclfftz-mi355x";

extern "C" {
FFTZ_FFTW_DEFINE(fftw_, double, fftw_complex)
FFTZ_FFTW_DEFINE(fftwf_, float, fftwf_complex)
}
