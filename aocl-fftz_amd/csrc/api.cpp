// C ABI of the MI355X FFT library: aoclfftz_setup_* / execute / execute_io /
// destroy / version (drop-in for the reference's api/aoclfftz_api.c) and the
// GPU extensions of aoclfftz_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/aoclfftz.h"
#include "../../include/aoclfftz_gpu.h"
#include "noalloc.h"
#include "plan.h"

namespace {

constexpr uint32_t kMagic = 0x46465a47u;  // "GZFF"

struct Handle {
    uint32_t magic = kMagic;
    fftz::Plan* plan = nullptr;
    std::mutex amu;                // serialises stream-mode calls
    fftz::Slot* aslot = nullptr;   // slot used in stream mode
    hipStream_t stream = nullptr;  // bound user stream (stream mode)
    std::string info;
};

void report_error(const char* func, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "[ERROR] : aoclfftz : %s : ", func);
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
}

const char* status_name(int s)
{
    switch (s) {
        case AOCLFFTZ_TIME_OUT: return "AOCLFFTZ_TIME_OUT";
        case AOCLFFTZ_MEMORY_FAILURE: return "AOCLFFTZ_MEMORY_FAILURE";
        case AOCLFFTZ_INVALID_INPUT: return "AOCLFFTZ_INVALID_INPUT";
        case AOCLFFTZ_SETUP_FAILURE: return "AOCLFFTZ_SETUP_FAILURE";
        case AOCLFFTZ_EXECUTION_FAILURE: return "AOCLFFTZ_EXECUTION_FAILURE";
        case AOCLFFTZ_SUCCESS: return "AOCLFFTZ_SUCCESS";
        default: return "Unrecognized error code";
    }
}

// Descriptor validation with the reference's rules and order
// (api/validate_problem.h:346-417).  Like the reference it normalises some
// control fields in place (logger_mode, measure_stats, opt_level,
// dynamic_load_model, num_threads) instead of rejecting them.
template <typename Desc, typename Dim>
int validate(Desc* pd)
{
    const char* F = "validate_problem";
    if (pd == nullptr) return AOCLFFTZ_INVALID_INPUT;
    if (pd->dim_rank <= 0) {
        report_error(F, "Dim rank must be atleast 1");
        return AOCLFFTZ_INVALID_INPUT;
    }
    if (pd->vec_rank <= 0) {
        report_error(F, "Vec rank must be atleast 1");
        return AOCLFFTZ_INVALID_INPUT;
    }
    if (pd->dims == nullptr) {
        report_error(F, "Dims cannot be NULL");
        return AOCLFFTZ_INVALID_INPUT;
    }
    for (int i = 0; i < pd->dim_rank; i++) {
        const Dim& d = pd->dims[i];
        if (d.n <= 0) {
            report_error(F, "Dimension[%d]: size must be atleast 1", i);
            return AOCLFFTZ_INVALID_INPUT;
        }
        if (d.in_stride <= 0) {
            report_error(F, "Dimension[%d]: in_stride must be greater than zero", i);
            return AOCLFFTZ_INVALID_INPUT;
        }
        if (d.out_stride <= 0) {
            report_error(F, "Dimension[%d]: out_stride must be greater than zero", i);
            return AOCLFFTZ_INVALID_INPUT;
        }
    }
    if (pd->vecs == nullptr) {
        report_error(F, "Vecs cannot be NULL");
        return AOCLFFTZ_INVALID_INPUT;
    }
    for (int i = 0; i < pd->vec_rank; i++) {
        const Dim& v = pd->vecs[i];
        if (v.n < 0) {
            report_error(F, "Vector[%d]: size cannot be negative", i);
            return AOCLFFTZ_INVALID_INPUT;
        }
        if (v.in_stride <= 0) {
            report_error(F, "Vector[%d]: in_stride must be greater than zero", i);
            return AOCLFFTZ_INVALID_INPUT;
        }
        if (v.out_stride <= 0) {
            report_error(F, "Vector[%d]: out_stride must be greater than zero", i);
            return AOCLFFTZ_INVALID_INPUT;
        }
    }
    aoclfftz_cntrl_params_t* c = &pd->cntrl_params;
    if ((int)c->logger_mode < 0 || (int)c->logger_mode > 3) c->logger_mode = AOCLFFTZ_LOG_NONE;
    if (c->measure_stats != 0) c->measure_stats = 0;
    if (!c->opt_off && (c->opt_level < 0 || c->opt_level > 3)) c->opt_level = 3;
    const aoclfftz_flags_t& fl = pd->flags;
    if (fl.fft_type > 1) {
        report_error(F, "fft_type can be Complex(0) or Real(1)");
        return AOCLFFTZ_INVALID_INPUT;
    }
    if (fl.fft_direction > 1) {
        report_error(F, "fft_direction can be Forward(0) or Backward(1)");
        return AOCLFFTZ_INVALID_INPUT;
    }
    if (fl.storage_order) {
        report_error(F, "Library does not support out-of-order outputs");
        return AOCLFFTZ_INVALID_INPUT;
    }
    if (fl.fft_placement > 1) {
        report_error(F, "fft_placement can be In-place(0) or Out-of-place(1)");
        return AOCLFFTZ_INVALID_INPUT;
    }
    if (fl.transpose_mode) {
        report_error(F, "Library does not support standalone transpose");
        return AOCLFFTZ_INVALID_INPUT;
    }
    if (fl.bit_reproducibility > 1) {
        report_error(F, "bit_reproducibility can be Disable(0) or Enable(1)");
        return AOCLFFTZ_INVALID_INPUT;
    }
    if (!fl.fft_placement) {
        if ((const void*)pd->in != (const void*)pd->out) {
            report_error(F, "Input and output buffer must be the same for in-place problems");
            return AOCLFFTZ_INVALID_INPUT;
        }
        if (!fl.fft_type) {
            for (int i = 0; i < pd->dim_rank; i++)
                if (pd->dims[i].in_stride != pd->dims[i].out_stride) {
                    report_error(F, "For dimension[%d]: in_stride (%d) != out_stride (%d)", i,
                                 (int)pd->dims[i].in_stride, (int)pd->dims[i].out_stride);
                    return AOCLFFTZ_INVALID_INPUT;
                }
            for (int i = 0; i < pd->vec_rank; i++)
                if (pd->vecs[i].in_stride != pd->vecs[i].out_stride) {
                    report_error(F, "For vecs[%d]: in_stride (%d) != out_stride (%d)", i,
                                 (int)pd->vecs[i].in_stride, (int)pd->vecs[i].out_stride);
                    return AOCLFFTZ_INVALID_INPUT;
                }
        } else {
            // R2C: real-side strides are twice the complex-side strides
            // (dims[0] may also be equal); C2R the other way round.
            const int64_t in_scale = fl.fft_direction == 0 ? 2 : 1;
            const int64_t out_scale = fl.fft_direction == 1 ? 2 : 1;
            for (int i = 0; i < pd->dim_rank; i++) {
                const int64_t is = pd->dims[i].in_stride, os = pd->dims[i].out_stride;
                if (is * out_scale != os * in_scale && !(i == 0 && is == os)) {
                    report_error(F, "For dimension[%d]: in-place real strides inconsistent", i);
                    return AOCLFFTZ_INVALID_INPUT;
                }
            }
            for (int i = 0; i < pd->vec_rank; i++) {
                const int64_t is = pd->vecs[i].in_stride, os = pd->vecs[i].out_stride;
                if (is * out_scale != os * in_scale && !(i == 0 && pd->vecs[i].n == 1)) {
                    report_error(F, "For vecs[%d]: in-place real strides inconsistent", i);
                    return AOCLFFTZ_INVALID_INPUT;
                }
            }
        }
    } else if ((const void*)pd->in == (const void*)pd->out) {
        report_error(F, "Input and output buffer cannot be the same for out-of-place problems");
        return AOCLFFTZ_INVALID_INPUT;
    }
    if (pd->pthr_fft.dynamic_load_model != 0 && pd->pthr_fft.dynamic_load_model != 1)
        pd->pthr_fft.dynamic_load_model = 1;
    if (pd->pthr_fft.num_threads < 1) pd->pthr_fft.num_threads = 1;
    return AOCLFFTZ_SUCCESS;
}

template <typename Desc>
fftz::ProblemSpec spec_of(const Desc* pd, int prec_bytes)
{
    fftz::ProblemSpec spec;
    spec.prec_bytes = prec_bytes;
    spec.dir = pd->flags.fft_direction;
    spec.inplace = pd->flags.fft_placement == 0;
    if (pd->flags.fft_type == 0)
        spec.type = fftz::FftType::C2C;
    else
        spec.type = spec.dir == 0 ? fftz::FftType::R2C : fftz::FftType::C2R;
    for (int i = 0; i < pd->dim_rank; i++)
        spec.dims.push_back({(int64_t)pd->dims[i].n, (int64_t)pd->dims[i].in_stride, (int64_t)pd->dims[i].out_stride});
    for (int i = 0; i < pd->vec_rank; i++)
        spec.vecs.push_back({(int64_t)pd->vecs[i].n, (int64_t)pd->vecs[i].in_stride, (int64_t)pd->vecs[i].out_stride});
    return spec;
}

template <typename Desc, typename Dim>
void* setup_common(Desc* pd, int prec_bytes, const char* fname)
{
    int rc = validate<Desc, Dim>(pd);
    if (rc != AOCLFFTZ_SUCCESS) {
        report_error(fname, "Problem descriptor validation failed with %s", status_name(rc));
        return nullptr;
    }
    fftz::ProblemSpec spec = spec_of(pd, prec_bytes);
    std::string err;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        report_error(fname, "no HIP device available");
        return nullptr;
    }
    fftz::Plan* plan = fftz::plan_create(spec, err);
    if (!plan) {
        report_error(fname, "plan creation failed: %s", err.c_str());
        return nullptr;
    }
    plan->setup_in = (void*)pd->in;
    plan->setup_out = (void*)pd->out;
    Handle* h = new Handle();
    h->plan = plan;
    return h;
}

Handle* as_handle(void* p)
{
    Handle* h = static_cast<Handle*>(p);
    if (h == nullptr || h->magic != kMagic) return nullptr;
    return h;
}

bool is_device_ptr(const void* p)
{
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

aoclfftz_error_type run(Handle* h, void* in, void* out, const char* fname)
{
    fftz::Plan* p = h->plan;
    if (p->noop) return AOCLFFTZ_SUCCESS;
    if (p->spec.inplace && in != out) {
        report_error(fname, "in-place plan executed with distinct buffers");
        return AOCLFFTZ_EXECUTION_FAILURE;
    }
    hipError_t e;
    if (h->stream) {
        // stream mode: the caller guarantees device buffers (aoclfftz_gpu.h);
        // no pointer queries, nothing waits.
        std::lock_guard<std::mutex> g(h->amu);
        if (!h->aslot) {
            h->aslot = fftz::plan_acquire_slot(p, &e);
            if (!h->aslot) {
                report_error(fname, "slot allocation failed: %s", hipGetErrorString(e));
                return AOCLFFTZ_MEMORY_FAILURE;
            }
        }
        e = fftz::plan_enqueue(p, h->aslot, in, out, h->stream);
        if (e != hipSuccess) {
            report_error(fname, "launch failed: %s", hipGetErrorString(e));
            return AOCLFFTZ_EXECUTION_FAILURE;
        }
        return AOCLFFTZ_SUCCESS;
    }
    const bool in_dev = is_device_ptr(in);
    const bool out_dev = is_device_ptr(out);
    fftz::Slot* s = fftz::plan_acquire_slot(p, &e);
    if (!s) {
        report_error(fname, "slot allocation failed: %s", hipGetErrorString(e));
        return AOCLFFTZ_MEMORY_FAILURE;
    }
    aoclfftz_error_type rc = AOCLFFTZ_SUCCESS;
    void* din = in;
    void* dout = out;
    do {
        if (!in_dev || !out_dev) {
            // host buffers: stage through device memory (reported separately
            // from the device-resident rate; PCIe-bound)
            const size_t need_in = (size_t)p->in_extent_bytes;
            const size_t need_out = (size_t)p->out_extent_bytes;
            const size_t need = p->spec.inplace ? std::max(need_in, need_out) : need_in;
            if (!in_dev) {
                if (s->stage_in_bytes < need) {
                    if (s->stage_in) (void)hipFree(s->stage_in);
                    s->stage_in = nullptr;
                    s->stage_in_bytes = 0;
                    if (hipMalloc(&s->stage_in, need) != hipSuccess) {
                        rc = AOCLFFTZ_MEMORY_FAILURE;
                        break;
                    }
                    s->stage_in_bytes = need;
                }
                din = s->stage_in;
                if (hipMemcpyAsync(din, in, need_in, hipMemcpyHostToDevice, s->stream) != hipSuccess) {
                    rc = AOCLFFTZ_EXECUTION_FAILURE;
                    break;
                }
            }
            if (p->spec.inplace) {
                dout = din;
            } else if (!out_dev) {
                if (s->stage_out_bytes < need_out) {
                    if (s->stage_out) (void)hipFree(s->stage_out);
                    s->stage_out = nullptr;
                    s->stage_out_bytes = 0;
                    if (hipMalloc(&s->stage_out, need_out) != hipSuccess) {
                        rc = AOCLFFTZ_MEMORY_FAILURE;
                        break;
                    }
                    s->stage_out_bytes = need_out;
                }
                dout = s->stage_out;
                // strided outputs: keep the caller's untouched elements
                if (hipMemcpyAsync(dout, out, need_out, hipMemcpyHostToDevice, s->stream) != hipSuccess) {
                    rc = AOCLFFTZ_EXECUTION_FAILURE;
                    break;
                }
            }
        }
        e = fftz::plan_enqueue(p, s, din, dout, s->stream);
        if (e != hipSuccess) {
            report_error(fname, "launch failed: %s", hipGetErrorString(e));
            rc = AOCLFFTZ_EXECUTION_FAILURE;
            break;
        }
        if (dout != out) {
            if (hipMemcpyAsync(out, dout, (size_t)p->out_extent_bytes, hipMemcpyDeviceToHost, s->stream) !=
                hipSuccess) {
                rc = AOCLFFTZ_EXECUTION_FAILURE;
                break;
            }
        }
        e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) {
            report_error(fname, "execution failed: %s", hipGetErrorString(e));
            rc = AOCLFFTZ_EXECUTION_FAILURE;
        }
    } while (0);
    if (rc != AOCLFFTZ_SUCCESS) (void)hipStreamSynchronize(s->stream);
    fftz::plan_release_slot(p, s);
    return rc;
}

}  // namespace

extern "C" {

FFTZ_VOID* aoclfftz_setup_f(aoclfftz_prob_desc_f* problem)
{
    return setup_common<aoclfftz_prob_desc_f, aoclfftz_dim_t>(problem, 4, "aoclfftz_setup_f");
}

FFTZ_VOID* aoclfftz_setup_d(aoclfftz_prob_desc_d* problem)
{
    return setup_common<aoclfftz_prob_desc_d, aoclfftz_dim_t>(problem, 8, "aoclfftz_setup_d");
}

FFTZ_VOID* aoclfftz_setup_f_64_(aoclfftz_prob_desc_f_64_* problem)
{
    return setup_common<aoclfftz_prob_desc_f_64_, aoclfftz_dim_t_64_>(problem, 4, "aoclfftz_setup_f_64_");
}

FFTZ_VOID* aoclfftz_setup_d_64_(aoclfftz_prob_desc_d_64_* problem)
{
    return setup_common<aoclfftz_prob_desc_d_64_, aoclfftz_dim_t_64_>(problem, 8, "aoclfftz_setup_d_64_");
}

aoclfftz_error_type aoclfftz_execute(FFTZ_VOID* handle)
{
    Handle* h = as_handle(handle);
    if (!h) return AOCLFFTZ_EXECUTION_FAILURE;
    if (h->plan->setup_in == nullptr || h->plan->setup_out == nullptr) {
        report_error("aoclfftz_execute", "One or both setup-time buffers are NULL; use aoclfftz_execute_io");
        return AOCLFFTZ_EXECUTION_FAILURE;
    }
    return run(h, h->plan->setup_in, h->plan->setup_out, "aoclfftz_execute");
}

aoclfftz_error_type aoclfftz_execute_io(FFTZ_VOID* handle, FFTZ_VOID* in, FFTZ_VOID* out)
{
    Handle* h = as_handle(handle);
    if (!h || in == nullptr || out == nullptr) return AOCLFFTZ_EXECUTION_FAILURE;
    return run(h, in, out, "aoclfftz_execute_io");
}

FFTZ_VOID aoclfftz_destroy(FFTZ_VOID* handle)
{
    Handle* h = as_handle(handle);
    if (!h) return;
    h->magic = 0;
    fftz::plan_destroy(h->plan);
    delete h;
}

const FFTZ_CHAR* aoclfftz_version(FFTZ_VOID)
{
    return AOCLFFTZ_LIBRARY_VERSION " MI355X gfx950";
}

aoclfftz_error_type aoclfftz_gpu_set_stream(FFTZ_VOID* handle, FFTZ_VOID* stream)
{
    Handle* h = as_handle(handle);
    if (!h) return AOCLFFTZ_EXECUTION_FAILURE;
    std::lock_guard<std::mutex> g(h->amu);
    h->stream = (hipStream_t)stream;
    if (!stream && h->aslot) {
        fftz::plan_release_slot(h->plan, h->aslot);
        h->aslot = nullptr;
    }
    return AOCLFFTZ_SUCCESS;
}

aoclfftz_error_type aoclfftz_gpu_synchronize(FFTZ_VOID* handle)
{
    Handle* h = as_handle(handle);
    if (!h) return AOCLFFTZ_EXECUTION_FAILURE;
    hipError_t e = hipSuccess;
    if (h->stream) e = hipStreamSynchronize(h->stream);
    std::lock_guard<std::mutex> g(h->plan->mu);
    for (fftz::Slot* s : h->plan->slots)
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? AOCLFFTZ_SUCCESS : AOCLFFTZ_EXECUTION_FAILURE;
}

aoclfftz_error_type aoclfftz_gpu_set_graph(FFTZ_VOID* handle, FFTZ_INT32 enable)
{
    Handle* h = as_handle(handle);
    if (!h) return AOCLFFTZ_EXECUTION_FAILURE;
    h->plan->use_graph = enable != 0;
    return AOCLFFTZ_SUCCESS;
}

aoclfftz_error_type aoclfftz_gpu_autotune(FFTZ_VOID* handle, FFTZ_VOID* in, FFTZ_VOID* out, FFTZ_INT32 reps)
{
    Handle* h = as_handle(handle);
    if (!h || !in || !out || reps <= 0) return AOCLFFTZ_EXECUTION_FAILURE;
    if (!is_device_ptr(in) || !is_device_ptr(out)) {
        report_error(__func__, "autotuning runs on device buffers");
        return AOCLFFTZ_INVALID_INPUT;
    }
    if (h->plan->spec.inplace && in != out) {
        report_error(__func__, "an in-place handle is tuned on one buffer (in == out)");
        return AOCLFFTZ_INVALID_INPUT;
    }
    std::lock_guard<std::mutex> g(h->amu);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    fftz::Plan* p = h->plan;
    if (p->noop) return AOCLFFTZ_SUCCESS;
    if (p->is_batch) {
        // workgroups per CU of the small and medium tile classes
        double best = 1e30;
        int bs = 0, bm = 0;
        for (int ws : {4, 8, 16})
            for (int wm : {1, 2}) {
                fftz::batch_set_grids(p, ws, wm);
                const double t = fftz::plan_time(p, in, out, reps);
                if (t >= 0 && t < best) {
                    best = t;
                    bs = ws;
                    bm = wm;
                }
            }
        if (!bs) return AOCLFFTZ_EXECUTION_FAILURE;
        fftz::batch_set_grids(p, bs, bm);
        fftz::TuneParams t;
        t.wg_small = bs;
        t.wg_medium = bm;
        fftz::tune_store(p->tune_key, t);
        return AOCLFFTZ_SUCCESS;
    }
    // single-size plans: row tile width, and column tile lines when the plan
    // has column passes
    bool has_col = false;
    for (const auto& ln : p->launches) has_col = has_col || (ln.d.flags & (fftz::kInCol | fftz::kOutCol));
    std::vector<fftz::TuneParams> cands;
    for (int rt : {64, 128, 256, 512, 1024}) {
        if (!has_col) {
            fftz::TuneParams t;
            t.row_t = rt;
            cands.push_back(t);
            continue;
        }
        for (int cc : {4, 8, 16, 32}) {
            fftz::TuneParams t;
            t.row_t = rt;
            t.col_c = cc;
            cands.push_back(t);
        }
    }
    fftz::Plan* best_plan = nullptr;
    double best = 1e30;
    auto trial = [&](const fftz::TuneParams& t) {
        fftz::ProblemSpec spec = p->spec;
        spec.tune = t;
        std::string err;
        fftz::Plan* q = fftz::plan_create(spec, err);
        if (!q) return;
        const double ms = fftz::plan_time(q, in, out, reps);
        if (ms >= 0 && ms < best) {
            if (best_plan) fftz::plan_destroy(best_plan);
            best_plan = q;
            best = ms;
        } else {
            fftz::plan_destroy(q);
        }
    };
    // 1. tile shapes
    for (const fftz::TuneParams& t : cands) trial(t);
    if (!best_plan) return AOCLFFTZ_EXECUTION_FAILURE;
    // 2. four-step split of the long axis (plans with a work buffer whose
    //    problem is one long 1-D axis)
    if (p->spec.dims.size() == 1 && p->work_bytes > 0 && p->spec.type == fftz::FftType::C2C) {
        const fftz::TuneParams base = best_plan->spec.tune;
        for (const auto& f : fftz::split_candidates(p->spec.dims[0].n, p->spec.prec_bytes, 6)) {
            fftz::TuneParams t = base;
            t.split = f;
            trial(t);
        }
    }
    // 3. radix order of each distinct pass length, one length at a time
    {
        std::vector<int> lengths;
        for (const auto& ln : best_plan->launches)
            if (ln.kind == fftz::KernelKind::Pass && ln.d.nst >= 2 && !ln.mfma &&
                std::find(lengths.begin(), lengths.end(), ln.d.L) == lengths.end())
                lengths.push_back(ln.d.L);
        for (int L : lengths) {
            const fftz::TuneParams base = best_plan->spec.tune;
            const auto orders = fftz::radix_orders(L, 6);
            for (size_t i = 1; i < orders.size(); i++) {  // [0] is the default order
                fftz::TuneParams t = base;
                t.radix_for.push_back({L, orders[i]});
                trial(t);
            }
        }
    }
    // 4. complex values per thread
    if (p->spec.mfma_mode == fftz::kMfmaNone) {
        const fftz::TuneParams base = best_plan->spec.tune;
        for (int e : {8, 16}) {
            fftz::TuneParams t = base;
            t.elems = e;
            trial(t);
        }
    }
    fftz::ProblemSpec key_spec = p->spec;
    key_spec.tune = fftz::TuneParams{};
    fftz::tune_store(fftz::problem_key(key_spec), best_plan->spec.tune);
    best_plan->setup_in = p->setup_in;
    best_plan->setup_out = p->setup_out;
    best_plan->use_graph = p->use_graph;
    best_plan->profile = p->profile;
    if (h->aslot) {
        fftz::plan_release_slot(p, h->aslot);
        h->aslot = nullptr;
    }
    fftz::plan_destroy(p);
    h->plan = best_plan;
    return AOCLFFTZ_SUCCESS;
}

FFTZ_INT32 aoclfftz_gpu_tune_export(const FFTZ_CHAR* path) { return fftz::tune_export(path); }
FFTZ_INT32 aoclfftz_gpu_tune_import(const FFTZ_CHAR* path) { return fftz::tune_import(path); }

FFTZ_CHAR* aoclfftz_gpu_tune_export_string(FFTZ_VOID)
{
    const std::string text = fftz::tune_text(nullptr);
    char* s = (char*)malloc(text.size() + 1);
    if (s) memcpy(s, text.c_str(), text.size() + 1);
    return s;
}

FFTZ_INT32 aoclfftz_gpu_tune_import_string(const FFTZ_CHAR* text)
{
    return text ? fftz::tune_parse(text) : -1;
}

FFTZ_VOID aoclfftz_gpu_tune_clear(FFTZ_VOID) { fftz::tune_clear(); }

aoclfftz_error_type aoclfftz_gpu_set_compute_mode(FFTZ_VOID* handle, FFTZ_INT32 mode)
{
    Handle* h = as_handle(handle);
    if (!h) return AOCLFFTZ_EXECUTION_FAILURE;
    if (mode < fftz::kMfmaNone || mode > fftz::kMfmaF16x2 || h->plan->is_batch) {
        report_error(__func__, "compute mode %d is not available for this handle", (int)mode);
        return AOCLFFTZ_INVALID_INPUT;
    }
    std::lock_guard<std::mutex> g(h->amu);
    fftz::Plan* old = h->plan;
    if (old->spec.mfma_mode == mode) return AOCLFFTZ_SUCCESS;
    fftz::ProblemSpec spec = old->spec;
    spec.mfma_mode = mode;
    std::string err;
    fftz::Plan* np = fftz::plan_create(spec, err);
    if (!np) {
        report_error(__func__, "re-planning failed: %s", err.c_str());
        return AOCLFFTZ_SETUP_FAILURE;
    }
    bool applies = mode == fftz::kMfmaNone;
    for (const auto& ln : np->launches) applies = applies || ln.mfma == mode;
    if (!applies) {
        fftz::plan_destroy(np);
        report_error(__func__, "compute mode %d applies to no pass of this plan (complex64, lengths 256 / 4096)",
                     (int)mode);
        return AOCLFFTZ_INVALID_INPUT;
    }
    np->setup_in = old->setup_in;
    np->setup_out = old->setup_out;
    np->use_graph = old->use_graph;
    np->profile = old->profile;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->aslot) {
        fftz::plan_release_slot(old, h->aslot);
        h->aslot = nullptr;
    }
    fftz::plan_destroy(old);
    h->plan = np;
    return AOCLFFTZ_SUCCESS;
}

const FFTZ_CHAR* aoclfftz_gpu_plan_info(FFTZ_VOID* handle)
{
    Handle* h = as_handle(handle);
    if (!h) return "";
    std::ostringstream os;
    fftz::Plan* p = h->plan;
    os << (p->spec.prec_bytes == 4 ? "fp32" : "fp64") << " " << p->description;
    if (!p->is_batch) os << "passes=" << p->launches.size();
    for (const auto& ln : p->launches) {
        static const char* kMfmaName[] = {"", "mfma-f32", "mfma-bf16x3", "mfma-f16x2"};
        os << " [" << (ln.kind == fftz::KernelKind::Bluestein ? "bluestein"
                       : ln.mfma ? kMfmaName[ln.mfma & 3]
                       : (ln.fixed ? "fixed" : "pass"))
           << " L=" << ln.d.L << " r=";
        for (int s = 0; s < ln.d.nst; s++) os << (s ? "x" : "") << ln.d.radix[s];
        os << " C=" << ln.d.C << " TL=" << ln.d.TL << " grid=" << ln.grid << " block=" << ln.block
           << " lds=" << ln.lds_bytes << ((ln.d.flags & fftz::kInCol) ? " incol" : "")
           << ((ln.d.flags & fftz::kOutCol) ? " outcol" : "") << ((ln.d.flags & fftz::kTw4) ? " tw4" : "") << "]";
    }
    os << " work=" << p->work_bytes;
    h->info = os.str();
    return h->info.c_str();
}

}  // extern "C"

namespace fftz {

bool spec_from_desc(const void* problem, int prec_bytes, bool ilp64, ProblemSpec& spec, const void** in, void** out,
                    const char* func)
{
    if (!problem || (prec_bytes != 4 && prec_bytes != 8)) return false;
    if (ilp64) {
        auto* pd = static_cast<aoclfftz_prob_desc_d_64_*>(const_cast<void*>(problem));
        int rc = validate<aoclfftz_prob_desc_d_64_, aoclfftz_dim_t_64_>(pd);
        if (rc != AOCLFFTZ_SUCCESS) {
            report_error(func, "Problem descriptor validation failed with %s", status_name(rc));
            return false;
        }
        spec = spec_of(pd, prec_bytes);
        *in = pd->in;
        *out = pd->out;
    } else {
        auto* pd = static_cast<aoclfftz_prob_desc_d*>(const_cast<void*>(problem));
        int rc = validate<aoclfftz_prob_desc_d, aoclfftz_dim_t>(pd);
        if (rc != AOCLFFTZ_SUCCESS) {
            report_error(func, "Problem descriptor validation failed with %s", status_name(rc));
            return false;
        }
        spec = spec_of(pd, prec_bytes);
        *in = pd->in;
        *out = pd->out;
    }
    return true;
}

void report_msg(const char* func, const char* msg) { report_error(func, "%s", msg); }

}  // namespace fftz

extern "C" {

FFTZ_INT64 aoclfftz_gpu_plan_json(FFTZ_VOID* problem, FFTZ_INT32 precision_bytes, FFTZ_INT32 ilp64,
                                  FFTZ_CHAR* buf, FFTZ_INT64 buflen)
{
    if (!problem || (precision_bytes != 4 && precision_bytes != 8)) return -1;
    fftz::ProblemSpec spec;
    if (ilp64) {
        auto* pd = static_cast<aoclfftz_prob_desc_d_64_*>(problem);
        if (validate<aoclfftz_prob_desc_d_64_, aoclfftz_dim_t_64_>(pd) != AOCLFFTZ_SUCCESS) return -1;
        spec = spec_of(pd, precision_bytes);
    } else {
        auto* pd = static_cast<aoclfftz_prob_desc_d*>(problem);
        if (validate<aoclfftz_prob_desc_d, aoclfftz_dim_t>(pd) != AOCLFFTZ_SUCCESS) return -1;
        spec = spec_of(pd, precision_bytes);
    }
    std::string err;
    fftz::Plan* plan = fftz::plan_create(spec, err, true);
    if (!plan) {
        report_error("aoclfftz_gpu_plan_json", "plan creation failed: %s", err.c_str());
        return -1;
    }
    std::string js = fftz::plan_json(plan);
    fftz::plan_destroy(plan);
    if (buf && buflen > 0) {
        const size_t n = std::min((size_t)(buflen - 1), js.size());
        memcpy(buf, js.data(), n);
        buf[n] = 0;
    }
    return (FFTZ_INT64)js.size();
}

aoclfftz_error_type aoclfftz_gpu_execute_many(FFTZ_VOID* handle, FFTZ_VOID** ins, FFTZ_VOID** outs,
                                             FFTZ_INT32 count)
{
    Handle* h = as_handle(handle);
    if (!h || count < 0 || (count > 0 && (!ins || !outs))) return AOCLFFTZ_EXECUTION_FAILURE;
    for (int i = 0; i < count; i++) {
        if (!ins[i] || !outs[i]) return AOCLFFTZ_EXECUTION_FAILURE;
        aoclfftz_error_type rc = run(h, ins[i], outs[i], "aoclfftz_gpu_execute_many");
        if (rc != AOCLFFTZ_SUCCESS) return rc;
    }
    return AOCLFFTZ_SUCCESS;
}

aoclfftz_error_type aoclfftz_gpu_execute_noalloc(FFTZ_VOID* handle, FFTZ_VOID* in, FFTZ_VOID* out)
{
    const char* fname = "aoclfftz_gpu_execute_noalloc";
    Handle* h = as_handle(handle);
    if (!h || in == nullptr || out == nullptr) return AOCLFFTZ_EXECUTION_FAILURE;
    fftz::Plan* p = h->plan;
    if (p->noop) return AOCLFFTZ_SUCCESS;
    if (p->spec.inplace && in != out) {
        report_error(fname, "in-place plan executed with distinct buffers");
        return AOCLFFTZ_EXECUTION_FAILURE;
    }
    const size_t counter_bytes = p->is_batch ? p->bclasses.size() * sizeof(unsigned) : 0;
    if (p->work_bytes > fftz::kNoAllocWorkBytes || counter_bytes > fftz::kNoAllocHeaderBytes) {
        report_error(fname, "plan needs %zu bytes of work memory; the allocation-free execute holds at most %zu",
                     p->work_bytes, fftz::kNoAllocWorkBytes);
        return AOCLFFTZ_INVALID_INPUT;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != p->device) {
        report_error(fname, "the current device (%d) is not the plan's device (%d)", dev, p->device);
        return AOCLFFTZ_INVALID_INPUT;
    }
    void* scratch = nullptr;
    hipError_t e = fftz::noalloc_scratch(dev, &scratch);
    if (e != hipSuccess) {
        report_error(fname, "device scratch lookup failed: %s", hipGetErrorString(e));
        return AOCLFFTZ_EXECUTION_FAILURE;
    }
    unsigned* counters = reinterpret_cast<unsigned*>(scratch);
    void* work = static_cast<unsigned char*>(scratch) + fftz::kNoAllocHeaderBytes;
    std::lock_guard<std::mutex> g(h->amu);
    if (h->stream) {
        e = fftz::plan_enqueue_scratch(p, work, counters, in, out, h->stream);
        if (e != hipSuccess) {
            report_error(fname, "launch failed: %s", hipGetErrorString(e));
            return AOCLFFTZ_EXECUTION_FAILURE;
        }
        return AOCLFFTZ_SUCCESS;
    }
    // synchronous form: device buffers only (staging host buffers would
    // allocate), launched on the null stream and waited for
    if (!is_device_ptr(in) || !is_device_ptr(out)) {
        report_error(fname, "in and out must be device memory");
        return AOCLFFTZ_INVALID_INPUT;
    }
    e = fftz::plan_enqueue_scratch(p, work, counters, in, out, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        report_error(fname, "execution failed: %s", hipGetErrorString(e));
        return AOCLFFTZ_EXECUTION_FAILURE;
    }
    return AOCLFFTZ_SUCCESS;
}

aoclfftz_error_type aoclfftz_gpu_profile(FFTZ_VOID* handle, FFTZ_INT32 enable)
{
    Handle* h = as_handle(handle);
    if (!h) return AOCLFFTZ_EXECUTION_FAILURE;
    h->plan->profile = enable != 0;
    return AOCLFFTZ_SUCCESS;
}

FFTZ_INT32 aoclfftz_gpu_profile_read(FFTZ_VOID* handle, float* ms, FFTZ_INT32 cap)
{
    Handle* h = as_handle(handle);
    if (!h) return -1;
    return fftz::plan_profile_read(h->plan, ms, cap);
}

FFTZ_INT32 aoclfftz_gpu_plan_launches(FFTZ_VOID* handle)
{
    Handle* h = as_handle(handle);
    if (!h) return -1;
    return (FFTZ_INT32)(h->plan->is_batch ? h->plan->bclasses.size() : h->plan->launches.size());
}

static bool batch_items(const aoclfftz_gpu_batch_item* items, FFTZ_INT64 count, std::vector<fftz::BatchItem>& v)
{
    if (!items || count < 0) return false;
    v.resize((size_t)count);
    for (FFTZ_INT64 i = 0; i < count; i++)
        v[(size_t)i] = {items[i].n, items[i].batch, items[i].in_offset, items[i].out_offset, items[i].dist};
    return true;
}

FFTZ_VOID* aoclfftz_gpu_batch_setup(const aoclfftz_gpu_batch_item* items, FFTZ_INT64 count,
                                    FFTZ_INT32 precision_bytes, FFTZ_INT32 direction)
{
    const char* F = "aoclfftz_gpu_batch_setup";
    std::vector<fftz::BatchItem> v;
    if (!batch_items(items, count, v) || (precision_bytes != 4 && precision_bytes != 8) ||
        (direction != 0 && direction != 1)) {
        report_error(F, "invalid arguments");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        report_error(F, "no HIP device available");
        return nullptr;
    }
    std::string err;
    fftz::Plan* plan = fftz::batch_plan_create(v, precision_bytes, direction, err);
    if (!plan) {
        report_error(F, "plan creation failed: %s", err.c_str());
        return nullptr;
    }
    Handle* h = new Handle();
    h->plan = plan;
    return h;
}

FFTZ_INT64 aoclfftz_gpu_batch_plan_json(const aoclfftz_gpu_batch_item* items, FFTZ_INT64 count,
                                        FFTZ_INT32 precision_bytes, FFTZ_INT32 direction, FFTZ_CHAR* buf,
                                        FFTZ_INT64 buflen)
{
    std::vector<fftz::BatchItem> v;
    if (!batch_items(items, count, v) || (precision_bytes != 4 && precision_bytes != 8)) return -1;
    std::string err;
    fftz::Plan* plan = fftz::batch_plan_create(v, precision_bytes, direction, err, true);
    if (!plan) {
        report_error("aoclfftz_gpu_batch_plan_json", "plan creation failed: %s", err.c_str());
        return -1;
    }
    std::string js = fftz::plan_json(plan);
    fftz::plan_destroy(plan);
    if (buf && buflen > 0) {
        const size_t n = std::min((size_t)(buflen - 1), js.size());
        memcpy(buf, js.data(), n);
        buf[n] = 0;
    }
    return (FFTZ_INT64)js.size();
}

FFTZ_DOUBLE aoclfftz_gpu_plan_flops(FFTZ_VOID* handle)
{
    Handle* h = as_handle(handle);
    return h ? h->plan->flops : 0.0;
}

FFTZ_INT64 aoclfftz_gpu_plan_device_bytes(FFTZ_VOID* handle)
{
    Handle* h = as_handle(handle);
    if (!h) return 0;
    return (FFTZ_INT64)h->plan->work_bytes;
}

}  // extern "C"
