// Serving runtime: many independent transform requests sharing one GPU.
//
// Requests (any descriptor the aoclfftz_setup_* calls accept, with device
// buffers) are submitted from any thread and return at once with an id.  A
// progress thread owns the device side:
//   * plans (and their twiddle / chirp tables) are built once per distinct
//     problem and stay resident in a plan cache;
//   * a request is admitted when fewer than max_inflight requests run and its
//     HBM footprint (the bytes it touches: input and output extents plus the
//     plan's work buffer) fits the budget next to everything already running;
//     admission is FIFO, and a request that arrives while others run is
//     admitted as soon as a slot and memory free up (continuous batching);
//   * each admitted request runs on an execution slot of its plan (own stream
//     and work buffer), so admitted requests overlap on the device;
//   * completion is detected with an event per request; latency is submit to
//     completion on the host clock.
// A request's result does not depend on what runs beside it: the same plan
// runs the same kernels on the same data whatever else is in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/aoclfftz.h"
#include "../../include/aoclfftz_gpu.h"
#include "plan.h"

namespace fftz {

bool spec_from_desc(const void* problem, int prec_bytes, bool ilp64, ProblemSpec& spec, const void** in, void** out,
                    const char* func);
void report_msg(const char* func, const char* msg);

namespace {

double now_s()
{
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

std::string spec_key(const ProblemSpec& s)
{
    std::ostringstream os;
    os << s.prec_bytes << ':' << (int)s.type << ':' << s.dir << ':' << s.inplace << ':' << s.mfma_mode << '|';
    for (const auto& a : s.dims) os << a.n << ',' << a.is << ',' << a.os << ';';
    os << '|';
    for (const auto& a : s.vecs) os << a.n << ',' << a.is << ',' << a.os << ';';
    return os.str();
}

enum ReqState : int { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Req {
    int64_t id = 0;
    Plan* plan = nullptr;
    const void* in = nullptr;
    void* out = nullptr;
    size_t footprint = 0;
    int64_t transforms = 0;
    Slot* slot = nullptr;
    hipEvent_t ev = nullptr;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int state = kQueued;
};

constexpr uint32_t kServerMagic = 0x53525646u;  // "FVRS"

struct Server {
    uint32_t magic = kServerMagic;
    size_t budget = 0;
    int max_inflight = 8;
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread progress;
    bool stop = false;
    std::mutex plan_mu;
    std::unordered_map<std::string, Plan*> plans;
    std::deque<Req*> queue;
    std::vector<Req*> running;
    std::unordered_map<int64_t, Req*> reqs;
    std::vector<hipEvent_t> event_pool;
    std::vector<double> latencies;
    int64_t next_id = 0, completed = 0, failed = 0, transforms_done = 0;
    size_t in_use = 0, peak = 0, queue_peak = 0;
    double t_first = -1.0, t_last = 0.0, flops_done = 0.0;
};

// Called with s.mu held.
void admit_and_reap(Server& s)
{
    for (size_t i = 0; i < s.running.size();) {
        Req* r = s.running[i];
        const hipError_t q = hipEventQuery(r->ev);
        if (q == hipErrorNotReady) {
            i++;
            continue;
        }
        r->t_done = now_s();
        r->state = q == hipSuccess ? kDone : kFailed;
        if (q != hipSuccess) (void)hipGetLastError();
        plan_release_slot(r->plan, r->slot);
        r->slot = nullptr;
        s.in_use -= r->footprint;
        if (r->state == kDone) {
            s.completed++;
            s.transforms_done += r->transforms;
            s.flops_done += r->plan->flops;
            s.latencies.push_back(r->t_done - r->t_submit);
            s.t_last = r->t_done;
        } else {
            s.failed++;
        }
        s.running[i] = s.running.back();
        s.running.pop_back();
        s.cv_done.notify_all();
    }
    while (!s.queue.empty() && (int)s.running.size() < s.max_inflight) {
        Req* r = s.queue.front();
        if (!s.running.empty() && s.in_use + r->footprint > s.budget) break;  // FIFO: wait for memory
        s.queue.pop_front();
        // a request that cannot start completes as failed, with its
        // completion time set so its reported latency stays meaningful
        auto fail = [&](Req* q) {
            q->t_done = now_s();
            q->state = kFailed;
            s.failed++;
            s.cv_done.notify_all();
        };
        hipError_t e = hipSuccess;
        r->slot = plan_acquire_slot(r->plan, &e);
        if (!r->slot) {
            fail(r);
            continue;
        }
        if (s.event_pool.empty()) {
            hipEvent_t ev;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
                plan_release_slot(r->plan, r->slot);
                fail(r);
                continue;
            }
            s.event_pool.push_back(ev);
        }
        r->ev = s.event_pool.back();
        s.event_pool.pop_back();
        e = plan_enqueue(r->plan, r->slot, r->in, r->out, r->slot->stream);
        if (e == hipSuccess) e = hipEventRecord(r->ev, r->slot->stream);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            (void)hipStreamSynchronize(r->slot->stream);
            s.event_pool.push_back(r->ev);
            r->ev = nullptr;
            plan_release_slot(r->plan, r->slot);
            fail(r);
            continue;
        }
        r->state = kRunning;
        r->t_admit = now_s();
        s.in_use += r->footprint;
        s.peak = std::max(s.peak, s.in_use);
        s.running.push_back(r);
    }
}

void progress_loop(Server* s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop || !s->running.empty() || !s->queue.empty()) {
        admit_and_reap(*s);
        if (s->running.empty() && s->queue.empty()) {
            if (s->stop) break;
            s->cv_work.wait(lk);
            continue;
        }
        lk.unlock();
        std::this_thread::yield();
        lk.lock();
    }
}

Server* as_server(void* p)
{
    Server* s = static_cast<Server*>(p);
    return (s && s->magic == kServerMagic) ? s : nullptr;
}

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace
}  // namespace fftz

using fftz::Server;

extern "C" {

FFTZ_VOID* aoclfftz_gpu_server_create(FFTZ_INT64 hbm_budget_bytes, FFTZ_INT32 max_inflight)
{
    if (hbm_budget_bytes <= 0 || max_inflight <= 0) {
        fftz::report_msg("aoclfftz_gpu_server_create", "budget and max_inflight must be positive");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        fftz::report_msg("aoclfftz_gpu_server_create", "no HIP device available");
        return nullptr;
    }
    Server* s = new Server();
    s->budget = (size_t)hbm_budget_bytes;
    s->max_inflight = max_inflight;
    (void)hipGetDevice(&s->device);
    s->progress = std::thread(fftz::progress_loop, s);
    return s;
}

}  // extern "C"

namespace fftz {
namespace {

// Queue one request whose arrival time (the start of its latency) is t_arrive.
int64_t submit_at(Server* s, void* problem, int precision_bytes, double t_arrive)
{
    fftz::ProblemSpec spec;
    const void* in = nullptr;
    void* out = nullptr;
    if (!fftz::spec_from_desc(problem, precision_bytes, true, spec, &in, &out, "aoclfftz_gpu_server_submit"))
        return -1;
    if (!in || !out) {
        fftz::report_msg("aoclfftz_gpu_server_submit", "request buffers must be set (device memory)");
        return -1;
    }
    fftz::Plan* plan = nullptr;
    {
        std::lock_guard<std::mutex> g(s->plan_mu);
        const std::string key = fftz::spec_key(spec);
        auto it = s->plans.find(key);
        if (it != s->plans.end()) {
            plan = it->second;
        } else {
            std::string err;
            plan = fftz::plan_create(spec, err);
            if (!plan) {
                fftz::report_msg("aoclfftz_gpu_server_submit", err.c_str());
                return -1;
            }
            plan->use_graph = false;  // buffers change with every request
            s->plans[key] = plan;
        }
    }
    auto* r = new fftz::Req();
    r->plan = plan;
    r->in = in;
    r->out = out;
    // in place, the one buffer spans the larger of the two extents (the
    // complex side of an in-place R2C / C2R is the larger)
    const size_t ext = spec.inplace ? (size_t)std::max(plan->in_extent_bytes, plan->out_extent_bytes)
                                    : (size_t)plan->in_extent_bytes + (size_t)plan->out_extent_bytes;
    r->footprint = ext + plan->work_bytes;
    r->transforms = 1;
    for (const auto& v : spec.vecs) r->transforms *= v.n;
    if (r->footprint > s->budget) {
        fftz::report_msg("aoclfftz_gpu_server_submit", "request footprint exceeds the server's HBM budget");
        delete r;
        return -2;
    }
    std::lock_guard<std::mutex> g(s->mu);
    r->id = s->next_id++;
    r->t_submit = t_arrive;
    if (s->t_first < 0) s->t_first = r->t_submit;
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    s->queue_peak = std::max(s->queue_peak, s->queue.size());
    s->cv_work.notify_one();
    return r->id;
}

}  // namespace
}  // namespace fftz

extern "C" {

FFTZ_INT64 aoclfftz_gpu_server_submit(FFTZ_VOID* server, FFTZ_VOID* problem, FFTZ_INT32 precision_bytes)
{
    Server* s = fftz::as_server(server);
    if (!s) return -1;
    return fftz::submit_at(s, problem, precision_bytes, fftz::now_s());
}

FFTZ_INT64 aoclfftz_gpu_server_submit_timed(FFTZ_VOID* server, FFTZ_VOID** problems,
                                           const FFTZ_INT32* precision_bytes, const FFTZ_DOUBLE* arrival_s,
                                           FFTZ_INT64 count, FFTZ_INT64* ids, FFTZ_DOUBLE* max_late_s)
{
    Server* s = fftz::as_server(server);
    if (!s || count < 0 || (count > 0 && (!problems || !precision_bytes || !arrival_s || !ids))) return -1;
    for (int64_t i = 1; i < count; i++)
        if (arrival_s[i] < arrival_s[i - 1]) {
            fftz::report_msg("aoclfftz_gpu_server_submit_timed", "arrival times must not decrease");
            return -1;
        }
    const double t0 = fftz::now_s();
    double late = 0.0;
    for (int64_t i = 0; i < count; i++) {
        const double due = t0 + arrival_s[i];
        double t = fftz::now_s();
        while (t < due) t = fftz::now_s();  // open loop: spin to the arrival time
        late = std::max(late, t - due);
        ids[i] = fftz::submit_at(s, problems[i], precision_bytes[i], due);
    }
    if (max_late_s) *max_late_s = late;
    return count;
}

FFTZ_INT32 aoclfftz_gpu_server_test(FFTZ_VOID* server, FFTZ_INT64 id)
{
    Server* s = fftz::as_server(server);
    if (!s) return -1;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return -1;
    return it->second->state >= fftz::kDone ? 1 : 0;
}

FFTZ_INT32 aoclfftz_gpu_server_wait(FFTZ_VOID* server, FFTZ_INT64 id, FFTZ_DOUBLE* latency_s)
{
    Server* s = fftz::as_server(server);
    if (!s) return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return -1;
    fftz::Req* r = it->second;
    s->cv_done.wait(lk, [&] { return r->state >= fftz::kDone; });
    if (latency_s) *latency_s = r->t_done - r->t_submit;
    const int rc = r->state == fftz::kDone ? 0 : -1;
    if (r->ev) s->event_pool.push_back(r->ev);
    s->reqs.erase(it);
    delete r;
    return rc;
}

FFTZ_INT32 aoclfftz_gpu_server_stats(FFTZ_VOID* server, FFTZ_DOUBLE* out, FFTZ_INT32 cap)
{
    Server* s = fftz::as_server(server);
    if (!s || !out) return -1;
    std::lock_guard<std::mutex> g(s->mu);
    const double span = s->t_first >= 0 && s->t_last > s->t_first ? s->t_last - s->t_first : 0.0;
    size_t nplans = 0;
    {
        std::lock_guard<std::mutex> pg(s->plan_mu);
        nplans = s->plans.size();
    }
    const double v[] = {(double)s->completed,
                        (double)s->failed,
                        fftz::percentile(s->latencies, 0.50) * 1e3,
                        fftz::percentile(s->latencies, 0.99) * 1e3,
                        s->latencies.empty() ? 0.0 : *std::max_element(s->latencies.begin(), s->latencies.end()) * 1e3,
                        span > 0 ? (double)s->transforms_done / span : 0.0,
                        span > 0 ? s->flops_done / span / 1e9 : 0.0,
                        (double)s->peak,
                        (double)nplans,
                        (double)s->queue_peak,
                        span};
    const int n = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < cap; i++) out[i] = v[i];
    return n;
}

FFTZ_VOID aoclfftz_gpu_server_reset_stats(FFTZ_VOID* server)
{
    Server* s = fftz::as_server(server);
    if (!s) return;
    std::lock_guard<std::mutex> g(s->mu);
    s->latencies.clear();
    s->completed = s->failed = s->transforms_done = 0;
    s->flops_done = 0.0;
    s->peak = s->in_use;
    s->queue_peak = s->queue.size();
    s->t_first = -1.0;
    s->t_last = 0.0;
}

FFTZ_VOID aoclfftz_gpu_server_destroy(FFTZ_VOID* server)
{
    Server* s = fftz::as_server(server);
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    s->progress.join();  // drains queued and running requests first
    for (auto& kv : s->reqs) {  // completed but never waited for
        if (kv.second->ev) (void)hipEventDestroy(kv.second->ev);
        delete kv.second;
    }
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    for (auto& kv : s->plans) fftz::plan_destroy(kv.second);
    s->magic = 0;
    delete s;
}

}  // extern "C"
