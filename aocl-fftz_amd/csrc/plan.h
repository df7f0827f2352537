// GPU plan: the MI355X replacement for the reference's selector + solver tree
// (selector/selector.c:1387-1620 prepare_and_setup_dft; core/solvers/**).
//
// A plan is a flat list of pass launches (see pass.h) plus one device arena
// holding every twiddle / chirp table, built once at setup.  Execution binds
// the caller's buffers to the launches and enqueues them on a HIP stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.h"
#include "pass.h"

namespace fftz {

enum class Role : int { In = 0, Out = 1, Work = 2 };

struct AxisSpec {
    int64_t n, is, os;  // length, input stride, output stride (elements)
};

enum class FftType : int { C2C = 0, R2C = 1, C2R = 2 };

// Plan choices the autotuner varies; 0 / empty keeps the planner's default.
struct TuneParams {
    int row_t = 0;      // threads per row tile (lines per tile = row_t / threads per line)
    int col_c = 0;      // most lines per column tile
    int wg_small = 0;   // uneven batch: workgroups per CU of the small class
    int wg_medium = 0;  // uneven batch: workgroups per CU of the medium class
    int elems = 0;      // complex values per thread per stage (8 or 16)
    // radix order per pass length: (L, radices in stage order)
    std::vector<std::pair<int, std::vector<int>>> radix_for;
    std::vector<int64_t> split;  // four-step factors of the long axis (2 or 3)
    bool empty() const
    {
        return !row_t && !col_c && !wg_small && !wg_medium && !elems && radix_for.empty() && split.empty();
    }
    const std::vector<int>* radices(int L) const
    {
        for (const auto& r : radix_for)
            if (r.first == L) return &r.second;
        return nullptr;
    }
};

struct ProblemSpec {
    int prec_bytes = 4;  // 4: float, 8: double
    FftType type = FftType::C2C;
    int dir = 0;         // 0 forward, 1 backward
    bool inplace = false;
    std::vector<AxisSpec> dims;  // dims[0] innermost
    std::vector<AxisSpec> vecs;  // batch axes
    int mfma_mode = 0;           // MfmaMode for complex64 power-of-16 passes
    TuneParams tune;             // empty: the tuned cache, else the defaults
};

// One uneven-batch entry: `batch` transforms of length n, `dist` elements
// apart, starting at element offsets in_off / out_off of the caller buffers.
struct BatchItem {
    int64_t n, batch, in_off, out_off, dist;
};

// A persistent-kernel class of the uneven batch (small or large tiles).
struct BatchClass {
    int E = 8, maxt = 256;
    uint32_t lds_bytes = 0;
    uint32_t grid = 0;
    int ntiles = 0;
    int64_t tiles_off = -1;  // BatchTile array in the plan arena
};

struct Launch {
    KernelKind kind = KernelKind::Pass;
    bool all_radix = false;
    int E = 16;        // complex values per thread per stage
    int maxt = 256;    // block-size bound of the kernel variant
    bool fixed = false;  // fixed-length kernel (kernels_fixed.hip)
    int mfma = 0;        // MfmaMode of a matrix-core pass (kernels_mfma.hip)
    int mfma_nb = 4;     // its column blocks per wavefront
    PassDesc d{};
    Role in_role = Role::In, out_role = Role::Out;
    int64_t in_off_bytes = 0, out_off_bytes = 0;
    uint32_t grid = 0, block = 0, lds_bytes = 0;
    // table offsets in the plan arena (-1: none), resolved after upload
    int64_t stw_off = -1, tw4_off = -1, chirp_off = -1, spec_off = -1, rtw_off = -1;
    int64_t frag_off = -1;  // MFMA DFT-16 operand fragments (64 lanes x 32 floats)
    const float* frag = nullptr;
};

// Per-call execution context: a stream, the work buffer and host staging
// buffers.  One slot per concurrent execute_io caller, so calls on a shared
// handle never share scratch (the GPU analogue of the reference's per-call
// scratch slab, core/common/memory_manager.h:53-175).
struct Slot {
    hipStream_t stream = nullptr;
    void* work = nullptr;
    size_t work_bytes = 0;
    void* stage_in = nullptr;
    void* stage_out = nullptr;
    size_t stage_in_bytes = 0, stage_out_bytes = 0;
    bool busy = false;
    unsigned* counters = nullptr;  // uneven-batch work-queue counters
    // graph cache for the last (in, out) pair run on this slot
    hipGraphExec_t graph = nullptr;
    const void* graph_in = nullptr;
    void* graph_out = nullptr;
};

struct Plan {
    ProblemSpec spec;
    int device = 0;
    bool noop = false;
    bool dry_run = false;  // launch list only: no device memory, zero-valued device-built tables
    std::vector<Launch> launches;
    size_t work_bytes = 0;
    int64_t in_extent_bytes = 0, out_extent_bytes = 0;
    std::vector<unsigned char> host_tables;
    void* tables = nullptr;
    void* setup_in = nullptr;
    void* setup_out = nullptr;
    std::mutex mu;
    std::vector<Slot*> slots;
    hipStream_t user_stream = nullptr;  // set by aoclfftz_gpu_set_stream
    bool use_graph = false;
    // per-launch timing: when enabled, an event is recorded before the first
    // launch and after each launch of every execute (graphs are bypassed)
    bool profile = false;
    std::vector<hipEvent_t> prof_events;  // recorded, not yet read
    std::vector<hipEvent_t> prof_pool;    // free events
    std::string description;            // human-readable plan summary
    // uneven batch (aoclfftz_gpu_batch_setup): size-class descriptors and
    // per-class tile queues in the arena
    bool is_batch = false;
    int64_t descs_off = -1;
    std::vector<BatchClass> bclasses;
    int64_t n_transforms = 0;
    std::string tune_key;               // uneven batch: its key in the tuned cache
    double flops = 0.0;                 // 5 N log2 N per transform x batch (C2C)
};

// Build a plan.  Returns nullptr and fills err on an unsupported problem.
// dry_run: build the launch list only (no device memory, no HIP calls), for
// host-side inspection of the plan geometry.
Plan* plan_create(const ProblemSpec& spec, std::string& err, bool dry_run = false);

// Uneven batch of 1-D C2C transforms in one persistent launch per tile class.
Plan* batch_plan_create(const std::vector<BatchItem>& items, int prec_bytes, int dir, std::string& err,
                        bool dry_run = false, const TuneParams* tune = nullptr);
// Re-derive the uneven-batch grids for per-CU workgroup counts (0: default).
void batch_set_grids(Plan* p, int wg_small, int wg_medium);

// Process-wide cache of tuned tile shapes, keyed by problem (precision, type,
// direction, placement, compute mode, dims, vecs) or by uneven-batch item list.
// plan_create / batch_plan_create consult it when the spec carries no tuning.
std::string problem_key(const ProblemSpec& spec);
std::string batch_key(const std::vector<BatchItem>& items, int prec_bytes, int dir);
bool tune_lookup(const std::string& key, TuneParams& out);
void tune_store(const std::string& key, const TuneParams& t);
int tune_export(const char* path);  // entries written, -1 on error
int tune_import(const char* path);  // entries read, -1 on error or a malformed line
std::string tune_text(int* entries);           // the cache as text, one problem per line
int tune_parse(const std::string& text);       // entries stored, -1 (nothing stored) if malformed
void tune_clear();

// Autotuner search spaces: radix orders of a pass length (the planner's
// default first, then other multisets of supported radices with at most one
// more stage, in descending and ascending order), and ordered 2- / 3-factor
// four-step splits of a long axis, most balanced first; at most maxn each.
std::vector<std::vector<int>> radix_orders(int L, int maxn);
std::vector<std::vector<int64_t>> split_candidates(int64_t N, int prec_bytes, int maxn);

// Kernel time of `reps` executes of p on (in, out) (device buffers) on a slot
// stream, in ms; negative on a HIP error.
double plan_time(Plan* p, const void* in, void* out, int reps);

// JSON description of every launch of a plan (geometry, strides, flags).
std::string plan_json(const Plan* p);
void plan_destroy(Plan* p);

// Enqueue the plan on the slot's stream (or the user stream) for device
// buffers.  Returns hipSuccess or the first HIP error.
hipError_t plan_enqueue(Plan* p, Slot* s, const void* in, void* out, hipStream_t stream);

// Enqueue the plan's kernels on `stream` with caller-provided work memory
// (four-step intermediates, plan->work_bytes) and queue counters (uneven
// batch, one per tile class): no slot, graph, event or allocation.
hipError_t plan_enqueue_scratch(Plan* p, void* work, unsigned* counters, const void* in, void* out,
                                hipStream_t stream);

// Read back per-launch durations (ms) recorded since the last read, in
// launch order (execute-major).  Returns the number of durations.
int plan_profile_read(Plan* p, float* ms, int cap);

// Acquire / release an execution slot (creating one if all are busy).
Slot* plan_acquire_slot(Plan* p, hipError_t* err);
void plan_release_slot(Plan* p, Slot* s);

// Device properties cached at first use.
struct DeviceInfo {
    int cus = 256;                  // MI355X defaults, used when no device
    size_t lds_per_block = 163840;  // is visible (dry-run planning)
    int warp = 64;
};
const DeviceInfo& device_info();

}  // namespace fftz
