// Fixed-length pass kernels: power-of-two L in [64, 4096] with the radix plan,
// threads per line (TL = L / 16) and LDS pitch known at compile time.
//
// Differences from the generic pass (kernels_body.h):
//   * stage 1 consumes the registers loaded from HBM: the load pattern
//     e = t + k*TL is exactly the first radix-16 stage's operand set
//     x[i*(L/16) + j], so the tile is not staged through LDS first;
//   * the last stage (ls = L/R, one group) writes y[i*ls + j] straight to
//     HBM, again in the store pattern e = t + k*TL;
//   * no per-element guards (TL * 16 == L exactly) and all index arithmetic
//     folds at compile time; only the line base offsets are runtime;
//   * each phase picks its own thread mapping: row (wavefront along a line)
//     or column (wavefront across adjacent lines), so column passes of the
//     four-step plan get the same register-direct first / last stage.
#include <hip/hip_runtime.h>

#include "codelets.h"
#include "kernels.h"
#include "pass.h"

namespace fftz {

namespace {

__device__ __forceinline__ uint32_t fdiv_u(uint32_t n, const FastDiv& f)
{
    return (__umulhi(n, f.mul) + n) >> f.shift;
}

struct LineOff {
    int64_t ioff, ooff;
    uint32_t l0;
};

__device__ __forceinline__ LineOff line_off(const PassDesc& p, uint32_t l)
{
    uint32_t r = fdiv_u(l, p.ln0);
    uint32_t l0 = l - r * p.ln0.d;
    uint32_t l2 = fdiv_u(r, p.ln1);
    uint32_t l1 = r - l2 * p.ln1.d;
    LineOff o;
    o.ioff = (int64_t)l0 * p.ils[0] + (int64_t)l1 * p.ils[1] + (int64_t)l2 * p.ils[2];
    o.ooff = (int64_t)l0 * p.ols[0] + (int64_t)l1 * p.ols[1] + (int64_t)l2 * p.ols[2];
    o.l0 = l0;
    return o;
}

// Line of tile slot c and whether it exists (flat tiles: l_base + c; ragged
// column tiles: PassDesc::vrow).
__device__ __forceinline__ bool slot_line_f(const PassDesc& p, int64_t l_base, int c, uint32_t& l)
{
    const int64_t v = l_base + c;
    if (p.vrow.d == 0) {
        l = (uint32_t)v;
        return v < p.nlines;
    }
    const uint32_t row = fdiv_u((uint32_t)v, p.vrow);
    const uint32_t cc = (uint32_t)v - row * p.vrow.d;
    l = row * p.ln0.d + cc;
    return cc < p.ln0.d && (int64_t)l < p.nlines;
}

// Radix plan of a power-of-two L: 16s first, remainder last.
template <int L>
struct Plan2 {
    static constexpr int lg = __builtin_ctz(L);
    static constexpr int S = (lg + 3) / 4;
    static constexpr int rem = lg - 4 * (S - 1);  // exponent of the last stage
    static constexpr int radix(int s) { return s < S - 1 ? 16 : (1 << rem); }
    static constexpr int ls(int s) { return s == 0 ? 1 : (1 << (4 * s)); }
};

template <int N>
__device__ __forceinline__ constexpr int pad_c(int e) { return e + (e >> 4); }

// Workgroup barrier that waits only for this wave's LDS traffic (no global
// memory fence: the stages exchange data through LDS only).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Stage exchange of a pass whose lines each live in one wavefront (row
// mapping on both sides, TL <= 64 threads per line): the wave's own LDS
// operations complete in order once counted down, so waiting for them is
// enough and the wavefronts of a workgroup never wait for each other.
__device__ __forceinline__ void wave_lds_sync()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}

template <int TL>
__device__ __forceinline__ void stage_sync(bool wave_local)
{
    if constexpr (TL <= 64) {
        if (wave_local) {
            wave_lds_sync();
            return;
        }
    }
    lds_barrier();
}

template <typename R, int RAD>
__device__ __forceinline__ void twiddle_legs(cplx<R>* v, const cplx<R>* w)
{
#pragma unroll
    for (int i = 1; i < RAD; i++) v[i] = cmul(v[i], w[i - 1]);
}

}  // namespace

// One LDS stage s (not the first, not the last): read K butterflies, barrier,
// twiddle + DFT, write, barrier.  (c, t) is this stage's thread mapping.
template <typename R, int L, int s>
__device__ __forceinline__ void fixed_mid_stage(const PassDesc& p, cplx<R>* lds, int c, int t, bool wave_local)
{
    using P = Plan2<L>;
    constexpr int RAD = P::radix(s);
    constexpr int ls = P::ls(s);
    constexpr int K = 16 / RAD;
    constexpr int TL = L / 16;
    constexpr int nbl = L / RAD;
    constexpr int LP = ((L - 1) + ((L - 1) >> 4) + 1) | 1;
    cplx<R>* line = lds + c * LP;
    cplx<R> v[K][RAD];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * TL;
#pragma unroll
        for (int i = 0; i < RAD; i++) v[k][i] = line[pad_c<L>(i * nbl + j)];
    }
    stage_sync<TL>(wave_local);
    const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(p.stw) + p.twoff[s];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * TL;
        const int g = j / ls;
        const int cc = j % ls;
        twiddle_legs<R, RAD>(v[k], tw + cc * (RAD - 1));
        Dft<RAD, R>::run(v[k]);
        const int base = g * RAD * ls + cc;
#pragma unroll
        for (int i = 0; i < RAD; i++) line[pad_c<L>(base + i * ls)] = v[k][i];
    }
    stage_sync<TL>(wave_local);
}

template <typename R, int L, int s, int S>
struct MidStages {
    static __device__ __forceinline__ void run(const PassDesc& p, cplx<R>* lds, int c, int t, bool wave_local)
    {
        if constexpr (s < S - 1) {
            fixed_mid_stage<R, L, s>(p, lds, c, t, wave_local);
            MidStages<R, L, s + 1, S>::run(p, lds, c, t, wave_local);
        }
    }
};

template <typename R, int L>
__device__ __forceinline__ void fixed_load(const PassDesc& p, int64_t l_base, int c, int t, cplx<R>* v)
{
    constexpr int TL = L / 16;
    uint32_t l;
    if (slot_line_f(p, l_base, c, l)) {
        LineOff lo = line_off(p, l);
        const cplx<R>* src = reinterpret_cast<const cplx<R>*>(p.in) + lo.ioff + (int64_t)t * p.ies;
        const int64_t step = (int64_t)TL * p.ies;
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = src[k * step];
    } else {
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = {R(0), R(0)};
    }
}

// One workgroup per tile of C lines (grid = number of tiles), at most MAXT
// threads (256, or 512 for wide column tiles).  A register prefetch of the
// next tile was measured slower: the extra 16 live values per lane cut
// occupancy from 5 to 2 waves/SIMD.
template <typename R, int L, int MAXT>
__global__ void __launch_bounds__(MAXT) k_fixed_pass(const PassDesc p)
{
    using P = Plan2<L>;
    constexpr int S = P::S;
    constexpr int TL = L / 16;
    constexpr int LP = ((L - 1) + ((L - 1) >> 4) + 1) | 1;
    constexpr int R0 = P::radix(0);
    constexpr int RL = P::radix(S - 1);
    static_assert(R0 == 16, "first stage is radix 16");
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    uint32_t tile = blockIdx.x;
    if (p.flags & kXcdRemap) {
        const uint32_t G = gridDim.x, per = G >> 3;
        if (tile < (per << 3)) tile = (tile & 7u) * per + (tile >> 3);
    }
    const int64_t l_base = (int64_t)tile * p.C;

    // load (input-side mapping) + stage 1 in registers
    const bool in_col = (p.flags & kInCol) != 0;
    const int ci = in_col ? (tid & (p.C - 1)) : (tid / TL);
    const int ti = in_col ? (tid >> p.c_shift) : (tid % TL);
    cplx<R> v[16];
    fixed_load<R, L>(p, l_base, ci, ti, v);
    if (p.flags & kConjIn) {
#pragma unroll
        for (int k = 0; k < 16; k++) v[k].y = -v[k].y;
    }
    // butterfly j = ti reads x[i*TL + ti]; writes y[ti*16 + i]
    Dft<16, R>::run(v);
    {
        cplx<R>* line = lds + ci * LP;
#pragma unroll
        for (int i = 0; i < 16; i++) line[pad_c<L>(ti * 16 + i)] = v[i];
    }
    // lines held by one wavefront each (row mapping throughout, TL <= 64)
    // exchange through LDS without workgroup barriers
    const bool wave_local = !(p.flags & (kInCol | kOutCol));
    stage_sync<TL>(wave_local);

    // middle stages (row mapping)
    MidStages<R, L, 1, S>::run(p, lds, tid / TL, tid % TL, wave_local);

    // last stage (output-side mapping) + store
    const bool out_col = (p.flags & kOutCol) != 0;
    const int co = out_col ? (tid & (p.C - 1)) : (tid / TL);
    const int to = out_col ? (tid >> p.c_shift) : (tid % TL);
    constexpr int KL = 16 / RL;
    constexpr int lsL = L / RL;
    cplx<R> w[KL][RL];
    {
        const cplx<R>* line = lds + co * LP;
#pragma unroll
        for (int k = 0; k < KL; k++) {
            const int j = to + k * TL;
#pragma unroll
            for (int i = 0; i < RL; i++) w[k][i] = line[pad_c<L>(i * lsL + j)];
        }
    }
    uint32_t l_out;
    if (!slot_line_f(p, l_base, co, l_out)) return;
    LineOff lo = line_off(p, l_out);
    const cplx<R>* twl = reinterpret_cast<const cplx<R>*>(p.stw) + p.twoff[S - 1];
    cplx<R>* dst = reinterpret_cast<cplx<R>*>(p.out) + lo.ooff;
    const cplx<double>* t4 = reinterpret_cast<const cplx<double>*>(p.tw4);
    const int64_t lo_n = (int64_t)1 << p.tw4_lobits;
    auto root = [&](int64_t q) -> cplx<double> {
        return cmul(t4[lo_n + (q >> p.tw4_lobits)], t4[q & (lo_n - 1)]);
    };
    const bool tw4 = (p.flags & kTw4) != 0;
    // four-step twiddle w_N^(l0 e) for e = i*lsL + j: one table product for
    // w^(l0 j) and one for the step w^(l0 lsL), then an FP64 recurrence over
    // i (at most RL-1 = 15 products, ~2e-15 relative)
    cplx<double> wstep = {1.0, 0.0};
    if (tw4) wstep = root((int64_t)lo.l0 * lsL);
#pragma unroll
    for (int k = 0; k < KL; k++) {
        const int j = to + k * TL;  // cc = j, g = 0
        twiddle_legs<R, RL>(w[k], twl + j * (RL - 1));
        Dft<RL, R>::run(w[k]);
        cplx<double> wc = {1.0, 0.0};
        if (tw4) wc = root((int64_t)lo.l0 * j);
#pragma unroll
        for (int i = 0; i < RL; i++) {
            const int e = i * lsL + j;
            cplx<R> y = w[k][i];
            if (tw4) {
                cplx<double> yd = {double(y.x), double(y.y)};
                yd = cmul(yd, wc);
                y = {R(yd.x), R(yd.y)};
                wc = cmul(wc, wstep);
            }
            if (p.flags & kScale) y = cscale(y, R(sizeof(R) == 4 ? p.scale_f : p.scale_d));
            if (p.flags & kConjOut) y.y = -y.y;
            dst[(int64_t)e * p.oes] = y;
        }
    }
}

template <typename R, int MAXT>
static const void* fixed_sym(int L)
{
    switch (L) {
        case 64: return MAXT == 256 ? (const void*)&k_fixed_pass<R, 64, 256> : nullptr;
        case 128: return MAXT == 256 ? (const void*)&k_fixed_pass<R, 128, 256> : nullptr;
        case 256: return (const void*)&k_fixed_pass<R, 256, MAXT>;
        case 512: return (const void*)&k_fixed_pass<R, 512, MAXT>;
        case 1024: return (const void*)&k_fixed_pass<R, 1024, MAXT>;
        case 2048: return (const void*)&k_fixed_pass<R, 2048, MAXT>;
        case 4096: return (const void*)&k_fixed_pass<R, 4096, MAXT>;
        default: return nullptr;
    }
}

const void* kernel_symbol_fixed(int prec_bytes, int L, int maxt)
{
    if (maxt != 256 && maxt != 512) return nullptr;
    if (prec_bytes == 4) return maxt == 256 ? fixed_sym<float, 256>(L) : fixed_sym<float, 512>(L);
    return maxt == 256 ? fixed_sym<double, 256>(L) : fixed_sym<double, 512>(L);
}

}  // namespace fftz
