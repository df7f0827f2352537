// LDS-resident FFT pass kernels for CDNA4 (gfx950).
//
// One workgroup owns a tile of C lines of length L.  The tile is loaded from
// HBM once (coalesced along the line for row passes, across adjacent lines for
// column passes), every Stockham stage of the length-L transform runs out of
// LDS through VGPRs, and the tile is written back once -- one HBM round trip
// per pass.  Element-wise work that the reference runs as separate sweeps
// (ele_mul, fused_twiddle_transpose, recombine/separate) is folded into the
// load or store of the adjacent pass.
//
// Stage s of radix R (Stockham autosort, the geometry of the reference's
// pow2_setup_stage_layout, core/solvers/complex/pow2_iterative_solver_dft.c:27-78):
//   butterfly j = g*ls + c reads  x[i*(L/R) + j],             i < R
//   multiplies leg i by w_{R*ls}^(i*c),
//   writes   y[g*R*ls + i*ls + c].
// Reads and writes happen in place with a workgroup barrier in between; each
// thread holds its K = ceil(E/R) butterflies in VGPRs across the barrier.
// Included by the instantiation units kernels_pass_{f32,f64}_{mixed,pow2}.hip
// and kernels_batch_{f32,f64}_{small,medium,large}.hip (compiled in
// parallel); kernels.hip holds the launch-table dispatch.
#pragma once

#include <hip/hip_runtime.h>

#include "codelets.h"
#include "kernels.h"
#include "pass.h"

namespace fftz {

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f)
{
    return (__umulhi(n, f.mul) + n) >> f.shift;
}

struct LineIdx {
    int64_t ioff, ooff;
    uint32_t l0;
    int64_t gin, gout;  // sequence index of the line's element 0 (kChirpFlags)
};

// The buffers and line set a pass body works on: the descriptor's own for a
// single-size pass, the queue entry's for the uneven-batch kernel (which reads
// its size class's descriptor in place and never patches it).
struct TileIo {
    const void* in;
    void* out;
    int64_t nlines;
    int64_t ils0, ols0;  // line strides of level 0
};

__device__ __forceinline__ TileIo io_of(const PassDesc& p)
{
    return TileIo{p.in, p.out, p.nlines, p.ils[0], p.ols[0]};
}

__device__ __forceinline__ LineIdx line_index(const PassDesc& p, const TileIo& io, uint32_t l)
{
    uint32_t r = fdiv(l, p.ln0);
    uint32_t l0 = l - r * p.ln0.d;
    uint32_t l2 = fdiv(r, p.ln1);
    uint32_t l1 = r - l2 * p.ln1.d;
    LineIdx li;
    li.ioff = (int64_t)l0 * io.ils0 + (int64_t)l1 * p.ils[1] + (int64_t)l2 * p.ils[2];
    li.ooff = (int64_t)l0 * io.ols0 + (int64_t)l1 * p.ols[1] + (int64_t)l2 * p.ols[2];
    li.l0 = l0;
    li.gin = li.gout = 0;
    if (p.flags & kChirpFlags) {
        li.gin = p.gin_base + (int64_t)l0 * p.gin[0] + (int64_t)l1 * p.gin[1] + (int64_t)l2 * p.gin[2];
        li.gout = p.gout_base + (int64_t)l0 * p.gout[0] + (int64_t)l1 * p.gout[1] + (int64_t)l2 * p.gout[2];
    }
    return li;
}

// Line held by tile slot c (l_base = tile index x C): line l_base + c for
// flat tiles; for ragged column tiles the slot's row and position in the row
// (PassDesc::vrow), live only below the row length.
struct SlotLine {
    uint32_t l;
    bool live;
};

__device__ __forceinline__ SlotLine slot_line(const PassDesc& p, const TileIo& io, int64_t l_base, int c)
{
    const int64_t v = l_base + c;
    if (p.vrow.d == 0) return SlotLine{(uint32_t)v, v < io.nlines};
    const uint32_t row = fdiv((uint32_t)v, p.vrow);
    const uint32_t cc = (uint32_t)v - row * p.vrow.d;
    const uint32_t l = row * p.ln0.d + cc;
    return SlotLine{l, cc < p.ln0.d && (int64_t)l < io.nlines};
}

template <typename R>
__device__ __forceinline__ cplx<R> to_prec(double x, double y) { return {R(x), R(y)}; }

// w_N^p from the two-level FP64 table: hi[p >> b] * lo[p & (2^b - 1)].
__device__ __forceinline__ cplx<double> tw4_root(const PassDesc& p, int64_t q)
{
    const cplx<double>* t = reinterpret_cast<const cplx<double>*>(p.tw4);
    const int64_t lo_n = (int64_t)1 << p.tw4_lobits;
    cplx<double> lo = t[q & (lo_n - 1)];
    cplx<double> hi = t[lo_n + (q >> p.tw4_lobits)];
    return cmul(hi, lo);
}

// ---------------------------------------------------------------------------
// Tile load / store.
// Row passes: thread -> (line c = tid / TL, lane t = tid % TL): a wavefront
// walks consecutive elements of one line.  Column passes: (c = tid % C,
// t = tid / C): a wavefront walks the same element of C adjacent lines.
// ---------------------------------------------------------------------------
template <typename R, int E>
__device__ __forceinline__ void load_tile(const PassDesc& p, const TileIo& io, cplx<R>* lds, int64_t l_base, int tid)
{
    const bool col = (p.flags & kInCol) != 0;
    const int c = col ? (tid & (p.C - 1)) : (tid >> p.tl_shift);
    const int t = col ? (tid >> p.c_shift) : (tid & (p.TL - 1));
    const SlotLine sl = slot_line(p, io, l_base, c);
    if (!sl.live) return;
    LineIdx li = line_index(p, io, sl.l);
    cplx<R>* line = lds + c * p.LP;
    if (p.flags & kRealIn) {
        // R2C: element j of the L-point complex line is (x[2j], x[2j+1])
        const R* src = reinterpret_cast<const R*>(io.in) + li.ioff;
        cplx<R> v[E];
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int e = t + k * p.TL;
            if (e < p.L) {
                v[k].x = src[(int64_t)(2 * e) * p.ies];
                v[k].y = src[(int64_t)(2 * e + 1) * p.ies];
            }
        }
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int e = t + k * p.TL;
            if (e < p.L) line[lds_pad(e)] = v[k];
        }
        return;
    }
    if (p.flags & (kRealPromote | kHermIn)) {
        // full-length real transforms: real x[e] as (x[e], 0), or the Hermitian
        // extension of bins 0..n/2 (n = nvalid_in)
        const int n = p.nvalid_in;
        cplx<R> v[E];
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int e = t + k * p.TL;
            v[k] = {R(0), R(0)};
            if (e < n) {
                if (p.flags & kRealPromote) {
                    const R* srcr = reinterpret_cast<const R*>(io.in) + li.ioff;
                    v[k].x = srcr[(int64_t)e * p.ies];
                } else {
                    const cplx<R>* srcc = reinterpret_cast<const cplx<R>*>(io.in) + li.ioff;
                    const int m = 2 * e <= n ? e : n - e;
                    v[k] = srcc[(int64_t)m * p.ies];
                    if (2 * e > n) v[k].y = -v[k].y;
                }
                if (p.flags & kConjIn) v[k].y = -v[k].y;
            }
        }
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int e = t + k * p.TL;
            if (e < p.L) line[lds_pad(e)] = v[k];
        }
        return;
    }
    const cplx<R>* src = reinterpret_cast<const cplx<R>*>(io.in) + li.ioff;
    const int nv = p.nvalid_in;
    const int ncover = (p.flags & kC2RPre) ? p.L + 1 : p.L;
    cplx<R> v[E + 1];
#pragma unroll
    for (int k = 0; k < E + 1; k++) {
        const int e = t + k * p.TL;
        v[k] = {R(0), R(0)};
        if (e < nv) v[k] = src[(int64_t)e * p.ies];
    }
    if (p.flags & kConjIn) {
#pragma unroll
        for (int k = 0; k < E + 1; k++) v[k].y = -v[k].y;
    }
#pragma unroll
    for (int k = 0; k < E + 1; k++) {
        const int e = t + k * p.TL;
        if (e < ncover) line[lds_pad(e)] = v[k];
    }
}

// Real-packed separate (C2R pre-processing), in LDS, on the L+1 loaded bins:
//   Z_k = E_k + i O_k,  E_k = X_k + conj(X_{L-k}),
//   O_k = (X_k - conj(X_{L-k})) exp(+2 pi i k / 2L);
// stored conjugated so the forward stages compute the inverse transform.
template <typename R>
__device__ __forceinline__ void c2r_separate(const PassDesc& p, cplx<R>* lds, int tid)
{
    const int c = tid >> p.tl_shift;
    const int t = tid & (p.TL - 1);
    if (c >= p.C) return;
    cplx<R>* line = lds + c * p.LP;
    const cplx<R>* rtw = reinterpret_cast<const cplx<R>*>(p.rtw);
    const int L = p.L;
    for (int k = t; k <= L / 2; k += p.TL) {
        cplx<R> xk = line[lds_pad(k)];
        cplx<R> xm = line[lds_pad(L - k)];
        cplx<R> w = rtw[k];  // exp(-2 pi i k / 2L)
        // k side
        cplx<R> ek = {xk.x + xm.x, xk.y - xm.y};
        cplx<R> dk = {xk.x - xm.x, xk.y + xm.y};
        cplx<R> ok = cmul(dk, cconj(w));
        cplx<R> zk = {ek.x - ok.y, ek.y + ok.x};
        // L-k side: E = conj(ek), O = (X_{L-k} - conj(X_k)) * (-w)
        cplx<R> em = {ek.x, -ek.y};
        cplx<R> dm = {xm.x - xk.x, xm.y + xk.y};
        cplx<R> om = cmul(dm, cplx<R>{-w.x, -w.y});
        cplx<R> zm = {em.x - om.y, em.y + om.x};
        if (k == 0) {
            line[lds_pad(0)] = cconj(zk);
        } else {
            line[lds_pad(k)] = cconj(zk);
            if (k != L - k) line[lds_pad(L - k)] = cconj(zm);
        }
    }
}

// Legs b0 + i * stride (i < RAD) of a line in the padded LDS layout.  When
// the stride is a multiple of 16 (LIN) the pad term carries nothing between
// legs: lds_pad(b0 + i stride) = lds_pad(b0) + i (stride + stride / 16), so
// each leg's slot is one add of a uniform step from the first.  Callers pick
// LIN with a uniform branch so the two address forms are separate code
// (complex64 only: the complex128 1024-thread variants, at their 128-VGPR
// bound, spill with both forms present).
template <typename R, int RAD, bool LIN>
__device__ __forceinline__ void lds_gather(const cplx<R>* line, int b0, int stride, cplx<R> (&v)[RAD])
{
    if constexpr (LIN) {
        const int step = stride + (stride >> 4);
        const cplx<R>* q = line + lds_pad(b0);
#pragma unroll
        for (int i = 0; i < RAD; i++) v[i] = q[i * step];
    } else {
#pragma unroll
        for (int i = 0; i < RAD; i++) v[i] = line[lds_pad(b0 + i * stride)];
    }
}

template <typename R, int RAD, bool LIN>
__device__ __forceinline__ void lds_scatter(cplx<R>* line, int b0, int stride, const cplx<R> (&v)[RAD])
{
    if constexpr (LIN) {
        const int step = stride + (stride >> 4);
        cplx<R>* q = line + lds_pad(b0);
#pragma unroll
        for (int i = 0; i < RAD; i++) q[i * step] = v[i];
    } else {
#pragma unroll
        for (int i = 0; i < RAD; i++) line[lds_pad(b0 + i * stride)] = v[i];
    }
}

// Stage twiddles held in registers across a stage's barrier: only while a
// thread's share stays within 32 VGPRs (K (R - 1) <= 16 complex64 or 8
// complex128 values), else they are fetched after it.
template <typename R, int E, int RAD>
constexpr bool kTwEarly = ((E + RAD - 1) / RAD) * (RAD - 1) * (int)sizeof(R) <= 64;

// One in-place Stockham stage out of LDS (LIN: both the read stride L / R and
// the write stride ls are multiples of 16).
template <typename R, int E, int RAD, bool LIN>
__device__ __forceinline__ void stage_impl(const PassDesc& p, cplx<R>* lds, int tid, int s)
{
    constexpr int K = (E + RAD - 1) / RAD;
    constexpr bool EARLY = kTwEarly<R, E, RAD>;
    const int c = tid >> p.tl_shift;
    const int t = tid & (p.TL - 1);
    const int nbl = p.L / RAD;
    const int ls = p.ls[s];
    cplx<R>* line = lds + c * p.LP;
    const bool act = c < p.C;  // batch kernel: threads beyond C * TL idle
    const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(p.stw) + p.twoff[s];
    cplx<R> v[K][RAD];
    cplx<R> w[EARLY ? K : 1][RAD - 1];
    if constexpr (EARLY) {
        // fetched before the LDS reads: the latency overlaps them and the barrier
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = t + k * p.TL;
            const int g = (int)fdiv((uint32_t)j, p.lsdiv[s]);
            const int cc = j - g * ls;
            if (act && j < nbl && ls > 1) {
#pragma unroll
                for (int i = 1; i < RAD; i++) w[k][i - 1] = tw[cc * (RAD - 1) + i - 1];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (act && j < nbl) lds_gather<R, RAD, LIN>(line, j, nbl, v[k]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (act && j < nbl) {
            const int g = (int)fdiv((uint32_t)j, p.lsdiv[s]);
            const int cc = j - g * ls;
            if (ls > 1) {
#pragma unroll
                for (int i = 1; i < RAD; i++) {
                    if constexpr (EARLY)
                        v[k][i] = cmul(v[k][i], w[k][i - 1]);
                    else
                        v[k][i] = cmul(v[k][i], tw[cc * (RAD - 1) + i - 1]);
                }
            }
            Dft<RAD, R>::run(v[k]);
            lds_scatter<R, RAD, LIN>(line, g * RAD * ls + cc, ls, v[k]);
        }
    }
    __syncthreads();
}

template <typename R, int E, int RAD>
__device__ __forceinline__ void stage(const PassDesc& p, cplx<R>* lds, int tid, int s)
{
    if (sizeof(R) == 4 && ((p.L / RAD) & 15) == 0 && (p.ls[s] & 15) == 0)
        stage_impl<R, E, RAD, true>(p, lds, tid, s);
    else
        stage_impl<R, E, RAD, false>(p, lds, tid, s);
}

template <typename R, bool ALLRAD, int E>
__device__ __forceinline__ void run_stages(const PassDesc& p, cplx<R>* lds, int tid, int s0, int s1)
{
    for (int s = s0; s < s1; s++) {
        switch (p.radix[s]) {
            case 2: stage<R, E, 2>(p, lds, tid, s); break;
            case 4: stage<R, E, 4>(p, lds, tid, s); break;
            case 8: stage<R, E, 8>(p, lds, tid, s); break;
            case 16: stage<R, E, 16>(p, lds, tid, s); break;
            default:
                if constexpr (ALLRAD) {
                    switch (p.radix[s]) {
                        case 3: stage<R, E, 3>(p, lds, tid, s); break;
                        case 5: stage<R, E, 5>(p, lds, tid, s); break;
                        case 6: stage<R, E, 6>(p, lds, tid, s); break;
                        case 7: stage<R, E, 7>(p, lds, tid, s); break;
                        case 9: stage<R, E, 9>(p, lds, tid, s); break;
                        case 10: stage<R, E, 10>(p, lds, tid, s); break;
                        case 11: stage<R, E, 11>(p, lds, tid, s); break;
                        case 12: stage<R, E, 12>(p, lds, tid, s); break;
                        case 13: stage<R, E, 13>(p, lds, tid, s); break;
                        default: break;
                    }
                }
                break;
        }
    }
}

// Real-packed recombine (R2C post-processing) straight from LDS to HBM:
// bins k and L-k of the N = 2L real transform from Z_k, Z_{L-k}
// (core/kernels/non_dft/real_packed/packed_rdft_common.h:19-60).
template <typename R>
__device__ __forceinline__ void r2c_recombine_store(const PassDesc& p, const TileIo& io, const cplx<R>* lds, int64_t l_base, int tid)
{
    const int c = tid >> p.tl_shift;
    const int t = tid & (p.TL - 1);
    const SlotLine sl = slot_line(p, io, l_base, c);
    if (!sl.live) return;
    LineIdx li = line_index(p, io, sl.l);
    const cplx<R>* line = lds + c * p.LP;
    cplx<R>* dst = reinterpret_cast<cplx<R>*>(io.out) + li.ooff;
    const cplx<R>* rtw = reinterpret_cast<const cplx<R>*>(p.rtw);
    const int L = p.L;
    const R half = R(0.5);
    for (int k = t; k <= L / 2; k += p.TL) {
        cplx<R> zk = line[lds_pad(k == 0 ? 0 : k)];
        cplx<R> zm = line[lds_pad(k == 0 ? 0 : L - k)];
        cplx<R> w = rtw[k];
        R sum_re = zk.x + zm.x, diff_im = zk.y - zm.y;
        R ra_re = zk.y + zm.y, ra_im = zm.x - zk.x;
        // rot = (0.5 w) * (ra_re + i ra_im)
        R rot_re = half * (w.x * ra_re - w.y * ra_im);
        R rot_im = half * (w.x * ra_im + w.y * ra_re);
        R h_re = half * sum_re, h_im = half * diff_im;
        cplx<R> xk = {h_re + rot_re, h_im + rot_im};
        cplx<R> xm = {h_re - rot_re, rot_im - h_im};
        if (k == 0) {
            xk = {zk.x + zk.y, R(0)};
            xm = {zk.x - zk.y, R(0)};
        }
        if (p.flags & kScale) {
            xk = cscale(xk, R(sizeof(R) == 4 ? p.scale_f : p.scale_d));
            xm = cscale(xm, R(sizeof(R) == 4 ? p.scale_f : p.scale_d));
        }
        dst[(int64_t)k * p.oes] = xk;
        if (k != L - k) dst[(int64_t)(L - k) * p.oes] = xm;
    }
}

template <typename R, int E>
__device__ __forceinline__ void store_tile(const PassDesc& p, const TileIo& io, const cplx<R>* lds, int64_t l_base, int tid)
{
    const bool col = (p.flags & kOutCol) != 0;
    const int c = col ? (tid & (p.C - 1)) : (tid >> p.tl_shift);
    const int t = col ? (tid >> p.c_shift) : (tid & (p.TL - 1));
    const SlotLine sl = slot_line(p, io, l_base, c);
    if (!sl.live) return;
    LineIdx li = line_index(p, io, sl.l);
    const cplx<R>* line = lds + c * p.LP;
    cplx<R> v[E];
#pragma unroll
    for (int k = 0; k < E; k++) {
        const int e = t + k * p.TL;
        if (e < p.nvalid_out) v[k] = line[lds_pad(e)];
    }
    if (p.flags & kTw4) {
        // w^(l0 e) for e = t + k TL: two table products, then an FP64
        // recurrence over k
        cplx<double> wc = tw4_root(p, (int64_t)li.l0 * t);
        const cplx<double> ws = tw4_root(p, (int64_t)li.l0 * p.TL);
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int e = t + k * p.TL;
            if (e < p.nvalid_out) {
                cplx<double> vd = {double(v[k].x), double(v[k].y)};
                vd = cmul(vd, wc);
                v[k] = {R(vd.x), R(vd.y)};
            }
            wc = cmul(wc, ws);
        }
    }
    if (p.flags & kScale) {
        const R s = R(sizeof(R) == 4 ? p.scale_f : p.scale_d);
#pragma unroll
        for (int k = 0; k < E; k++) v[k] = cscale(v[k], s);
    }
    if (p.flags & kConjOut) {
#pragma unroll
        for (int k = 0; k < E; k++) v[k].y = -v[k].y;
    }
    if (p.flags & kRealPart) {
        R* dst = reinterpret_cast<R*>(io.out) + li.ooff;
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int e = t + k * p.TL;
            if (e < p.nvalid_out) dst[(int64_t)e * p.oes] = v[k].x;
        }
        return;
    }
    if (p.flags & kRealOut) {
        // C2R: element j is (x[2j], x[2j+1]); stored conjugated-domain value
        // was already flipped back by kConjOut.
        R* dst = reinterpret_cast<R*>(io.out) + li.ooff;
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int e = t + k * p.TL;
            if (e < p.nvalid_out) {
                dst[(int64_t)(2 * e) * p.oes] = v[k].x;
                dst[(int64_t)(2 * e + 1) * p.oes] = v[k].y;
            }
        }
        return;
    }
    cplx<R>* dst = reinterpret_cast<cplx<R>*>(io.out) + li.ooff;
#pragma unroll
    for (int k = 0; k < E; k++) {
        const int e = t + k * p.TL;
        if (e < p.nvalid_out) dst[(int64_t)e * p.oes] = v[k];
    }
}

// The pass descriptor read in place from the kernel-argument segment (the
// kernels take it as their only argument, at offset 0).  A by-value copy
// whose arrays are indexed by the runtime stage number would be materialised
// in per-lane scratch memory.
__device__ __forceinline__ const PassDesc& kernarg_desc()
{
    return *(const PassDesc*)__builtin_amdgcn_kernarg_segment_ptr();  // address-space cast
}

__device__ __forceinline__ uint32_t tile_of_block(const PassDesc& p)
{
    uint32_t tile = blockIdx.x;
    if (p.flags & kXcdRemap) {
        // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD
        // a contiguous run of tiles so column tiles that share cache lines
        // meet in the same L2.
        const uint32_t G = gridDim.x;
        const uint32_t per = G >> 3;
        if (tile < (per << 3)) tile = (tile & 7u) * per + (tile >> 3);
    }
    return tile;
}

// ---------------------------------------------------------------------------
// Register-direct first and last stages (complex in / complex or real-pair
// out, no real-packed step).  Stage 0 (ls = 1) of butterfly j reads
// x[i*nbl + j], i < R -- coalesced across j -- so a thread loads exactly its
// butterflies' operands from HBM, runs the DFT and writes y[j*R + i] to LDS.
// The last stage (ls = nbl, one group) writes y[i*nbl + j] straight to HBM
// with the fused four-step twiddle, scale and conjugation.  This drops the
// LDS round trip and barrier of a staged load and of a staged store.
// (c, t) is the side's thread mapping (row or column).
// ---------------------------------------------------------------------------
template <typename R>
struct LineIo {
    const cplx<R>* src;
    cplx<R>* dst;
    uint32_t l0;
    bool live_in, live_out;
};

// Element e of a line whose element 0 has sequence index gbase, as loaded:
// with kChirpIn, (conjugated if backward) x * chirp[g] for g < nchirp and 0
// beyond (the input holds only nchirp elements); else x for e < nvalid_in.
template <typename R>
__device__ __forceinline__ cplx<R> load_elem(const PassDesc& p, const cplx<R>* src, int e, int64_t gbase)
{
    cplx<R> v = {R(0), R(0)};
    if (p.flags & kChirpIn) {
        const int64_t g = gbase + (int64_t)e * p.gin[3];
        if (g < p.nchirp) {
            v = src[(int64_t)e * p.ies];
            if (p.flags & kConjIn) v.y = -v.y;
            v = cmul(v, reinterpret_cast<const cplx<R>*>(p.chirp)[g]);
        }
    } else if (e < p.nvalid_in) {
        v = src[(int64_t)e * p.ies];
    }
    return v;
}

template <typename R, int E, int RAD>
__device__ __forceinline__ void first_stage_direct(const PassDesc& p, cplx<R>* lds, int c, int t, const cplx<R>* src,
                                                   bool live, int64_t gbase)
{
    constexpr int K = (E + RAD - 1) / RAD;
    const int nbl = p.L / RAD;
    cplx<R> v[K][RAD];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (live && j < nbl) {
#pragma unroll
            for (int i = 0; i < RAD; i++) v[k][i] = load_elem<R>(p, src, i * nbl + j, gbase);
        }
    }
    cplx<R>* line = lds + c * p.LP;
    // (with kChirpIn the conjugation happened before the chirp product)
    const bool cj = (p.flags & (kConjIn | kChirpIn)) == kConjIn;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (live && j < nbl) {
            if (cj) {
#pragma unroll
                for (int i = 0; i < RAD; i++) v[k][i].y = -v[k][i].y;
            }
            Dft<RAD, R>::run(v[k]);
#pragma unroll
            for (int i = 0; i < RAD; i++) line[lds_pad(j * RAD + i)] = v[k][i];
        }
    }
}

// Output side of the last stage for one butterfly: element e = i*nbl + j.
template <typename R, int RAD>
__device__ __forceinline__ void store_butterfly(const PassDesc& p, cplx<R>* v, int j, int nbl, cplx<R>* dst,
                                                uint32_t l0, int64_t gbase)
{
    if (p.flags & kTw4) {
        // w_N^(l0 e), e = i*nbl + j: two table products, then an FP64
        // recurrence over i
        cplx<double> wc = tw4_root(p, (int64_t)l0 * j);
        const cplx<double> ws = tw4_root(p, (int64_t)l0 * nbl);
#pragma unroll
        for (int i = 0; i < RAD; i++) {
            cplx<double> vd = {double(v[i].x), double(v[i].y)};
            vd = cmul(vd, wc);
            v[i] = {R(vd.x), R(vd.y)};
            wc = cmul(wc, ws);
        }
    }
    if (p.flags & kScale) {
        const R sc = R(sizeof(R) == 4 ? p.scale_f : p.scale_d);
#pragma unroll
        for (int i = 0; i < RAD; i++) v[i] = cscale(v[i], sc);
    }
    if (p.flags & (kSpecOut | kChirpOut)) {
        // multi-pass Bluestein: Z = conj(A) * spec after the first M-point
        // transform; conj(chirp) * W after the second (stored for g < n only)
#pragma unroll
        for (int i = 0; i < RAD; i++) {
            const int64_t g = gbase + (int64_t)(i * nbl + j) * p.gout[3];
            if (p.flags & kSpecOut) {
                v[i] = cmul(cconj(v[i]), reinterpret_cast<const cplx<R>*>(p.spec)[g]);
            } else if (g < p.nchirp) {
                v[i] = cmul(cconj(reinterpret_cast<const cplx<R>*>(p.chirp)[g]), v[i]);
            }
        }
    }
    if (p.flags & kConjOut) {
#pragma unroll
        for (int i = 0; i < RAD; i++) v[i].y = -v[i].y;
    }
    if (p.flags & kChirpOut) {
#pragma unroll
        for (int i = 0; i < RAD; i++) {
            const int e = i * nbl + j;
            if (gbase + (int64_t)e * p.gout[3] < p.nchirp) dst[(int64_t)e * p.oes] = v[i];
        }
        return;
    }
    const int nvo = p.nvalid_out;
    if (p.flags & kRealOut) {
        R* d = reinterpret_cast<R*>(dst);
#pragma unroll
        for (int i = 0; i < RAD; i++) {
            const int e = i * nbl + j;
            if (e < nvo) {
                d[(int64_t)(2 * e) * p.oes] = v[i].x;
                d[(int64_t)(2 * e + 1) * p.oes] = v[i].y;
            }
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < RAD; i++) {
        const int e = i * nbl + j;
        if (e < nvo) dst[(int64_t)e * p.oes] = v[i];
    }
}

template <typename R, int E, int RAD, bool LIN>
__device__ __forceinline__ void last_stage_direct_impl(const PassDesc& p, const cplx<R>* lds, int c, int t,
                                                       cplx<R>* dst, uint32_t l0, bool live, int s, int64_t gbase)
{
    constexpr int K = (E + RAD - 1) / RAD;
    const int nbl = p.L / RAD;  // = ls of the last stage; g = 0, cc = j
    const cplx<R>* line = lds + c * p.LP;
    cplx<R> v[K][RAD];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (live && j < nbl) lds_gather<R, RAD, LIN>(line, j, nbl, v[k]);
    }
    const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(p.stw) + p.twoff[s];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (live && j < nbl) {
            if (nbl > 1 && s > 0) {
                const cplx<R>* w = tw + j * (RAD - 1);
#pragma unroll
                for (int i = 1; i < RAD; i++) v[k][i] = cmul(v[k][i], w[i - 1]);
            }
            Dft<RAD, R>::run(v[k]);
            store_butterfly<R, RAD>(p, v[k], j, nbl, dst, l0, gbase);
        }
    }
}

template <typename R, int E, int RAD>
__device__ __forceinline__ void last_stage_direct(const PassDesc& p, const cplx<R>* lds, int c, int t, cplx<R>* dst,
                                                  uint32_t l0, bool live, int s, int64_t gbase)
{
    if (sizeof(R) == 4 && ((p.L / RAD) & 15) == 0)
        last_stage_direct_impl<R, E, RAD, true>(p, lds, c, t, dst, l0, live, s, gbase);
    else
        last_stage_direct_impl<R, E, RAD, false>(p, lds, c, t, dst, l0, live, s, gbase);
}

// One-stage lengths (L <= 16): load, DFT, store, no LDS.
template <typename R, int E, int RAD>
__device__ __forceinline__ void single_stage_direct(const PassDesc& p, int t, const cplx<R>* src, cplx<R>* dst,
                                                    uint32_t l0, bool live, int64_t gin, int64_t gout)
{
    if (!live || t != 0) return;
    cplx<R> v[RAD];
    const bool cj = (p.flags & (kConjIn | kChirpIn)) == kConjIn;
#pragma unroll
    for (int i = 0; i < RAD; i++) {
        v[i] = load_elem<R>(p, src, i, gin);
        if (cj) v[i].y = -v[i].y;
    }
    Dft<RAD, R>::run(v);
    store_butterfly<R, RAD>(p, v, 0, 1, dst, l0, gout);
}

#define FFTZ_RADIX_SWITCH(r, ALLRAD, CALL)                 \
    switch (r) {                                           \
        case 2: { constexpr int RAD = 2; CALL; } break;    \
        case 4: { constexpr int RAD = 4; CALL; } break;    \
        case 8: { constexpr int RAD = 8; CALL; } break;    \
        case 16: { constexpr int RAD = 16; CALL; } break;  \
        default:                                           \
            if constexpr (ALLRAD) {                        \
                switch (r) {                               \
                    case 3: { constexpr int RAD = 3; CALL; } break;   \
                    case 5: { constexpr int RAD = 5; CALL; } break;   \
                    case 6: { constexpr int RAD = 6; CALL; } break;   \
                    case 7: { constexpr int RAD = 7; CALL; } break;   \
                    case 9: { constexpr int RAD = 9; CALL; } break;   \
                    case 10: { constexpr int RAD = 10; CALL; } break; \
                    case 11: { constexpr int RAD = 11; CALL; } break; \
                    case 12: { constexpr int RAD = 12; CALL; } break; \
                    case 13: { constexpr int RAD = 13; CALL; } break; \
                    default: break;                        \
                }                                          \
            }                                              \
            break;                                         \
    }

// Complex-in pass with register-direct first / last stages.
template <typename R, bool ALLRAD, int E>
__device__ __forceinline__ void pass_body_direct(const PassDesc& p, const TileIo& io, cplx<R>* lds, int64_t l_base, int tid)
{
    const bool in_col = (p.flags & kInCol) != 0, out_col = (p.flags & kOutCol) != 0;
    const int ci = in_col ? (tid & (p.C - 1)) : (tid >> p.tl_shift);
    const int ti = in_col ? (tid >> p.c_shift) : (tid & (p.TL - 1));
    const int co = out_col ? (tid & (p.C - 1)) : (tid >> p.tl_shift);
    const int to = out_col ? (tid >> p.c_shift) : (tid & (p.TL - 1));
    const SlotLine sli = slot_line(p, io, l_base, ci);
    const SlotLine slo = slot_line(p, io, l_base, co);
    const bool live_in = ci < p.C && sli.live;
    const bool live_out = co < p.C && slo.live;
    const cplx<R>* src = nullptr;
    int64_t gin = 0;
    if (live_in) {
        LineIdx li = line_index(p, io, sli.l);
        src = reinterpret_cast<const cplx<R>*>(io.in) + li.ioff;
        gin = li.gin;
    }
    cplx<R>* dst = nullptr;
    uint32_t l0 = 0;
    int64_t gout = 0;
    if (live_out) {
        LineIdx li = line_index(p, io, slo.l);
        dst = reinterpret_cast<cplx<R>*>(io.out) + li.ooff;
        l0 = li.l0;
        gout = li.gout;
    }
    const int S = p.nst;
    if (S == 1) {
        // the loading thread stores: use the input-side line for both
        cplx<R>* d1 = nullptr;
        uint32_t l01 = 0;
        int64_t gout1 = 0;
        if (live_in) {
            LineIdx li = line_index(p, io, sli.l);
            d1 = reinterpret_cast<cplx<R>*>(io.out) + li.ooff;
            l01 = li.l0;
            gout1 = li.gout;
        }
        FFTZ_RADIX_SWITCH(p.radix[0], ALLRAD,
                          (single_stage_direct<R, E, RAD>(p, ti, src, d1, l01, live_in, gin, gout1)))
        return;
    }
    FFTZ_RADIX_SWITCH(p.radix[0], ALLRAD, (first_stage_direct<R, E, RAD>(p, lds, ci, ti, src, live_in, gin)))
    __syncthreads();
    run_stages<R, ALLRAD, E>(p, lds, tid, 1, S - 1);
    FFTZ_RADIX_SWITCH(p.radix[S - 1], ALLRAD,
                      (last_stage_direct<R, E, RAD>(p, lds, co, to, dst, l0, live_out, S - 1, gout)))
}

template <typename R, bool ALLRAD, int E>
__device__ __forceinline__ void pass_body(const PassDesc& p, const TileIo& io, cplx<R>* lds, int64_t l_base, int tid)
{
    // (a zero-stage plan, L = 1, is a strided copy: the staged path below)
    if (!(p.flags & kStagedFlags) && p.nst >= 1) {
        pass_body_direct<R, ALLRAD, E>(p, io, lds, l_base, tid);
        return;
    }
    load_tile<R, E>(p, io, lds, l_base, tid);
    __syncthreads();
    if (p.flags & kC2RPre) {
        c2r_separate<R>(p, lds, tid);
        __syncthreads();
    }
    run_stages<R, ALLRAD, E>(p, lds, tid, 0, p.nst);
    if (p.flags & kR2CPost) {
        r2c_recombine_store<R>(p, io, lds, l_base, tid);
    } else {
        store_tile<R, E>(p, io, lds, l_base, tid);
    }
}

template <typename R, bool ALLRAD, int E, int MAXT>
__global__ void __launch_bounds__(MAXT) k_fft_pass(const PassDesc p_arg)
{
    (void)p_arg;
    const PassDesc& p = kernarg_desc();
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int64_t l_base = (int64_t)tile_of_block(p) * p.C;
    pass_body<R, ALLRAD, E>(p, io_of(p), lds, l_base, tid);
}

// ---------------------------------------------------------------------------
// Bluestein (chirp-z) for one or more lines of prime length n, fused into a
// single pass over an M-point LDS tile (M = smooth length >= 2n-1, the
// reference's get_extended_length, core/common/bluestein_utils.c:39-70):
//   a = x * conj(B) zero-padded to M     (load)
//   A = FFT_M(a)                         (stages)
//   Z = conj(A) * conj(FFT_M(B)) / M      (in LDS)
//   W = FFT_M(Z) = conj(IFFT_M(A * FFT_M(B)) / M)
//   X = conj(B * W)                       (store)
// which is core/solvers/complex/bluestein_solver_dft.c:137-234 with its five
// sweeps (pre_mul, memset, FFT, mul, IFFT, post_mul) in one HBM round trip.
// ---------------------------------------------------------------------------
template <typename R, bool ALLRAD, int E>
__device__ __forceinline__ void bluestein_body(const PassDesc& p, const TileIo& io, cplx<R>* lds, int64_t l_base, int tid)
{
    const int c = tid >> p.tl_shift;
    const int t = tid & (p.TL - 1);
    const SlotLine sl = slot_line(p, io, l_base, c);
    const bool live = sl.live;
    LineIdx li = {0, 0, 0};
    if (live) li = line_index(p, io, sl.l);
    cplx<R>* line = lds + c * p.LP;
    const cplx<R>* chirp = reinterpret_cast<const cplx<R>*>(p.chirp);
    const cplx<R>* spec = reinterpret_cast<const cplx<R>*>(p.spec);
    const int n = p.nvalid_in;
    const int M = p.L;
    if (live) {
        const cplx<R>* src = reinterpret_cast<const cplx<R>*>(io.in) + li.ioff;
        for (int e = t; e < M; e += p.TL) {
            cplx<R> a = {R(0), R(0)};
            if (e < n) {
                cplx<R> x = src[(int64_t)e * p.ies];
                if (p.flags & kConjIn) x.y = -x.y;
                a = cmul(x, chirp[e]);
            }
            line[lds_pad(e)] = a;
        }
    }
    __syncthreads();
    run_stages<R, ALLRAD, E>(p, lds, tid, 0, p.nst);
    if (c < p.C) {
        for (int e = t; e < M; e += p.TL) {
            cplx<R> a = line[lds_pad(e)];
            line[lds_pad(e)] = cmul(cconj(a), spec[e]);
        }
    }
    __syncthreads();
    run_stages<R, ALLRAD, E>(p, lds, tid, 0, p.nst);
    if (live) {
        cplx<R>* dst = reinterpret_cast<cplx<R>*>(io.out) + li.ooff;
        for (int k = t; k < n; k += p.TL) {
            cplx<R> w = line[lds_pad(k)];
            cplx<R> b = cconj(chirp[k]);
            cplx<R> y = cmul(b, w);  // B * W
            if (!(p.flags & kConjOut)) y.y = -y.y;  // forward result is conj(B W)
            dst[(int64_t)k * p.oes] = y;
        }
    }
}

// Register-direct Bluestein stages (M-point plan with at least two stages):
//   bs_first:  stage 0 of FFT_M(a) on operands a_e = x_e conj(B_e) (e < n,
//              zero beyond), loaded straight from HBM;
//   bs_spec:   last stage of FFT_M(a); each thread rewrites exactly the
//              positions it read with conj(A_e) * spec_e (no barrier needed);
//   bs_final:  last stage of the second FFT_M, X_k = conj(B_k W_k) (forward),
//              stored straight to HBM for k < n.
template <typename R, int E, int RAD>
__device__ __forceinline__ void bs_first(const PassDesc& p, cplx<R>* lds, int c, int t, const cplx<R>* src, bool live)
{
    constexpr int K = (E + RAD - 1) / RAD;
    const int nbl = p.L / RAD;
    const int n = p.nvalid_in;
    const cplx<R>* chirp = reinterpret_cast<const cplx<R>*>(p.chirp);
    const bool cj = (p.flags & kConjIn) != 0;
    cplx<R> v[K][RAD];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (live && j < nbl) {
#pragma unroll
            for (int i = 0; i < RAD; i++) {
                const int e = i * nbl + j;
                v[k][i] = {R(0), R(0)};
                if (e < n) {
                    cplx<R> x;
                    if (p.flags & kRealPromote) {
                        x = {reinterpret_cast<const R*>(src)[(int64_t)e * p.ies], R(0)};
                    } else if (p.flags & kHermIn) {
                        const int m = 2 * e <= n ? e : n - e;
                        x = src[(int64_t)m * p.ies];
                        if (2 * e > n) x.y = -x.y;
                    } else {
                        x = src[(int64_t)e * p.ies];
                    }
                    if (cj) x.y = -x.y;
                    v[k][i] = cmul(x, chirp[e]);
                }
            }
        }
    }
    cplx<R>* line = lds + c * p.LP;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (live && j < nbl) {
            Dft<RAD, R>::run(v[k]);
#pragma unroll
            for (int i = 0; i < RAD; i++) line[lds_pad(j * RAD + i)] = v[k][i];
        }
    }
}

template <typename R, int E, int RAD, bool LIN>
__device__ __forceinline__ void bs_last_common(const PassDesc& p, const cplx<R>* line, int t, bool live, int s,
                                               cplx<R> (&v)[(E + RAD - 1) / RAD][RAD])
{
    constexpr int K = (E + RAD - 1) / RAD;
    const int nbl = p.L / RAD;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (live && j < nbl) lds_gather<R, RAD, LIN>(line, j, nbl, v[k]);
    }
    const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(p.stw) + p.twoff[s];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (live && j < nbl) {
            if (nbl > 1) {
                const cplx<R>* w = tw + j * (RAD - 1);
#pragma unroll
                for (int i = 1; i < RAD; i++) v[k][i] = cmul(v[k][i], w[i - 1]);
            }
            Dft<RAD, R>::run(v[k]);
        }
    }
}

template <typename R, int E, int RAD, bool LIN>
__device__ __forceinline__ void bs_spec_impl(const PassDesc& p, cplx<R>* lds, int c, int t, bool live, int s)
{
    constexpr int K = (E + RAD - 1) / RAD;
    const int nbl = p.L / RAD;
    cplx<R>* line = lds + c * p.LP;
    cplx<R> v[K][RAD];
    bs_last_common<R, E, RAD, LIN>(p, line, t, live, s, v);
    const cplx<R>* spec = reinterpret_cast<const cplx<R>*>(p.spec);
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (live && j < nbl) {
#pragma unroll
            for (int i = 0; i < RAD; i++) v[k][i] = cmul(cconj(v[k][i]), spec[i * nbl + j]);
            lds_scatter<R, RAD, LIN>(line, j, nbl, v[k]);
        }
    }
}

template <typename R, int E, int RAD>
__device__ __forceinline__ void bs_spec(const PassDesc& p, cplx<R>* lds, int c, int t, bool live, int s)
{
    if (sizeof(R) == 4 && ((p.L / RAD) & 15) == 0)
        bs_spec_impl<R, E, RAD, true>(p, lds, c, t, live, s);
    else
        bs_spec_impl<R, E, RAD, false>(p, lds, c, t, live, s);
}

template <typename R, int E, int RAD, bool LIN>
__device__ __forceinline__ void bs_final_impl(const PassDesc& p, const cplx<R>* lds, int c, int t, cplx<R>* dst,
                                              bool live, int s)
{
    constexpr int K = (E + RAD - 1) / RAD;
    const int nbl = p.L / RAD;
    const int n = p.nvalid_out;
    const cplx<R>* line = lds + c * p.LP;
    cplx<R> v[K][RAD];
    bs_last_common<R, E, RAD, LIN>(p, line, t, live, s, v);
    const cplx<R>* chirp = reinterpret_cast<const cplx<R>*>(p.chirp);
    const bool fwd = !(p.flags & kConjOut);
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (live && j < nbl) {
#pragma unroll
            for (int i = 0; i < RAD; i++) {
                const int e = i * nbl + j;
                if (e < n) {
                    cplx<R> y = cmul(cconj(chirp[e]), v[k][i]);  // B * W
                    if (fwd) y.y = -y.y;                          // forward result is conj(B W)
                    if (p.flags & kRealPart)
                        reinterpret_cast<R*>(dst)[(int64_t)e * p.oes] = y.x;
                    else
                        dst[(int64_t)e * p.oes] = y;
                }
            }
        }
    }
}

template <typename R, int E, int RAD>
__device__ __forceinline__ void bs_final(const PassDesc& p, const cplx<R>* lds, int c, int t, cplx<R>* dst, bool live,
                                         int s)
{
    if (sizeof(R) == 4 && ((p.L / RAD) & 15) == 0)
        bs_final_impl<R, E, RAD, true>(p, lds, c, t, dst, live, s);
    else
        bs_final_impl<R, E, RAD, false>(p, lds, c, t, dst, live, s);
}

template <typename R, bool ALLRAD, int E>
__device__ __forceinline__ void bluestein_body_direct(const PassDesc& p, const TileIo& io, cplx<R>* lds, int64_t l_base, int tid)
{
    const int c = tid >> p.tl_shift;
    const int t = tid & (p.TL - 1);
    const SlotLine sl = slot_line(p, io, l_base, c);
    const bool live = c < p.C && sl.live;
    const cplx<R>* src = nullptr;
    cplx<R>* dst = nullptr;
    if (live) {
        LineIdx li = line_index(p, io, sl.l);
        // real-side lines are offset in real elements (and then only read /
        // written through real pointers)
        src = (p.flags & kRealPromote)
                  ? reinterpret_cast<const cplx<R>*>(reinterpret_cast<const R*>(io.in) + li.ioff)
                  : reinterpret_cast<const cplx<R>*>(io.in) + li.ioff;
        dst = (p.flags & kRealPart) ? reinterpret_cast<cplx<R>*>(reinterpret_cast<R*>(io.out) + li.ooff)
                                    : reinterpret_cast<cplx<R>*>(io.out) + li.ooff;
    }
    const int S = p.nst;
    FFTZ_RADIX_SWITCH(p.radix[0], ALLRAD, (bs_first<R, E, RAD>(p, lds, c, t, src, live)))
    __syncthreads();
    run_stages<R, ALLRAD, E>(p, lds, tid, 1, S - 1);
    FFTZ_RADIX_SWITCH(p.radix[S - 1], ALLRAD, (bs_spec<R, E, RAD>(p, lds, c, t, live, S - 1)))
    __syncthreads();
    run_stages<R, ALLRAD, E>(p, lds, tid, 0, S - 1);
    FFTZ_RADIX_SWITCH(p.radix[S - 1], ALLRAD, (bs_final<R, E, RAD>(p, lds, c, t, dst, live, S - 1)))
}

template <typename R, bool ALLRAD, int E, int MAXT>
__global__ void __launch_bounds__(MAXT) k_bluestein(const PassDesc p_arg)
{
    (void)p_arg;
    const PassDesc& p = kernarg_desc();
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem_raw);
    const TileIo io = io_of(p);
    if (p.nst >= 2)
        bluestein_body_direct<R, ALLRAD, E>(p, io, lds, (int64_t)blockIdx.x * p.C, threadIdx.x);
    else
        bluestein_body<R, ALLRAD, E>(p, io, lds, (int64_t)blockIdx.x * p.C, threadIdx.x);
}

// Length-1 items of an uneven batch: the transform is the identity (the
// backward direction's two conjugations cancel).
template <typename R>
__device__ __forceinline__ void batch_identity(const PassDesc& p, const TileIo& io, int64_t l_base, int tid)
{
    const int c = tid >> p.tl_shift;
    const SlotLine sl = slot_line(p, io, l_base, c);
    if (c >= p.C || !sl.live || (tid & (p.TL - 1)) != 0) return;
    const LineIdx li = line_index(p, io, sl.l);
    reinterpret_cast<cplx<R>*>(io.out)[li.ooff] = reinterpret_cast<const cplx<R>*>(io.in)[li.ioff];
}

// ---------------------------------------------------------------------------
// Uneven batches in one persistent launch.  A device work queue lists tiles
// (size class, first line, line count, offsets); each workgroup takes the
// next tile from an atomic counter, copies the size class's pass descriptor
// into LDS, patches the tile's buffers in, and runs the same pass or
// Bluestein body as the single-size kernels.  Tiles are sorted by cost on
// the host, largest first.  The arithmetic of a tile does not depend on the
// workgroup that runs it, so results are bitwise identical run to run.
// ---------------------------------------------------------------------------
template <typename R, int E, int MAXT>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 2)))
k_batch(const BatchArgs a)
{
    // dynamic LDS: [tile index | tile data].  The tile index is made uniform
    // with readfirstlane, and the queue entry and its size class's descriptor
    // are read in place from the plan tables through the constant address
    // space (read-only for the whole launch), i.e. with scalar loads.
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    int& s_tile = *reinterpret_cast<int*>(smem_raw);
    cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem_raw + kBatchHdrBytes);
    const int tid = threadIdx.x;
    const int esz = 2 * (int)sizeof(R);
    if (tid == 0) s_tile = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    for (;;) {
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        if (t >= a.ntiles) break;
        // the next tile's index is fetched now and published at the end, so
        // the atomic's round trip overlaps this tile's work
        int nxt = 0;
        if (tid == 0) nxt = (int)atomicAdd(a.counter, 1u);
        const BatchTile& bt = *(const BatchTile*)((const __attribute__((address_space(4))) BatchTile*)a.tiles + t);
        const PassDesc& p = *(const PassDesc*)((const __attribute__((address_space(4))) PassDesc*)a.descs + bt.desc);
        const TileIo io{reinterpret_cast<const char*>(a.in) + bt.in_off * esz,
                        reinterpret_cast<char*>(a.out) + bt.out_off * esz, (int64_t)bt.nlines, bt.dist, bt.dist};
        const bool blue = bt.bluestein != 0;
        const int C = p.C;
        // a queue entry covers one or more rounds of C lines
        for (int64_t lb = 0; lb < io.nlines; lb += C) {
            // only the register-direct bodies: batch items are plain complex
            // transforms (no staged real-packed steps), and a Bluestein
            // extended length (M >= 33) always has two or more stages
            if (blue)
                bluestein_body_direct<R, true, E>(p, io, lds, lb, tid);
            else if (p.nst >= 1)
                pass_body_direct<R, true, E>(p, io, lds, lb, tid);
            else
                batch_identity<R>(p, io, lb, tid);
            __syncthreads();
        }
        if (tid == 0) s_tile = nxt;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Launch table: precision x radix set x elements per thread x block limit.
// ---------------------------------------------------------------------------
template <typename R, bool A, int E, int T>
inline const void* sym(KernelKind kind)
{
    return kind == KernelKind::Pass ? (const void*)&k_fft_pass<R, A, E, T> : (const void*)&k_bluestein<R, A, E, T>;
}

template <typename R, bool A>
inline const void* sym_et(KernelKind kind, int E, int maxt)
{
    if (E == 8) return maxt <= 256 ? sym<R, A, 8, 256>(kind) : sym<R, A, 8, 1024>(kind);
    return maxt <= 256 ? sym<R, A, 16, 256>(kind) : sym<R, A, 16, 1024>(kind);
}

}  // namespace fftz
