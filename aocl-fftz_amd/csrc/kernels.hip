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
};

__device__ __forceinline__ LineIdx line_index(const PassDesc& p, uint32_t l)
{
    uint32_t r = fdiv(l, p.ln0);
    uint32_t l0 = l - r * p.ln0.d;
    uint32_t l2 = fdiv(r, p.ln1);
    uint32_t l1 = r - l2 * p.ln1.d;
    LineIdx li;
    li.ioff = (int64_t)l0 * p.ils[0] + (int64_t)l1 * p.ils[1] + (int64_t)l2 * p.ils[2];
    li.ooff = (int64_t)l0 * p.ols[0] + (int64_t)l1 * p.ols[1] + (int64_t)l2 * p.ols[2];
    li.l0 = l0;
    return li;
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
__device__ __forceinline__ void load_tile(const PassDesc& p, cplx<R>* lds, int64_t l_base, int tid)
{
    const bool col = (p.flags & kInCol) != 0;
    const int c = col ? (tid & (p.C - 1)) : (tid >> p.tl_shift);
    const int t = col ? (tid >> p.c_shift) : (tid & (p.TL - 1));
    const int64_t l = l_base + c;
    if (l >= p.nlines) return;
    LineIdx li = line_index(p, (uint32_t)l);
    cplx<R>* line = lds + c * p.LP;
    if (p.flags & kRealIn) {
        // R2C: element j of the L-point complex line is (x[2j], x[2j+1])
        const R* src = reinterpret_cast<const R*>(p.in) + li.ioff;
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
    const cplx<R>* src = reinterpret_cast<const cplx<R>*>(p.in) + li.ioff;
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

template <typename R, int E, int RAD>
__device__ __forceinline__ void stage(const PassDesc& p, cplx<R>* lds, int tid, int s)
{
    constexpr int K = (E + RAD - 1) / RAD;
    const int c = tid >> p.tl_shift;
    const int t = tid & (p.TL - 1);
    const int nbl = p.L / RAD;
    const int ls = p.ls[s];
    cplx<R>* line = lds + c * p.LP;
    const bool act = c < p.C;  // batch kernel: threads beyond C * TL idle
    cplx<R> v[K][RAD];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (act && j < nbl) {
#pragma unroll
            for (int i = 0; i < RAD; i++) v[k][i] = line[lds_pad(i * nbl + j)];
        }
    }
    __syncthreads();
    const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(p.stw) + p.twoff[s];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = t + k * p.TL;
        if (act && j < nbl) {
            const int g = (int)fdiv((uint32_t)j, p.lsdiv[s]);
            const int cc = j - g * ls;
            if (ls > 1) {
                const cplx<R>* w = tw + cc * (RAD - 1);
#pragma unroll
                for (int i = 1; i < RAD; i++) v[k][i] = cmul(v[k][i], w[i - 1]);
            }
            Dft<RAD, R>::run(v[k]);
            const int base = g * RAD * ls + cc;
#pragma unroll
            for (int i = 0; i < RAD; i++) line[lds_pad(base + i * ls)] = v[k][i];
        }
    }
    __syncthreads();
}

template <typename R, bool ALLRAD, int E>
__device__ __forceinline__ void run_stages(const PassDesc& p, cplx<R>* lds, int tid)
{
    for (int s = 0; s < p.nst; s++) {
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
__device__ __forceinline__ void r2c_recombine_store(const PassDesc& p, const cplx<R>* lds, int64_t l_base, int tid)
{
    const int c = tid >> p.tl_shift;
    const int t = tid & (p.TL - 1);
    const int64_t l = l_base + c;
    if (l >= p.nlines) return;
    LineIdx li = line_index(p, (uint32_t)l);
    const cplx<R>* line = lds + c * p.LP;
    cplx<R>* dst = reinterpret_cast<cplx<R>*>(p.out) + li.ooff;
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
__device__ __forceinline__ void store_tile(const PassDesc& p, const cplx<R>* lds, int64_t l_base, int tid)
{
    const bool col = (p.flags & kOutCol) != 0;
    const int c = col ? (tid & (p.C - 1)) : (tid >> p.tl_shift);
    const int t = col ? (tid >> p.c_shift) : (tid & (p.TL - 1));
    const int64_t l = l_base + c;
    if (l >= p.nlines) return;
    LineIdx li = line_index(p, (uint32_t)l);
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
    if (p.flags & kRealOut) {
        // C2R: element j is (x[2j], x[2j+1]); stored conjugated-domain value
        // was already flipped back by kConjOut.
        R* dst = reinterpret_cast<R*>(p.out) + li.ooff;
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
    cplx<R>* dst = reinterpret_cast<cplx<R>*>(p.out) + li.ooff;
#pragma unroll
    for (int k = 0; k < E; k++) {
        const int e = t + k * p.TL;
        if (e < p.nvalid_out) dst[(int64_t)e * p.oes] = v[k];
    }
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

template <typename R, bool ALLRAD, int E>
__device__ __forceinline__ void pass_body(const PassDesc& p, cplx<R>* lds, int64_t l_base, int tid)
{
    load_tile<R, E>(p, lds, l_base, tid);
    __syncthreads();
    if (p.flags & kC2RPre) {
        c2r_separate<R>(p, lds, tid);
        __syncthreads();
    }
    run_stages<R, ALLRAD, E>(p, lds, tid);
    if (p.flags & kR2CPost) {
        r2c_recombine_store<R>(p, lds, l_base, tid);
    } else {
        store_tile<R, E>(p, lds, l_base, tid);
    }
}

template <typename R, bool ALLRAD, int E, int MAXT>
__global__ void __launch_bounds__(MAXT) k_fft_pass(const PassDesc p)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem_raw);
    const int tid = threadIdx.x;
    const int64_t l_base = (int64_t)tile_of_block(p) * p.C;
    pass_body<R, ALLRAD, E>(p, lds, l_base, tid);
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
__device__ __forceinline__ void bluestein_body(const PassDesc& p, cplx<R>* lds, int64_t l_base, int tid)
{
    const int c = tid >> p.tl_shift;
    const int t = tid & (p.TL - 1);
    const int64_t l = l_base + c;
    const bool live = l < p.nlines;
    LineIdx li = {0, 0, 0};
    if (live) li = line_index(p, (uint32_t)l);
    cplx<R>* line = lds + c * p.LP;
    const cplx<R>* chirp = reinterpret_cast<const cplx<R>*>(p.chirp);
    const cplx<R>* spec = reinterpret_cast<const cplx<R>*>(p.spec);
    const int n = p.nvalid_in;
    const int M = p.L;
    if (live) {
        const cplx<R>* src = reinterpret_cast<const cplx<R>*>(p.in) + li.ioff;
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
    run_stages<R, ALLRAD, E>(p, lds, tid);
    if (c < p.C) {
        for (int e = t; e < M; e += p.TL) {
            cplx<R> a = line[lds_pad(e)];
            line[lds_pad(e)] = cmul(cconj(a), spec[e]);
        }
    }
    __syncthreads();
    run_stages<R, ALLRAD, E>(p, lds, tid);
    if (live) {
        cplx<R>* dst = reinterpret_cast<cplx<R>*>(p.out) + li.ooff;
        for (int k = t; k < n; k += p.TL) {
            cplx<R> w = line[lds_pad(k)];
            cplx<R> b = cconj(chirp[k]);
            cplx<R> y = cmul(b, w);  // B * W
            if (!(p.flags & kConjOut)) y.y = -y.y;  // forward result is conj(B W)
            dst[(int64_t)k * p.oes] = y;
        }
    }
}

template <typename R, bool ALLRAD, int E, int MAXT>
__global__ void __launch_bounds__(MAXT) k_bluestein(const PassDesc p)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem_raw);
    bluestein_body<R, ALLRAD, E>(p, lds, (int64_t)blockIdx.x * p.C, threadIdx.x);
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
__global__ void __launch_bounds__(MAXT) k_batch(const BatchArgs a)
{
    // one dynamic LDS region, laid out explicitly: [descriptor | tile index |
    // tile data], each part 16-byte aligned (static __shared__ variables next
    // to a dynamic array would leave the array misaligned for 8/16-byte LDS
    // accesses)
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    PassDesc& pd = *reinterpret_cast<PassDesc*>(smem_raw);
    int& s_tile = *reinterpret_cast<int*>(smem_raw + kBatchHdrBytes - 16);
    cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem_raw + kBatchHdrBytes);
    const int tid = threadIdx.x;
    constexpr int kWords = sizeof(PassDesc) / 4;
    for (;;) {
        if (tid == 0) s_tile = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int t = s_tile;
        if (t >= a.ntiles) break;
        const BatchTile bt = a.tiles[t];
        // word-parallel copy through memcpy (char-typed accesses may alias
        // the descriptor's pointer / int64 / double fields; plain uint32_t
        // stores would let the compiler keep the previous tile's values)
        const char* src = reinterpret_cast<const char*>(a.descs + bt.desc);
        char* dstb = reinterpret_cast<char*>(&pd);
        for (int w = tid; w < kWords; w += MAXT) __builtin_memcpy(dstb + 4 * w, src + 4 * w, 4);
        __syncthreads();
        if (tid == 0) {
            const int esz = 2 * (int)sizeof(R);
            pd.in = reinterpret_cast<const char*>(a.in) + bt.in_off * esz;
            pd.out = reinterpret_cast<char*>(a.out) + bt.out_off * esz;
            pd.nlines = bt.nlines;
            pd.ils[0] = bt.dist;
            pd.ols[0] = bt.dist;
        }
        __syncthreads();
        if (bt.bluestein) {
            bluestein_body<R, true, E>(pd, lds, 0, tid);
        } else {
            pass_body<R, true, E>(pd, lds, 0, tid);
        }
        __syncthreads();
    }
}

__global__ void k_reset_counters(unsigned* c, int n)
{
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// Launch table: precision x radix set x elements per thread x block limit.
// ---------------------------------------------------------------------------
template <typename R, bool A, int E, int T>
static const void* sym(KernelKind kind)
{
    return kind == KernelKind::Pass ? (const void*)&k_fft_pass<R, A, E, T> : (const void*)&k_bluestein<R, A, E, T>;
}

template <typename R, bool A>
static const void* sym_et(KernelKind kind, int E, int maxt)
{
    if (E == 8) return maxt <= 256 ? sym<R, A, 8, 256>(kind) : sym<R, A, 8, 1024>(kind);
    return maxt <= 256 ? sym<R, A, 16, 256>(kind) : sym<R, A, 16, 1024>(kind);
}

const void* kernel_symbol_batch(int prec_bytes, int E, int maxt)
{
    if (prec_bytes == 4) {
        if (E == 8) return maxt <= 256 ? (const void*)&k_batch<float, 8, 256> : (const void*)&k_batch<float, 8, 1024>;
        return maxt <= 256 ? (const void*)&k_batch<float, 16, 256> : (const void*)&k_batch<float, 16, 1024>;
    }
    if (E == 8) return maxt <= 256 ? (const void*)&k_batch<double, 8, 256> : (const void*)&k_batch<double, 8, 1024>;
    return maxt <= 256 ? (const void*)&k_batch<double, 16, 256> : (const void*)&k_batch<double, 16, 1024>;
}

const void* kernel_symbol_reset() { return (const void*)&k_reset_counters; }

const void* kernel_symbol(KernelKind kind, int prec_bytes, bool all_radix, int E, int maxt)
{
    if (prec_bytes == 4)
        return all_radix ? sym_et<float, true>(kind, E, maxt) : sym_et<float, false>(kind, E, maxt);
    return all_radix ? sym_et<double, true>(kind, E, maxt) : sym_et<double, false>(kind, E, maxt);
}

}  // namespace fftz
