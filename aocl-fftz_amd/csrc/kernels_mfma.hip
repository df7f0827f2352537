// Split-precision matrix-core pass for power-of-16 lengths (L = 256, 4096),
// complex64 data.
//
// Every radix-16 Stockham stage is a small DFT-matrix GEMM on the MFMA units:
// for a block of 16 butterflies (columns) with inputs X (16 complex values
// each),
//   [Yr]   [Fr  -Fi] [Xr]
//   [Yi] = [Fi   Fr] [Xi],    F[k][j] = exp(-2 pi i j k / 16),
// a real 32x32 by 32x16 product: two v_mfma_f32_16x16x32 per product term
// (rows 0-15 give Yr, rows 16-31 give Yi), K = 32 = [Xr; Xi].
//
// Operand precision (MODE):
//   kMfmaF32   -- plain FP32 operands on v_mfma_f32_16x16x4_f32 (8 K-steps);
//   kMfmaBf16x3 -- F and X split into bf16 hi + lo, three products
//                  (Fl Xh + Fh Xl + Fh Xh), FP32 accumulate;
//   kMfmaF16x2  -- X scaled per column by a power of two so its largest entry
//                  is below 2^14, F and X split into fp16 hi + lo, the same
//                  three products, FP32 accumulate, scale undone exactly.
// The DFT-matrix fragments (already split, exact FP64 products rounded once)
// come from a per-lane table built on the host.
//
// Data path per tile (16 column blocks, 4 per wavefront):
//   stage 0 gathers its operands straight from HBM;
//   later stages read theirs from LDS and multiply them by the stage twiddle
//   (FP32 stage table, w_{16 ls}^(i c), loaded before the LDS reads);
//   operands are split into hi / lo a pair at a time (packed conversions);
//   every stage but the last writes its outputs to LDS in Stockham order,
//   the last stores straight to HBM.
// Hardware operand layout of the 16x16x32 forms: lane l holds A[l % 16][8 (l/16) + t]
// and B[8 (l/16) + t][l % 16], t < 8; result D[4 (l/16) + r][l % 16], r < 4.
// The 16x16x4 FP32 form: A[l % 16][l / 16], B[l / 16][l % 16], same D.  (Both
// checked on gfx950 with exact small-integer operands.)
#include <hip/hip_runtime.h>

#include "codelets.h"
#include "kernels.h"
#include "pass.h"

namespace fftz {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// hi + lo split of 8 floats, a pair at a time (one packed conversion for
// each half of each pair): hi = round(x), lo = round(x - hi).
__device__ __forceinline__ void split_bf16(const float (&x)[8], bf16x8& hi, bf16x8& lo)
{
#pragma unroll
    for (int t = 0; t < 8; t += 2) {
        const f32x2 v = {x[t], x[t + 1]};
        const bf16x2 h = __builtin_convertvector(v, bf16x2);
        const f32x2 r = v - __builtin_convertvector(h, f32x2);
        const bf16x2 l = __builtin_convertvector(r, bf16x2);
        hi[t] = h[0];
        hi[t + 1] = h[1];
        lo[t] = l[0];
        lo[t + 1] = l[1];
    }
}

__device__ __forceinline__ void split_f16(const float (&x)[8], f16x8& hi, f16x8& lo)
{
#pragma unroll
    for (int t = 0; t < 8; t += 2) {
        const f32x2 v = {x[t], x[t + 1]};
        const f16x2 h = __builtin_convertvector(v, f16x2);
        const f32x2 r = v - __builtin_convertvector(h, f32x2);
        const f16x2 l = __builtin_convertvector(r, f16x2);
        hi[t] = h[0];
        hi[t + 1] = h[1];
        lo[t] = l[0];
        lo[t + 1] = l[1];
    }
}

__device__ __forceinline__ void mfma_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

__device__ __forceinline__ uint32_t fdiv_m(uint32_t n, const FastDiv& f)
{
    return (__umulhi(n, f.mul) + n) >> f.shift;
}

struct LineOffM {
    int64_t ioff, ooff;
};

__device__ __forceinline__ LineOffM line_off_m(const PassDesc& p, uint32_t l)
{
    uint32_t r = fdiv_m(l, p.ln0);
    uint32_t l0 = l - r * p.ln0.d;
    uint32_t l2 = fdiv_m(r, p.ln1);
    uint32_t l1 = r - l2 * p.ln1.d;
    LineOffM o;
    o.ioff = (int64_t)l0 * p.ils[0] + (int64_t)l1 * p.ils[1] + (int64_t)l2 * p.ils[2];
    o.ooff = (int64_t)l0 * p.ols[0] + (int64_t)l1 * p.ols[1] + (int64_t)l2 * p.ols[2];
    return o;
}

// Operand layout (all modes).  The GEMM's K index is permuted so every lane
// holds four whole complex inputs of its column: lane group g = lane / 16
// holds rows rho(g, q) = 4 q + g, q < 4, as the 8-vector
//   t = 2 q + part:  part 0 -> Re x[rho], part 1 -> Im x[rho],
// i.e. the natural (re, im) memory order.  The host builds the DFT-matrix
// fragments in the same K order (A_top[t] = Fr or -Fi, A_bot[t] = Fi or Fr).
// For the FP32 form the 8-vector is consumed one element per K-step.
__device__ __forceinline__ int rho(int g, int q) { return 4 * q + g; }

template <int MODE>
struct Frag;
template <>
struct Frag<kMfmaF32> {
    float top[8], bot[8];
    __device__ __forceinline__ void init(const float* a)
    {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            top[t] = a[t];
            bot[t] = a[8 + t];
        }
    }
};
template <>
struct Frag<kMfmaBf16x3> {
    bf16x8 th, tl, bh, bl;
    __device__ __forceinline__ void init(const float* a)
    {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            th[t] = (__bf16)a[t];
            tl[t] = (__bf16)a[8 + t];
            bh[t] = (__bf16)a[16 + t];
            bl[t] = (__bf16)a[24 + t];
        }
    }
};
template <>
struct Frag<kMfmaF16x2> {
    f16x8 th, tl, bh, bl;
    __device__ __forceinline__ void init(const float* a)
    {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            th[t] = (_Float16)a[t];
            tl[t] = (_Float16)a[8 + t];
            bh[t] = (_Float16)a[16 + t];
            bl[t] = (_Float16)a[24 + t];
        }
    }
};

// Y = F X for NB column blocks at once (independent MFMA chains interleaved);
// yr[b][r], yi[b][r] = Re, Im of Y[4 g + r] of block b.
template <int MODE, int NB>
__device__ __forceinline__ void dft16_blocks(const float (&x)[NB][8], const Frag<MODE>& A, f32x4 (&yr)[NB],
                                             f32x4 (&yi)[NB])
{
#pragma unroll
    for (int b = 0; b < NB; b++) {
        yr[b] = f32x4{0.f, 0.f, 0.f, 0.f};
        yi[b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    if constexpr (MODE == kMfmaF32) {
#pragma unroll
        for (int kk = 0; kk < 8; kk++)
#pragma unroll
            for (int b = 0; b < NB; b++) {
                yr[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(A.top[kk], x[b][kk], yr[b], 0, 0, 0);
                yi[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(A.bot[kk], x[b][kk], yi[b], 0, 0, 0);
            }
    } else if constexpr (MODE == kMfmaBf16x3) {
        bf16x8 xh[NB], xl[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) split_bf16(x[b], xh[b], xl[b]);
#pragma unroll
        for (int b = 0; b < NB; b++) {
            yr[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.tl, xh[b], yr[b], 0, 0, 0);
            yi[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.bl, xh[b], yi[b], 0, 0, 0);
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            yr[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.th, xl[b], yr[b], 0, 0, 0);
            yi[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.bh, xl[b], yi[b], 0, 0, 0);
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            yr[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.th, xh[b], yr[b], 0, 0, 0);
            yi[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.bh, xh[b], yi[b], 0, 0, 0);
        }
    } else {
        // operands already scaled into fp16 range by the caller (see the
        // kernel: per-line power of two, stage twiddles carrying 1/16)
        f16x8 xh[NB], xl[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) split_f16(x[b], xh[b], xl[b]);
#pragma unroll
        for (int b = 0; b < NB; b++) {
            yr[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.tl, xh[b], yr[b], 0, 0, 0);
            yi[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.bl, xh[b], yi[b], 0, 0, 0);
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            yr[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.th, xl[b], yr[b], 0, 0, 0);
            yi[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.bh, xl[b], yi[b], 0, 0, 0);
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            yr[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.th, xh[b], yr[b], 0, 0, 0);
            yi[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.bh, xh[b], yi[b], 0, 0, 0);
        }
    }
}

// Largest |value| over the wavefront's 64 lanes.
__device__ __forceinline__ float wave_max(float m)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) m = fmaxf(m, __shfl_xor(m, o));
    return m;
}

// 2^(14 - e) for m < 2^e: maps a line's largest component to below 2^14.
__device__ __forceinline__ float fp16_scale(float m)
{
    int e = 0;
    (void)frexpf(m, &e);
    e = e < -100 ? -100 : (e > 100 ? 100 : e);
    return ldexpf(1.0f, 14 - e);
}

}  // namespace

// One workgroup per tile of C lines, C * L / 256 = 16 column blocks per stage,
// NB per wavefront (16 / NB wavefronts per workgroup).  Stage s (ls = 16^s)
// reads butterfly j's inputs x[i * nbl + j], multiplies input i by the stage
// twiddle w_{16 ls}^(i (j mod ls)) (table layout [j mod ls][i - 1], fetched
// before the LDS reads), runs the DFT-16 GEMM and writes output k to
// y[(j / ls) 16 ls + k ls + j mod ls] (Stockham order; HBM at the last stage).
template <int MODE, int L, int NB>
__global__ void __launch_bounds__(64 * 16 / NB) k_mfma_pass(const PassDesc p, const float* __restrict__ frag)
{
    constexpr int S = L == 256 ? 2 : 3;
    constexpr int nbl = L / 16;           // butterflies per line per stage
    constexpr int CPL = L / 256;          // column blocks per line
    constexpr int LP = ((L - 1) + ((L - 1) >> 4) + 1) | 1;
    constexpr int RSTRIDE = nbl + nbl / 16;  // LDS distance of inputs i and i + 1 (padded)
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cplx<float>* lds = reinterpret_cast<cplx<float>*>(smem_raw);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int cl = lane & 15, g = lane >> 4;
    const int64_t l_base = (int64_t)blockIdx.x * p.C;

    Frag<MODE> A;
    {
        float af[32];
        const float4* fp = reinterpret_cast<const float4*>(frag + lane * 32);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const float4 v = fp[q];
            af[4 * q] = v.x;
            af[4 * q + 1] = v.y;
            af[4 * q + 2] = v.z;
            af[4 * q + 3] = v.w;
        }
        A.init(af);
    }
    const float cj = (p.flags & kConjIn) ? -1.0f : 1.0f;
    // lane-ordered stage twiddles (p.rtw): stage s >= 1 at offset
    // 16 (16 + ... + 16^(s-1)), entry [c][g][q] = w_{16 ls}^((4 q + g) c)
    const float4* twq = reinterpret_cast<const float4*>(p.rtw);

    // per-block constants: line c in the tile, butterfly j in the line
    int bc[NB], bj[NB];
    bool live[NB];
    float lsc[NB];  // fp16 operands: the block's line scale
#pragma unroll
    for (int b = 0; b < NB; b++) lsc[b] = 1.0f;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int blk = wv * NB + b;
        bc[b] = blk / CPL;
        bj[b] = (blk % CPL) * 16 + cl;
        live[b] = (l_base + bc[b]) < p.nlines;
    }

#pragma unroll
    for (int s = 0; s < S; s++) {
        const int ls = s == 0 ? 1 : (s == 1 ? 16 : 256);
        float x[NB][8];
        if (s == 0) {
#pragma unroll
            for (int b = 0; b < NB; b++) {
                cplx<float> v[4] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
                if (live[b]) {
                    const int64_t ioff = line_off_m(p, (uint32_t)(l_base + bc[b])).ioff;
                    const int64_t rstep = (int64_t)nbl * p.ies;
                    // rows 4 q + g: one base for the lane group, then 4-row steps
                    const cplx<float>* src =
                        reinterpret_cast<const cplx<float>*>(p.in) + ioff + (int64_t)bj[b] * p.ies + g * rstep;
                    const int64_t step4 = 4 * rstep;
#pragma unroll
                    for (int q = 0; q < 4; q++) v[q] = src[q * step4];
                }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    x[b][2 * q] = v[q].x;
                    x[b][2 * q + 1] = v[q].y * cj;
                }
            }
            if constexpr (MODE == kMfmaF16x2) {
                // Scaled fp16 operands: every line is scaled by a power of two
                // that maps its largest component below 2^14 (exact, undone at
                // the last store); later stages' twiddles carry 1/16, so the
                // inputs of a stage stay below 2^14 * sqrt(2)^s, inside fp16.
                // A line is one block (L = 256) or the whole tile (L = 4096).
                if constexpr (CPL == 1) {
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        float m = 0.f;
#pragma unroll
                        for (int t = 0; t < 8; t++) m = fmaxf(m, fabsf(x[b][t]));
                        lsc[b] = fp16_scale(wave_max(m));
                    }
                } else {
                    float m = 0.f;
#pragma unroll
                    for (int b = 0; b < NB; b++)
#pragma unroll
                        for (int t = 0; t < 8; t++) m = fmaxf(m, fabsf(x[b][t]));
                    m = wave_max(m);
                    float* slot = reinterpret_cast<float*>(lds + p.C * LP);  // past the tile
                    if (lane == 0) slot[wv] = m;
                    mfma_barrier();
                    for (int w2 = 0; w2 < 16 / NB; w2++) m = fmaxf(m, slot[w2]);
                    const float sc = fp16_scale(m);
#pragma unroll
                    for (int b = 0; b < NB; b++) lsc[b] = sc;
                }
#pragma unroll
                for (int b = 0; b < NB; b++)
#pragma unroll
                    for (int t = 0; t < 8; t++) x[b][t] *= lsc[b];
            }
        } else {
            // twiddles first: w_{16 ls}^(i cc) for the lane's inputs i = 4 q + g,
            // 32 contiguous bytes per block (one set for all blocks at stage
            // 1, where cc = j mod 16 is the lane's column)
            cplx<float> w[NB][4];
            // stage offset in rows of 16 entries: stage 1 has 16 rows (c < 16),
            // stage 2 follows it
            const int soff = s == 1 ? 0 : 16;
#pragma unroll
            for (int b = 0; b < NB; b++) {
                if (s == 1 && b > 0) {
#pragma unroll
                    for (int q = 0; q < 4; q++) w[b][q] = w[0][q];
                    continue;
                }
                const int cc = bj[b] & (ls - 1);
                const float4* tq = twq + ((soff + cc) * 16 + 4 * g) / 2;
                const float4 a0 = tq[0], a1 = tq[1];
                w[b][0] = {a0.x, a0.y};
                w[b][1] = {a0.z, a0.w};
                w[b][2] = {a1.x, a1.y};
                w[b][3] = {a1.z, a1.w};
            }
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const cplx<float>* line = lds + bc[b] * LP + bj[b] + (bj[b] >> 4);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const cplx<float> v = cmul(line[rho(g, q) * RSTRIDE], w[b][q]);
                    x[b][2 * q] = v.x;
                    x[b][2 * q + 1] = v.y;
                }
            }
            mfma_barrier();  // every read of this stage done before its writes
        }
        f32x4 yr[NB], yi[NB];
        dft16_blocks<MODE, NB>(x, A, yr, yi);
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int j = bj[b];
            if (s < S - 1) {
                // output k = 4 g + r to y[gg 16 ls + k ls + cc]; padded index
                // pos + pos / 16 is linear in r
                const int gg = j / ls, cc = j & (ls - 1);
                const int pos0 = gg * 16 * ls + 4 * g * ls + cc;
                cplx<float>* line = lds + bc[b] * LP;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int pos = pos0 + r * ls;
                    line[pos + (pos >> 4)] = cplx<float>{yr[b][r], yi[b][r]};
                }
            } else if (live[b]) {
                const int64_t ooff = line_off_m(p, (uint32_t)(l_base + bc[b])).ooff;
                const int64_t rstep = (int64_t)nbl * p.oes;
                // outputs 4 g + r (last stage: gg = 0, cc = j)
                cplx<float>* dst =
                    reinterpret_cast<cplx<float>*>(p.out) + ooff + (int64_t)j * p.oes + (int64_t)(4 * g) * rstep;
                // undo the fp16 line scale and the 1/16 of each later stage
                const float unscale = MODE == kMfmaF16x2 ? (S == 3 ? 256.0f : 16.0f) / lsc[b] : 1.0f;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    cplx<float> y = {yr[b][r], yi[b][r]};
                    if constexpr (MODE == kMfmaF16x2) y = cscale(y, unscale);
                    if (p.flags & kScale) y = cscale(y, p.scale_f);
                    if (p.flags & kConjOut) y.y = -y.y;
                    dst[r * rstep] = y;
                }
            }
        }
        if (s < S - 1) mfma_barrier();
    }
}

template <int MODE, int NB>
static const void* mfma_sym(int L)
{
    return L == 256 ? (const void*)&k_mfma_pass<MODE, 256, NB> : (const void*)&k_mfma_pass<MODE, 4096, NB>;
}

const void* kernel_symbol_mfma(int mode, int L, int nb)
{
    if (nb == 4) {
        switch (mode) {
            case kMfmaF32: return mfma_sym<kMfmaF32, 4>(L);
            case kMfmaBf16x3: return mfma_sym<kMfmaBf16x3, 4>(L);
            case kMfmaF16x2: return mfma_sym<kMfmaF16x2, 4>(L);
            default: return nullptr;
        }
    }
    switch (mode) {
        case kMfmaF32: return mfma_sym<kMfmaF32, 2>(L);
        case kMfmaBf16x3: return mfma_sym<kMfmaBf16x3, 2>(L);
        case kMfmaF16x2: return mfma_sym<kMfmaF16x2, 2>(L);
        default: return nullptr;
    }
}

}  // namespace fftz
