// Pass descriptor: one kernel launch that performs length-L FFTs on a set of
// "lines" (1-D sub-sequences of a strided tensor), entirely inside LDS.
//
// A pass reads a tile of C lines from HBM, runs every radix stage of the
// length-L transform in LDS / VGPRs, applies the fused element-wise ops
// (conjugation for the backward direction, the four-step inter-pass twiddle,
// Bluestein chirp and spectrum products, real-packed recombine / separate) and
// writes the tile back.  Every plan is a short list of passes; the reference's
// solver tree (core/solvers/**) maps onto it as:
//   pow2-iterative / direct / batched / CT-L1-direct  -> 1 pass per axis
//   four-step (+ fused twiddle transpose)             -> 2-3 passes, twiddle
//                                                       fused into the column
//                                                       pass, transpose fused
//                                                       into the row pass
//   Bluestein                                          -> 1 fused pass
//   N-D (row-column)                                   -> 1 pass per axis
//
// Shared by host (plan.cpp) and device (kernels_body.h etc.): plain POD only.
#pragma once

#include <stdint.h>

namespace fftz {

constexpr int kMaxStages = 12;

enum PassFlags : uint32_t {
    kConjIn = 1u << 0,     // conjugate after load (backward direction)
    kConjOut = 1u << 1,    // conjugate before store (backward direction)
    kTw4 = 1u << 2,        // multiply element k of line l by w_N^(l0 * k)
    kInCol = 1u << 3,      // tile load: lanes walk adjacent lines (column pass)
    kOutCol = 1u << 4,     // tile store: lanes walk adjacent lines
    kScale = 1u << 5,      // multiply output by a real scale
    kXcdRemap = 1u << 6,   // map adjacent tiles onto the same XCD (shared L2)
    kR2CPost = 1u << 7,    // real-packed recombine after the L-point FFT
    kC2RPre = 1u << 8,     // real-packed separate before the L-point FFT
    kRealIn = 1u << 9,     // R2C: input is real, read as L complex pairs
    kRealOut = 1u << 10,   // C2R: output is real, written as L complex pairs
    kRealPromote = 1u << 11,  // R2C (any length): real element j loaded as (x[j], 0)
    kHermIn = 1u << 12,       // C2R (any length): bins 0..n/2 stored, bin k > n/2 is conj(bin n-k)
    kRealPart = 1u << 13,     // C2R (any length): store only the real part, as real elements
    // Bluestein over several passes (transform length M above one LDS tile):
    // element-wise products fused into the first / last pass of the two
    // M-point four-step transforms.  g is the element's index in the whole
    // M-point sequence (gin / gout below).
    kChirpIn = 1u << 14,   // load: x[g] * chirp[g] for g < nchirp, 0 beyond
    kSpecOut = 1u << 15,   // store: conj(y[g]) * spec[g]
    kChirpOut = 1u << 16,  // store: conj(chirp[g]) * y[g], only g < nchirp
};
// Flags of the multi-pass Bluestein products (generic kernel only).
constexpr uint32_t kChirpFlags = kChirpIn | kSpecOut | kChirpOut;
// Flags whose passes stage the tile through LDS before the first stage.
constexpr uint32_t kStagedFlags = kRealIn | kC2RPre | kR2CPost | kRealPromote | kHermIn | kRealPart;

// Operand precision of the matrix-core radix-16 pass (kernels_mfma.hip).
enum MfmaMode : int {
    kMfmaNone = 0,    // plain FP32 / FP64 Stockham codelets on the vector ALUs
    kMfmaF32 = 1,     // DFT-16 GEMM with FP32 operands (v_mfma_f32_16x16x4_f32)
    kMfmaBf16x3 = 2,  // bf16 hi + lo split, three products, FP32 accumulate
    kMfmaF16x2 = 3,   // per-column power-of-two scale, fp16 hi + lo, three products
};

// n / d and n % d for 0 <= n < 2^31 with one mul_hi (Granlund-Montgomery).
struct FastDiv {
    uint32_t mul;
    uint32_t shift;
    uint32_t d;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((1ull << l) < d) l++;
    f.shift = l;
    f.mul = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    if (d == 1) { f.mul = 0; f.shift = 0; }
    return f;
}

struct PassDesc {
    const void* in;
    void* out;
    const void* stw;       // stage twiddles, plan precision
    const void* tw4;       // four-step twiddles: double2 lo[1<<tw4_lobits] then hi[]
    const void* chirp;     // Bluestein: conj(B_k), k < nvalid (plan precision)
    const void* spec;      // Bluestein: conj(FFT_M(B)) / M (plan precision)
    const void* rtw;       // real-packed twiddles exp(-2 pi i k / 2L), k <= L/2;
                           // matrix-core pass: its lane-ordered stage twiddles
    int64_t ies, oes;      // element strides along the transform axis
    int64_t ils[3], ols[3];// line strides for line levels l0, l1, l2
    int64_t nlines;        // total lines
    // index of an element in the whole sequence of a multi-pass transform:
    // g = gbase + l0 * g[0] + l1 * g[1] + l2 * g[2] + e * g[3]
    // (input side gin, output side gout; only read with kChirpFlags)
    int64_t gin[4], gout[4];
    int64_t gin_base, gout_base;
    int64_t nchirp;        // Bluestein: n (valid input / output elements)
    int64_t tw4_n;         // four-step twiddle base N
    FastDiv ln0, ln1;      // line level counts l0, l1 (l2 = rest)
    // Ragged column tiles (a row of ln0 lines that C does not divide): the
    // tile index times C is a virtual line index v; row = v / vrow.d with
    // vrow.d = tiles per row x C, and slot v % vrow.d is line l0 of that row
    // when below ln0 (idle otherwise).  vrow.d == 0: flat tiles of C lines.
    FastDiv vrow;
    int L;                 // FFT length of this pass
    int LP;                // padded LDS pitch of one line (complex elements)
    int C;                 // lines per tile (power of 2)
    int TL;                // threads per line (power of 2)
    int tl_shift;          // log2(TL)
    int c_shift;           // log2(C)
    int nvalid_in;         // elements loaded per line (< L: zero padded)
    int nvalid_out;        // elements stored per line
    int nst;
    int tw4_lobits;
    uint32_t flags;
    float scale_f;
    double scale_d;
    int radix[kMaxStages];
    int ls[kMaxStages];
    FastDiv lsdiv[kMaxStages];
    int twoff[kMaxStages];
};

// One tile of an uneven batch: `nlines` consecutive transforms of size class
// `desc`, `dist` elements apart, starting at in_off / out_off (elements).
struct BatchTile {
    uint32_t desc;
    uint32_t nlines;
    uint32_t bluestein;
    uint32_t pad_;
    int64_t in_off, out_off;
    int64_t dist;
};

// Bytes at the start of the batch kernel's dynamic LDS that hold the tile
// index (rounded to 256 so the tile data stays aligned).
constexpr int kBatchHdrBytes = 256;

struct BatchArgs {
    const PassDesc* descs;
    const BatchTile* tiles;
    int ntiles;
    unsigned* counter;
    const void* in;
    void* out;
};

// LDS index of element e inside one line: one pad slot every 16 elements so
// the stride-R writes of the first Stockham stages spread over the banks.
__host__ __device__ inline int lds_pad(int e) { return e + (e >> 4); }

}  // namespace fftz
