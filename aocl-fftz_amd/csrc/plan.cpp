// GPU planner and executor.
//
// The reference's fixed-mode selector (selector/selector.c:509-789) walks a
// gate list -- pow2-iterative, four-step, batched CT-L1 direct, batched,
// N-D, Bluestein, direct, CT -- and builds a tree of CPU solvers.  On MI355X
// the same decisions collapse into a few shapes of LDS pass:
//
//   axis length n with all prime factors <= 13 and n <= single-pass limit
//       -> one pass along that axis (Stockham stages in LDS)
//   1-D n above the limit
//       -> four-step over p = 2 or 3 factors: column passes with the inter-pass
//          twiddle fused into their store (pow2_fourstep_solver_dft.c:700-763,
//          fused_twiddle_transpose_c.c:20-110), then a row pass whose store is
//          the transpose
//   axis length with a prime factor > 13
//       -> one fused Bluestein pass (bluestein_solver_dft.c:137-234)
//   R2C / C2R with even innermost n
//       -> real-packed pass with recombine / separate fused
//          (packed_solver_rdft.c:154-186)
//   N-D -> one pass per axis (ndim_solver_dft.c:96-121), batch = line axes.
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <thread>
#include <atomic>

namespace fftz {

// ---------------------------------------------------------------------------
// Device info
// ---------------------------------------------------------------------------
const DeviceInfo& device_info()
{
    static DeviceInfo info;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return;
        info.cus = prop.multiProcessorCount;
        info.lds_per_block = prop.sharedMemPerBlock;
        info.warp = prop.warpSize;
        // Allow every kernel to use the full LDS of a workgroup.
        for (int prec : {4, 8})
            for (bool all : {false, true})
                for (KernelKind k : {KernelKind::Pass, KernelKind::Bluestein})
                    for (int e : {8, 16})
                        for (int t : {256, 1024})
                            (void)hipFuncSetAttribute(kernel_symbol(k, prec, all, e, t),
                                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                                      (int)info.lds_per_block);
        for (int prec : {4, 8})
            for (int L = 64; L <= 4096; L *= 2)
                for (int t : {256, 512})
                    if (const void* f = kernel_symbol_fixed(prec, L, t))
                        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  (int)info.lds_per_block);
        for (int m : {kMfmaF32, kMfmaBf16x3, kMfmaF16x2})
            for (int L : {256, 4096})
                for (int nb : {2, 4})
                    (void)hipFuncSetAttribute(kernel_symbol_mfma(m, L, nb), hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)info.lds_per_block);
    });
    return info;
}

// ---------------------------------------------------------------------------
// Roots of unity in double: exp(-2 pi i p / n), exact integer reduction and a
// fold onto [0, pi/4] before sin/cos.
// ---------------------------------------------------------------------------
struct d2 {
    double x, y;
};

static d2 root_fwd(int64_t p, int64_t n)
{
    p %= n;
    if (p < 0) p += n;
    const __int128 p8 = (__int128)p * 8;
    const int64_t oct = (int64_t)(p8 / n);         // octant 0..7
    const int64_t rem = (int64_t)(p8 - (__int128)oct * n);  // angle = (pi/4)(oct + rem/n)
    double c, s;
    const double a = (M_PI / 4.0) * ((double)rem / (double)n);
    const double b = (M_PI / 4.0) * ((double)(n - rem) / (double)n);
    // base angle theta = (pi/4) * (oct + rem/n); use the complementary
    // angle on odd octants so the argument stays <= pi/4
    double ct, st;  // cos, sin of theta
    switch (oct) {
        case 0: ct = cos(a); st = sin(a); break;
        case 1: ct = sin(b); st = cos(b); break;
        case 2: ct = -sin(a); st = cos(a); break;
        case 3: ct = -cos(b); st = sin(b); break;
        case 4: ct = -cos(a); st = -sin(a); break;
        case 5: ct = -sin(b); st = -cos(b); break;
        case 6: ct = sin(a); st = -cos(a); break;
        default: ct = cos(b); st = -sin(b); break;
    }
    c = ct;
    s = st;
    return {c, -s};
}

// ---------------------------------------------------------------------------
// Host FP64 FFT (recursive mixed radix) used only at setup, for the Bluestein
// spectrum table.  The roots of the top-level length are tabulated once
// (root_fwd, exact integer reduction); a sub-transform of length m reads
// root(j, m) as tab[j * (n / m)], which is the same double as root_fwd(j, m)
// (both reduce to the same correctly rounded fraction of the circle).
// ---------------------------------------------------------------------------
static void host_fft_rec(int64_t n, const d2* in, int64_t is, d2* out, const d2* tab, int64_t tstride)
{
    if (n == 1) { out[0] = in[0]; return; }
    int64_t r = 0;
    for (int64_t c : {4, 2, 3, 5, 7, 11, 13})
        if (n % c == 0) { r = c; break; }
    if (r == 0) {
        for (int64_t k = 0; k < n; k++) {
            d2 s = {0, 0};
            for (int64_t j = 0; j < n; j++) {
                const d2 w = tab[((j * k) % n) * tstride];
                const d2 x = in[j * is];
                s.x += x.x * w.x - x.y * w.y;
                s.y += x.x * w.y + x.y * w.x;
            }
            out[k] = s;
        }
        return;
    }
    const int64_t m = n / r;
    std::vector<d2> y((size_t)n);
    for (int64_t q = 0; q < r; q++) host_fft_rec(m, in + q * is, is * r, y.data() + q * m, tab, tstride * r);
    for (int64_t k1 = 0; k1 < m; k1++)
        for (int64_t k2 = 0; k2 < r; k2++) {
            d2 s = {0, 0};
            for (int64_t q = 0; q < r; q++) {
                const d2 w = tab[((q * (k1 + m * k2)) % n) * tstride];
                const d2 x = y[(size_t)(q * m + k1)];
                s.x += x.x * w.x - x.y * w.y;
                s.y += x.x * w.y + x.y * w.x;
            }
            out[k1 + m * k2] = s;
        }
}

static void host_fft(int64_t n, const d2* in, d2* out)
{
    std::vector<d2> tab((size_t)n);
    for (int64_t j = 0; j < n; j++) tab[(size_t)j] = root_fwd(j, n);
    host_fft_rec(n, in, 1, out, tab.data(), 1);
}

// ---------------------------------------------------------------------------
// Factorisation helpers
// ---------------------------------------------------------------------------
static const int kRadices[] = {16, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2};

static bool smooth13(int64_t n)
{
    for (int64_t p : {2, 3, 5, 7, 11, 13})
        while (n % p == 0) n /= p;
    return n == 1;
}

// Fewest radix stages from kRadices (largest first on ties).
static bool factor_radices(int L, std::vector<int>& out)
{
    out.clear();
    if (L == 1) return true;
    std::map<int, std::pair<int, int>> memo;
    std::function<int(int)> solve = [&](int n) -> int {
        if (n == 1) return 0;
        auto it = memo.find(n);
        if (it != memo.end()) return it->second.first;
        int best = 1 << 30, br = 0;
        for (int r : kRadices) {
            if (n % r) continue;
            int c = solve(n / r);
            if (c + 1 < best) { best = c + 1; br = r; }
        }
        memo[n] = {best, br};
        return best;
    };
    if (solve(L) >= (1 << 30)) return false;
    for (int n = L; n > 1;) {
        int r = memo[n].second;
        out.push_back(r);
        n /= r;
    }
    std::sort(out.begin(), out.end(), std::greater<int>());
    return (int)out.size() <= kMaxStages;
}

// A radix order the kernels can run: stages from kRadices whose product is L.
static bool radix_order_ok(int L, const std::vector<int>& r)
{
    if (r.empty() || (int)r.size() > kMaxStages) return false;
    int64_t prod = 1;
    for (int v : r) {
        if (std::find(std::begin(kRadices), std::end(kRadices), v) == std::end(kRadices)) return false;
        prod *= v;
    }
    return prod == L;
}

std::vector<std::vector<int>> radix_orders(int L, int maxn)
{
    // Multisets of radices with product L and at most one stage more than
    // the fewest; each in descending and ascending order.  The default order
    // (factor_radices) comes first.
    std::vector<std::vector<int>> out;
    std::vector<int> dflt;
    if (!factor_radices(L, dflt) || dflt.size() < 2) return out;
    out.push_back(dflt);
    const size_t smax = dflt.size() + 1;
    std::vector<std::vector<int>> sets;
    std::vector<int> cur;
    std::function<void(int64_t, size_t)> rec = [&](int64_t n, size_t from) {
        if (n == 1) {
            if (cur.size() >= 2) sets.push_back(cur);
            return;
        }
        if (cur.size() >= smax) return;
        for (size_t i = from; i < sizeof(kRadices) / sizeof(kRadices[0]); i++) {
            const int r = kRadices[i];
            if (n % r) continue;
            cur.push_back(r);
            rec(n / r, i);
            cur.pop_back();
        }
    };
    rec(L, 0);
    std::stable_sort(sets.begin(), sets.end(),
                     [](const std::vector<int>& a, const std::vector<int>& b) { return a.size() < b.size(); });
    for (const auto& m : sets) {
        std::vector<int> d = m, a = m;
        std::sort(d.begin(), d.end(), std::greater<int>());
        std::sort(a.begin(), a.end());
        for (const auto& o : {d, a})
            if ((int)out.size() < maxn && std::find(out.begin(), out.end(), o) == out.end()) out.push_back(o);
    }
    return out;
}

static int next_pow2(int64_t v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

static int ilog2(int v)
{
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

// ---------------------------------------------------------------------------
// Tables
// ---------------------------------------------------------------------------
struct TableBuilder {
    std::vector<unsigned char>& buf;
    int64_t add(const void* data, size_t bytes)
    {
        size_t off = (buf.size() + 255) & ~(size_t)255;
        buf.resize(off + bytes);
        memcpy(buf.data() + off, data, bytes);
        return (int64_t)off;
    }
    int64_t add_complex(const std::vector<d2>& v, int prec)
    {
        if (prec == 8) return add(v.data(), v.size() * sizeof(d2));
        std::vector<float> f(v.size() * 2);
        for (size_t i = 0; i < v.size(); i++) {
            f[2 * i] = (float)v[i].x;
            f[2 * i + 1] = (float)v[i].y;
        }
        return add(f.data(), f.size() * sizeof(float));
    }
};

// Stage twiddles for radices r[0..]: per stage, for c < ls, i in [1, R):
// w_{R ls}^{i c} -- the per-column layout of core/common/twiddle.c:46-66.
static int64_t stage_twiddles(TableBuilder& tb, int prec, const std::vector<int>& radix, int* twoff, int* ls_out)
{
    std::vector<d2> tw;
    int ls = 1;
    for (size_t s = 0; s < radix.size(); s++) {
        const int R = radix[s];
        twoff[s] = (int)tw.size();
        ls_out[s] = ls;
        for (int c = 0; c < ls; c++)
            for (int i = 1; i < R; i++) tw.push_back(root_fwd((int64_t)i * c, (int64_t)R * ls));
        ls *= R;
    }
    if (tw.empty()) tw.push_back({1.0, 0.0});
    return tb.add_complex(tw, prec);
}

// Two-level FP64 table for w_N^p, p < N: lo[q] = w_N^q (q < 2^b), hi[h] = w_N^(h 2^b).
static int64_t fourstep_twiddles(TableBuilder& tb, int64_t N, int* lobits)
{
    int b = 0;
    while (((int64_t)1 << (2 * b)) < N) b++;
    const int64_t lo_n = (int64_t)1 << b;
    const int64_t hi_n = (N + lo_n - 1) / lo_n;
    std::vector<d2> t((size_t)(lo_n + hi_n));
    for (int64_t q = 0; q < lo_n; q++) t[(size_t)q] = root_fwd(q, N);
    for (int64_t h = 0; h < hi_n; h++) t[(size_t)(lo_n + h)] = root_fwd(h * lo_n, N);
    *lobits = b;
    return tb.add(t.data(), t.size() * sizeof(d2));
}

// ---------------------------------------------------------------------------
// Pass construction
// ---------------------------------------------------------------------------
struct LineAxis {
    int64_t n, is, os;
    // coefficients of this axis's index in the sequence index g of the
    // input / output side (multi-pass Bluestein products; 0 for batch axes)
    int64_t gi = 0, go = 0;
};

struct PassReq {
    KernelKind kind = KernelKind::Pass;
    int L = 1;                 // transform length (M for Bluestein)
    int nvalid_in = -1, nvalid_out = -1;
    int64_t ies = 1, oes = 1;
    std::vector<LineAxis> lines;
    uint32_t flags = 0;
    bool want_in_col = false, want_out_col = false;
    Role in_role = Role::In, out_role = Role::Out;
    int64_t tw4_n = 0;
    int64_t chirp_off = -1, spec_off = -1, rtw_off = -1;
    int extra_lds_elems = 0;   // C2R needs L+1 slots
    int64_t in_base = 0, out_base = 0;  // byte offsets added to the role pointers
    int64_t gie = 0, goe = 0;  // element coefficients of the sequence index (kChirpFlags)
    int64_t nchirp = 0;
};

static int env_int(const char* name, int dflt)
{
    const char* v = getenv(name);
    return v ? atoi(v) : dflt;
}

// Merge line axes that form one dense run for both input and output.
static std::vector<LineAxis> merge_axes(std::vector<LineAxis> ax)
{
    ax.erase(std::remove_if(ax.begin(), ax.end(), [](const LineAxis& a) { return a.n == 1; }), ax.end());
    bool changed = true;
    while (changed && ax.size() > 1) {
        changed = false;
        for (size_t i = 0; i < ax.size() && !changed; i++)
            for (size_t j = 0; j < ax.size() && !changed; j++) {
                if (i == j) continue;
                if (ax[j].is == ax[i].is * ax[i].n && ax[j].os == ax[i].os * ax[i].n &&
                    ax[j].gi == ax[i].gi * ax[i].n && ax[j].go == ax[i].go * ax[i].n) {
                    ax[i].n *= ax[j].n;
                    ax.erase(ax.begin() + (long)j);
                    changed = true;
                }
            }
    }
    return ax;
}

// Round x to `bits` significant bits, round-to-nearest-even, with a floor
// on the quantum exponent (subnormal spacing): bf16 = (8, -133), fp16 = (11, -24).
static double round_sig(double x, int bits, int qmin)
{
    if (x == 0.0) return 0.0;
    int e = 0;
    (void)std::frexp(x, &e);  // |x| in [2^(e-1), 2^e)
    const int q = std::max(e - bits, qmin);
    return std::ldexp(std::nearbyint(std::ldexp(x, -q)), q);
}

// Per-lane DFT-16 operand fragments for the matrix-core pass.  The real
// 32 x 32 operator is A = [[Fr, -Fi], [Fi, Fr]] with F[k][j] = w_16^(jk)
// (forward); rows 0-15 ("top") produce Re Y, rows 16-31 ("bottom") Im Y;
// column k < 16 multiplies Re x[k], column 16 + k multiplies Im x[k].
// The kernel permutes K so that lane group g = l / 16 supplies, as its
// element t < 8, part t & 1 (re / im) of input row rho = 4 (t >> 1) + g.
// Lane l (output row r = l % 16), 32 floats:
//   kMfmaF32:  [0..7] top[r][K(g, t)], [8..15] bottom[r][K(g, t)];
//   split:     [0..7] hi(top), [8..15] lo(top), [16..23] hi(bottom),
//              [24..31] lo(bottom); hi = round(a), lo = round(a - hi) in bf16
//              (kMfmaBf16x3) or fp16 (kMfmaF16x2), stored as the exactly
//              representable float values.
static int64_t mfma_fragments(TableBuilder& tb, int mode)
{
    std::vector<float> f(64 * 32, 0.0f);
    auto A = [](int half, int r, int k) -> double {
        const d2 w = root_fwd((int64_t)r * (k & 15), 16);  // (Fr, Fi)
        const bool im = k >= 16;
        if (half == 0) return im ? -w.y : w.x;
        return im ? w.x : w.y;
    };
    const int bits = mode == kMfmaBf16x3 ? 8 : 11;
    const int qmin = mode == kMfmaBf16x3 ? -133 : -24;
    for (int l = 0; l < 64; l++) {
        const int r = l & 15, g = l >> 4;
        float* o = &f[(size_t)l * 32];
        auto kidx = [g](int t) { return 16 * (t & 1) + 4 * (t >> 1) + g; };
        if (mode == kMfmaF32) {
            for (int t = 0; t < 8; t++) {
                o[t] = (float)A(0, r, kidx(t));
                o[8 + t] = (float)A(1, r, kidx(t));
            }
            continue;
        }
        for (int t = 0; t < 8; t++)
            for (int half = 0; half < 2; half++) {
                const double a = A(half, r, kidx(t));
                const double hi = round_sig(a, bits, qmin);
                const double lo = round_sig(a - hi, bits, qmin);
                o[16 * half + t] = (float)hi;
                o[16 * half + 8 + t] = (float)lo;
            }
    }
    return tb.add(f.data(), f.size() * sizeof(float));
}

// Stage twiddles of the matrix-core pass in the order its lanes read them:
// for stage s >= 1 (ls = 16^s) and c = j mod ls, 16 entries [g][q] =
// w_{16 ls}^((4 q + g) c) -- lane group g's four inputs i = 4 q + g of
// butterfly j, entry i = 0 being 1 -- stage after stage (FP64 roots rounded
// to FP32).  A lane reads its four twiddles of a block as 32 contiguous bytes.
// With scaled fp16 operands the entries carry an exact factor 1/16, which
// keeps each stage's inputs inside the fp16 range (the kernel undoes it at
// the last store).
static int64_t mfma_stage_twiddles(TableBuilder& tb, int L, int mode)
{
    std::vector<d2> t;
    const int S = L == 256 ? 2 : 3;
    const double f = mode == kMfmaF16x2 ? 1.0 / 16.0 : 1.0;
    for (int s = 1, ls = 16; s < S; s++, ls *= 16)
        for (int c = 0; c < ls; c++)
            for (int g = 0; g < 4; g++)
                for (int q = 0; q < 4; q++) {
                    const d2 w = root_fwd((int64_t)(4 * q + g) * c, 16 * ls);
                    t.push_back({w.x * f, w.y * f});
                }
    return tb.add_complex(t, 4);
}

static bool build_pass(Plan* plan, TableBuilder& tb, const PassReq& q, std::string& err)
{
    const int prec = plan->spec.prec_bytes;
    const int esize = 2 * prec;
    const DeviceInfo& dev = device_info();
    Launch ln;
    ln.kind = q.kind;
    PassDesc& d = ln.d;
    d.L = q.L;
    d.nvalid_in = q.nvalid_in < 0 ? q.L : q.nvalid_in;
    d.nvalid_out = q.nvalid_out < 0 ? q.L : q.nvalid_out;
    d.ies = q.ies;
    d.oes = q.oes;
    d.flags = q.flags;
    d.scale_f = 1.0f;
    d.scale_d = 1.0;

    std::vector<int> radix;
    if (!factor_radices(q.L, radix)) {
        err = "length " + std::to_string(q.L) + " has no radix plan";
        return false;
    }
    // a tuned radix order for this length (plain passes; the matrix-core
    // pass needs the all-16 plan)
    if (const std::vector<int>* tr = plan->spec.tune.radices(q.L))
        if (q.kind == KernelKind::Pass && plan->spec.mfma_mode == kMfmaNone && radix_order_ok(q.L, *tr)) radix = *tr;
    d.nst = (int)radix.size();
    ln.all_radix = false;
    for (size_t s = 0; s < radix.size(); s++) {
        d.radix[s] = radix[s];
        if (radix[s] != 2 && radix[s] != 4 && radix[s] != 8 && radix[s] != 16) ln.all_radix = true;
    }
    ln.stw_off = stage_twiddles(tb, prec, radix, d.twoff, d.ls);
    for (int s = 0; s < d.nst; s++) d.lsdiv[s] = make_fastdiv((uint32_t)d.ls[s]);

    // elements per thread: 16 for power-of-two radices; 8 for the mixed
    // radix variant (fewer live VGPRs) unless that needs > 512 threads/line
    int E = 16;
    if (ln.all_radix && next_pow2((q.L + 7) / 8) <= 512) E = 8;
    // complex128 outside the fixed-kernel range: 8 values per thread keeps
    // the 1024-thread variant within its 128 VGPRs (16 would spill)
    const bool fixed_range = q.L >= 64 && q.L <= 4096 && (q.L & (q.L - 1)) == 0;
    if (prec == 8 && !fixed_range && next_pow2((q.L + 7) / 8) <= 1024) E = 8;
    {
        const int e = plan->spec.tune.elems ? plan->spec.tune.elems : env_int("FFTZ_E", 0);
        if ((e == 8 || e == 16) && next_pow2((q.L + e - 1) / e) <= 1024) E = e;
    }
    ln.E = E;
    // threads per line
    int TL = next_pow2((q.L + E - 1) / E);
    if (TL < 1) TL = 1;
    d.TL = TL;
    d.tl_shift = ilog2(TL);
    int need = q.L + q.extra_lds_elems;
    int LP = lds_pad(need - 1) + 1;
    if ((LP & 1) == 0) LP++;
    d.LP = LP;

    // fixed-length kernel: power-of-two L in [64, 4096], plain C2C pass
    const bool pow2L = q.L >= 64 && q.L <= 4096 && (q.L & (q.L - 1)) == 0;
    if ((q.flags & kChirpFlags) && (q.kind != KernelKind::Pass || (q.flags & (kStagedFlags | kRealOut)))) {
        err = "internal: Bluestein products on a staged pass";
        return false;
    }
    const bool fixed_ok = q.kind == KernelKind::Pass && pow2L && E == 16 && d.nvalid_in == q.L &&
                          d.nvalid_out == q.L &&
                          !(q.flags & (kStagedFlags | kRealOut | kChirpFlags)) && env_int("FFTZ_FIXED", 1);
    bool plan_match = fixed_ok;
    if (fixed_ok) {
        // the fixed kernels hard-code 16, 16, ..., 2^rem (kernels_fixed.hip)
        const int lg = ilog2(q.L), S = (lg + 3) / 4, rem = lg - 4 * (S - 1);
        plan_match = (int)radix.size() == S;
        for (int s = 0; plan_match && s < S; s++)
            plan_match = radix[s] == (s < S - 1 ? 16 : (1 << rem));
    }
    // fixed kernels come in 256- and 512-thread variants (512 for wide column
    // tiles of L >= 256)
    const int tmax = plan_match ? (q.L >= 256 ? 512 : 256) : 1024;
    // matrix-core pass: complex64, L = 16^2 or 16^3 with the plain 16-radix
    // plan, row mapping, no fused twiddle / real-packed step
    const int mm = plan->spec.mfma_mode;
    const bool use_mfma = mm != kMfmaNone && prec == 4 && plan_match && (q.L == 256 || q.L == 4096) &&
                          !q.want_in_col && !q.want_out_col &&
                          !(q.flags & (kTw4 | kStagedFlags | kRealOut | kChirpFlags));

    // line axes: choose l0
    std::vector<LineAxis> ax = merge_axes(q.lines);
    if (ax.empty()) ax.push_back({1, 0, 0});
    size_t i0 = 0;
    bool in_col = false, out_col = false;
    if (q.want_in_col || q.want_out_col) {
        for (size_t i = 0; i < ax.size(); i++) {
            if ((q.want_in_col && ax[i].is == 1) || (!q.want_in_col && q.want_out_col && ax[i].os == 1)) {
                i0 = i;
                break;
            }
        }
    } else {
        for (size_t i = 1; i < ax.size(); i++)
            if (std::llabs(ax[i].is) < std::llabs(ax[i0].is)) i0 = i;
    }
    std::swap(ax[0], ax[i0]);
    // remaining axes sorted by input stride so l1 is the inner one
    std::sort(ax.begin() + 1, ax.end(), [](const LineAxis& a, const LineAxis& b) { return a.is < b.is; });
    int64_t n0 = ax[0].n;
    in_col = q.want_in_col && ax[0].is == 1 && q.ies != 1;
    out_col = q.want_out_col && ax[0].os == 1 && q.oes != 1;

    // lines per tile
    const size_t lds_cap = dev.lds_per_block;
    const size_t line_bytes = (size_t)LP * esize;
    if (line_bytes > lds_cap) {
        err = "pass length " + std::to_string(q.L) + " exceeds LDS";
        return false;
    }
    int C = 1;
    if (use_mfma) {
        in_col = out_col = false;
        C = q.L == 4096 ? 1 : 16;  // 16 column blocks of 16 butterflies per tile
        TL = 16;
        d.TL = TL;
        d.tl_shift = 4;
    } else if (in_col || out_col) {
        const int cmax = plan->spec.tune.col_c ? plan->spec.tune.col_c : env_int("FFTZ_COL_C", prec == 8 ? 16 : 32);
        const size_t budget = std::min(lds_cap, (size_t)env_int("FFTZ_COL_LDS", 96 * 1024));
        // a row of n0 lines that C does not divide is tiled raggedly (the
        // last tile of each row partly idle) rather than falling back to
        // element-strided row tiles
        while (C * 2 <= cmax && (size_t)(C * 2) * line_bytes <= budget && C * 2 <= n0 && C * 2 * TL <= tmax)
            C *= 2;
        if (C < 2) { in_col = false; out_col = false; C = 1; }
    }
    if (!(in_col || out_col) && !use_mfma) {
        const int tgt = plan->spec.tune.row_t ? plan->spec.tune.row_t : env_int("FFTZ_ROW_T", 256);
        const size_t budget = std::min(lds_cap, (size_t)env_int("FFTZ_ROW_LDS", 64 * 1024));
        int64_t nl_total = 1;
        for (auto& a : ax) nl_total *= a.n;
        while (C * TL < tgt && (size_t)(C * 2) * line_bytes <= budget && C * 2 * TL <= tmax && C < nl_total) C *= 2;
    }
    if (C * TL > 1024) {
        err = "tile too large";
        return false;
    }
    d.C = C;
    d.c_shift = ilog2(C);
    if (in_col) d.flags |= kInCol;
    if (out_col) d.flags |= kOutCol;

    // line levels: l0, l1, l2 (+ host-looped outer axes)
    std::vector<LineAxis> host_axes;
    while (ax.size() > 3) {
        host_axes.push_back(ax.back());
        ax.pop_back();
    }
    while (ax.size() < 3) ax.push_back({1, 0, 0});
    int64_t nl = ax[0].n * ax[1].n * ax[2].n;
    if (nl >= ((int64_t)1 << 31)) {
        err = "too many lines in one pass";
        return false;
    }
    d.nlines = nl;
    d.ln0 = make_fastdiv((uint32_t)ax[0].n);
    d.ln1 = make_fastdiv((uint32_t)ax[1].n);
    for (int k = 0; k < 3; k++) {
        d.ils[k] = ax[k].is;
        d.ols[k] = ax[k].os;
        d.gin[k] = ax[k].gi;
        d.gout[k] = ax[k].go;
    }
    d.gin[3] = q.gie;
    d.gout[3] = q.goe;
    d.nchirp = q.nchirp;
    if (q.flags & kTw4) {
        d.tw4_n = q.tw4_n;
        ln.tw4_off = fourstep_twiddles(tb, q.tw4_n, &d.tw4_lobits);
    }
    ln.chirp_off = q.chirp_off;
    ln.spec_off = q.spec_off;
    ln.rtw_off = q.rtw_off;
    ln.in_role = q.in_role;
    ln.out_role = q.out_role;
    ln.block = (uint32_t)(C * TL);
    ln.fixed = plan_match && ln.block <= 512 && !use_mfma;
    ln.maxt = ln.block <= 256 ? 256 : (ln.fixed ? 512 : 1024);
    if (use_mfma) {
        ln.mfma = mm;
        ln.mfma_nb = env_int("FFTZ_MFMA_NB", 4) == 2 ? 2 : 4;
        ln.block = (uint32_t)(1024 / ln.mfma_nb);
        ln.frag_off = mfma_fragments(tb, mm);
        ln.rtw_off = mfma_stage_twiddles(tb, q.L, mm);  // the pass's d.rtw: see kernels_mfma.hip
    }
    int64_t tiles = (nl + C - 1) / C;
    if ((in_col || out_col) && n0 % C != 0) {
        // ragged rows: ceil(n0 / C) tiles per row of n0 lines
        const int64_t tpr = (n0 + C - 1) / C;
        tiles = tpr * (nl / n0);
        if (tiles * C >= ((int64_t)1 << 31)) {
            err = "too many lines in one pass";
            return false;
        }
        d.vrow = make_fastdiv((uint32_t)(tpr * C));
    }
    ln.grid = (uint32_t)tiles;
    ln.lds_bytes = (uint32_t)((size_t)C * line_bytes);
    if (use_mfma) ln.lds_bytes += 64;  // per-wavefront maxima of the fp16 line scale
    // XCD-contiguous tile order: measured to help passes that read rows and
    // store columns (the four-step's transposing last pass) and to cost ~5%
    // on passes that read columns
    // Also for column tiles whose segments straddle cache lines (an element
    // stride that is not a multiple of 128 bytes, e.g. the 257-wide
    // half-complex rows of a 3-D R2C): neighbouring tiles share those lines.
    const bool straddle = in_col && ((q.ies * esize) % 128) != 0;
    if (((out_col && !in_col) || straddle) && tiles >= 64 && env_int("FFTZ_XCD_REMAP", 1)) d.flags |= kXcdRemap;

    // expand host-looped axes
    std::vector<int64_t> ioffs{q.in_base}, ooffs{q.out_base}, gis{0}, gos{0};
    // real-side strides count real elements (packed R2C / C2R and the
    // full-length promoted / real-part forms alike)
    const int64_t in_scale = ((q.flags & (kRealIn | kRealPromote)) ? prec : esize);
    const int64_t out_scale = ((q.flags & (kRealOut | kRealPart)) ? prec : esize);
    for (auto& h : host_axes) {
        std::vector<int64_t> ni, no, ngi, ngo;
        for (size_t k = 0; k < ioffs.size(); k++)
            for (int64_t j = 0; j < h.n; j++) {
                ni.push_back(ioffs[k] + j * h.is * in_scale);
                no.push_back(ooffs[k] + j * h.os * out_scale);
                ngi.push_back(gis[k] + j * h.gi);
                ngo.push_back(gos[k] + j * h.go);
            }
        ioffs.swap(ni);
        ooffs.swap(no);
        gis.swap(ngi);
        gos.swap(ngo);
    }
    for (size_t k = 0; k < ioffs.size(); k++) {
        Launch l2 = ln;
        l2.in_off_bytes = ioffs[k];
        l2.out_off_bytes = ooffs[k];
        l2.d.gin_base = gis[k];
        l2.d.gout_base = gos[k];
        plan->launches.push_back(l2);
    }
    return true;
}

// ---------------------------------------------------------------------------
// Bluestein tables: chirp conj(B_k) (k < n) and spectrum conj(FFT_M(b)) / M.
// B_k = exp(+i pi (k^2 mod 2n) / n) (core/common/bluestein_utils.c:88-158).
// ---------------------------------------------------------------------------
static int64_t bluestein_length(int64_t n)
{
    for (int64_t m = 2 * n - 1;; m++)
        if (smooth13(m)) return m;
}

// Estimated vector instructions per element for one generic LDS stage of
// radix r: the in-register DFT-r (real operations per element of the codelet
// in codelets.h), the stage-twiddle product ((r - 1) / r complex products)
// and the LDS addressing of one read and one write.
static double radix_stage_cost(int r)
{
    switch (r) {
        case 2: return 17.5;
        case 3: return 19.5;
        case 4: return 19.5;
        case 5: return 22.0;
        case 6: return 22.5;
        case 7: return 25.0;
        case 8: return 22.0;
        case 9: return 26.0;
        case 10: return 25.5;
        case 11: return 29.0;
        case 12: return 26.5;
        case 13: return 31.0;
        default: return 25.0;  // 16
    }
}

// Extended length of a Bluestein transform that runs in one LDS tile.  Any
// 13-smooth M >= 2n - 1 gives the same transform; the reference takes the
// smallest (bluestein_length above, which the multi-pass form keeps).  In one
// tile the work is two M-point transforms through the LDS stages of M's
// fewest-stage radix plan, so this takes the M with the least
// M x sum(radix_stage_cost) among the 13-smooth lengths from the smallest up
// to the next power of two (the same thread count and tile class) and the
// tile limit.  E.g. n = 4099: 8232 = 12 7 7 7 2 becomes 8448 = 16 16 11 3.
// Lengths beyond the limit are returned as the smallest (the caller plans
// them multi-pass).
static double transform_stage_cost(int64_t m)
{
    std::vector<int> r;
    if (!smooth13(m) || !factor_radices((int)m, r)) return -1.0;
    double per = 0.0;
    for (int v : r) per += radix_stage_cost(v);
    return (double)m * per;
}

// transform_stage_cost of every length up to 16384 (-1 where there is no
// radix plan), built once per process
static const std::vector<double>& stage_cost_table()
{
    static std::vector<double> t;
    static std::once_flag once;
    std::call_once(once, [] {
        t.assign(16385, -1.0);
        for (int64_t m = 2; m <= 16384; m++) t[(size_t)m] = transform_stage_cost(m);
    });
    return t;
}

static int64_t bluestein_tile_length(int64_t n, int64_t limit)
{
    const int64_t m0 = bluestein_length(n);
    if (m0 > limit || !env_int("FFTZ_BLUESTEIN_COST", 1)) return m0;
    int64_t hi = 1;
    while (hi < m0) hi <<= 1;
    hi = std::min(hi, limit);
    const std::vector<double>& tab = stage_cost_table();
    int64_t best = m0;
    double best_cost = 0.0;
    for (int64_t m = m0; m <= hi; m++) {
        const double cost = m < (int64_t)tab.size() ? tab[(size_t)m] : transform_stage_cost(m);
        if (cost < 0.0) continue;
        if (m == m0 || cost < best_cost) {
            best = m;
            best_cost = cost;
        }
    }
    return best;
}

Plan* plan_create(const ProblemSpec& spec, std::string& err, bool dry_run);
hipError_t plan_enqueue(Plan* p, Slot* s, const void* in, void* out, hipStream_t stream);

// In-place FP64 forward FFT of v (13-smooth length) on the device, through a
// plan of this library: the spectrum table of a long Bluestein transform
// (M up to ~2^26), where the host recursion would take tens of seconds.
static bool device_fft_fp64(std::vector<d2>& v)
{
    ProblemSpec ss;
    ss.prec_bytes = 8;
    ss.type = FftType::C2C;
    ss.dir = 0;
    ss.inplace = true;
    ss.dims = {{(int64_t)v.size(), 1, 1}};
    std::string err;
    Plan* sub = plan_create(ss, err, false);
    if (!sub) return false;
    void* buf = nullptr;
    const size_t bytes = v.size() * sizeof(d2);
    bool ok = hipMalloc(&buf, bytes) == hipSuccess &&
              hipMemcpy(buf, v.data(), bytes, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipError_t e = hipSuccess;
        Slot* s = plan_acquire_slot(sub, &e);
        ok = s && plan_enqueue(sub, s, buf, buf, s->stream) == hipSuccess &&
             hipStreamSynchronize(s->stream) == hipSuccess;
        if (s) plan_release_slot(sub, s);
    }
    ok = ok && hipMemcpy(v.data(), buf, bytes, hipMemcpyDeviceToHost) == hipSuccess;
    if (buf) (void)hipFree(buf);
    plan_destroy(sub);
    return ok;
}

// dry_run: only the table sizes matter (no device, zero spectrum).
static void bluestein_tables(TableBuilder& tb, int prec, int64_t n, int64_t M, int64_t* chirp_off, int64_t* spec_off,
                             bool dry_run)
{
    std::vector<d2> b((size_t)M, d2{0, 0});
    std::vector<d2> cb((size_t)n);
    for (int64_t k = 0; k < n; k++) {
        const int64_t sq = (int64_t)(((__int128)k * k) % (2 * n));
        d2 w = root_fwd(sq, 2 * n);  // exp(-i pi sq / n) = conj(B_k)
        cb[(size_t)k] = w;
        d2 bk = {w.x, -w.y};
        b[(size_t)k] = bk;
        if (k > 0) b[(size_t)(M - k)] = bk;
    }
    std::vector<d2> fb;
    if (dry_run) {
        fb.assign((size_t)M, d2{0, 0});
    } else if (M > (1 << 18) && device_fft_fp64(b)) {
        fb.swap(b);
    } else {
        fb.resize((size_t)M);
        host_fft(M, b.data(), fb.data());
    }
    const double inv = 1.0 / (double)M;
    for (auto& v : fb) v = {v.x * inv, -v.y * inv};
    *chirp_off = tb.add_complex(cb, prec);
    *spec_off = tb.add_complex(fb, prec);
}

// ---------------------------------------------------------------------------
// Plan construction
// ---------------------------------------------------------------------------
static int64_t single_pass_limit(int prec)
{
    return prec == 4 ? env_int("FFTZ_LMAX_F", 16384) : env_int("FFTZ_LMAX_D", 8192);
}

static int64_t extent(const std::vector<AxisSpec>& dims, const std::vector<AxisSpec>& vecs, bool use_in,
                      int64_t n0_override)
{
    int64_t e = 0;
    for (size_t i = 0; i < dims.size(); i++) {
        int64_t n = (i == 0 && n0_override > 0) ? n0_override : dims[i].n;
        e += (n - 1) * std::llabs(use_in ? dims[i].is : dims[i].os);
    }
    for (auto& v : vecs) e += (v.n - 1) * std::llabs(use_in ? v.is : v.os);
    return e + 1;
}

// Balanced factorisation of a large 1-D length into p factors, each <= fmax
// and a product of radices.
static bool split_factors(int64_t N, int64_t fmax, std::vector<int64_t>& f)
{
    if (const char* forced = getenv("FFTZ_FACTORS")) {
        // explicit factor list (tuning experiments): "a,b[,c]" with product N
        std::vector<int64_t> g;
        int64_t prod = 1;
        for (const char* c = forced; *c;) {
            int64_t v = strtoll(c, (char**)&c, 10);
            if (v > 1) { g.push_back(v); prod *= v; }
            if (*c == ',') c++;
            else if (*c) break;
        }
        if (prod == N && g.size() >= 2 && g.size() <= 3) {
            f = g;
            return true;
        }
    }
    for (int p = 2; p <= 3; p++) {
        // candidate first factors near N^(1/p)
        const double target = std::pow((double)N, 1.0 / p);
        std::vector<int64_t> divs;
        for (int64_t d = 2; d <= fmax && d <= N; d++)
            if (N % d == 0 && smooth13(d)) divs.push_back(d);
        std::sort(divs.begin(), divs.end(),
                  [&](int64_t a, int64_t b) { return std::fabs(std::log((double)a / target)) < std::fabs(std::log((double)b / target)); });
        for (int64_t d0 : divs) {
            int64_t rest = N / d0;
            if (p == 2) {
                if (rest <= fmax && smooth13(rest)) {
                    f = {d0, rest};
                    return true;
                }
            } else {
                const double t2 = std::sqrt((double)rest);
                std::vector<int64_t> d2s;
                for (int64_t d = 2; d <= fmax && d <= rest; d++)
                    if (rest % d == 0 && smooth13(d) && rest / d <= fmax && smooth13(rest / d)) d2s.push_back(d);
                if (d2s.empty()) continue;
                std::sort(d2s.begin(), d2s.end(),
                          [&](int64_t a, int64_t b) { return std::fabs(std::log((double)a / t2)) < std::fabs(std::log((double)b / t2)); });
                f = {d0, d2s[0], rest / d2s[0]};
                return true;
            }
        }
    }
    return false;
}

static bool split_ok(int64_t N, const std::vector<int64_t>& f, int64_t fmax)
{
    if (f.size() < 2 || f.size() > 3) return false;
    int64_t prod = 1;
    for (int64_t v : f) {
        if (v < 2 || v > fmax || !smooth13(v)) return false;
        prod *= v;
    }
    return prod == N;
}

std::vector<std::vector<int64_t>> split_candidates(int64_t N, int prec, int maxn)
{
    // Ordered 2- and 3-factor splits within one LDS pass per factor, most
    // balanced first (sum of |log f - log N^(1/p)|), three factors before two.
    const int64_t fmax = single_pass_limit(prec);
    std::vector<std::pair<double, std::vector<int64_t>>> c;
    std::vector<int64_t> divs;
    for (int64_t d = 16; d <= fmax && d <= N; d++)
        if (N % d == 0 && smooth13(d)) divs.push_back(d);
    for (int p = 3; p >= 2; p--) {
        const double t = std::log((double)N) / p;
        for (int64_t a : divs)
            for (int64_t b : divs) {
                if (N % (a * b)) continue;
                const int64_t r = N / (a * b);
                std::vector<int64_t> f = p == 3 ? std::vector<int64_t>{a, b, r} : std::vector<int64_t>{a, N / a};
                if (p == 2 && b != divs[0]) continue;
                if (!split_ok(N, f, fmax) || f.back() < 16) continue;
                double score = (3 - p) * 100.0;
                for (int64_t v : f) score += std::fabs(std::log((double)v) - t);
                c.push_back({score, f});
            }
    }
    std::stable_sort(c.begin(), c.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    std::vector<std::vector<int64_t>> out;
    for (const auto& e : c)
        if ((int)out.size() < maxn && std::find(out.begin(), out.end(), e.second) == out.end()) out.push_back(e.second);
    return out;
}

// A length-N transform along one axis of a tensor, through the work buffer,
// as a four-step over p = 2 or 3 factors (the p = 2 form is
// core/solvers/complex/pow2_fourstep_solver_dft.c:700-763, with the twiddle
// and transpose of fused_twiddle_transpose_c.c:20-110 folded into the passes'
// stores): the first pass reads the source and writes the dense intermediate
// [batch][N] in the work buffer with the inter-pass twiddle applied, middle
// passes work in place on it, and the last pass writes the destination with
// the transpose folded into its store.  Any number of other axes ride along
// as batch lines (the reference loops them, batched_solver_dft.c:33-110).
struct LongAxis {
    int64_t N = 0;
    int64_t is = 1, os = 1;               // element strides in the source / destination
    std::vector<AxisSpec> batch;          // other axes: n, source stride, destination stride
    Role src = Role::In, dst = Role::Out;
    int64_t src_base = 0, dst_base = 0;   // bytes into the source / destination buffer
    int64_t work_base = 0;                // bytes into the work buffer of the intermediate
    bool conj_in = false, conj_out = false;
    uint32_t first_flags = 0, last_flags = 0;  // Bluestein products (kChirpFlags)
    int64_t nchirp = 0, chirp_off = -1, spec_off = -1;
};

static bool plan_long_axis(Plan* plan, TableBuilder& tb, const LongAxis& a, std::string& err)
{
    const int prec = plan->spec.prec_bytes;
    const int64_t N = a.N;
    // Largest pass length: 256 keeps column tiles 16 lines wide (256-byte
    // FP64 / 128-byte FP32 segments) with 256-thread workgroups; grow it only
    // when three passes of 256 cannot cover N.
    int64_t fmax = env_int(prec == 8 ? "FFTZ_FMAX_D" : "FFTZ_FMAX_F", 256);
    while (fmax < 4096 && fmax * fmax * fmax < N) fmax *= 2;
    std::vector<int64_t> f;
    const std::vector<int64_t>& tuned = plan->spec.tune.split;
    const bool use_tuned = split_ok(N, tuned, single_pass_limit(prec));
    if (use_tuned) {
        f = tuned;
    } else if (!split_factors(N, fmax, f)) {
        err = "no four-step factorisation for length " + std::to_string(N);
        return false;
    }
    // power-of-two factors last: the final (transposing) pass then runs on
    // the fixed-length kernel
    if (!use_tuned && !getenv("FFTZ_FACTORS"))
        std::stable_sort(f.begin(), f.end(), [](int64_t x, int64_t y) {
            return ((x & (x - 1)) != 0) > ((y & (y - 1)) != 0);
        });
    // dense intermediate: batch axis i at stride N * prod_{j<i} n_j
    std::vector<int64_t> wst(a.batch.size());
    int64_t B = 1;
    for (size_t i = 0; i < a.batch.size(); i++) {
        wst[i] = N * B;
        B *= a.batch[i].n;
    }
    const size_t p = f.size();
    std::vector<int64_t> M(p);  // M[s] = N / (f0...fs)
    int64_t acc = N;
    for (size_t s = 0; s < p; s++) {
        acc /= f[s];
        M[s] = acc;
    }
    plan->work_bytes = std::max(plan->work_bytes, (size_t)(a.work_base + N * B * 2 * prec));
    // step 0: columns of the source (element g = l0 + e M0), twiddle on store
    {
        PassReq q;
        q.L = (int)f[0];
        q.ies = M[0] * a.is;
        q.oes = M[0];
        LineAxis l0{M[0], a.is, 1};
        l0.gi = 1;
        q.lines.push_back(l0);
        for (size_t i = 0; i < a.batch.size(); i++) q.lines.push_back({a.batch[i].n, a.batch[i].is, wst[i]});
        q.flags = kTw4 | (a.conj_in ? kConjIn : 0) | (a.first_flags & kChirpIn);
        q.tw4_n = N;
        q.want_in_col = (a.is == 1);
        q.want_out_col = true;
        q.in_role = a.src;
        q.out_role = Role::Work;
        q.in_base = a.src_base;
        q.out_base = a.work_base;
        q.gie = M[0];
        q.nchirp = a.nchirp;
        q.chirp_off = a.chirp_off;
        if (!build_pass(plan, tb, q, err)) return false;
    }
    // middle steps (p == 3)
    for (size_t s = 1; s + 1 < p; s++) {
        PassReq q;
        q.L = (int)f[s];
        q.ies = M[s];
        q.oes = M[s];
        q.lines.push_back({M[s], 1, 1});
        const int64_t rows = N / M[s - 1];
        q.lines.push_back({rows, M[s - 1], M[s - 1]});
        for (size_t i = 0; i < a.batch.size(); i++) q.lines.push_back({a.batch[i].n, wst[i], wst[i]});
        q.flags = kTw4;
        q.tw4_n = M[s - 1];
        q.want_in_col = q.want_out_col = true;
        q.in_role = q.out_role = Role::Work;
        q.in_base = q.out_base = a.work_base;
        if (!build_pass(plan, tb, q, err)) return false;
    }
    // last step: rows of length f[p-1]; row (k0, k1) element e is output
    // k = k0 + f0 k1 + e (N / f[p-1]), stored at k * os
    {
        PassReq q;
        q.L = (int)f[p - 1];
        q.ies = 1;
        const int64_t outer = N / f[p - 1];
        q.oes = a.os * outer;
        q.goe = outer;
        LineAxis k0{f[0], M[0], a.os};
        k0.go = 1;
        q.lines.push_back(k0);
        if (p == 3) {
            LineAxis k1{f[1], M[1], a.os * f[0]};
            k1.go = f[0];
            q.lines.push_back(k1);
        }
        for (size_t i = 0; i < a.batch.size(); i++) q.lines.push_back({a.batch[i].n, wst[i], a.batch[i].os});
        q.flags = (a.conj_out ? kConjOut : 0) | (a.last_flags & (kSpecOut | kChirpOut));
        q.want_in_col = false;
        q.want_out_col = (a.os == 1);
        q.in_role = Role::Work;
        q.out_role = a.dst;
        q.in_base = a.work_base;
        q.out_base = a.dst_base;
        q.nchirp = a.nchirp;
        q.chirp_off = a.chirp_off;
        q.spec_off = a.spec_off;
        if (!build_pass(plan, tb, q, err)) return false;
    }
    std::ostringstream os;
    os << "four-step";
    for (auto v : f) os << " " << v;
    plan->description += os.str() + "; ";
    return true;
}

// Bluestein for a length n whose extended length M exceeds one LDS tile
// (core/solvers/complex/bluestein_solver_dft.c:137-234): the chirp product
// x conj(B) rides on the first pass of a four-step FFT_M, the spectrum
// product conj(A) conj(FFT_M(b)) / M on its last store, a second four-step
// FFT_M reads that and its last store forms conj(B W) (forward) or B W
// (backward, whose input was conjugated on load) for the n outputs.  Work
// buffer: [intermediate of both transforms | product Z], M x batch each.
static bool plan_bluestein_long(Plan* plan, TableBuilder& tb, int64_t n, int64_t is, int64_t os,
                                const std::vector<AxisSpec>& batch, Role src, Role dst, bool bwd,
                                std::string& err)
{
    const int prec = plan->spec.prec_bytes;
    const int64_t M = bluestein_length(n);
    int64_t B = 1;
    for (auto& b : batch) B *= b.n;
    int64_t chirp_off = -1, spec_off = -1;
    bluestein_tables(tb, prec, n, M, &chirp_off, &spec_off, plan->dry_run);
    const int64_t zbase = M * B * 2 * prec;  // region of the product Z
    std::vector<AxisSpec> b1 = batch, b2 = batch;
    int64_t dense = M;
    for (size_t i = 0; i < batch.size(); i++) {
        b1[i].os = dense;
        b2[i].is = dense;
        dense *= batch[i].n;
    }
    LongAxis a1;
    a1.N = M;
    a1.is = is;
    a1.os = 1;
    a1.batch = b1;
    a1.src = src;
    a1.dst = Role::Work;
    a1.dst_base = zbase;
    a1.conj_in = bwd;
    a1.first_flags = kChirpIn;
    a1.last_flags = kSpecOut;
    a1.nchirp = n;
    a1.chirp_off = chirp_off;
    a1.spec_off = spec_off;
    if (!plan_long_axis(plan, tb, a1, err)) return false;
    LongAxis a2;
    a2.N = M;
    a2.is = 1;
    a2.os = os;
    a2.batch = b2;
    a2.src = Role::Work;
    a2.src_base = zbase;
    a2.dst = dst;
    a2.conj_out = !bwd;
    a2.last_flags = kChirpOut;
    a2.nchirp = n;
    a2.chirp_off = chirp_off;
    if (!plan_long_axis(plan, tb, a2, err)) return false;
    plan->work_bytes = std::max(plan->work_bytes, (size_t)(2 * zbase));
    plan->description += "Bluestein n=" + std::to_string(n) + " M=" + std::to_string(M) + " (multi-pass); ";
    return true;
}

// One complete 1-D transform along axis `ax` of a tensor as a list of passes:
// one LDS pass when the length fits a tile (a fused Bluestein pass for a
// length with a prime factor above 13), else the four-step or multi-pass
// Bluestein through the work buffer.  The input is read with the caller's
// input strides, or with the output strides when the axis runs in place on
// the output (use_out_strides_for_in).
static bool plan_axis_c2c(Plan* plan, TableBuilder& tb, const std::vector<AxisSpec>& dims,
                          const std::vector<AxisSpec>& vecs, size_t axis, bool conj_in, bool conj_out,
                          Role in_role, Role out_role, bool use_out_strides_for_in, std::string& err)
{
    const int prec = plan->spec.prec_bytes;
    const int64_t n = dims[axis].n;
    PassReq q;
    q.in_role = in_role;
    q.out_role = out_role;
    q.ies = use_out_strides_for_in ? dims[axis].os : dims[axis].is;
    q.oes = dims[axis].os;
    std::vector<AxisSpec> batch;
    for (size_t j = 0; j < dims.size(); j++)
        if (j != axis) batch.push_back({dims[j].n, use_out_strides_for_in ? dims[j].os : dims[j].is, dims[j].os});
    for (auto& v : vecs) batch.push_back({v.n, use_out_strides_for_in ? v.os : v.is, v.os});
    for (auto& b : batch) q.lines.push_back({b.n, b.is, b.os});
    if (conj_in) q.flags |= kConjIn;
    if (conj_out) q.flags |= kConjOut;
    q.want_in_col = q.ies != 1;
    q.want_out_col = q.oes != 1;
    if (smooth13(n)) {
        if (n > single_pass_limit(prec)) {
            LongAxis a;
            a.N = n;
            a.is = q.ies;
            a.os = q.oes;
            a.batch = batch;
            a.src = in_role;
            a.dst = out_role;
            a.conj_in = conj_in;
            a.conj_out = conj_out;
            return plan_long_axis(plan, tb, a, err);
        }
        q.L = (int)n;
        return build_pass(plan, tb, q, err);
    }
    // Bluestein
    const int64_t M = bluestein_tile_length(n, single_pass_limit(prec));
    if (M > single_pass_limit(prec))
        return plan_bluestein_long(plan, tb, n, q.ies, q.oes, batch, in_role, out_role, conj_in, err);
    q.kind = KernelKind::Bluestein;
    q.L = (int)M;
    q.nvalid_in = (int)n;
    q.nvalid_out = (int)n;
    bluestein_tables(tb, prec, n, M, &q.chirp_off, &q.spec_off, plan->dry_run);
    q.want_in_col = q.want_out_col = false;
    return build_pass(plan, tb, q, err);
}

static bool plan_c2c(Plan* plan, TableBuilder& tb, std::vector<AxisSpec> dims, const std::vector<AxisSpec>& vecs,
                     std::string& err)
{
    const bool bwd = plan->spec.dir == 1;
    for (size_t a = 0; a < dims.size(); a++) {
        const bool first = (a == 0);
        if (!plan_axis_c2c(plan, tb, dims, vecs, a, bwd, bwd, first ? Role::In : Role::Out, Role::Out, !first, err))
            return false;
    }
    plan->description += "row-column " + std::to_string(dims.size()) + "-D; ";
    return true;
}

// Full-length pass for a real axis of length n (odd n, or an even n whose half
// has no radix plan): the n-point complex transform of the real line promoted
// to complex (R2C), or of the Hermitian extension of the n/2+1 stored bins
// with only the real part stored (C2R).  Bluestein when n has a prime factor
// above 13.  Twice the arithmetic of the packed form; same HBM traffic.
static bool real_full_pass(Plan* plan, TableBuilder& tb, PassReq q, int64_t n, std::string& err)
{
    const int prec = plan->spec.prec_bytes;
    const bool r2c = (q.flags & kRealPromote) != 0;
    q.nvalid_in = (int)n;
    q.nvalid_out = r2c ? (int)(n / 2 + 1) : (int)n;
    q.want_in_col = q.want_out_col = false;
    if (smooth13(n)) {
        if (n > single_pass_limit(prec)) {
            err = "real length " + std::to_string(n) + " above the single-pass limit";
            return false;
        }
        q.L = (int)n;
        return build_pass(plan, tb, q, err);
    }
    const int64_t M = bluestein_tile_length(n, single_pass_limit(prec));
    if (M > single_pass_limit(prec)) {
        err = "Bluestein length " + std::to_string(M) + " for real length " + std::to_string(n) +
              " exceeds the fused single-pass limit";
        return false;
    }
    q.kind = KernelKind::Bluestein;
    q.L = (int)M;
    bluestein_tables(tb, prec, n, M, &q.chirp_off, &q.spec_off, plan->dry_run);
    return build_pass(plan, tb, q, err);
}

static bool plan_real(Plan* plan, TableBuilder& tb, std::vector<AxisSpec> dims, const std::vector<AxisSpec>& vecs,
                      std::string& err)
{
    const int prec = plan->spec.prec_bytes;
    const bool r2c = plan->spec.type == FftType::R2C;
    const int64_t n0 = dims[0].n;
    const int64_t L = n0 / 2;
    // real-packed (n0/2-point complex transform + recombine / separate) when
    // n0 is even and its half has a single-pass radix plan; otherwise the
    // full-length promoted / Hermitian pass
    const bool packed = n0 % 2 == 0 && smooth13(L) && L <= single_pass_limit(prec) && L >= 1;
    const int64_t nbins = n0 / 2 + 1;
    int64_t rtw_off = -1;
    if (packed) {
        std::vector<d2> rtw((size_t)(L / 2 + 1));
        for (int64_t k = 0; k <= L / 2; k++) rtw[(size_t)k] = root_fwd(k, n0);
        rtw_off = tb.add_complex(rtw, prec);
    }
    // complex geometry of the half-spectrum side
    std::vector<AxisSpec> cd = dims;  // n of dim 0 becomes L+1 bins on the complex side
    if (r2c) {
        PassReq q;
        q.L = (int)L;
        q.ies = dims[0].is;  // real elements
        q.oes = dims[0].os;
        for (size_t j = 1; j < dims.size(); j++) q.lines.push_back({dims[j].n, dims[j].is, dims[j].os});
        for (auto& v : vecs) q.lines.push_back({v.n, v.is, v.os});
        q.in_role = Role::In;
        q.out_role = Role::Out;
        if (packed) {
            q.flags = kRealIn | kR2CPost;
            q.rtw_off = rtw_off;
            if (!build_pass(plan, tb, q, err)) return false;
        } else {
            q.flags = kRealPromote;
            if (!real_full_pass(plan, tb, q, n0, err)) return false;
        }
        // remaining axes: C2C forward in place on the output
        cd[0].n = nbins;
        for (size_t a = 1; a < dims.size(); a++) {
            std::vector<AxisSpec> od = cd;
            if (!plan_axis_c2c(plan, tb, od, vecs, a, false, false, Role::Out, Role::Out, true, err))
                return false;
        }
        plan->description += packed ? "real-packed R2C; " : "real R2C (full length); ";
        return true;
    }
    // C2R: outer axes backward into the work buffer, then separate + FFT.
    Role src = Role::In;
    std::vector<AxisSpec> wdims = dims;
    std::vector<AxisSpec> wvecs = vecs;
    if (dims.size() > 1) {
        // dense work layout: [vec][dims reversed], dim0 has L+1 bins
        int64_t stride = 1;
        std::vector<AxisSpec> dd = dims;
        dd[0].n = nbins;
        for (size_t j = 0; j < dd.size(); j++) {
            wdims[j] = {dd[j].n, dd[j].is, stride};
            stride *= dd[j].n;
        }
        int64_t per = stride;
        for (size_t v = 0; v < wvecs.size(); v++) {
            wvecs[v] = {vecs[v].n, vecs[v].is, stride};
            stride *= vecs[v].n;
        }
        plan->work_bytes = std::max(plan->work_bytes, (size_t)(stride * 2 * prec));
        (void)per;
        // copy-transform axis 1 from In to Work, then the rest in place on Work
        for (size_t a = 1; a < dims.size(); a++) {
            const bool first = (a == 1);
            std::vector<AxisSpec> od = wdims;
            if (!plan_axis_c2c(plan, tb, od, wvecs, a, true, true, first ? Role::In : Role::Work,
                               Role::Work, !first, err))
                return false;
        }
        src = Role::Work;
        for (size_t j = 0; j < wdims.size(); j++) wdims[j].is = wdims[j].os;
        for (size_t v = 0; v < wvecs.size(); v++) wvecs[v].is = wvecs[v].os;
    }
    PassReq q;
    q.L = (int)L;
    q.ies = wdims[0].is;
    q.oes = dims[0].os;  // real elements
    for (size_t j = 1; j < dims.size(); j++) q.lines.push_back({dims[j].n, wdims[j].is, dims[j].os});
    for (size_t v = 0; v < vecs.size(); v++) q.lines.push_back({vecs[v].n, wvecs[v].is, vecs[v].os});
    q.in_role = src;
    q.out_role = Role::Out;
    if (!packed) {
        // backward = conjugate in and out around the forward transform
        q.flags = kHermIn | kConjIn | kConjOut | kRealPart;
        if (!real_full_pass(plan, tb, q, n0, err)) return false;
        plan->description += "real C2R (full length); ";
        return true;
    }
    q.nvalid_in = (int)(L + 1);
    q.extra_lds_elems = 1;
    q.flags = kC2RPre | kConjOut | kRealOut;
    q.rtw_off = rtw_off;
    if (!build_pass(plan, tb, q, err)) return false;
    plan->description += "real-packed C2R; ";
    return true;
}

// ---------------------------------------------------------------------------
// Uneven batches (many lengths, one persistent launch per tile class).
// The reference runs such a list as independent plans, one setup/execute per
// (n, batch) item; here every distinct length becomes a pass descriptor in the
// plan arena and every item a run of tiles in a device work queue.
// ---------------------------------------------------------------------------
namespace {
struct SizeInfo {
    int64_t n = 0;
    bool bluestein = false;
    int L = 0;
    int cls = 0;
    int C = 1;
    PassDesc d{};
};
}  // namespace

static bool batch_size_desc(TableBuilder& tb, int prec, int64_t n, int E, int maxt, size_t lds_budget,
                            SizeInfo& si, std::string& err)
{
    si.n = n;
    si.bluestein = !smooth13(n);
    const int64_t L = si.bluestein ? bluestein_tile_length(n, single_pass_limit(prec)) : n;
    si.L = (int)L;
    PassDesc& d = si.d;
    d = PassDesc{};
    std::vector<int> radix;
    if (!factor_radices((int)L, radix)) {
        err = "no radix plan for batch length " + std::to_string(L);
        return false;
    }
    d.L = (int)L;
    d.nst = (int)radix.size();
    if (si.bluestein && d.nst < 2) {
        // the batch kernel runs the register-direct Bluestein body only
        err = "internal: Bluestein batch length " + std::to_string(L) + " with one stage";
        return false;
    }
    for (size_t s = 0; s < radix.size(); s++) d.radix[s] = radix[s];
    const int64_t stw = stage_twiddles(tb, prec, radix, d.twoff, d.ls);
    for (int s = 0; s < d.nst; s++) d.lsdiv[s] = make_fastdiv((uint32_t)d.ls[s]);
    int TL = next_pow2((L + E - 1) / E);
    if (TL < 1) TL = 1;
    if (TL > maxt) {
        err = "batch length " + std::to_string(L) + " too long for its tile class";
        return false;
    }
    d.TL = TL;
    d.tl_shift = ilog2(TL);
    int LP = lds_pad((int)L - 1) + 1;
    if ((LP & 1) == 0) LP++;
    d.LP = LP;
    const size_t line_bytes = (size_t)LP * 2 * prec;
    if (line_bytes > lds_budget) {
        err = "batch length " + std::to_string(L) + " exceeds LDS";
        return false;
    }
    int C = 1;
    const int cmax = env_int("FFTZ_BATCH_CMAX", 64);
    while (C * 2 * TL <= maxt && (size_t)(C * 2) * line_bytes <= lds_budget && C * 2 <= cmax) C *= 2;
    si.C = C;
    d.C = C;
    d.c_shift = ilog2(C);
    d.ies = 1;
    d.oes = 1;
    d.ln0 = make_fastdiv(1u << 30);  // l0 = line index (< 2^30)
    d.ln1 = make_fastdiv(1);
    d.ils[1] = d.ils[2] = d.ols[1] = d.ols[2] = 0;
    d.nvalid_in = si.bluestein ? (int)n : (int)L;
    d.nvalid_out = si.bluestein ? (int)n : (int)L;
    d.scale_f = 1.0f;
    d.scale_d = 1.0;
    // table pointers hold arena offsets until upload (resolved in place)
    d.stw = (const void*)(intptr_t)stw;
    if (si.bluestein) {
        int64_t co, so;
        bluestein_tables(tb, prec, n, L, &co, &so, false);
        d.chirp = (const void*)(intptr_t)co;
        d.spec = (const void*)(intptr_t)so;
    }
    return true;
}

Plan* batch_plan_create(const std::vector<BatchItem>& items, int prec, int dir, std::string& err, bool dry_run,
                        const TuneParams* tune)
{
    Plan* plan = new Plan();
    plan->dry_run = dry_run;
    plan->spec.prec_bytes = prec;
    plan->spec.dir = dir;
    plan->spec.type = FftType::C2C;
    plan->is_batch = true;
    if (!dry_run) {
        (void)hipGetDevice(&plan->device);
        (void)device_info();
    }
    const DeviceInfo& dev = device_info();
    const int esize = 2 * prec;
    TableBuilder tb{plan->host_tables};
    // size classes: small tiles (E = 8, 256 threads, <= 64 KiB LDS, several
    // workgroups per CU) and large tiles (E = 16, 1024 threads, whole LDS)
    const size_t static_lds = kBatchHdrBytes;
    const size_t small_budget = (size_t)env_int("FFTZ_BATCH_SMALL_LDS", 64 * 1024);
    const size_t big_budget = dev.lds_per_block - static_lds;
    // medium tiles: 512 threads, at most half the LDS, two workgroups per CU
    const size_t medium_budget = dev.lds_per_block / 2 - static_lds;
    std::map<int64_t, int> size_index;
    std::vector<SizeInfo> sizes;
    int64_t in_ext = 0, out_ext = 0;
    double flops = 0.0;
    for (const BatchItem& it : items) {
        if (it.n <= 0 || it.batch < 0 || it.in_off < 0 || it.out_off < 0 || it.dist < 0) {
            err = "invalid batch item";
            delete plan;
            return nullptr;
        }
        if (it.batch == 0) continue;
        const int64_t dist = it.dist ? it.dist : it.n;
        in_ext = std::max(in_ext, it.in_off + (it.batch - 1) * dist + it.n);
        out_ext = std::max(out_ext, it.out_off + (it.batch - 1) * dist + it.n);
        flops += 5.0 * (double)it.n * std::log2(std::max<double>((double)it.n, 2.0)) * (double)it.batch;
        plan->n_transforms += it.batch;
        if (size_index.count(it.n)) continue;
        SizeInfo si;
        si.n = it.n;
        const int64_t L = smooth13(it.n) ? it.n : bluestein_tile_length(it.n, single_pass_limit(prec));
        const bool small = next_pow2((L + 7) / 8) <= 256 && (size_t)(lds_pad((int)L) + 2) * esize <= small_budget / 2;
        const bool medium = !small && next_pow2((L + 15) / 16) <= 512 &&
                            (size_t)(lds_pad((int)L) + 2) * esize <= medium_budget;
        si.cls = small ? 0 : (medium ? 1 : 2);
        size_index[it.n] = (int)sizes.size();
        sizes.push_back(si);
    }
    // Per-length descriptors and tables (stage twiddles, Bluestein chirp and
    // spectrum) are independent: build them on worker threads into separate
    // arenas, then append each arena to the plan's and rebase its offsets.
    {
        const int cE[3] = {8, 16, 16}, cT[3] = {256, 512, 1024};
        const size_t cB[3] = {small_budget, medium_budget, big_budget};
        std::vector<std::vector<unsigned char>> arenas(sizes.size());
        std::vector<std::string> errs(sizes.size());
        std::vector<char> ok(sizes.size(), 0);
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t i; (i = next.fetch_add(1)) < sizes.size();) {
                TableBuilder ltb{arenas[i]};
                SizeInfo& si = sizes[i];
                const int cls = si.cls;
                ok[i] = batch_size_desc(ltb, prec, si.n, cE[cls], cT[cls], cB[cls], si, errs[i]) ? 1 : 0;
                si.cls = cls;
            }
        };
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const size_t nthr = std::min<size_t>(hw, (sizes.size() + 7) / 8);
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nthr; t++) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
        for (size_t i = 0; i < sizes.size(); i++) {
            if (!ok[i]) {
                err = errs[i];
                delete plan;
                return nullptr;
            }
            const int64_t base = tb.add(arenas[i].data(), arenas[i].size());
            PassDesc& d = sizes[i].d;
            auto rebase = [base](const void*& ptr) { ptr = (const void*)((intptr_t)ptr + (intptr_t)base); };
            rebase(d.stw);
            if (sizes[i].bluestein) {
                rebase(d.chirp);
                rebase(d.spec);
            }
        }
    }
    plan->in_extent_bytes = in_ext * esize;
    plan->out_extent_bytes = out_ext * esize;
    plan->flops = flops;
    // tiles per class, cost-sorted (largest first)
    // a queue entry covers up to `rounds` rounds of C lines (about 128 KiB of
    // input; 256 KiB in the large class, whose single-line rounds are up to
    // 128 KiB each: measured 2.6% faster there, no gain in the others) so the
    // per-entry cost (queue atomic, descriptor loads, barriers) is amortised
    const int64_t entry_env = env_int("FFTZ_BATCH_ENTRY_BYTES", 0);
    const int64_t entry_cls[3] = {128 * 1024, 128 * 1024, 256 * 1024};
    std::vector<std::vector<std::pair<double, BatchTile>>> tiles(3);
    for (const BatchItem& it : items) {
        if (it.batch == 0) continue;
        const SizeInfo& si = sizes[(size_t)size_index[it.n]];
        const int64_t dist = it.dist ? it.dist : it.n;
        const int64_t entry_bytes = entry_env > 0 ? entry_env : entry_cls[si.cls];
        const int64_t per_round = (int64_t)si.C * si.n * esize;
        const int64_t rounds = std::max<int64_t>(1, std::min<int64_t>(16, entry_bytes / std::max<int64_t>(1, per_round)));
        const int64_t step = (int64_t)si.C * rounds;
        for (int64_t l0 = 0; l0 < it.batch; l0 += step) {
            BatchTile bt{};
            bt.desc = (uint32_t)size_index[it.n];
            bt.nlines = (uint32_t)std::min<int64_t>(step, it.batch - l0);
            bt.bluestein = si.bluestein ? 1u : 0u;
            bt.in_off = it.in_off + l0 * dist;
            bt.out_off = it.out_off + l0 * dist;
            bt.dist = dist;
            const double cost = (double)bt.nlines * si.L * (std::log2((double)si.L + 1.0) + 8.0) *
                                (si.bluestein ? 2.5 : 1.0);
            tiles[(size_t)si.cls].push_back({cost, bt});
        }
    }
    // direction: every tile is a complete transform; backward conjugates in
    // and out (the Bluestein body handles its own conjugation)
    for (SizeInfo& si : sizes)
        if (dir == 1) si.d.flags |= kConjIn | kConjOut;
    std::vector<PassDesc> descs;
    for (const SizeInfo& si : sizes) descs.push_back(si.d);
    plan->descs_off = tb.add(descs.data(), std::max<size_t>(1, descs.size()) * sizeof(PassDesc));
    for (int c = 0; c < 3; c++) {
        auto& tl = tiles[(size_t)c];
        std::stable_sort(tl.begin(), tl.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        BatchClass bc;
        bc.E = c == 0 ? 8 : 16;
        bc.maxt = c == 0 ? 256 : (c == 1 ? 512 : 1024);
        bc.ntiles = (int)tl.size();
        size_t lds = 0;
        for (const SizeInfo& si : sizes)
            if (si.cls == c) lds = std::max(lds, (size_t)si.C * si.d.LP * esize);
        bc.lds_bytes = (uint32_t)(lds + kBatchHdrBytes);
        bc.grid = 1;
        if (bc.ntiles) {
            std::vector<BatchTile> v;
            v.reserve(tl.size());
            for (auto& t : tl) v.push_back(t.second);
            bc.tiles_off = tb.add(v.data(), v.size() * sizeof(BatchTile));
        }
        plan->bclasses.push_back(bc);
    }
    std::ostringstream os;
    {
        plan->tune_key = batch_key(items, prec, dir);
        TuneParams t;
        if (tune) t = *tune;
        else (void)tune_lookup(plan->tune_key, t);
        batch_set_grids(plan, t.wg_small, t.wg_medium);
    }
    static const char* kClassName[3] = {"small", "medium", "large"};
    os << "uneven batch: " << plan->n_transforms << " transforms, " << sizes.size() << " lengths; tiles";
    for (int c = 0; c < 3; c++)
        os << " " << kClassName[c] << "=" << plan->bclasses[(size_t)c].ntiles << " (grid " << plan->bclasses[(size_t)c].grid
           << ", lds " << plan->bclasses[(size_t)c].lds_bytes << ")";
    os << "; ";
    plan->description = os.str();
    if (dry_run) return plan;
    if (hipMalloc(&plan->tables, plan->host_tables.size()) != hipSuccess) {
        err = "device allocation for batch tables failed";
        delete plan;
        return nullptr;
    }
    // resolve descriptor table offsets to device pointers, then upload
    char* base = (char*)plan->tables;
    PassDesc* hd = reinterpret_cast<PassDesc*>(plan->host_tables.data() + plan->descs_off);
    for (size_t i = 0; i < sizes.size(); i++) {
        hd[i].stw = base + (intptr_t)hd[i].stw;
        if (sizes[i].bluestein) {
            hd[i].chirp = base + (intptr_t)hd[i].chirp;
            hd[i].spec = base + (intptr_t)hd[i].spec;
        }
    }
    if (hipMemcpy(plan->tables, plan->host_tables.data(), plan->host_tables.size(), hipMemcpyHostToDevice) !=
        hipSuccess) {
        err = "upload of batch tables failed";
        (void)hipFree(plan->tables);
        delete plan;
        return nullptr;
    }
    for (size_t c = 0; c < plan->bclasses.size(); c++)
        (void)hipFuncSetAttribute(kernel_symbol_batch(prec, plan->bclasses[c].E, plan->bclasses[c].maxt),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)dev.lds_per_block);
    plan->host_tables.clear();
    plan->host_tables.shrink_to_fit();
    plan->use_graph = env_int("FFTZ_BATCH_GRAPH", 1) != 0;
    return plan;
}

void batch_set_grids(Plan* p, int wg_small, int wg_medium)
{
    const DeviceInfo& dev = device_info();
    for (size_t c = 0; c < p->bclasses.size(); c++) {
        BatchClass& bc = p->bclasses[c];
        const int per_cu = c == 0 ? (wg_small ? wg_small : env_int("FFTZ_BATCH_WG_SMALL", 8))
                           : c == 1 ? (wg_medium ? wg_medium : env_int("FFTZ_BATCH_WG_MEDIUM", 2))
                                    : env_int("FFTZ_BATCH_WG_BIG", 1);
        bc.grid = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(bc.ntiles, (int64_t)dev.cus * per_cu));
    }
    // captured graphs hold the old grids
    for (Slot* s : p->slots) {
        if (s->graph) (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
        s->graph_in = nullptr;
        s->graph_out = nullptr;
    }
}

// ---------------------------------------------------------------------------
// Tuned-parameter cache
// ---------------------------------------------------------------------------
namespace {
std::mutex g_tune_mu;
std::map<std::string, TuneParams> g_tune;
}  // namespace

std::string problem_key(const ProblemSpec& s)
{
    std::ostringstream os;
    os << "p" << s.prec_bytes << ':' << (int)s.type << ':' << s.dir << ':' << s.inplace << ':' << s.mfma_mode << '|';
    for (const auto& a : s.dims) os << a.n << ',' << a.is << ',' << a.os << ';';
    os << '|';
    for (const auto& a : s.vecs) os << a.n << ',' << a.is << ',' << a.os << ';';
    return os.str();
}

std::string batch_key(const std::vector<BatchItem>& items, int prec, int dir)
{
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the item list
    auto mix = [&](int64_t v) {
        for (int b = 0; b < 8; b++) {
            h ^= (uint64_t)((v >> (8 * b)) & 0xff);
            h *= 1099511628211ull;
        }
    };
    for (const BatchItem& it : items) {
        mix(it.n);
        mix(it.batch);
        mix(it.in_off);
        mix(it.out_off);
        mix(it.dist);
    }
    std::ostringstream os;
    os << "b" << prec << ':' << dir << ':' << items.size() << ':' << std::hex << h;
    return os.str();
}

bool tune_lookup(const std::string& key, TuneParams& out)
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    out = it->second;
    return true;
}

void tune_store(const std::string& key, const TuneParams& t)
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[key] = t;
}

std::string tune_text(int* entries)
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    std::ostringstream os;
    os << "# fftz tuned plans: key row_t col_c wg_small wg_medium [e=elems] [r=L:r.r.r/L:r.r] [f=f.f.f]\n";
    for (const auto& kv : g_tune) {
        const TuneParams& t = kv.second;
        os << kv.first << ' ' << t.row_t << ' ' << t.col_c << ' ' << t.wg_small << ' ' << t.wg_medium;
        if (t.elems) os << " e=" << t.elems;
        if (!t.radix_for.empty()) {
            os << " r=";
            for (size_t i = 0; i < t.radix_for.size(); i++) {
                os << (i ? "/" : "") << t.radix_for[i].first << ':';
                for (size_t j = 0; j < t.radix_for[i].second.size(); j++)
                    os << (j ? "." : "") << t.radix_for[i].second[j];
            }
        }
        if (!t.split.empty()) {
            os << " f=";
            for (size_t i = 0; i < t.split.size(); i++) os << (i ? "." : "") << t.split[i];
        }
        os << '\n';
    }
    if (entries) *entries = (int)g_tune.size();
    return os.str();
}

// "1.2.3" -> {1, 2, 3}; false on anything but positive integers and dots.
static bool parse_dotted(const std::string& v, std::vector<int64_t>& out)
{
    out.clear();
    size_t i = 0;
    while (i < v.size()) {
        size_t j = v.find('.', i);
        if (j == std::string::npos) j = v.size();
        const std::string tok = v.substr(i, j - i);
        if (tok.empty() || tok.size() > 12 || tok.find_first_not_of("0123456789") != std::string::npos) return false;
        out.push_back(std::stoll(tok));
        if (out.back() <= 0) return false;
        i = j + 1;
        if (j == v.size()) break;
        if (i == v.size()) return false;  // trailing dot
    }
    return !out.empty();
}

// Optional fields of a tuned-cache line: e=8, r=1024:16.8.8/256:16.16, f=256.256.256.
static bool parse_tune_token(const std::string& tok, TuneParams& t)
{
    if (tok.size() < 3 || tok[1] != '=') return false;
    const std::string v = tok.substr(2);
    std::vector<int64_t> nums;
    switch (tok[0]) {
        case 'e':
            if (v != "8" && v != "16") return false;
            t.elems = std::stoi(v);
            return true;
        case 'f':
            if (!parse_dotted(v, nums) || nums.size() < 2 || nums.size() > 3) return false;
            for (int64_t f : nums)
                if (f < 2) return false;
            t.split = nums;
            return true;
        case 'r': {
            size_t i = 0;
            while (i <= v.size()) {
                size_t j = v.find('/', i);
                if (j == std::string::npos) j = v.size();
                const std::string item = v.substr(i, j - i);
                const size_t c = item.find(':');
                if (c == std::string::npos || c == 0) return false;
                const std::string ls = item.substr(0, c);
                if (ls.size() > 9 || ls.find_first_not_of("0123456789") != std::string::npos) return false;
                const int L = std::stoi(ls);
                if (!parse_dotted(item.substr(c + 1), nums)) return false;
                std::vector<int> r(nums.begin(), nums.end());
                if (!radix_order_ok(L, r)) return false;
                t.radix_for.push_back({L, r});
                i = j + 1;
                if (j == v.size()) break;
            }
            return !t.radix_for.empty();
        }
        default:
            return false;
    }
}

int tune_parse(const std::string& text)
{
    // All or nothing: a malformed line leaves the cache untouched.
    std::vector<std::pair<std::string, TuneParams>> got;
    std::istringstream is(text);
    std::string line;
    while (std::getline(is, line)) {
        const size_t b = line.find_first_not_of(" \t\r");
        if (b == std::string::npos || line[b] == '#') continue;
        std::istringstream ls(line);
        std::string key, extra;
        TuneParams t;
        if (!(ls >> key >> t.row_t >> t.col_c >> t.wg_small >> t.wg_medium)) return -1;
        for (int v : {t.row_t, t.col_c, t.wg_small, t.wg_medium})
            if (v < 0 || v > 4096) return -1;
        while (ls >> extra)
            if (!parse_tune_token(extra, t)) return -1;
        got.emplace_back(key, t);
    }
    for (const auto& kv : got) tune_store(kv.first, kv.second);
    return (int)got.size();
}

void tune_clear()
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune.clear();
}

int tune_export(const char* path)
{
    FILE* f = path ? fopen(path, "w") : nullptr;
    if (!f) return -1;
    int n = 0;
    const std::string text = tune_text(&n);
    const bool ok = fwrite(text.data(), 1, text.size(), f) == text.size();
    return (fclose(f) == 0 && ok) ? n : -1;
}

int tune_import(const char* path)
{
    FILE* f = path ? fopen(path, "r") : nullptr;
    if (!f) return -1;
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    fclose(f);
    return tune_parse(text);
}

Plan* plan_create(const ProblemSpec& spec_in, std::string& err, bool dry_run)
{
    Plan* plan = new Plan();
    plan->dry_run = dry_run;
    plan->spec = spec_in;
    ProblemSpec& spec = plan->spec;
    if (spec.tune.empty()) (void)tune_lookup(problem_key(spec), spec.tune);
    if (!dry_run) {
        (void)hipGetDevice(&plan->device);
        (void)device_info();
    }

    // Drop size-1 dims (SHRINK_DIM_RANK, selector/selector.h:194-214) and
    // size-1 batch axes; a zero-length batch axis is a no-op plan
    // (selector/selector.c:1375-1385).
    std::vector<AxisSpec> dims, vecs;
    for (size_t i = 0; i < spec.dims.size(); i++)
        if (spec.dims[i].n != 1 || (i == 0 && spec.type != FftType::C2C)) dims.push_back(spec.dims[i]);
    if (dims.empty()) dims.push_back({1, 1, 1});
    for (auto& v : spec.vecs) {
        if (v.n == 0) plan->noop = true;
        if (v.n > 1) vecs.push_back(v);
    }
    // extents for host staging
    if (spec.type == FftType::C2C) {
        plan->in_extent_bytes = extent(spec.dims, spec.vecs, true, -1) * 2 * spec.prec_bytes;
        plan->out_extent_bytes = extent(spec.dims, spec.vecs, false, -1) * 2 * spec.prec_bytes;
    } else if (spec.type == FftType::R2C) {
        plan->in_extent_bytes = extent(spec.dims, spec.vecs, true, -1) * spec.prec_bytes;
        plan->out_extent_bytes = extent(spec.dims, spec.vecs, false, spec.dims[0].n / 2 + 1) * 2 * spec.prec_bytes;
    } else {
        plan->in_extent_bytes = extent(spec.dims, spec.vecs, true, spec.dims[0].n / 2 + 1) * 2 * spec.prec_bytes;
        plan->out_extent_bytes = extent(spec.dims, spec.vecs, false, -1) * spec.prec_bytes;
    }
    if (plan->noop) return plan;

    // algorithmic flop count (5 N log2 N for C2C, 2.5 N log2 N for real)
    double ntot = 1.0, nb = 1.0;
    for (auto& d : spec.dims) ntot *= (double)d.n;
    for (auto& v : spec.vecs) nb *= (double)v.n;
    plan->flops = (spec.type == FftType::C2C ? 5.0 : 2.5) * ntot * std::log2(std::max(ntot, 2.0)) * nb;

    TableBuilder tb{plan->host_tables};
    bool ok;
    if (spec.type == FftType::C2C)
        ok = plan_c2c(plan, tb, dims, vecs, err);
    else
        ok = plan_real(plan, tb, dims, vecs, err);
    if (!ok) {
        delete plan;
        return nullptr;
    }
    if (dry_run) return plan;
    // upload tables and resolve pointers
    if (!plan->host_tables.empty()) {
        if (hipMalloc(&plan->tables, plan->host_tables.size()) != hipSuccess ||
            hipMemcpy(plan->tables, plan->host_tables.data(), plan->host_tables.size(), hipMemcpyHostToDevice) !=
                hipSuccess) {
            err = "device allocation for plan tables failed";
            if (plan->tables) (void)hipFree(plan->tables);
            delete plan;
            return nullptr;
        }
    }
    char* base = (char*)plan->tables;
    for (auto& ln : plan->launches) {
        if (ln.stw_off >= 0) ln.d.stw = base + ln.stw_off;
        if (ln.tw4_off >= 0) ln.d.tw4 = base + ln.tw4_off;
        if (ln.chirp_off >= 0) ln.d.chirp = base + ln.chirp_off;
        if (ln.spec_off >= 0) ln.d.spec = base + ln.spec_off;
        if (ln.rtw_off >= 0) ln.d.rtw = base + ln.rtw_off;
        if (ln.frag_off >= 0) ln.frag = reinterpret_cast<const float*>(base + ln.frag_off);
    }
    plan->host_tables.clear();
    plan->host_tables.shrink_to_fit();
    plan->use_graph = env_int("FFTZ_GRAPH", 0) != 0;
    return plan;
}

static std::string batch_json(const Plan* p)
{
    std::ostringstream os;
    os << "{\"batch\":1,\"prec\":" << p->spec.prec_bytes << ",\"in_extent_bytes\":" << p->in_extent_bytes
       << ",\"out_extent_bytes\":" << p->out_extent_bytes << ",\"sizes\":[";
    const PassDesc* hd = p->host_tables.empty()
                             ? nullptr
                             : reinterpret_cast<const PassDesc*>(p->host_tables.data() + p->descs_off);
    size_t nsz = 0;
    for (const BatchClass& bc : p->bclasses) {
        const BatchTile* t = hd ? reinterpret_cast<const BatchTile*>(p->host_tables.data() + bc.tiles_off) : nullptr;
        for (int i = 0; hd && i < bc.ntiles; i++) nsz = std::max(nsz, (size_t)t[i].desc + 1);
    }
    for (size_t i = 0; hd && i < nsz; i++) {
        const PassDesc& d = hd[i];
        os << (i ? "," : "") << "{\"L\":" << d.L << ",\"nvalid\":" << d.nvalid_in << ",\"C\":" << d.C
           << ",\"TL\":" << d.TL << ",\"LP\":" << d.LP << ",\"flags\":" << d.flags << ",\"radix\":[";
        for (int s = 0; s < d.nst; s++) os << (s ? "," : "") << d.radix[s];
        os << "]}";
    }
    os << "],\"classes\":[";
    for (size_t c = 0; c < p->bclasses.size(); c++) {
        const BatchClass& bc = p->bclasses[c];
        os << (c ? "," : "") << "{\"E\":" << bc.E << ",\"maxt\":" << bc.maxt << ",\"lds\":" << bc.lds_bytes
           << ",\"grid\":" << bc.grid << ",\"tiles\":[";
        const BatchTile* t = hd ? reinterpret_cast<const BatchTile*>(p->host_tables.data() + bc.tiles_off) : nullptr;
        for (int i = 0; t && i < bc.ntiles; i++)
            os << (i ? "," : "") << "[" << t[i].desc << "," << t[i].nlines << "," << t[i].bluestein << ","
               << t[i].in_off << "," << t[i].out_off << "," << t[i].dist << "]";
        os << "]}";
    }
    os << "]}";
    return os.str();
}

std::string plan_json(const Plan* p)
{
    if (p->is_batch) return batch_json(p);
    std::ostringstream os;
    os << "{\"noop\":" << (p->noop ? 1 : 0) << ",\"work_bytes\":" << p->work_bytes
       << ",\"in_extent_bytes\":" << p->in_extent_bytes << ",\"out_extent_bytes\":" << p->out_extent_bytes
       << ",\"prec\":" << p->spec.prec_bytes << ",\"launches\":[";
    bool first = true;
    for (const Launch& ln : p->launches) {
        const PassDesc& d = ln.d;
        os << (first ? "" : ",") << "{\"kind\":\"" << (ln.kind == KernelKind::Bluestein ? "bluestein" : "pass")
           << "\",\"L\":" << d.L << ",\"nvalid_in\":" << d.nvalid_in << ",\"nvalid_out\":" << d.nvalid_out
           << ",\"ies\":" << d.ies << ",\"oes\":" << d.oes << ",\"ils\":[" << d.ils[0] << "," << d.ils[1]
           << "," << d.ils[2] << "],\"ols\":[" << d.ols[0] << "," << d.ols[1] << "," << d.ols[2]
           << "],\"nlines\":" << d.nlines << ",\"ln0\":" << d.ln0.d << ",\"ln1\":" << d.ln1.d
           << ",\"flags\":" << d.flags << ",\"C\":" << d.C << ",\"TL\":" << d.TL << ",\"LP\":" << d.LP
           << ",\"vrow\":" << d.vrow.d << ",\"maxt\":" << ln.maxt
           << ",\"E\":" << ln.E << ",\"fixed\":" << (ln.fixed ? 1 : 0) << ",\"mfma\":" << ln.mfma << ",\"grid\":" << ln.grid << ",\"block\":" << ln.block
           << ",\"lds\":" << ln.lds_bytes << ",\"in_role\":" << (int)ln.in_role
           << ",\"out_role\":" << (int)ln.out_role << ",\"in_off\":" << ln.in_off_bytes
           << ",\"out_off\":" << ln.out_off_bytes << ",\"tw4_n\":" << d.tw4_n << ",\"nchirp\":" << d.nchirp
           << ",\"gin\":[" << d.gin[0] << "," << d.gin[1] << "," << d.gin[2] << "," << d.gin[3] << "],\"gout\":["
           << d.gout[0] << "," << d.gout[1] << "," << d.gout[2] << "," << d.gout[3] << "],\"gin_base\":" << d.gin_base
           << ",\"gout_base\":" << d.gout_base << ",\"radix\":[";
        for (int s = 0; s < d.nst; s++) os << (s ? "," : "") << d.radix[s];
        os << "],\"ls\":[";
        for (int s = 0; s < d.nst; s++) os << (s ? "," : "") << d.ls[s];
        os << "]}";
        first = false;
    }
    os << "]}";
    return os.str();
}

static void slot_free(Slot* s)
{
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    if (s->work) (void)hipFree(s->work);
    if (s->stage_in) (void)hipFree(s->stage_in);
    if (s->stage_out) (void)hipFree(s->stage_out);
    if (s->counters) (void)hipFree(s->counters);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

void plan_destroy(Plan* p)
{
    if (!p) return;
    for (Slot* s : p->slots) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        slot_free(s);
    }
    if (p->tables) (void)hipFree(p->tables);
    for (hipEvent_t e : p->prof_events) (void)hipEventDestroy(e);
    for (hipEvent_t e : p->prof_pool) (void)hipEventDestroy(e);
    delete p;
}

Slot* plan_acquire_slot(Plan* p, hipError_t* err)
{
    std::lock_guard<std::mutex> g(p->mu);
    *err = hipSuccess;
    for (Slot* s : p->slots)
        if (!s->busy) {
            s->busy = true;
            return s;
        }
    Slot* s = new Slot();
    // a blocking stream: a synchronous execute is ordered after work still
    // pending on the null stream (where most callers produced their buffers)
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamDefault);
    if (e == hipSuccess && p->work_bytes) {
        e = hipMalloc(&s->work, p->work_bytes);
        s->work_bytes = p->work_bytes;
    }
    if (e == hipSuccess && p->is_batch) e = hipMalloc((void**)&s->counters, 64);
    if (e != hipSuccess) {
        *err = e;
        slot_free(s);
        return nullptr;
    }
    s->busy = true;
    p->slots.push_back(s);
    return s;
}

void plan_release_slot(Plan* p, Slot* s)
{
    std::lock_guard<std::mutex> g(p->mu);
    s->busy = false;
}

static hipError_t record_prof(Plan* p, hipStream_t stream)
{
    hipEvent_t ev;
    {
        std::lock_guard<std::mutex> g(p->mu);
        if (p->prof_pool.empty()) {
            hipError_t e = hipEventCreate(&ev);
            if (e != hipSuccess) return e;
        } else {
            ev = p->prof_pool.back();
            p->prof_pool.pop_back();
        }
        p->prof_events.push_back(ev);
    }
    return hipEventRecord(ev, stream);
}

// Launches of one execute.  work: the work buffer (four-step and multi-axis
// C2R intermediates), counters: the uneven batch's queue counters; prof:
// record timing events (which allocates them on first use).
static hipError_t launch_batch(Plan* p, unsigned* counters, const void* in, void* out, hipStream_t stream,
                               bool prof)
{
    // counters are cleared by a kernel (not a memset node) so the sequence
    // can be captured in a graph and replayed
    int nc = (int)p->bclasses.size();
    void* rargs[] = {&counters, &nc};
    hipError_t e = hipLaunchKernel(kernel_symbol_reset(), dim3(1), dim3(64), rargs, 0, stream);
    if (e != hipSuccess) return e;
    if (prof && (e = record_prof(p, stream)) != hipSuccess) return e;
    const char* base = (const char*)p->tables;
    for (size_t c = 0; c < p->bclasses.size(); c++) {
        const BatchClass& bc = p->bclasses[c];
        if (bc.ntiles > 0) {
            BatchArgs a;
            a.descs = reinterpret_cast<const PassDesc*>(base + p->descs_off);
            a.tiles = reinterpret_cast<const BatchTile*>(base + bc.tiles_off);
            a.ntiles = bc.ntiles;
            a.counter = counters + c;
            a.in = in;
            a.out = out;
            void* args[] = {&a};
            e = hipLaunchKernel(kernel_symbol_batch(p->spec.prec_bytes, bc.E, bc.maxt), dim3(bc.grid),
                                dim3(bc.maxt), args, bc.lds_bytes, stream);
            if (e != hipSuccess) return e;
        }
        if (prof && (e = record_prof(p, stream)) != hipSuccess) return e;
    }
    return hipGetLastError();
}

static hipError_t launch_all(Plan* p, void* work, unsigned* counters, const void* in, void* out, hipStream_t stream,
                             bool prof)
{
    if (p->is_batch) return launch_batch(p, counters, in, out, stream, prof);
    if (prof) {
        hipError_t e = record_prof(p, stream);
        if (e != hipSuccess) return e;
    }
    for (const Launch& ln : p->launches) {
        PassDesc d = ln.d;
        auto role_ptr = [&](Role r) -> char* {
            switch (r) {
                case Role::In: return (char*)in;
                case Role::Out: return (char*)out;
                default: return (char*)work;
            }
        };
        d.in = role_ptr(ln.in_role) + ln.in_off_bytes;
        d.out = role_ptr(ln.out_role) + ln.out_off_bytes;
        const float* frag = ln.frag;
        void* args[] = {&d, &frag};
        const void* fn = ln.mfma ? kernel_symbol_mfma(ln.mfma, ln.d.L, ln.mfma_nb)
                         : ln.fixed ? kernel_symbol_fixed(p->spec.prec_bytes, ln.d.L, ln.maxt)
                                    : kernel_symbol(ln.kind, p->spec.prec_bytes, ln.all_radix, ln.E, ln.maxt);
        hipError_t e = hipLaunchKernel(fn, dim3(ln.grid),
                                       dim3(ln.block), args, ln.lds_bytes, stream);
        if (e != hipSuccess) return e;
        if (prof) {
            e = record_prof(p, stream);
            if (e != hipSuccess) return e;
        }
    }
    return hipGetLastError();
}

hipError_t plan_enqueue_scratch(Plan* p, void* work, unsigned* counters, const void* in, void* out,
                                hipStream_t stream)
{
    if (p->noop) return hipSuccess;
    return launch_all(p, work, counters, in, out, stream, false);
}

int plan_profile_read(Plan* p, float* ms, int cap)
{
    std::vector<hipEvent_t> evs;
    {
        std::lock_guard<std::mutex> g(p->mu);
        evs.swap(p->prof_events);
    }
    const size_t per = (p->is_batch ? p->bclasses.size() : p->launches.size()) + 1;  // events per execute
    int n = 0;
    for (size_t base = 0; base + per <= evs.size(); base += per) {
        for (size_t i = 0; i + 1 < per; i++) {
            float t = 0.f;
            if (hipEventSynchronize(evs[base + i + 1]) != hipSuccess ||
                hipEventElapsedTime(&t, evs[base + i], evs[base + i + 1]) != hipSuccess)
                t = -1.f;
            if (ms && n < cap) ms[n] = t;
            n++;
        }
    }
    std::lock_guard<std::mutex> g(p->mu);
    for (hipEvent_t e : evs) p->prof_pool.push_back(e);
    return n;
}

hipError_t plan_enqueue(Plan* p, Slot* s, const void* in, void* out, hipStream_t stream)
{
    if (p->noop) return hipSuccess;
    if (!p->use_graph || p->profile) return launch_all(p, s->work, s->counters, in, out, stream, p->profile);
    // One hipGraph per (slot, in, out): captured once, replayed afterwards.
    if (!s->graph || s->graph_in != in || s->graph_out != out) {
        if (s->graph) {
            (void)hipGraphExecDestroy(s->graph);
            s->graph = nullptr;
        }
        hipGraph_t g;
        hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        e = launch_all(p, s->work, s->counters, in, out, s->stream, false);
        hipError_t e2 = hipStreamEndCapture(s->stream, &g);
        if (e != hipSuccess) return e;
        if (e2 != hipSuccess) return e2;
        e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return e;
        s->graph_in = in;
        s->graph_out = out;
    }
    return hipGraphLaunch(s->graph, stream);
}

}  // namespace fftz

namespace fftz {

double plan_time(Plan* p, const void* in, void* out, int reps)
{
    hipError_t e = hipSuccess;
    Slot* s = plan_acquire_slot(p, &e);
    if (!s) return -1.0;
    hipEvent_t a = nullptr, b = nullptr;
    double ms = -1.0;
    if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
        e = plan_enqueue(p, s, in, out, s->stream);  // warm (and graph capture)
        if (e == hipSuccess) e = hipEventRecord(a, s->stream);
        for (int r = 0; r < reps && e == hipSuccess; r++) e = plan_enqueue(p, s, in, out, s->stream);
        if (e == hipSuccess) e = hipEventRecord(b, s->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float t = 0.f;
        if (e == hipSuccess && hipEventElapsedTime(&t, a, b) == hipSuccess) ms = (double)t / reps;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (e != hipSuccess) (void)hipGetLastError();
    plan_release_slot(p, s);
    return ms;
}

}  // namespace fftz
