// In-register DFT codelets for CDNA4 (gfx950).
//
// Each codelet computes the forward DFT  y[k] = sum_j v[j] exp(-2 pi i j k / R)
// in place on an array of R complex values held in VGPRs.  The backward
// transform is never coded separately: the passes conjugate on load and on
// store (conj(FFT(conj(x))) = unnormalised inverse), the same identity the
// reference uses when it swaps real/imaginary pointers for the backward
// codelets (core/kernels/dft/c2c/c/fft8c.c:51-63).
//
// Prime radices use the symmetric-pair form (about R^2/2 real FMAs); composite
// radices are Cooley-Tukey splits done entirely in registers with constant
// twiddles, so every multiply by a trivial root (1, -i, -1, i, (1-i)/sqrt2 ...)
// folds away at compile time.
#pragma once

#include <hip/hip_runtime.h>

#include "root_tables.h"

namespace fftz {

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cplx<R> cconj(cplx<R> a) { return {a.x, -a.y}; }
template <typename R>
__device__ __forceinline__ cplx<R> cscale(cplx<R> a, R s) { return {a.x * s, a.y * s}; }
// a * (-i)
template <typename R>
__device__ __forceinline__ cplx<R> mul_mi(cplx<R> a) { return {a.y, -a.x}; }

// a * exp(-2 pi i m / N), m and N compile-time after unrolling.
template <int N, typename R>
__device__ __forceinline__ cplx<R> rot(cplx<R> a, int m_in)
{
    const int m = ((m_in % N) + N) % N;
    if (m == 0) return a;
    if (4 * m == N) return {a.y, -a.x};
    if (2 * m == N) return {-a.x, -a.y};
    if (4 * m == 3 * N) return {-a.y, a.x};
    const R h = R(0.707106781186547524400844362104849039);
    if (8 * m == N) return {(a.x + a.y) * h, (a.y - a.x) * h};
    if (8 * m == 3 * N) return {(a.y - a.x) * h, -(a.x + a.y) * h};
    if (8 * m == 5 * N) return {-(a.x + a.y) * h, (a.x - a.y) * h};
    if (8 * m == 7 * N) return {(a.x - a.y) * h, (a.x + a.y) * h};
    const R c = R(root_cos<N>(m));
    const R s = R(-root_sin<N>(m));
    return {a.x * c - a.y * s, a.x * s + a.y * c};
}

template <int R_, typename R>
struct Dft;

template <typename R>
struct Dft<2, R> {
    static __device__ __forceinline__ void run(cplx<R>* v)
    {
        cplx<R> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename R>
struct Dft<3, R> {
    static __device__ __forceinline__ void run(cplx<R>* v)
    {
        const R c = R(0.5), s = R(0.866025403784438646763723170752936183);
        cplx<R> sm = cadd(v[1], v[2]);
        cplx<R> df = csub(v[1], v[2]);
        cplx<R> y0 = cadd(v[0], sm);
        cplx<R> m = {v[0].x - c * sm.x, v[0].y - c * sm.y};
        // -i * s * df
        cplx<R> t = {s * df.y, -s * df.x};
        v[0] = y0;
        v[1] = cadd(m, t);
        v[2] = csub(m, t);
    }
};

template <typename R>
struct Dft<4, R> {
    static __device__ __forceinline__ void run(cplx<R>* v)
    {
        cplx<R> t0 = cadd(v[0], v[2]);
        cplx<R> t1 = csub(v[0], v[2]);
        cplx<R> t2 = cadd(v[1], v[3]);
        cplx<R> t3 = mul_mi(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

// Odd prime P: pairs s_j = v_j + v_{P-j}, d_j = v_j - v_{P-j};
// y_k = v_0 + sum_j cos(2 pi jk/P) s_j  -/+  i sum_j sin(2 pi jk/P) d_j.
template <int P, typename R>
struct DftPrime {
    static __device__ __forceinline__ void run(cplx<R>* v)
    {
        constexpr int H = (P - 1) / 2;
        cplx<R> s[H], d[H];
#pragma unroll
        for (int j = 1; j <= H; j++) {
            s[j - 1] = cadd(v[j], v[P - j]);
            d[j - 1] = csub(v[j], v[P - j]);
        }
        cplx<R> y0 = v[0];
#pragma unroll
        for (int j = 0; j < H; j++) y0 = cadd(y0, s[j]);
        cplx<R> out[P];
        out[0] = y0;
#pragma unroll
        for (int k = 1; k <= H; k++) {
            cplx<R> a = v[0];
            cplx<R> b = {R(0), R(0)};
#pragma unroll
            for (int j = 1; j <= H; j++) {
                const R c = R(root_cos<P>(j * k));
                const R sn = R(root_sin<P>(j * k));
                a.x += c * s[j - 1].x;
                a.y += c * s[j - 1].y;
                b.x += sn * d[j - 1].x;
                b.y += sn * d[j - 1].y;
            }
            // y_k = a - i b ; y_{P-k} = a + i b
            out[k] = {a.x + b.y, a.y - b.x};
            out[P - k] = {a.x - b.y, a.y + b.x};
        }
#pragma unroll
        for (int k = 0; k < P; k++) v[k] = out[k];
    }
};

template <typename R> struct Dft<5, R> : DftPrime<5, R> {};
template <typename R> struct Dft<7, R> : DftPrime<7, R> {};
template <typename R> struct Dft<11, R> : DftPrime<11, R> {};
template <typename R> struct Dft<13, R> : DftPrime<13, R> {};

// Composite R = R1 * R2 in registers:
//   u[j2][k1] = DFT_R1 over j1 of v[R2 j1 + j2];  u *= w_R^{j2 k1};
//   y[k1 + R1 k2] = DFT_R2 over j2 of u[j2][k1].
template <int R1, int R2, typename R>
struct DftCT {
    static constexpr int N = R1 * R2;
    static __device__ __forceinline__ void run(cplx<R>* v)
    {
        cplx<R> u[R2][R1];
#pragma unroll
        for (int j2 = 0; j2 < R2; j2++) {
#pragma unroll
            for (int j1 = 0; j1 < R1; j1++) u[j2][j1] = v[R2 * j1 + j2];
            Dft<R1, R>::run(u[j2]);
#pragma unroll
            for (int k1 = 0; k1 < R1; k1++) u[j2][k1] = rot<N>(u[j2][k1], j2 * k1);
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) {
            cplx<R> w[R2];
#pragma unroll
            for (int j2 = 0; j2 < R2; j2++) w[j2] = u[j2][k1];
            Dft<R2, R>::run(w);
#pragma unroll
            for (int k2 = 0; k2 < R2; k2++) v[k1 + R1 * k2] = w[k2];
        }
    }
};

template <typename R> struct Dft<6, R> : DftCT<3, 2, R> {};
template <typename R> struct Dft<8, R> : DftCT<4, 2, R> {};
template <typename R> struct Dft<9, R> : DftCT<3, 3, R> {};
template <typename R> struct Dft<10, R> : DftCT<5, 2, R> {};
template <typename R> struct Dft<12, R> : DftCT<4, 3, R> {};
template <typename R> struct Dft<16, R> : DftCT<4, 4, R> {};

}  // namespace fftz
