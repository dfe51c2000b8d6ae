// Register + LDS Stockham FFT of one complex line of length N (power of two,
// 16 <= N <= 4096), N/E threads per line, E = 8 or 16 elements per thread.
//
// Passes are radix E (radix 16 as a 4x4 decomposition in registers) except
// the last, which is radix 2, 4 or 8 when log2(N) is not a multiple of
// log2(E).  Before the first and after the last pass thread t holds the
// elements at positions t + m*N/E, m = 0..E-1 ("strided" layout), so loads
// and stores of whole lines are coalesced and pointwise work between a
// forward and an inverse transform needs no data exchange.  Between passes
// the line goes through LDS as separate real and imaginary arrays, padded
// (fft_lds) so that both the E-strided writes and the unit-stride reads of
// every exchange are free of bank conflicts and every thread addresses its
// E words from one base register with immediate offsets.
//
// Twiddles come from a table of exp(-2 pi i t/N) (FP64, host-computed in
// long double): per butterfly the FP64 transform reads one entry and forms
// the rest by squarings and complex products; the FP32 transform reads
// three (four for radix 16).
#pragma once

#include <hip/hip_runtime.h>

namespace mhd2d {

template <typename T>
struct cx {
    T re, im;
};
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> w)
{
    return {a.re * w.re - a.im * w.im, a.re * w.im + a.im * w.re};
}
template <typename T>
__device__ __forceinline__ cx<T> mul_mi(cx<T> a) { return {a.im, -a.re}; }  // times -i
template <typename T>
__device__ __forceinline__ void bfly(cx<T>& a, cx<T>& b)
{
    const cx<T> t = a;
    a = t + b;
    b = t - b;
}
// v * (i s) and v * s for a real s
template <typename T>
__device__ __forceinline__ cx<T> mul_i_scal(cx<T> a, T s) { return {-a.im * s, a.re * s}; }
template <typename T>
__device__ __forceinline__ cx<T> mul_scal(cx<T> a, T s) { return {a.re * s, a.im * s}; }

// ---- small in-register DFTs, forward sign, natural order in and out
template <typename T>
__device__ __forceinline__ void dft2(cx<T>* v) { bfly(v[0], v[1]); }

template <typename T>
__device__ __forceinline__ void dft4(cx<T>* v)
{
    bfly(v[0], v[2]);
    bfly(v[1], v[3]);
    v[3] = mul_mi(v[3]);
    bfly(v[0], v[1]);
    bfly(v[2], v[3]);
    const cx<T> t = v[1];  // bit reversal: (0,2,1,3) -> natural
    v[1] = v[2];
    v[2] = t;
}

template <typename T>
__device__ __forceinline__ void dft8(cx<T>* v)
{
    const T s = (T)0.70710678118654752440084436210485;
    bfly(v[0], v[4]);
    bfly(v[1], v[5]);
    bfly(v[2], v[6]);
    bfly(v[3], v[7]);
    v[5] = cx<T>{s * (v[5].re + v[5].im), s * (v[5].im - v[5].re)};   // * e^{-i pi/4}
    v[6] = mul_mi(v[6]);                                               // * e^{-i pi/2}
    v[7] = cx<T>{s * (v[7].im - v[7].re), -s * (v[7].re + v[7].im)};  // * e^{-3i pi/4}
    bfly(v[0], v[2]);
    bfly(v[1], v[3]);
    bfly(v[4], v[6]);
    bfly(v[5], v[7]);
    v[3] = mul_mi(v[3]);
    v[7] = mul_mi(v[7]);
    bfly(v[0], v[1]);
    bfly(v[2], v[3]);
    bfly(v[4], v[5]);
    bfly(v[6], v[7]);
    cx<T> t = v[1];  // bit reversal
    v[1] = v[4];
    v[4] = t;
    t = v[3];
    v[3] = v[6];
    v[6] = t;
}

// 16-point DFT as 4 x 4: X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1)
//                                        sum_n1 W4^(n1 k1) x[4 n1 + n2]
template <typename T>
__device__ __forceinline__ void dft16(cx<T>* v)
{
    const T c = (T)0.92387953251128675612818318939678829;  // cos(pi/8)
    const T s = (T)0.38268343236508977172845998403039887;  // sin(pi/8)
    const T h = (T)0.70710678118654752440084436210484904;  // cos(pi/4)
    cx<T> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        cx<T> a[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
        dft4(a);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) y[n2][k1] = a[k1];
    }
    // W16^(n2 k1) for n2, k1 in 1..3 (exponents 1,2,3,2,4,6,3,6,9)
    y[1][1] = cmul(y[1][1], cx<T>{c, -s});
    y[1][2] = cmul(y[1][2], cx<T>{h, -h});
    y[1][3] = cmul(y[1][3], cx<T>{s, -c});
    y[2][1] = cmul(y[2][1], cx<T>{h, -h});
    y[2][2] = mul_mi(y[2][2]);
    y[2][3] = cmul(y[2][3], cx<T>{-h, -h});
    y[3][1] = cmul(y[3][1], cx<T>{s, -c});
    y[3][2] = cmul(y[3][2], cx<T>{-h, -h});
    y[3][3] = cmul(y[3][3], cx<T>{-c, s});
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        cx<T> b[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        dft4(b);
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = b[k2];
    }
}

template <typename T, int R>
__device__ __forceinline__ void dftR(cx<T>* v)
{
    if constexpr (R == 16) dft16(v);
    else if constexpr (R == 8) dft8(v);
    else if constexpr (R == 4) dft4(v);
    else dft2(v);
}

constexpr int fft_ilog2(int n) { return n <= 1 ? 0 : 1 + fft_ilog2(n / 2); }
constexpr int fft_ipow(int b, int p) { return p == 0 ? 1 : b * fft_ipow(b, p - 1); }

// E elements per thread (8 or 16): passes of radix E, then one of radix
// 2^(log2 N mod log2 E) if needed.
template <int N, int E>
struct FFTShape {
    static constexpr int M = fft_ilog2(N);
    static constexpr int LE = fft_ilog2(E);
    static constexpr int PE = M / LE, REM = M % LE;
    static constexpr int NP = PE + (REM ? 1 : 0);
    static constexpr int TPR = N / E;          // threads per line
    static constexpr int LDS = N + N / E;      // words per array (padded)
    static constexpr int radix(int p) { return p < PE ? E : (1 << REM); }
};

// LDS word of line position p in the exchange that follows a pass with
// sub-transform length NS (1, E, E^2, ...).  The LDS serves G lanes per
// cycle: 16 lanes of 8-byte words (FP64) or 32 lanes of 4-byte words (FP32),
// one word per slot of G.  Such a group writes positions
// (t / NS) NS E + t % NS + r NS: runs of NS consecutive positions E NS
// apart (stride E for NS == 1), and reads G consecutive positions.
//   NS == 1:  p + p / G          stride-E writes land on distinct slots
//   NS <  G:  p + NS p / (E NS)  the G / NS runs of a group shift apart
//   NS >= G:  p                  writes are already consecutive
// Every term is additive in the per-thread base and the compile-time
// offsets r NS (writes) and b N/E + r N/R (reads), so each thread forms one
// address per exchange side.  Injective, below N + N/E, conflict-free and
// additive for every exchange of N = 16..4096 with E = 8 or 16 (checked
// exhaustively by the CPU tests).
template <int E, int NS, int G>
__host__ __device__ constexpr int fft_lds(int p)
{
    constexpr int LG = fft_ilog2(G);
    constexpr int LEN = fft_ilog2(E * NS);
    if constexpr (NS == 1) return p + (p >> LG);
    else if constexpr (NS < G) return p + ((p >> LEN) * NS);
    else return p;
}

// w^2 of a unit complex number: (re + im)(re - im) keeps cos 2a accurate
template <typename T>
__device__ __forceinline__ cx<T> csq(cx<T> w)
{
    return {(w.re + w.im) * (w.re - w.im), (T)2 * (w.re * w.im)};
}

// v[r] *= W^(r k STEP), r = 1..R-1 (idx1 = k STEP).  FP64: one table read
// per butterfly, W^2k and W^4k by squaring and the rest by complex
// products (a few ulp; the table reads were a quarter of the y-sweep's
// transform time); FP32: W, W^2, W^4 (and W^8) from the FP64 table.
template <typename T, int R>
__device__ __forceinline__ void twiddle(cx<T>* v, int idx1, const double2* __restrict__ tw)
{
    // The factors are the same for every transform of the line; an opaque
    // index keeps the compiler from hoisting them all (tens of registers)
    // out of the callers' loops.  The empty asm only moves a register.
    asm volatile("" : "+v"(idx1));
    constexpr bool SQ = sizeof(T) == 8;
    const double2 a1 = tw[idx1];
    const cx<T> w1{(T)a1.x, (T)a1.y};
    v[1] = cmul(v[1], w1);
    if constexpr (R > 2) {
        cx<T> w2;
        if constexpr (SQ) {
            w2 = csq(w1);
        } else {
            const double2 a2 = tw[2 * idx1];
            w2 = cx<T>{(T)a2.x, (T)a2.y};
        }
        const cx<T> w3 = cmul(w2, w1);
        v[2] = cmul(v[2], w2);
        v[3] = cmul(v[3], w3);
        if constexpr (R > 4) {
            cx<T> w4;
            if constexpr (SQ) {
                w4 = csq(w2);
            } else {
                const double2 a4 = tw[4 * idx1];
                w4 = cx<T>{(T)a4.x, (T)a4.y};
            }
            v[4] = cmul(v[4], w4);
            v[5] = cmul(v[5], cmul(w4, w1));
            v[6] = cmul(v[6], cmul(w4, w2));
            v[7] = cmul(v[7], cmul(w4, w3));
            if constexpr (R > 8) {
                const double2 a8 = tw[8 * idx1];
                const cx<T> w8{(T)a8.x, (T)a8.y};
                v[8] = cmul(v[8], w8);
                v[9] = cmul(v[9], cmul(w8, w1));
                v[10] = cmul(v[10], cmul(w8, w2));
                v[11] = cmul(v[11], cmul(w8, w3));
                const cx<T> w12 = cmul(w8, w4);
                v[12] = cmul(v[12], w12);
                v[13] = cmul(v[13], cmul(w12, w1));
                v[14] = cmul(v[14], cmul(w12, w2));
                v[15] = cmul(v[15], cmul(w12, w3));
            }
        }
    }
}

template <typename T, int N, int E, int P>
__device__ __forceinline__ void fft_pass(cx<T> (&v)[E], int t, T* lre, T* lim,
                                         const double2* __restrict__ tw)
{
    using S = FFTShape<N, E>;
    constexpr int R = S::radix(P);
    constexpr int B = E / R;  // butterflies per thread
    constexpr int TPR = S::TPR;
    constexpr int NS = fft_ipow(E, P);
#pragma unroll
    for (int b = 0; b < B; b++) {
        if constexpr (NS > 1) {
            const int j = t + b * TPR;
            constexpr int STEP = N / (NS * R);
            twiddle<T, R>(&v[b * R], (j & (NS - 1)) * STEP, tw);
        }
        dftR<T, R>(&v[b * R]);
    }
    if constexpr (P + 1 < S::NP) {
        // intermediate passes are radix E (B == 1); one base word per side,
        // the rest are immediate offsets.  The bases are recomputed from an
        // opaque t in every exchange instead of being hoisted into registers
        // for the whole kernel.
        asm volatile("" : "+v"(t));
        constexpr int R2 = S::radix(P + 1);
        constexpr int B2 = E / R2;
        constexpr int G = 128 / sizeof(T);  // lanes per LDS cycle
        T* const wre = lre + fft_lds<E, NS, G>((t / NS) * NS * R + (t & (NS - 1)));
        T* const wim = lim + (wre - lre);
#pragma unroll
        for (int r = 0; r < R; r++) {
            wre[fft_lds<E, NS, G>(r * NS)] = v[r].re;
            wim[fft_lds<E, NS, G>(r * NS)] = v[r].im;
        }
        __syncthreads();
        const T* const rre = lre + fft_lds<E, NS, G>(t);
        const T* const rim = lim + fft_lds<E, NS, G>(t);
#pragma unroll
        for (int b = 0; b < B2; b++) {
#pragma unroll
            for (int r = 0; r < R2; r++) {
                const int o = fft_lds<E, NS, G>(b * TPR + r * (N / R2));
                v[b * R2 + r] = cx<T>{rre[o], rim[o]};
            }
        }
        __syncthreads();
        fft_pass<T, N, E, P + 1>(v, t, lre, lim, tw);
    } else if constexpr (B > 1) {
        // last pass: element (b, r) sits at t + TPR * (b + r * B)
        cx<T> tmp[E];
#pragma unroll
        for (int m = 0; m < E; m++) tmp[m] = v[(m % B) * R + m / B];
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = tmp[m];
    }
}

template <typename T, int N, int E>
__device__ __forceinline__ void fft_fwd(cx<T> (&v)[E], int t, T* lre, T* lim,
                                        const double2* __restrict__ tw)
{
    fft_pass<T, N, E, 0>(v, t, lre, lim, tw);
}

// unnormalised inverse: conj(FFT(conj(x)))
template <typename T, int N, int E>
__device__ __forceinline__ void fft_inv(cx<T> (&v)[E], int t, T* lre, T* lim,
                                        const double2* __restrict__ tw)
{
#pragma unroll
    for (int m = 0; m < E; m++) v[m].im = -v[m].im;
    fft_pass<T, N, E, 0>(v, t, lre, lim, tw);
#pragma unroll
    for (int m = 0; m < E; m++) v[m].im = -v[m].im;
}

}  // namespace mhd2d
