// Mixed-radix (2,3,4,5) complex FFT on data held in LDS, for gfx950.
//
// Stockham autosort, in place: each stage has every thread compute its butterflies into
// registers, barrier, write back, barrier.  Transforms are grouped into rounds so that a
// round holds whole transforms (in-place stages would otherwise overwrite inputs another
// round still needs).  Twiddles come from a table tw[k] = exp(-2 pi i k / n) (host-computed
// in long double, rounded once); sign = -1 forward (e^{-i}), +1 backward (e^{+i}), no scaling
// -- the FFTW convention the reference uses (fft_grid_serial.hpp:209-223).
#pragma once
#include <hip/hip_runtime.h>

namespace q3x {

template <typename T>
struct cx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) { return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by i*s (s = +-1)
template <typename T>
__device__ __forceinline__ cx<T> cmuli(cx<T> a, T s) { return {-a.y * s, a.x * s}; }

struct FftPlan {
    int n;
    int nstage;
    int radix[8];
};

// Small DFTs: y_q = sum_r x_r exp(sgn * 2 pi i q r / R)
template <typename T, int R>
__device__ __forceinline__ void dft(cx<T>* v, T sgn);

template <typename T>
struct DftConst {
    static constexpr T s3 = T(0.86602540378443864676);   // sin(2pi/3)
    static constexpr T c51 = T(0.30901699437494742410);  // cos(2pi/5)
    static constexpr T c52 = T(-0.80901699437494742410); // cos(4pi/5)
    static constexpr T s51 = T(0.95105651629515357212);  // sin(2pi/5)
    static constexpr T s52 = T(0.58778525229247312917);  // sin(4pi/5)
};

template <>
__device__ __forceinline__ void dft<double, 2>(cx<double>* v, double) {
    cx<double> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void dft<float, 2>(cx<float>* v, float) {
    cx<float> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename T>
__device__ __forceinline__ void dft4(cx<T>* v, T sgn) {
    cx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    cx<T> c = cadd(v[1], v[3]), d = cmuli(csub(v[1], v[3]), sgn);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
}
template <>
__device__ __forceinline__ void dft<double, 4>(cx<double>* v, double s) { dft4(v, s); }
template <>
__device__ __forceinline__ void dft<float, 4>(cx<float>* v, float s) { dft4(v, s); }

template <typename T>
__device__ __forceinline__ void dft3(cx<T>* v, T sgn) {
    cx<T> t1 = cadd(v[1], v[2]);
    cx<T> t2 = csub(v[1], v[2]);
    cx<T> m = {v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
    cx<T> r = cmuli(t2, sgn * DftConst<T>::s3);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}
template <>
__device__ __forceinline__ void dft<double, 3>(cx<double>* v, double s) { dft3(v, s); }
template <>
__device__ __forceinline__ void dft<float, 3>(cx<float>* v, float s) { dft3(v, s); }

template <typename T>
__device__ __forceinline__ void dft5(cx<T>* v, T sgn) {
    using K = DftConst<T>;
    cx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    cx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    cx<T> a1 = {v[0].x + K::c51 * t1.x + K::c52 * t2.x, v[0].y + K::c51 * t1.y + K::c52 * t2.y};
    cx<T> a2 = {v[0].x + K::c52 * t1.x + K::c51 * t2.x, v[0].y + K::c52 * t1.y + K::c51 * t2.y};
    cx<T> b1 = {sgn * (K::s51 * t3.x + K::s52 * t4.x), sgn * (K::s51 * t3.y + K::s52 * t4.y)};
    cx<T> b2 = {sgn * (K::s52 * t3.x - K::s51 * t4.x), sgn * (K::s52 * t3.y - K::s51 * t4.y)};
    v[0] = cadd(v[0], cadd(t1, t2));
    cx<T> ib1 = cmuli(b1, T(1)), ib2 = cmuli(b2, T(1));
    v[1] = cadd(a1, ib1);
    v[4] = csub(a1, ib1);
    v[2] = cadd(a2, ib2);
    v[3] = csub(a2, ib2);
}
template <>
__device__ __forceinline__ void dft<double, 5>(cx<double>* v, double s) { dft5(v, s); }
template <>
__device__ __forceinline__ void dft<float, 5>(cx<float>* v, float s) { dft5(v, s); }

// One Stockham stage of radix R over ntrans transforms of length n stored at
// buf[t * ts + e * es].  BPT = butterflies per thread per round.
template <typename T, int R, int BPT>
__device__ void fft_stage(cx<T>* buf, int n, int ns, const cx<T>* tw, T sgn, int ntrans, int es, int ts,
                          int tid, int nthr) {
    const int m = n / R;
    int per_round = (nthr * BPT) / m;
    if (per_round < 1) per_round = 1;  // host guarantees m <= nthr * BPT
    const int twstep = n / (ns * R);
    for (int t0 = 0; t0 < ntrans; t0 += per_round) {
        const int tcount = min(per_round, ntrans - t0);
        const int work = tcount * m;
        cx<T> v[BPT][R];
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int w = tid + b * nthr;
            if (w < work) {
                const int t = t0 + w / m, j = w % m;
                const int jn = j % ns;
                const cx<T>* src = buf + t * ts;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    cx<T> x = src[(j + r * m) * es];
                    const int e = r * jn * twstep;
                    if (e) {
                        const cx<T> we = tw[e];
                        x = sgn < T(0) ? cmul(x, we) : cmulc(x, we);
                    }
                    v[b][r] = x;
                }
                dft<T, R>(v[b], sgn);
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int w = tid + b * nthr;
            if (w < work) {
                const int t = t0 + w / m, j = w % m;
                const int jn = j % ns;
                cx<T>* dst = buf + t * ts;
                const int base = (j / ns) * ns * R + jn;
#pragma unroll
                for (int q = 0; q < R; ++q) dst[(base + q * ns) * es] = v[b][q];
            }
        }
        __syncthreads();
    }
}

template <typename T, int BPT>
__device__ void fft_lds(cx<T>* buf, const FftPlan& p, const cx<T>* tw, int sign, int ntrans, int es, int ts,
                        int tid, int nthr) {
    const T sgn = sign < 0 ? T(-1) : T(1);
    int ns = 1;
    for (int s = 0; s < p.nstage; ++s) {
        const int R = p.radix[s];
        switch (R) {
            case 4: fft_stage<T, 4, BPT>(buf, p.n, ns, tw, sgn, ntrans, es, ts, tid, nthr); break;
            case 2: fft_stage<T, 2, BPT>(buf, p.n, ns, tw, sgn, ntrans, es, ts, tid, nthr); break;
            case 3: fft_stage<T, 3, BPT>(buf, p.n, ns, tw, sgn, ntrans, es, ts, tid, nthr); break;
            case 5: fft_stage<T, 5, BPT>(buf, p.n, ns, tw, sgn, ntrans, es, ts, tid, nthr); break;
            default: break;
        }
        ns *= R;
    }
}

}  // namespace q3x
