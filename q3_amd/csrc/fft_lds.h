// Mixed-radix complex FFT on data held in LDS, for gfx950.
//
// Stockham autosort, in place.  A stage of radix R (R in {2,3,4,5,6,8}; larger composite
// butterflies exist below but push FP64 kernels past the register file) maps
// thread tid to butterfly j = tid mod m (m = n/R) and transform group tid / m; the index
// arithmetic and the R-1 twiddles of a thread are computed once per stage, and the butterfly
// itself is an unrolled register DFT (composite radices as two levels of small DFTs with
// constant inner twiddles).  Every thread loads its butterflies of a round into registers,
// barrier, writes back, barrier; a round holds whole transforms only.
// Twiddles come from a stage-major table (FftPlan::twoff; host-computed in long double);
// sign = -1 forward (e^{-i}), +1 backward (e^{+i}), no scaling -- the FFTW convention the
// reference uses (fft_grid_serial.hpp:209-223).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_twiddle_consts.h"

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

// Stages in order with their radices; the twiddles of stage s (stride ns = product of the
// earlier radices) are stored stage-major at twoff[s] as [r-1][jn] = exp(-2 pi i r jn / (ns R)),
// so that neighbouring butterflies read neighbouring entries.  The table holds n - 1 entries.
struct FftPlan {
    int n;
    int nstage;
    int radix[8];
    int twoff[8];
};

// LDS padding of a transform's storage: element e lives at e + (e >> PAD).  The first
// Stockham stages write with a stride of R elements, and transposed accesses walk across
// transforms; without padding both hit the same banks on every lane.  One pad slot per 128
// bytes of data: PAD = 3 for complex double (16 B), 4 for complex float; PAD = 31 disables it.
template <int PAD>
__device__ __forceinline__ int pidx(int e) {
    return e + (e >> PAD);
}
template <typename T>
constexpr int lds_pad_shift() {
    return sizeof(T) == 8 ? 3 : 4;
}
// storage length of one padded n-point transform; the extra slot staggers consecutive
// transforms across banks
__host__ __device__ constexpr int padded_len(int n, int shift) { return n + (n >> shift) + 1; }

// ---- small prime-size DFTs: y_q = sum_r x_r exp(sgn * 2 pi i q r / R), in place
template <typename T>
__device__ __forceinline__ void dft2(cx<T>* v) {
    const cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <typename T>
__device__ __forceinline__ void dft3(cx<T>* v, T sgn) {
    const T s3 = T(0.86602540378443864676);
    const cx<T> t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    const cx<T> m = {v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
    const cx<T> r = cmuli(t2, sgn * s3);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}
template <typename T>
__device__ __forceinline__ void dft4(cx<T>* v, T sgn) {
    const cx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    const cx<T> c = cadd(v[1], v[3]), d = cmuli(csub(v[1], v[3]), sgn);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
}
template <typename T>
__device__ __forceinline__ void dft5(cx<T>* v, T sgn) {
    const T c51 = T(0.30901699437494742410), c52 = T(-0.80901699437494742410);
    const T s51 = T(0.95105651629515357212), s52 = T(0.58778525229247312917);
    const cx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const cx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const cx<T> a1 = {v[0].x + c51 * t1.x + c52 * t2.x, v[0].y + c51 * t1.y + c52 * t2.y};
    const cx<T> a2 = {v[0].x + c52 * t1.x + c51 * t2.x, v[0].y + c52 * t1.y + c51 * t2.y};
    const cx<T> b1 = {sgn * (s51 * t3.x + s52 * t4.x), sgn * (s51 * t3.y + s52 * t4.y)};
    const cx<T> b2 = {sgn * (s52 * t3.x - s51 * t4.x), sgn * (s52 * t3.y - s51 * t4.y)};
    v[0] = cadd(v[0], cadd(t1, t2));
    const cx<T> ib1 = cmuli(b1, T(1)), ib2 = cmuli(b2, T(1));
    v[1] = cadd(a1, ib1);
    v[4] = csub(a1, ib1);
    v[2] = cadd(a2, ib2);
    v[3] = csub(a2, ib2);
}

template <typename T, int R>
struct Dft;
template <typename T>
struct Dft<T, 2> {
    static __device__ __forceinline__ void run(cx<T>* v, T) { dft2(v); }
};
template <typename T>
struct Dft<T, 3> {
    static __device__ __forceinline__ void run(cx<T>* v, T s) { dft3(v, s); }
};
template <typename T>
struct Dft<T, 4> {
    static __device__ __forceinline__ void run(cx<T>* v, T s) { dft4(v, s); }
};
template <typename T>
struct Dft<T, 5> {
    static __device__ __forceinline__ void run(cx<T>* v, T s) { dft5(v, s); }
};

// Composite R = A * B: r = A r2 + r1, k = k1 + B k2;
// B-point DFTs over the stride-A subsequences, twiddle W_R^{r1 k1}, A-point DFTs.
template <typename T, int A, int B>
struct DftComposite {
    static __device__ __forceinline__ void run(cx<T>* v, T sgn) {
        constexpr int R = A * B;
        cx<T> y[A][B];
#pragma unroll
        for (int r1 = 0; r1 < A; ++r1) {
#pragma unroll
            for (int r2 = 0; r2 < B; ++r2) y[r1][r2] = v[A * r2 + r1];
            Dft<T, B>::run(y[r1], sgn);
        }
#pragma unroll
        for (int r1 = 1; r1 < A; ++r1)
#pragma unroll
            for (int k1 = 1; k1 < B; ++k1) {
                const int p = (r1 * k1) % R;
                const cx<T> w = {T(TwR<R>::c[p]), sgn * T(TwR<R>::s[p])};
                y[r1][k1] = cmul(y[r1][k1], w);
            }
#pragma unroll
        for (int k1 = 0; k1 < B; ++k1) {
            cx<T> z[A];
#pragma unroll
            for (int r1 = 0; r1 < A; ++r1) z[r1] = y[r1][k1];
            Dft<T, A>::run(z, sgn);
#pragma unroll
            for (int k2 = 0; k2 < A; ++k2) v[k1 + B * k2] = z[k2];
        }
    }
};
template <typename T>
struct Dft<T, 6> : DftComposite<T, 2, 3> {};
template <typename T>
struct Dft<T, 8> : DftComposite<T, 2, 4> {};
template <typename T>
struct Dft<T, 9> : DftComposite<T, 3, 3> {};
template <typename T>
struct Dft<T, 10> : DftComposite<T, 2, 5> {};
template <typename T>
struct Dft<T, 12> : DftComposite<T, 3, 4> {};
template <typename T>
struct Dft<T, 15> : DftComposite<T, 3, 5> {};
template <typename T>
struct Dft<T, 16> : DftComposite<T, 4, 4> {};

// Synchronisation between the read and write halves of a stage.  Block scope: barrier.
// Wave scope (one wave64 owns the transforms): LDS operations of one wavefront execute in
// issue order, so only the compiler must be kept from moving LDS accesses across the point.
template <bool WAVE>
__device__ __forceinline__ void stage_sync() {
    if (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

// One Stockham stage of radix R over ntrans transforms of length n stored at
// buf[t * ts + pidx<PAD>(e * es)].  BPT = transforms per thread per round.  Requires
// m = n/R <= nthr.
template <typename T, int R, int BPT, bool WAVE = false, int PAD = 31>
__device__ void fft_stage(cx<T>* buf, int n, int ns, const cx<T>* tw, T sgn, int ntrans, int es, int ts, int tid,
                          int nthr) {
    const int m = n / R;
    const int groups = nthr / m;
    const int tg = tid / m;
    const int j = tid - tg * m;
    const bool active = tg < groups;
    const int jn = j % ns;
    const int base = (j - jn) * R + jn;  // (j / ns) * ns * R + jn
    const cx<T>* tws = tw + jn;          // stage table: entry (r - 1) * ns + jn
    for (int t0 = 0; t0 < ntrans; t0 += groups * BPT) {
        cx<T> v[BPT][R];
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int t = t0 + tg + b * groups;
            if (active && t < ntrans) {
                const cx<T>* src = buf + t * ts;
                v[b][0] = src[pidx<PAD>(j * es)];
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const cx<T> w = tws[(r - 1) * ns];  // LDS twiddle table, fetched per use
                    const cx<T> x = src[pidx<PAD>((j + r * m) * es)];
                    v[b][r] = sgn < T(0) ? cmul(x, w) : cmulc(x, w);
                }
                Dft<T, R>::run(v[b], sgn);
            }
        }
        stage_sync<WAVE>();
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int t = t0 + tg + b * groups;
            if (active && t < ntrans) {
                cx<T>* dst = buf + t * ts;
#pragma unroll
                for (int q = 0; q < R; ++q) dst[pidx<PAD>((base + q * ns) * es)] = v[b][q];
            }
        }
        stage_sync<WAVE>();
    }
}


// Runs the plan's stages.  WAVE = true: the calling wavefront alone owns these transforms
// (tid = lane, nthr = 64) and no block barrier is used.
template <typename T, int BPT, bool WAVE = false, int PAD = 31>
__device__ void fft_lds(cx<T>* buf, const FftPlan& p, const cx<T>* tw, int sign, int ntrans, int es, int ts,
                        int tid, int nthr) {
    const T sgn = sign < 0 ? T(-1) : T(1);
    int ns = 1;
    for (int s = 0; s < p.nstage; ++s) {
        const int R = p.radix[s];
        const cx<T>* ts_tw = tw + p.twoff[s];
        switch (R) {
            case 2: fft_stage<T, 2, BPT, WAVE, PAD>(buf, p.n, ns, ts_tw, sgn, ntrans, es, ts, tid, nthr); break;
            case 3: fft_stage<T, 3, BPT, WAVE, PAD>(buf, p.n, ns, ts_tw, sgn, ntrans, es, ts, tid, nthr); break;
            case 4: fft_stage<T, 4, BPT, WAVE, PAD>(buf, p.n, ns, ts_tw, sgn, ntrans, es, ts, tid, nthr); break;
            case 5: fft_stage<T, 5, BPT, WAVE, PAD>(buf, p.n, ns, ts_tw, sgn, ntrans, es, ts, tid, nthr); break;
            case 6: fft_stage<T, 6, BPT, WAVE, PAD>(buf, p.n, ns, ts_tw, sgn, ntrans, es, ts, tid, nthr); break;
            case 8: fft_stage<T, 8, BPT, WAVE, PAD>(buf, p.n, ns, ts_tw, sgn, ntrans, es, ts, tid, nthr); break;
            default: break;
        }
        ns *= R;
    }
}

}  // namespace q3x
