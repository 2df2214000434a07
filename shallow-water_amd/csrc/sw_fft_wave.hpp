// Wave-per-sequence FFT passes for lengths N = 64 E (E = 2, 4, 8: N = 128, 256, 512), FP64 or FP32.
//
// Same contract as fft_fwd_pass / fft_inv_pass (sw_fft.hpp): one pass transforms the rows of a batch
// of n x n slabs, two real rows packed as one complex sequence, with prologue/epilogue functors and
// either a straight or a transposed (LDS-staged, coalesced) store.  What changes is the mapping: one
// wave owns one complex sequence, lane j holding elements j + 64 r (r < E).  The transform is P radix-E
// Stockham passes (N = E^P) with every index a compile-time shift or mask.  The loaded layout is the
// first pass's input and the last pass leaves lane j holding outputs j + 64 r, so only the P - 1
// exchanges between passes (plus the Hermitian pack/unpack) go through LDS, each wave in its own
// padded region.  Straight passes run one wave per block; transposing passes group NW waves (2 NW
// rows) so that the transposed store writes 2 NW consecutive values per output row.
#pragma once
#include <hip/hip_runtime.h>

#include "sw_fft.hpp"

namespace sw {

template <int N>
struct WaveShape;
template <>
struct WaveShape<128> {
    static constexpr int E = 2, P = 7;
};
template <>
struct WaveShape<256> {
    static constexpr int E = 4, P = 4;
};
template <>
struct WaveShape<512> {
    static constexpr int E = 8, P = 3;
};

// complex slots of one wave's LDS region: one pad slot after every 8 (spreads the stride-E writes
// over the banks)
template <int N>
constexpr int wave_region() { return N + N / 8; }
__device__ __forceinline__ int wpad(int i) { return i + (i >> 3); }

template <int N, int NW, typename T>
constexpr size_t wave_lds_bytes() {
    size_t ex = (size_t)NW * wave_region<N>() * sizeof(C2<T>);
    size_t stg = (size_t)N * 2 * NW * sizeof(T);
    return ex > stg ? ex : stg;
}

// P radix-E Stockham passes over the sequence held as v[r] = x[lane + 64 r]; afterwards
// v[r] = X[lane + 64 r].  buf is this wave's LDS region.  Every thread of the block calls this.
template <int N, bool INV, typename T>
__device__ __forceinline__ void wave_stockham(C2<T> (&v)[WaveShape<N>::E], C2<T>* buf, const C2<T>* __restrict__ tw,
                                              int lane) {
    constexpr int E = WaveShape<N>::E, P = WaveShape<N>::P;
    int Ns = 1;
#pragma unroll
    for (int q = 0; q < P; q++) {
        const int k = lane & (Ns - 1);
        if (q > 0) {
            const int stride = N / (Ns * E);
#pragma unroll
            for (int r = 1; r < E; r++) {
                C2<T> w = tw[r * k * stride];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dftR<E, INV>(v);
        if (q < P - 1) {
            __syncthreads();  // every lane has read the previous contents of buf
#pragma unroll
            for (int r = 0; r < E; r++) buf[wpad((lane - k) * E + k + r * Ns)] = v[r];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < E; r++) v[r] = buf[wpad(lane + 64 * r)];
        }
        Ns *= E;
    }
}

// Forward (real -> Hermitian) pass; pro(s, r, j) gives row r, element j.  TRANS: element (r, f) goes
// to out[r + n f + n n s]; otherwise epi(s, r, f, value).
template <typename T, int N, int NW, bool TRANS, class Pro, class Epi>
__global__ void __launch_bounds__(64 * NW) fftw_fwd_pass(const C2<T>* __restrict__ tw, int nslab, Pro pro, Epi epi,
                                                        T* tout) {
    constexpr int E = WaveShape<N>::E, BPS = N / (2 * NW);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    C2<T>* buf = reinterpret_cast<C2<T>*>(lds_raw) + wv * wave_region<N>();
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * 2 * NW, ra = r0 + 2 * wv;
    if (s >= nslab) return;

    C2<T> v[E];
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int j = lane + 64 * r;
        v[r] = C2<T>{pro(s, ra, j), pro(s, ra + 1, j)};
    }
    wave_stockham<N, false>(v, buf, tw, lane);

    // split the two real transforms: A_k = (Z_k + conj Z_{N-k}) / 2, B_k = (Z_k - conj Z_{N-k}) / 2i
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; r++) buf[wpad(lane + 64 * r)] = v[r];
    __syncthreads();
    const T sc = T(0.5) / sqrt(T(N));
    constexpr int NT = E / 2 + 1;  // k = lane + 64 t <= N/2
    T ar[NT], ai[NT], br[NT], bi[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int k = lane + 64 * t;
        if (t < NT - 1 || lane == 0) {
            const int kc = (N - k) & (N - 1);
            C2<T> z = buf[wpad(k)], w = buf[wpad(kc)];
            ar[t] = (z.x + w.x) * sc;
            ai[t] = (z.y - w.y) * sc;
            br[t] = (z.y + w.y) * sc;
            bi[t] = (w.x - z.x) * sc;
        }
    }
    if (!TRANS) {
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int k = lane + 64 * t;
            if (t < NT - 1 || lane == 0) {
                epi(s, ra, k, ar[t]);
                epi(s, ra + 1, k, br[t]);
                if (k > 0 && 2 * k < N) {
                    epi(s, ra, N - k, ai[t]);
                    epi(s, ra + 1, N - k, bi[t]);
                }
            }
        }
    } else {
        constexpr int RB = 2 * NW;  // rows of the block
        T* stg = reinterpret_cast<T*>(lds_raw);  // [f][row]
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int k = lane + 64 * t;
            if (t < NT - 1 || lane == 0) {
                stg[k * RB + 2 * wv] = ar[t];
                stg[k * RB + 2 * wv + 1] = br[t];
                if (k > 0 && 2 * k < N) {
                    stg[(N - k) * RB + 2 * wv] = ai[t];
                    stg[(N - k) * RB + 2 * wv + 1] = bi[t];
                }
            }
        }
        __syncthreads();
        T* o = tout + (size_t)s * N * N + r0;
#pragma unroll 4
        for (int e = threadIdx.x; e < RB * N; e += 64 * NW) {
            const int f = e / RB, rr = e - f * RB;
            o[(size_t)f * N + rr] = stg[e];
        }
    }
}

// Inverse (Hermitian -> real) pass; pro(s, r, f) gives the Hermitian row r at packed position f.
// TRANS: element (r, x) goes to out[r + n x + n n s]; otherwise epi(s, r, x, value).
template <typename T, int N, int NW, bool TRANS, class Pro, class Epi>
__global__ void __launch_bounds__(64 * NW) fftw_inv_pass(const C2<T>* __restrict__ tw, int nslab, Pro pro, Epi epi,
                                                        T* tout) {
    constexpr int E = WaveShape<N>::E, BPS = N / (2 * NW);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    C2<T>* buf = reinterpret_cast<C2<T>*>(lds_raw) + wv * wave_region<N>();
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * 2 * NW, ra = r0 + 2 * wv;
    if (s >= nslab) return;

    // pack: Z_k = (a_k - b'_k, a'_k + b_k) and Z_{N-k} = (a_k + b'_k, b_k - a'_k) from the cosine (a, b)
    // and sine (a', b') coefficients of the two rows at k <= N/2
    constexpr int NT = E / 2 + 1;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int k = lane + 64 * t;
        if (t < NT - 1 || lane == 0) {
            const bool has_s = k > 0 && 2 * k < N;
            const T ca = pro(s, ra, k), cb = pro(s, ra + 1, k);
            const T sa = has_s ? pro(s, ra, N - k) : T(0), sb = has_s ? pro(s, ra + 1, N - k) : T(0);
            buf[wpad(k)] = C2<T>{ca - sb, sa + cb};
            if (has_s) buf[wpad(N - k)] = C2<T>{ca + sb, cb - sa};
        }
    }
    __syncthreads();
    C2<T> v[E];
#pragma unroll
    for (int r = 0; r < E; r++) v[r] = buf[wpad(lane + 64 * r)];
    wave_stockham<N, true>(v, buf, tw, lane);

    const T sc = T(1) / sqrt(T(N));
    if (!TRANS) {
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = lane + 64 * r;
            epi(s, ra, x, v[r].x * sc);
            epi(s, ra + 1, x, v[r].y * sc);
        }
    } else {
        constexpr int RB = 2 * NW;
        T* stg = reinterpret_cast<T*>(lds_raw);  // [x][row]
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = lane + 64 * r;
            stg[x * RB + 2 * wv] = v[r].x * sc;
            stg[x * RB + 2 * wv + 1] = v[r].y * sc;
        }
        __syncthreads();
        T* o = tout + (size_t)s * N * N + r0;
#pragma unroll 4
        for (int e = threadIdx.x; e < RB * N; e += 64 * NW) {
            const int x = e / RB, rr = e - x * RB;
            o[(size_t)x * N + rr] = stg[e];
        }
    }
}

}  // namespace sw
