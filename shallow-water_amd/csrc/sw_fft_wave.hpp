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

// Ordering of LDS accesses among the lanes of one wave (each wave's exchange region is private): a
// wave's LDS instructions execute in issue order, so only the compiler must be kept from moving memory
// operations across this point (the fences at wavefront scope), no s_barrier is needed.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// f(integral_constant<int, I>) for I = B .. E - 1, in order (a loop whose index must be a constant
// expression, e.g. to pick compile-time register indices)
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Row stride of the LDS stage [f][row] through which a block of NW waves writes (or reads) 2 NW rows
// transposed: one pad value per f, so that the lanes of a wave, which access the stage at an odd
// stride of 2 NW + 1 values, spread over the banks instead of all hitting one, and the stage fits in
// the block's exchange regions (2 NW + 1 <= 18 N / 8 values per wave).
template <int NW>
constexpr int stage_stride() { return 2 * NW + 1; }

// Dynamic LDS of a block of NW waves: the waves' exchange regions, which a transposing block (NW > 1)
// reuses as its stage; straight passes (NW = 1) stage nothing.
template <int N, int NW, typename T>
constexpr size_t wave_lds_bytes() {
    size_t ex = (size_t)NW * wave_region<N>() * sizeof(C2<T>);
    size_t stg = NW > 1 ? (size_t)N * stage_stride<NW>() * sizeof(T) : 0;
    return ex > stg ? ex : stg;
}

// P radix-E Stockham passes over the sequence held as v[r] = x[lane + 64 r]; afterwards
// v[r] = X[lane + 64 r].  buf is this wave's LDS region.  Every thread of the block calls this.
// TWS: the table holds exp(-2 pi i k / (TWS N)) (a longer transform's table, every TWS-th entry used).
template <int N, bool INV, typename T, int TWS = 1>
__device__ __forceinline__ void wave_stockham(C2<T> (&v)[WaveShape<N>::E], C2<T>* buf, const C2<T>* __restrict__ tw,
                                              int lane) {
    constexpr int E = WaveShape<N>::E, P = WaveShape<N>::P;
    // The twiddle loads stay next to their use: an empty asm makes the table address opaque, so the
    // compiler cannot hoist the loads out of a caller's loop and hold (E - 1) (P - 1) complex values in
    // registers for the whole kernel.
    const C2<T>* twp = tw;
    asm volatile("" : "+s"(twp));
    int Ns = 1;
#pragma unroll
    for (int q = 0; q < P; q++) {
        const int k = lane & (Ns - 1);
        if (q > 0) {
            const int stride = N / (Ns * E);
#pragma unroll
            for (int r = 1; r < E; r++) {
                C2<T> w = twp[r * k * stride * TWS];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dftR<E, INV>(v);
        if (q < P - 1) {
            wave_lds_sync();  // every lane has read the previous contents of buf
#pragma unroll
            for (int r = 0; r < E; r++) buf[wpad((lane - k) * E + k + r * Ns)] = v[r];
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < E; r++) v[r] = buf[wpad(lane + 64 * r)];
        }
        Ns *= E;
    }
}

// The cosine (a, b) and sine (a', b') coefficients of two Hermitian rows at k = lane + 64 t <= N/2,
// fetched into registers ahead of wave_pack_herm; ld(h, f) gives row h (0 or 1) at packed position f.
template <int N, typename T>
struct HermPair {
    static constexpr int NT = WaveShape<N>::E / 2 + 1;
    T ca[NT], cb[NT], sa[NT], sb[NT];
};
// CUT: wavenumbers k > kcut are taken as zero and not loaded.
template <int N, bool CUT = false, typename T, class Ld>
__device__ __forceinline__ void wave_fetch_herm(HermPair<N, T>& h, int lane, Ld ld, int kcut = N / 2) {
    constexpr int NT = HermPair<N, T>::NT;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int k = lane + 64 * t;
        if (t < NT - 1 || lane == 0) {
            const bool in = !CUT || k <= kcut, has_s = in && k > 0 && 2 * k < N;
            h.ca[t] = in ? ld(0, k) : T(0);
            h.cb[t] = in ? ld(1, k) : T(0);
            h.sa[t] = has_s ? ld(0, N - k) : T(0);
            h.sb[t] = has_s ? ld(1, N - k) : T(0);
        }
    }
}
// Pack fetched rows as one complex sequence ready for the inverse transform, Z_k = (a_k - b'_k,
// a'_k + b_k), Z_{N-k} = (a_k + b'_k, b_k - a'_k), through this wave's LDS region into v.  DERIV: pack
// the derivative along the transform axis of the fetched rows instead (d[f] = dmul(f) x[mir(f)]: -k
// times the sine part at k, k times the cosine part at N - k, zero at k = 0 and N/2).
template <int N, bool DERIV = false, typename T>
__device__ __forceinline__ void wave_pack_herm(C2<T> (&v)[WaveShape<N>::E], C2<T>* buf, int lane, const HermPair<N, T>& h) {
    constexpr int E = WaveShape<N>::E, NT = HermPair<N, T>::NT;
    wave_lds_sync();  // every lane has read the previous contents of buf
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int k = lane + 64 * t;
        if (t < NT - 1 || lane == 0) {
            T ca = h.ca[t], cb = h.cb[t], sa = h.sa[t], sb = h.sb[t];
            if (DERIV) {
                const T m = (k == 0 || 2 * k == N) ? T(0) : T(k);
                ca = -m * h.sa[t];
                cb = -m * h.sb[t];
                sa = m * h.ca[t];
                sb = m * h.cb[t];
            }
            buf[wpad(k)] = C2<T>{ca - sb, sa + cb};
            if (k > 0 && 2 * k < N) buf[wpad(N - k)] = C2<T>{ca + sb, cb - sa};
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < E; r++) v[r] = buf[wpad(lane + 64 * r)];
}
template <int N, bool CUT = false, typename T, class Ld>
__device__ __forceinline__ void wave_load_herm_pair(C2<T> (&v)[WaveShape<N>::E], C2<T>* buf, int lane, Ld ld,
                                                    int kcut = N / 2) {
    HermPair<N, T> h;
    wave_fetch_herm<N, CUT>(h, lane, ld, kcut);
    wave_pack_herm<N>(v, buf, lane, h);
}

// After the forward transform of a packed pair (v[r] = Z[lane + 64 r]), split the two real transforms,
// A_k = (Z_k + conj Z_{N-k}) / 2, B_k = (Z_k - conj Z_{N-k}) / 2i, scaled by 1/sqrt(N): cosine parts
// ar, br and sine parts ai, bi at k = lane + 64 t <= N/2 (t < E/2 + 1; the last t only on lane 0).
template <int N, typename T>
__device__ __forceinline__ void wave_split_pair(const C2<T> (&v)[WaveShape<N>::E], C2<T>* buf, int lane,
                                                T (&ar)[WaveShape<N>::E / 2 + 1], T (&ai)[WaveShape<N>::E / 2 + 1],
                                                T (&br)[WaveShape<N>::E / 2 + 1], T (&bi)[WaveShape<N>::E / 2 + 1]) {
    constexpr int E = WaveShape<N>::E, NT = E / 2 + 1;
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < E; r++) buf[wpad(lane + 64 * r)] = v[r];
    wave_lds_sync();
    const T sc = T(0.5) / sqrt(T(N));
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
}

// Transposed store of the 2 NW rows r0 .. r0 + 2 NW - 1 of a block (wave wv holds rows r0 + 2 wv and
// r0 + 2 wv + 1 as split by wave_split_pair): element (r, f) goes to o[r + N f], written through an LDS
// stage [f][row] so that each output row gets a run of 2 NW consecutive values.  o = slab base + r0.
// CUT: output rows f whose wavenumber min(f, N - f) exceeds kcut are not stored (a band that the next
// pass treats as zero).
template <int N, int NW, bool CUT = false, typename T, typename O>
__device__ __forceinline__ void wave_store_trans(const T (&ar)[WaveShape<N>::E / 2 + 1], const T (&ai)[WaveShape<N>::E / 2 + 1],
                                                 const T (&br)[WaveShape<N>::E / 2 + 1], const T (&bi)[WaveShape<N>::E / 2 + 1],
                                                 T* stg, int lane, int wv, O* o, int kcut = N / 2) {
    constexpr int E = WaveShape<N>::E, NT = E / 2 + 1, RB = 2 * NW, RS = stage_stride<NW>();
    static_assert(NW > 1, "a transposed store stages 2 NW rows through the block's LDS");
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int k = lane + 64 * t;
        if (t < NT - 1 || lane == 0) {
            stg[k * RS + 2 * wv] = ar[t];
            stg[k * RS + 2 * wv + 1] = br[t];
            if (k > 0 && 2 * k < N) {
                stg[(N - k) * RS + 2 * wv] = ai[t];
                stg[(N - k) * RS + 2 * wv + 1] = bi[t];
            }
        }
    }
    __syncthreads();
    if constexpr (CUT) {
        const int f1 = kcut + 1, f2 = N - kcut > f1 ? N - kcut : f1;  // rows [0, f1) and [f2, N)
#pragma unroll 4
        for (int e = threadIdx.x; e < f1 * RB; e += 64 * NW) {
            const int f = e / RB, rr = e - f * RB;
            o[(size_t)f * N + rr] = stg[f * RS + rr];
        }
#pragma unroll 4
        for (int e = f2 * RB + threadIdx.x; e < N * RB; e += 64 * NW) {
            const int f = e / RB, rr = e - f * RB;
            o[(size_t)f * N + rr] = stg[f * RS + rr];
        }
    } else {
#pragma unroll 4
        for (int e = threadIdx.x; e < RB * N; e += 64 * NW) {
            const int f = e / RB, rr = e - f * RB;
            o[(size_t)f * N + rr] = stg[f * RS + rr];
        }
    }
}

// Forward (real -> Hermitian) pass; pro(s, r, j) gives row r, element j.  TRANS: element (r, f) goes
// to out[r + n f + n n s] (with CUT, rows f with min(f, N - f) > kcut left unwritten: see
// wave_store_trans); otherwise epi(s, r, f, value).
template <typename T, int N, int NW, bool TRANS, class Pro, class Epi, bool CUT = false, typename TO = T>
__global__ void __launch_bounds__(64 * NW) fftw_fwd_pass(const C2<T>* __restrict__ tw, int nslab, Pro pro, Epi epi,
                                                        TO* tout, int kcut) {
    constexpr int E = WaveShape<N>::E, BPS = N / (2 * NW), NT = E / 2 + 1;
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
    T ar[NT], ai[NT], br[NT], bi[NT];
    wave_split_pair<N>(v, buf, lane, ar, ai, br, bi);
    if constexpr (!TRANS) {
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
        wave_store_trans<N, NW, CUT>(ar, ai, br, bi, reinterpret_cast<T*>(lds_raw), lane, wv,
                                     tout + (size_t)s * N * N + r0, kcut);
    }
}

// Forward pass of rows that are a spectral combination of two real inputs: row r of slab s has inputs
// pro(s, r, j) = {a_j, c_j}, transformed together as one complex sequence, split, and combined as
// A + sig(s) D(C), D the derivative along the transform axis (d[f] = dmul(f) x[mir(f)]: cosine part
// -k C'_k, sine part k C_k, zero at k = 0 and N/2).  A wave owns rows ra and ra + 1, one transform
// each.  TRANS: element (r, f) goes to out[r + n f + n n s]; otherwise epi(s, r, f, value).
template <typename T, int N, int NW, bool TRANS, class Pro, class Sig, class Epi, typename TO = T>
__global__ void __launch_bounds__(64 * NW) fftw_fwd_dcomb(const C2<T>* __restrict__ tw, int nslab, Pro pro, Sig sig,
                                                         Epi epi, TO* tout) {
    constexpr int E = WaveShape<N>::E, BPS = N / (2 * NW), NT = E / 2 + 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    C2<T>* buf = reinterpret_cast<C2<T>*>(lds_raw) + wv * wave_region<N>();
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * 2 * NW, ra = r0 + 2 * wv;
    if (s >= nslab) return;
    const T sg = sig(s);

    T oc[2][NT], os[2][NT];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        C2<T> v[E];
#pragma unroll
        for (int r = 0; r < E; r++) v[r] = pro(s, ra + h, lane + 64 * r);
        wave_stockham<N, false>(v, buf, tw, lane);
        T ar[NT], ai[NT], cr[NT], ci[NT];
        wave_split_pair<N>(v, buf, lane, ar, ai, cr, ci);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int k = lane + 64 * t;
            const T m = (k == 0 || 2 * k == N) ? T(0) : sg * T(k);
            oc[h][t] = ar[t] - m * ci[t];
            os[h][t] = ai[t] + m * cr[t];
        }
    }
    if constexpr (!TRANS) {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int k = lane + 64 * t;
                if (t < NT - 1 || lane == 0) {
                    epi(s, ra + h, k, oc[h][t]);
                    if (k > 0 && 2 * k < N) epi(s, ra + h, N - k, os[h][t]);
                }
            }
    } else {
        wave_store_trans<N, NW>(oc[0], os[0], oc[1], os[1], reinterpret_cast<T*>(lds_raw), lane, wv,
                                tout + (size_t)s * N * N + r0);
    }
}

// Inverse (Hermitian -> real) pass; pro(s, r, f) gives the Hermitian row r at packed position f (with
// CUT, the positions of wavenumbers above kcut taken as zero and not read).  TRANS: element (r, x) goes
// to out[r + n x + n n s]; otherwise epi(s, r, x, value).
template <typename T, int N, int NW, bool TRANS, class Pro, class Epi, bool CUT = false>
__global__ void __launch_bounds__(64 * NW) fftw_inv_pass(const C2<T>* __restrict__ tw, int nslab, Pro pro, Epi epi,
                                                        T* tout, Gate gate, int kcut) {
    constexpr int E = WaveShape<N>::E, BPS = N / (2 * NW);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    C2<T>* buf = reinterpret_cast<C2<T>*>(lds_raw) + wv * wave_region<N>();
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * 2 * NW, ra = r0 + 2 * wv;
    if (s >= nslab || gate.off(s)) return;  // gate: a finished device-controlled iteration

    C2<T> v[E];
    wave_load_herm_pair<N, CUT>(v, buf, lane, [&](int h, int f) { return pro(s, ra + h, f); }, kcut);
    wave_stockham<N, true>(v, buf, tw, lane);

    const T sc = T(1) / sqrt(T(N));
    if constexpr (!TRANS) {
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = lane + 64 * r;
            epi(s, ra, x, v[r].x * sc);
            epi(s, ra + 1, x, v[r].y * sc);
        }
    } else {
        constexpr int RB = 2 * NW, RS = stage_stride<NW>();
        static_assert(!TRANS || NW > 1, "a transposed store stages 2 NW rows through the block's LDS");
        T* stg = reinterpret_cast<T*>(lds_raw);  // [x][row], row stride RS
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = lane + 64 * r;
            stg[x * RS + 2 * wv] = v[r].x * sc;
            stg[x * RS + 2 * wv + 1] = v[r].y * sc;
        }
        __syncthreads();
        T* o = tout + (size_t)s * N * N + r0;
#pragma unroll 4
        for (int e = threadIdx.x; e < RB * N; e += 64 * NW) {
            const int x = e / RB, rr = e - x * RB;
            o[(size_t)x * N + rr] = stg[x * RS + rr];
        }
    }
}

// Inverse pass, pointwise combination and forward pass along the same axis in one launch.  For slab
// group s (block-uniform) the block's rows r are the inverse transforms of K = comb.count(s) Hermitian
// inputs, pro(s, k, r, f) = input k, row r, packed position f (or, when comb.deriv_of_prev(s, k), the
// derivative of input k - 1 along the transform axis); each transformed value val (row r,
// position j) is folded into a per-element accumulator, a = a + comb.coef(s, k, r, j) * val, which
// starts from comb.init(s, r, j); the accumulated real row is transformed forward again and stored
// transposed: element (r, f) at out[r + n f + n n s].  Nothing between the two transforms goes
// through HBM.
//
// TILED: the inputs are read in the tiled layout instead (tile_pos with RB = Pro::RB, a multiple of 2 NW,
// so that the block's rows lie in one tile): pro.tile(s, k, r0) points at row r0, position 0 of input k,
// and each wave's two rows are adjacent at every position f (one 2-value load each).
template <typename T, int N, int NW, class Pro, class Comb, bool TILED = false>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 5 : 3)))
fftw_inv_comb_fwd(const C2<T>* __restrict__ tw, int nslab, Pro pro, Comb comb, T* tout, Gate gate) {
    constexpr int E = WaveShape<N>::E, BPS = N / (2 * NW), NT = E / 2 + 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    C2<T>* buf = reinterpret_cast<C2<T>*>(lds_raw) + wv * wave_region<N>();
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * 2 * NW, ra = r0 + 2 * wv;
    if (s >= nslab || gate.off(s)) return;

    C2<T> acc[E];  // rows ra (.x) and ra + 1 (.y) at j = lane + 64 r
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int j = lane + 64 * r;
        acc[r] = C2<T>{comb.init(s, ra, j), comb.init(s, ra + 1, j)};
    }
    const int K = comb.count(s);
    const T sc = T(1) / sqrt(T(N));
    // Register budget: at most 168 VGPRs so that three waves share a SIMD (the launch uses 4-wave
    // blocks, three per CU).  Hence no prefetch of the next input, and the lane index laundered
    // through an empty asm in the loop and after it: otherwise the compiler hoists every lane-dependent
    // address of the loop (loads, LDS slots, twiddles) out of it and keeps them live (~60 VGPRs).
    HermPair<N, T> h;
    C2<T> v[E];
#pragma unroll 1
    for (int k = 0; k < K; k++) {
        int ln = lane;
        asm volatile("" : "+v"(ln));
        // an input that is the derivative along the transform axis of the previous one is packed from
        // the previous input's registers instead of being fetched
        const bool dcur = k > 0 && comb.deriv_of_prev(s, k);
        if constexpr (TILED) {
            if (!dcur) {
                // the wave's two rows are adjacent in every tile row f: one 2-value load per position
                constexpr int RB = Pro::RB;
                static_assert(RB % (2 * NW) == 0, "a block's rows lie in one tile");
                const C2<T>* __restrict__ src = reinterpret_cast<const C2<T>*>(pro.tile(s, k, r0) + 2 * wv);
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const int kk = ln + 64 * t;
                    if (t < NT - 1 || ln == 0) {
                        const bool has_s = kk > 0 && 2 * kk < N;
                        const C2<T> c = src[(size_t)kk * (RB / 2)];
                        const C2<T> sn = has_s ? src[(size_t)(N - kk) * (RB / 2)] : C2<T>{T(0), T(0)};
                        h.ca[t] = c.x;
                        h.cb[t] = c.y;
                        h.sa[t] = sn.x;
                        h.sb[t] = sn.y;
                    }
                }
            }
        } else {
            if (!dcur) wave_fetch_herm<N>(h, ln, [&](int hh, int f) { return pro(s, k, ra + hh, f); });
        }
        if (dcur) wave_pack_herm<N, true>(v, buf, ln, h);
        else wave_pack_herm<N>(v, buf, ln, h);
        wave_stockham<N, true>(v, buf, tw, ln);
        // every coefficient load of the input issued before the first is used
        T cx[E], cy[E];
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int j = ln + 64 * r;
            cx[r] = comb.coef(s, k, ra, j);
            cy[r] = comb.coef(s, k, ra + 1, j);
        }
#pragma unroll
        for (int r = 0; r < E; r++) {
            acc[r].x = acc[r].x + cx[r] * (v[r].x * sc);
            acc[r].y = acc[r].y + cy[r] * (v[r].y * sc);
        }
    }
    int lp = lane;
    asm volatile("" : "+v"(lp));
    wave_stockham<N, false>(acc, buf, tw, lp);
    T ar[NT], ai[NT], br[NT], bi[NT];
    wave_split_pair<N>(acc, buf, lp, ar, ai, br, bi);
    wave_store_trans<N, NW>(ar, ai, br, bi, reinterpret_cast<T*>(lds_raw), lp, wv, tout + (size_t)s * N * N + r0);
}

// Forward pass, spectral multiply and inverse pass along the same axis in one launch: the rows r of
// slab s (pro(s, r, j), a real row) are transformed forward, each spectral coefficient at packed position
// f is multiplied by mul(s, r, f) (and, when spec is not null, stored at spec[s n n + r n + f]), and the
// rows are transformed back and stored transposed: element (r, j) at out[r + n j + n n s].  This is the
// middle of a de-aliasing 2-D transform pair (forward pass 2, filter, inverse pass 1), with no spectral
// intermediate in HBM.  kcut: mul is zero on every row r with min(r, N - r) > kcut, whose input rows the
// pass before did not store; such rows are not read (taken as zero), and without spec a block whose rows
// all lie there stores nothing (the pass after reads none of it).
template <typename T, int N, int NW, class Pro, class Mul>
__global__ void __launch_bounds__(64 * NW) fftw_fwd_mul_inv(const C2<T>* __restrict__ tw, int nslab, Pro pro, Mul mul,
                                                           T* spec, T* tout, int kcut) {
    constexpr int E = WaveShape<N>::E, BPS = N / (2 * NW), NT = E / 2 + 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    C2<T>* buf = reinterpret_cast<C2<T>*>(lds_raw) + wv * wave_region<N>();
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * 2 * NW, ra = r0 + 2 * wv;
    if (s >= nslab) return;
    auto cut = [&](int r) { return (r < N - r ? r : N - r) > kcut; };
    // the cut rows form one interval, so a block lies in it when its first and last rows do
    if (!spec && cut(r0) && cut(r0 + 2 * NW - 1)) return;
    const bool z0 = cut(ra), z1 = cut(ra + 1);

    C2<T> v[E];
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int j = lane + 64 * r;
        v[r] = C2<T>{z0 ? T(0) : pro(s, ra, j), z1 ? T(0) : pro(s, ra + 1, j)};
    }
    wave_stockham<N, false>(v, buf, tw, lane);
    T ar[NT], ai[NT], br[NT], bi[NT];
    wave_split_pair<N>(v, buf, lane, ar, ai, br, bi);
    HermPair<N, T> h;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int k = lane + 64 * t;
        if (t < NT - 1 || lane == 0) {
            const bool has_s = k > 0 && 2 * k < N;
            h.ca[t] = ar[t] * mul(s, ra, k);
            h.cb[t] = br[t] * mul(s, ra + 1, k);
            h.sa[t] = has_s ? ai[t] * mul(s, ra, N - k) : T(0);
            h.sb[t] = has_s ? bi[t] * mul(s, ra + 1, N - k) : T(0);
            if (spec) {
                T* o = spec + (size_t)s * N * N + (size_t)ra * N;
                o[k] = h.ca[t];
                o[N + k] = h.cb[t];
                if (has_s) {
                    o[N - k] = h.sa[t];
                    o[2 * N - k] = h.sb[t];
                }
            }
        }
    }
    wave_pack_herm<N>(v, buf, lane, h);
    wave_stockham<N, true>(v, buf, tw, lane);
    const T sc = T(1) / sqrt(T(N));
    constexpr int RB = 2 * NW, RS = stage_stride<NW>();
    static_assert(NW > 1, "a transposed store stages 2 NW rows through the block's LDS");
    T* stg = reinterpret_cast<T*>(lds_raw);  // [x][row], row stride RS
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int x = lane + 64 * r;
        stg[x * RS + 2 * wv] = v[r].x * sc;
        stg[x * RS + 2 * wv + 1] = v[r].y * sc;
    }
    __syncthreads();
    T* o = tout + (size_t)s * N * N + r0;
#pragma unroll 4
    for (int e = threadIdx.x; e < RB * N; e += 64 * NW) {
        const int x = e / RB, rr = e - x * RB;
        o[(size_t)x * N + rr] = stg[x * RS + rr];
    }
}

// Derivative along one axis of real N x N slabs (physical layout y + N x) by a 1-D transform pair along
// that axis alone: for output slab s the input io.in(s) is transformed forward along the axis, each
// spectrum is replaced by its derivative (d[f] = dmul(f) x[mir(f)], wave_pack_herm<DERIV>) and
// transformed back into io.out(s).  The transforms along the other axis of a 2-D pair cancel, so this
// is the 2-D spectral derivative without them.  io.along_x(s) (block-uniform) picks the axis: rows x
// (along y, contiguous) are loaded and stored straight, one pair per wave; columns y (along x) go
// through an LDS stage [x][column] of the block's 2 NW columns (row stride stage_stride<NW>()) and are
// read and written in runs of 2 NW.  LDS: wave_lds_bytes<N, NW, T>().
template <typename T, int N, int NW, class Io>
__global__ void __launch_bounds__(64 * NW) fftw_deriv_lines(const C2<T>* __restrict__ tw, int nslab, Io io) {
    constexpr int E = WaveShape<N>::E, BPS = N / (2 * NW), RB = 2 * NW, RS = stage_stride<NW>();
    static_assert(NW > 1, "columns are staged through the block's LDS");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    C2<T>* buf = reinterpret_cast<C2<T>*>(lds_raw) + wv * wave_region<N>();
    T* stg = reinterpret_cast<T*>(lds_raw);  // column mode: [x][column], row stride RS
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * RB, ra = r0 + 2 * wv;
    if (s >= nslab) return;
    const bool ax = io.along_x(s);
    const T* __restrict__ in = io.in(s);
    auto* __restrict__ out = io.out(s);  // T, or a narrower type the values are rounded to

    C2<T> v[E];
    if (ax) {
#pragma unroll 4
        for (int e = threadIdx.x; e < RB * N; e += 64 * NW) {
            const int x = e / RB, rr = e - x * RB;
            stg[x * RS + rr] = in[(size_t)x * N + r0 + rr];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = lane + 64 * r;
            v[r] = C2<T>{stg[x * RS + 2 * wv], stg[x * RS + 2 * wv + 1]};
        }
        __syncthreads();  // the stage becomes the waves' exchange regions
    } else {
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int j = lane + 64 * r;
            v[r] = C2<T>{in[(size_t)ra * N + j], in[(size_t)(ra + 1) * N + j]};
        }
    }
    wave_stockham<N, false>(v, buf, tw, lane);
    HermPair<N, T> h;
    wave_split_pair<N>(v, buf, lane, h.ca, h.sa, h.cb, h.sb);
    wave_pack_herm<N, true>(v, buf, lane, h);
    wave_stockham<N, true>(v, buf, tw, lane);
    const T sc = T(1) / sqrt(T(N));
    if (ax) {
        __syncthreads();  // every wave is done with its exchange region
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = lane + 64 * r;
            stg[x * RS + 2 * wv] = v[r].x * sc;
            stg[x * RS + 2 * wv + 1] = v[r].y * sc;
        }
        __syncthreads();
#pragma unroll 4
        for (int e = threadIdx.x; e < RB * N; e += 64 * NW) {
            const int x = e / RB, rr = e - x * RB;
            out[(size_t)x * N + r0 + rr] = stg[x * RS + rr];
        }
    } else {
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int j = lane + 64 * r;
            out[(size_t)ra * N + j] = v[r].x * sc;
            out[(size_t)(ra + 1) * N + j] = v[r].y * sc;
        }
    }
}

// Levels of fftw_rows_cols_inv's column store that have LDS of their own: those beyond 2 NWV
template <int NWV, int LV>
constexpr int rows_cols_x() { return LV > 2 * NWV ? LV - 2 * NWV : 0; }
template <int N, int NWV, int LV, typename T>
constexpr size_t rows_cols_inv_lds_bytes() {
    constexpr int X = rows_cols_x<NWV, LV>();
    size_t ex = (size_t)NWV * wave_region<N>() * sizeof(C2<T>);
    size_t cs = (size_t)(LV - X) * N * sizeof(T);
    return (size_t)X * N * sizeof(T) + (ex > cs ? ex : cs);
}

// Position of element (x, f) of a slab in the tiled layout that the column-solve launch below writes
// and the staged inputs of fftw_inv_comb_fwd read: tiles of RB consecutive x (the rows of one
// transposing block of that launch) by all N values of f, each tile contiguous with x fastest, so that
// a block reads its tile as one contiguous run and a row f of the writer stores runs of RB values.
template <int N, int RB>
__device__ __forceinline__ size_t tile_pos(int x, int f) {
    return (size_t)(x / RB) * (RB * N) + (size_t)f * RB + (x % RB);
}

// Forward pass along x of every level of one row, a per-column functor across the levels, then, in the
// same launch, inverse transforms along kx of ND rows derived from the columns.  Block b = mb N + ky
// (member mb of a batch) does nothing once its member's gate flag is set.  Its NWV waves transform the LV
// levels of row ky (ld(mb, l, ky, x)) as level pairs into an LDS column store cs[l N + kx]; then
// col(mb, kx, ky, cs + kx, pc) runs per column (at most one per thread) and leaves the column's new values
// in the registers pc[LV].  Then, in rounds of NWV row pairs, every thread writes the raw values
// der.raw<s>(pc) of the round's rows s at its column into the exchange region of the wave that owns the
// row's pair, and each wave packs its pair as Hermitian rows (a row with der.xd(s) takes the derivative
// along kx on the way: d[f] = dmul(f) raw[mir(f)]), transforms it inverse along kx and stores element
// (x, ky) of row s at dout[(mb ND + s) N N + tile_pos<N, RB>(x, ky)].  The Acc<2> results of col are
// summed per block, in a fixed order, into part[2 b + c].  Nothing between the column solve and the
// derived rows goes through HBM.
//
// SOLVE = false: no forward pass and no column functor; col.load(mb, kx, ky, pc) fills each column's
// registers (from HBM) and nothing is written to part (the derived rows of a given column field).
template <typename T, int N, int NWV, int LV, int ND, int RB, class Ld, class Col, class Der, bool SOLVE = true>
__global__ void __launch_bounds__(64 * NWV) __attribute__((amdgpu_waves_per_eu(4)))
fftw_rows_cols_inv(const C2<T>* __restrict__ tw, Ld ld, Col col, Der der, T* __restrict__ dout, double* part, Gate gate) {
    const int mb = blockIdx.x / N, ky = blockIdx.x - mb * N;
    if (gate.member_off(mb)) return;
    constexpr int E = WaveShape<N>::E, NT = E / 2 + 1;
    // levels [0, X) of the column store have LDS of their own ahead of the exchange regions and receive
    // their spectra as soon as they are transformed (as XP pairs, the first a single level when X is odd);
    // levels [X, LV), at most 2 NWV, overlay the regions and wait in registers, one pair per wave, until
    // every wave is done with its region
    constexpr int X = SOLVE ? rows_cols_x<NWV, LV>() : 0, XP = (X + 1) / 2, MP = (LV - X + 1) / 2;
    constexpr int NDP = (ND + 1) / 2, ROUNDS = (NDP + NWV - 1) / NWV;
    static_assert(N <= 64 * NWV, "one column per thread");
    static_assert(!SOLVE || MP <= NWV, "one pair of the overlaid levels per wave");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    C2<T>* const regions = reinterpret_cast<C2<T>*>(lds_raw + (size_t)X * N * sizeof(T));
    C2<T>* buf = regions + wv * wave_region<N>();
    T* cs = reinterpret_cast<T*>(lds_raw);

    // transform the pair of levels (l0, l1) (l1 < 0: l0 alone) of row ky along x
    auto fwd_pair = [&](int l0, int l1, T (&ar)[NT], T (&ai)[NT], T (&br)[NT], T (&bi)[NT]) {
        C2<T> v[E];
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = lane + 64 * r;
            v[r] = C2<T>{ld(mb, l0, ky, x), l1 >= 0 ? ld(mb, l1, ky, x) : T(0)};
        }
        wave_stockham<N, false>(v, buf, tw, lane);
        wave_split_pair<N>(v, buf, lane, ar, ai, br, bi);
    };
    auto put_pair = [&](int l0, int l1, const T (&ar)[NT], const T (&ai)[NT], const T (&br)[NT], const T (&bi)[NT]) {
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int k = lane + 64 * t;
            if (t < NT - 1 || lane == 0) {
                cs[l0 * N + k] = ar[t];
                if (l1 >= 0) cs[l1 * N + k] = br[t];
                if (k > 0 && 2 * k < N) {
                    cs[l0 * N + N - k] = ai[t];
                    if (l1 >= 0) cs[l1 * N + N - k] = bi[t];
                }
            }
        }
    };
    if constexpr (SOLVE) {
        T ar[NT], ai[NT], br[NT], bi[NT];
        for (int j = wv; j < XP; j += NWV) {  // wave-uniform
            const int l0 = (X & 1) ? (j == 0 ? 0 : 2 * j - 1) : 2 * j, l1 = (X & 1) && j == 0 ? -1 : l0 + 1;
            fwd_pair(l0, l1, ar, ai, br, bi);
            put_pair(l0, l1, ar, ai, br, bi);
        }
        const int l0 = X + 2 * wv, l1 = l0 + 1 < LV ? l0 + 1 : -1;
        if (wv < MP) fwd_pair(l0, l1, ar, ai, br, bi);
        __syncthreads();  // every wave is done with its exchange region (the overlaid levels cover them)
        if (wv < MP) put_pair(l0, l1, ar, ai, br, bi);
        __syncthreads();
    }
    const int kx = threadIdx.x;
    const bool has_col = kx < N;
    T pc[LV];
    double s0 = 0.0, s1 = 0.0;
    if constexpr (SOLVE) {
        if (has_col) {
            const auto a = col(mb, kx, ky, cs + kx, pc);
            s0 = a.v[0];
            s1 = a.v[1];
        }
        for (int off = 32; off > 0; off >>= 1) {
            s0 += __shfl_down(s0, off, 64);
            s1 += __shfl_down(s1, off, 64);
        }
    } else {
        if (has_col) col.load(mb, kx, ky, pc);
    }
    const T sc = T(1) / sqrt(T(N));
    T* const db = dout + (size_t)mb * ND * N * N;
    static_for<0, ROUNDS>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        __syncthreads();  // the column store, or the previous round's regions, are no longer read
        if (has_col) {
            static_for<0, NWV>([&](auto wc) {
                constexpr int pi = q * NWV + decltype(wc)::value, sa = 2 * pi, sb = 2 * pi + 1;
                if constexpr (pi < NDP) {
                    C2<T>* reg = regions + decltype(wc)::value * wave_region<N>();
                    T vb = T(0);
                    if constexpr (sb < ND) vb = der.template raw<sb>(pc);
                    reg[wpad(kx)] = C2<T>{der.template raw<sa>(pc), vb};
                }
            });
        }
        __syncthreads();
        const int pi = q * NWV + wv;  // wave-uniform
        if (pi < NDP) {
            const int sa = 2 * pi, sb = 2 * pi + 1;
            const bool xa = der.xd(sa), xb = sb < ND && der.xd(sb);
            // pack in place: the lane that reads positions k and N - k (k = lane + 64 t <= N/2) is the
            // only one that writes them
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int k = lane + 64 * t;
                if (t < NT - 1 || lane == 0) {
                    const bool has_s = k > 0 && 2 * k < N;
                    const C2<T> c = buf[wpad(k)];
                    const C2<T> sn = has_s ? buf[wpad(N - k)] : C2<T>{T(0), T(0)};
                    const T m = (k == 0 || 2 * k == N) ? T(0) : T(k);
                    // derivative along kx: -k times the sine part at k, k times the cosine part at N - k
                    const T ca = xa ? -m * sn.x : c.x, sa_ = xa ? m * c.x : sn.x;
                    const T cb = xb ? -m * sn.y : c.y, sb_ = xb ? m * c.y : sn.y;
                    buf[wpad(k)] = C2<T>{ca - sb_, sa_ + cb};
                    if (has_s) buf[wpad(N - k)] = C2<T>{ca + sb_, cb - sa_};
                }
            }
            wave_lds_sync();
            C2<T> v[E];
#pragma unroll
            for (int r = 0; r < E; r++) v[r] = buf[wpad(lane + 64 * r)];
            wave_stockham<N, true>(v, buf, tw, lane);
            T* oa = db + (size_t)sa * N * N;
#pragma unroll
            for (int r = 0; r < E; r++) {
                const int x = lane + 64 * r;
                const size_t o = tile_pos<N, RB>(x, ky);
                oa[o] = v[r].x * sc;
                if (sb < ND) oa[(size_t)N * N + o] = v[r].y * sc;
            }
        }
    });
    if constexpr (!SOLVE) return;
    __syncthreads();  // the exchange regions are no longer read
    double* red = reinterpret_cast<double*>(lds_raw);
    if (lane == 0) red[2 * wv] = s0, red[2 * wv + 1] = s1;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a0 = red[0], a1 = red[1];
        for (int w = 1; w < NWV; w++) a0 += red[2 * w], a1 += red[2 * w + 1];
        part[2 * (size_t)blockIdx.x] = a0;
        part[2 * (size_t)blockIdx.x + 1] = a1;
    }
}


// ---------------------------------------------------------------- length 3 N (N = 512: 1536)
// A complex sequence of N3 = 3 N values is transformed by a group of three waves: wave j of the group
// transforms the decimated subsequence z[3 m + j] (m = lane + 64 r) with wave_stockham<N> (the N3 table,
// every third entry), twiddles it by W^(j k) (W = exp(-+2 pi i / N3)), and the three are combined by a
// radix-3 butterfly through the group's LDS (its three exchange regions, which also hold the whole
// sequence when it is exchanged: element n at region n / N, slot wpad(n mod N)).  A block holds NG
// groups, each owning a pair of real rows, as the wave-per-sequence passes above.
template <int N>
__device__ __forceinline__ int g3pos(int n) {
    return (n / N) * wave_region<N>() + wpad(n % N);
}
template <int N, int NG, typename T>
constexpr size_t wave3_lds_bytes() {
    size_t ex = (size_t)NG * 3 * wave_region<N>() * sizeof(C2<T>);
    size_t stg = NG > 1 ? (size_t)3 * N * stage_stride<NG>() * sizeof(T) : 0;
    return ex > stg ? ex : stg;
}

// v[r] = z[3 (lane + 64 r) + wj] on entry; Z[wj N + lane + 64 r] on exit (INV: the inverse transform).
// grp: the group's three regions; no wave of the group may still read them on entry.  Every thread of
// the block calls this (it holds block barriers).
template <int N, bool INV, typename T>
__device__ __forceinline__ void wave3_dit(C2<T> (&v)[WaveShape<N>::E], C2<T>* grp, int wj, const C2<T>* __restrict__ tw3,
                                          int lane) {
    constexpr int E = WaveShape<N>::E;
    C2<T>* reg = grp + wj * wave_region<N>();
    wave_stockham<N, INV, T, 3>(v, reg, tw3, lane);
    const C2<T>* twp = tw3;
    asm volatile("" : "+s"(twp));
    if (wj > 0) {
#pragma unroll
        for (int r = 0; r < E; r++) {
            C2<T> w = twp[wj * (lane + 64 * r)];
            if (INV) w.y = -w.y;
            v[r] = cmul(v[r], w);
        }
    }
    wave_lds_sync();  // the transform's last reads of reg are done
#pragma unroll
    for (int r = 0; r < E; r++) reg[wpad(lane + 64 * r)] = v[r];
    __syncthreads();
    // Z[k + N q] = b0 + w3^q b1 + w3^(2 q) b2, w3 = exp(-+2 pi i / 3) = c + i sn
    const T c = T(-0.5), sn = INV ? T(0.86602540378443864676) : T(-0.86602540378443864676);
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int k = wpad(lane + 64 * r);
        const C2<T> b0 = grp[k], b1 = grp[wave_region<N>() + k], b2 = grp[2 * wave_region<N>() + k];
        const C2<T> t1 = cadd(b1, b2), t2 = csub(b1, b2);
        if (wj == 0) {
            v[r] = cadd(b0, t1);
        } else {
            // w3 b1 + w3^2 b2 = c t1 + i sn t2; w3^2 b1 + w3 b2 = c t1 - i sn t2
            const T sg = wj == 1 ? sn : -sn;
            v[r] = C2<T>{b0.x + c * t1.x - sg * t2.y, b0.y + c * t1.y + sg * t2.x};
        }
    }
    __syncthreads();  // every wave has read the regions
}

// Forward (real -> Hermitian) pass of length N3 over rows of N3 x N3 slabs, as fftw_fwd_pass: group g of
// the block owns rows ra = r0 + 2 g and ra + 1 (pro(s, r, j)); TRANS: element (r, f) to
// tout[r + N3 f + N3 N3 s] through an LDS stage (runs of 2 NG values per output row); otherwise epi(s, r,
// f, value).  A gated slab group (gate.off(s)) does nothing.
template <typename T, int N, int NG, bool TRANS, class Pro, class Epi>
__global__ void __launch_bounds__(192 * NG) fftw3_fwd_pass(const C2<T>* __restrict__ tw3, int nslab, Pro pro, Epi epi,
                                                          T* tout, Gate gate) {
    constexpr int E = WaveShape<N>::E, N3 = 3 * N, BPS = N3 / (2 * NG), HT = E / 2;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = wv / 3, wj = wv - 3 * g;
    C2<T>* grp = reinterpret_cast<C2<T>*>(lds_raw) + g * 3 * wave_region<N>();
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * 2 * NG, ra = r0 + 2 * g;
    if (s >= nslab || gate.off(s)) return;

    C2<T> v[E];
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int j = 3 * (lane + 64 * r) + wj;
        v[r] = C2<T>{pro(s, ra, j), pro(s, ra + 1, j)};
    }
    wave3_dit<N, false>(v, grp, wj, tw3, lane);
#pragma unroll
    for (int r = 0; r < E; r++) grp[wj * wave_region<N>() + wpad(lane + 64 * r)] = v[r];
    __syncthreads();
    // split the two real transforms at k = wj N/2 + lane + 64 t (and N3/2 on wave 2, lane 0)
    const T sc = T(0.5) / sqrt(T(N3));
    T ar[HT + 1], ai[HT + 1], br[HT + 1], bi[HT + 1];
#pragma unroll
    for (int t = 0; t <= HT; t++) {
        const int k = wj * (N / 2) + lane + 64 * t;
        if (t < HT || (wj == 2 && lane == 0)) {
            const int kc = k == 0 ? 0 : N3 - k;
            const C2<T> z = grp[g3pos<N>(k)], w = grp[g3pos<N>(kc)];
            ar[t] = (z.x + w.x) * sc;
            ai[t] = (z.y - w.y) * sc;
            br[t] = (z.y + w.y) * sc;
            bi[t] = (w.x - z.x) * sc;
        }
    }
    if constexpr (!TRANS) {
#pragma unroll
        for (int t = 0; t <= HT; t++) {
            const int k = wj * (N / 2) + lane + 64 * t;
            if (t < HT || (wj == 2 && lane == 0)) {
                epi(s, ra, k, ar[t]);
                epi(s, ra + 1, k, br[t]);
                if (k > 0 && 2 * k < N3) {
                    epi(s, ra, N3 - k, ai[t]);
                    epi(s, ra + 1, N3 - k, bi[t]);
                }
            }
        }
    } else {
        static_assert(!TRANS || NG > 1, "a transposed store stages 2 NG rows through the block's LDS");
        constexpr int RB = 2 * NG, RS = stage_stride<NG>();
        T* stg = reinterpret_cast<T*>(lds_raw);  // [f][row], row stride RS
        __syncthreads();                         // every group is done with its regions
#pragma unroll
        for (int t = 0; t <= HT; t++) {
            const int k = wj * (N / 2) + lane + 64 * t;
            if (t < HT || (wj == 2 && lane == 0)) {
                stg[k * RS + 2 * g] = ar[t];
                stg[k * RS + 2 * g + 1] = br[t];
                if (k > 0 && 2 * k < N3) {
                    stg[(N3 - k) * RS + 2 * g] = ai[t];
                    stg[(N3 - k) * RS + 2 * g + 1] = bi[t];
                }
            }
        }
        __syncthreads();
        T* o = tout + (size_t)s * N3 * N3 + r0;
#pragma unroll 4
        for (int e = threadIdx.x; e < RB * N3; e += 192 * NG) {
            const int f = e / RB, rr = e - f * RB;
            o[(size_t)f * N3 + rr] = stg[f * RS + rr];
        }
    }
}

// Inverse (Hermitian -> real) pass of length N3, as fftw_inv_pass: pro(s, r, f) gives Hermitian row r at
// packed position f; TRANS: element (r, x) to tout[r + N3 x + N3 N3 s]; otherwise epi(s, r, x, value).
template <typename T, int N, int NG, bool TRANS, class Pro, class Epi>
__global__ void __launch_bounds__(192 * NG) fftw3_inv_pass(const C2<T>* __restrict__ tw3, int nslab, Pro pro, Epi epi,
                                                          T* tout, Gate gate) {
    constexpr int E = WaveShape<N>::E, N3 = 3 * N, BPS = N3 / (2 * NG), HT = E / 2;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = wv / 3, wj = wv - 3 * g;
    C2<T>* grp = reinterpret_cast<C2<T>*>(lds_raw) + g * 3 * wave_region<N>();
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * 2 * NG, ra = r0 + 2 * g;
    if (s >= nslab || gate.off(s)) return;

    // the pair at k = wj N/2 + lane + 64 t (and N3/2 on wave 2, lane 0), every load issued first
    T ca[HT + 1], cb[HT + 1], sa[HT + 1], sb[HT + 1];
#pragma unroll
    for (int t = 0; t <= HT; t++) {
        const int k = wj * (N / 2) + lane + 64 * t;
        if (t < HT || (wj == 2 && lane == 0)) {
            const bool has_s = k > 0 && 2 * k < N3;
            ca[t] = pro(s, ra, k);
            cb[t] = pro(s, ra + 1, k);
            sa[t] = has_s ? pro(s, ra, N3 - k) : T(0);
            sb[t] = has_s ? pro(s, ra + 1, N3 - k) : T(0);
        }
    }
#pragma unroll
    for (int t = 0; t <= HT; t++) {
        const int k = wj * (N / 2) + lane + 64 * t;
        if (t < HT || (wj == 2 && lane == 0)) {
            grp[g3pos<N>(k)] = C2<T>{ca[t] - sb[t], sa[t] + cb[t]};
            if (k > 0 && 2 * k < N3) grp[g3pos<N>(N3 - k)] = C2<T>{ca[t] + sb[t], cb[t] - sa[t]};
        }
    }
    __syncthreads();
    C2<T> v[E];
#pragma unroll
    for (int r = 0; r < E; r++) v[r] = grp[g3pos<N>(3 * (lane + 64 * r) + wj)];
    __syncthreads();  // every wave has its subsequence before the regions are reused
    wave3_dit<N, true>(v, grp, wj, tw3, lane);

    const T sc = T(1) / sqrt(T(N3));
    if constexpr (!TRANS) {
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = wj * N + lane + 64 * r;
            epi(s, ra, x, v[r].x * sc);
            epi(s, ra + 1, x, v[r].y * sc);
        }
    } else {
        static_assert(!TRANS || NG > 1, "a transposed store stages 2 NG rows through the block's LDS");
        constexpr int RB = 2 * NG, RS = stage_stride<NG>();
        T* stg = reinterpret_cast<T*>(lds_raw);  // [x][row], row stride RS
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = wj * N + lane + 64 * r;
            stg[x * RS + 2 * g] = v[r].x * sc;
            stg[x * RS + 2 * g + 1] = v[r].y * sc;
        }
        __syncthreads();
        T* o = tout + (size_t)s * N3 * N3 + r0;
#pragma unroll 4
        for (int e = threadIdx.x; e < RB * N3; e += 192 * NG) {
            const int x = e / RB, rr = e - x * RB;
            o[(size_t)x * N3 + rr] = stg[x * RS + rr];
        }
    }
}

// fftw_inv_comb_fwd for length N3: for slab group s the block's groups own pairs of rows r; the K =
// comb.count(s) Hermitian inputs pro(s, k, r, f) (input k the derivative along the transform axis of input
// k - 1 when comb.deriv_of_prev(s, k)) are transformed inverse, each folded at once into a = a +
// comb.coef(s, k, r, j) val from a = comb.init(s, r, j), and the accumulated rows transformed forward and
// stored transposed: element (r, f) at tout[r + N3 f + N3 N3 s].  In FP32 the registers are held to 80
// (a few spill), so that two 12-wave blocks share a CU: 888 -> 752 us per launch at 1536x24, against one
// block per CU at the 94 the compiler picks otherwise.
template <typename T, int N, int NG, class Pro, class Comb>
__global__ void __launch_bounds__(192 * NG) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 1)))
fftw3_inv_comb_fwd(const C2<T>* __restrict__ tw3, int nslab, Pro pro, Comb comb,
                                                              T* tout, Gate gate) {
    constexpr int E = WaveShape<N>::E, N3 = 3 * N, BPS = N3 / (2 * NG), HT = E / 2;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = wv / 3, wj = wv - 3 * g;
    C2<T>* grp = reinterpret_cast<C2<T>*>(lds_raw) + g * 3 * wave_region<N>();
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * 2 * NG, ra = r0 + 2 * g;
    if (s >= nslab || gate.off(s)) return;

    C2<T> acc[E];  // rows ra (.x) and ra + 1 (.y) at j = wj N + lane + 64 r
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int j = wj * N + lane + 64 * r;
        acc[r] = C2<T>{comb.init(s, ra, j), comb.init(s, ra + 1, j)};
    }
    const int K = comb.count(s);
    const T sc = T(1) / sqrt(T(N3));
    // the current input's Hermitian pair at k = wj N/2 + lane + 64 t (and N3/2 on wave 2, lane 0)
    T hca[HT + 1], hcb[HT + 1], hsa[HT + 1], hsb[HT + 1];
    C2<T> v[E];
    // the lane index passes through an empty asm in the loop and after it, so that the compiler does not
    // hoist every lane-dependent address of the loop out of it and hold them in registers
#pragma unroll 1
    for (int k = 0; k < K; k++) {
        int lane = threadIdx.x & 63;
        asm volatile("" : "+v"(lane));
        const bool dcur = k > 0 && comb.deriv_of_prev(s, k);
        if (!dcur) {
#pragma unroll
            for (int t = 0; t <= HT; t++) {
                const int kk = wj * (N / 2) + lane + 64 * t;
                if (t < HT || (wj == 2 && lane == 0)) {
                    const bool has_s = kk > 0 && 2 * kk < N3;
                    hca[t] = pro(s, k, ra, kk);
                    hcb[t] = pro(s, k, ra + 1, kk);
                    hsa[t] = has_s ? pro(s, k, ra, N3 - kk) : T(0);
                    hsb[t] = has_s ? pro(s, k, ra + 1, N3 - kk) : T(0);
                }
            }
        }
        // pack (the derivative along the transform axis when dcur: -k times the sine part at k, k
        // times the cosine part at N3 - k)
#pragma unroll
        for (int t = 0; t <= HT; t++) {
            const int kk = wj * (N / 2) + lane + 64 * t;
            if (t < HT || (wj == 2 && lane == 0)) {
                T ca = hca[t], cb = hcb[t], sa = hsa[t], sb = hsb[t];
                if (dcur) {
                    const T m = (kk == 0 || 2 * kk == N3) ? T(0) : T(kk);
                    ca = -m * hsa[t];
                    cb = -m * hsb[t];
                    sa = m * hca[t];
                    sb = m * hcb[t];
                }
                grp[g3pos<N>(kk)] = C2<T>{ca - sb, sa + cb};
                if (kk > 0 && 2 * kk < N3) grp[g3pos<N>(N3 - kk)] = C2<T>{ca + sb, cb - sa};
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; r++) v[r] = grp[g3pos<N>(3 * (lane + 64 * r) + wj)];
        __syncthreads();
        wave3_dit<N, true>(v, grp, wj, tw3, lane);
        T cx[E], cy[E];
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int j = wj * N + lane + 64 * r;
            cx[r] = comb.coef(s, k, ra, j);
            cy[r] = comb.coef(s, k, ra + 1, j);
        }
#pragma unroll
        for (int r = 0; r < E; r++) {
            acc[r].x = acc[r].x + cx[r] * (v[r].x * sc);
            acc[r].y = acc[r].y + cy[r] * (v[r].y * sc);
        }
    }
    // forward transform of the accumulated rows: decimate them through the group's regions
    int lp = threadIdx.x & 63;
    asm volatile("" : "+v"(lp));
    const int lane_out = lp;
#pragma unroll
    for (int r = 0; r < E; r++) grp[wj * wave_region<N>() + wpad(lane_out + 64 * r)] = acc[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; r++) v[r] = grp[g3pos<N>(3 * (lane_out + 64 * r) + wj)];
    __syncthreads();
    wave3_dit<N, false>(v, grp, wj, tw3, lane_out);
#pragma unroll
    for (int r = 0; r < E; r++) grp[wj * wave_region<N>() + wpad(lane_out + 64 * r)] = v[r];
    __syncthreads();
    const T sh = T(0.5) * sc;
    T ar[HT + 1], ai[HT + 1], br[HT + 1], bi[HT + 1];
#pragma unroll
    for (int t = 0; t <= HT; t++) {
        const int kk = wj * (N / 2) + lane_out + 64 * t;
        if (t < HT || (wj == 2 && lane_out == 0)) {
            const int kc = kk == 0 ? 0 : N3 - kk;
            const C2<T> z = grp[g3pos<N>(kk)], w = grp[g3pos<N>(kc)];
            ar[t] = (z.x + w.x) * sh;
            ai[t] = (z.y - w.y) * sh;
            br[t] = (z.y + w.y) * sh;
            bi[t] = (w.x - z.x) * sh;
        }
    }
    static_assert(NG > 1, "a transposed store stages 2 NG rows through the block's LDS");
    constexpr int RB = 2 * NG, RS = stage_stride<NG>();
    T* stg = reinterpret_cast<T*>(lds_raw);  // [f][row], row stride RS
    __syncthreads();                         // every group is done with its regions
#pragma unroll
    for (int t = 0; t <= HT; t++) {
        const int kk = wj * (N / 2) + lane_out + 64 * t;
        if (t < HT || (wj == 2 && lane_out == 0)) {
            stg[kk * RS + 2 * g] = ar[t];
            stg[kk * RS + 2 * g + 1] = br[t];
            if (kk > 0 && 2 * kk < N3) {
                stg[(N3 - kk) * RS + 2 * g] = ai[t];
                stg[(N3 - kk) * RS + 2 * g + 1] = bi[t];
            }
        }
    }
    __syncthreads();
    T* o = tout + (size_t)s * N3 * N3 + r0;
#pragma unroll 4
    for (int e = threadIdx.x; e < RB * N3; e += 192 * NG) {
        const int f = e / RB, rr = e - f * RB;
        o[(size_t)f * N3 + rr] = stg[f * RS + rr];
    }
}

// fftw_deriv_lines for length N3: the derivative along one axis of real N3 x N3 slabs (physical layout
// y + N3 x) by a 1-D transform pair along that axis, one group per pair of lines; io.along_x(s) picks
// the axis (columns y through the block's LDS stage, runs of 2 NG values; rows x straight).
template <typename T, int N, int NG, class Io>
__global__ void __launch_bounds__(192 * NG) fftw3_deriv_lines(const C2<T>* __restrict__ tw3, int nslab, Io io) {
    constexpr int E = WaveShape<N>::E, N3 = 3 * N, BPS = N3 / (2 * NG), RB = 2 * NG, RS = stage_stride<NG>(), HT = E / 2;
    static_assert(NG > 1, "columns are staged through the block's LDS");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = wv / 3, wj = wv - 3 * g;
    C2<T>* grp = reinterpret_cast<C2<T>*>(lds_raw) + g * 3 * wave_region<N>();
    T* stg = reinterpret_cast<T*>(lds_raw);  // column mode: [x][column], row stride RS
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * RB, ra = r0 + 2 * g;
    if (s >= nslab) return;
    const bool ax = io.along_x(s);
    const T* __restrict__ in = io.in(s);
    auto* __restrict__ out = io.out(s);  // T, or a narrower type the values are rounded to

    C2<T> v[E];
    if (ax) {
#pragma unroll 4
        for (int e = threadIdx.x; e < RB * N3; e += 192 * NG) {
            const int x = e / RB, rr = e - x * RB;
            stg[x * RS + rr] = in[(size_t)x * N3 + r0 + rr];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = 3 * (lane + 64 * r) + wj;
            v[r] = C2<T>{stg[x * RS + 2 * g], stg[x * RS + 2 * g + 1]};
        }
        __syncthreads();  // the stage becomes the groups' regions
    } else {
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int j = 3 * (lane + 64 * r) + wj;
            v[r] = C2<T>{in[(size_t)ra * N3 + j], in[(size_t)(ra + 1) * N3 + j]};
        }
    }
    wave3_dit<N, false>(v, grp, wj, tw3, lane);
#pragma unroll
    for (int r = 0; r < E; r++) grp[wj * wave_region<N>() + wpad(lane + 64 * r)] = v[r];
    __syncthreads();
    // split at k = wj N/2 + lane + 64 t (and N3/2 on wave 2, lane 0), then pack the derivative along the
    // axis: -k times the sine part at k, k times the cosine part at N3 - k
    const T sh = T(0.5) / sqrt(T(N3));
    T ca[HT + 1], sa[HT + 1], cb[HT + 1], sb[HT + 1];
#pragma unroll
    for (int t = 0; t <= HT; t++) {
        const int kk = wj * (N / 2) + lane + 64 * t;
        if (t < HT || (wj == 2 && lane == 0)) {
            const int kc = kk == 0 ? 0 : N3 - kk;
            const C2<T> z = grp[g3pos<N>(kk)], w = grp[g3pos<N>(kc)];
            ca[t] = (z.x + w.x) * sh;
            sa[t] = (z.y - w.y) * sh;
            cb[t] = (z.y + w.y) * sh;
            sb[t] = (w.x - z.x) * sh;
        }
    }
    __syncthreads();  // every wave has read the spectrum
#pragma unroll
    for (int t = 0; t <= HT; t++) {
        const int kk = wj * (N / 2) + lane + 64 * t;
        if (t < HT || (wj == 2 && lane == 0)) {
            const T m = (kk == 0 || 2 * kk == N3) ? T(0) : T(kk);
            const T dca = -m * sa[t], dcb = -m * sb[t], dsa = m * ca[t], dsb = m * cb[t];
            grp[g3pos<N>(kk)] = C2<T>{dca - dsb, dsa + dcb};
            if (kk > 0 && 2 * kk < N3) grp[g3pos<N>(N3 - kk)] = C2<T>{dca + dsb, dcb - dsa};
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; r++) v[r] = grp[g3pos<N>(3 * (lane + 64 * r) + wj)];
    __syncthreads();
    wave3_dit<N, true>(v, grp, wj, tw3, lane);
    const T sc = T(1) / sqrt(T(N3));
    if (ax) {
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int x = wj * N + lane + 64 * r;
            stg[x * RS + 2 * g] = v[r].x * sc;
            stg[x * RS + 2 * g + 1] = v[r].y * sc;
        }
        __syncthreads();
#pragma unroll 4
        for (int e = threadIdx.x; e < RB * N3; e += 192 * NG) {
            const int x = e / RB, rr = e - x * RB;
            out[(size_t)x * N3 + r0 + rr] = stg[x * RS + rr];
        }
    } else {
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int j = wj * N + lane + 64 * r;
            out[(size_t)ra * N3 + j] = v[r].x * sc;
            out[(size_t)(ra + 1) * N3 + j] = v[r].y * sc;
        }
    }
}

// fftw_fwd_mul_inv for length N3: rows r of slab s (pro(s, r, j)) transformed forward, each coefficient at
// packed position f multiplied by mul(s, r, f) (and stored at spec[s N3 N3 + r N3 + f] when spec is not
// null), transformed back and stored transposed: element (r, j) at out[r + N3 j + N3 N3 s].
template <typename T, int N, int NG, class Pro, class Mul>
__global__ void __launch_bounds__(192 * NG) fftw3_fwd_mul_inv(const C2<T>* __restrict__ tw3, int nslab, Pro pro, Mul mul,
                                                             T* spec, T* tout) {
    constexpr int E = WaveShape<N>::E, N3 = 3 * N, BPS = N3 / (2 * NG), RB = 2 * NG, RS = stage_stride<NG>(), HT = E / 2;
    static_assert(NG > 1, "a transposed store stages 2 NG rows through the block's LDS");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = wv / 3, wj = wv - 3 * g;
    C2<T>* grp = reinterpret_cast<C2<T>*>(lds_raw) + g * 3 * wave_region<N>();
    const int s = blockIdx.x / BPS, r0 = (blockIdx.x - s * BPS) * RB, ra = r0 + 2 * g;
    if (s >= nslab) return;

    C2<T> v[E];
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int j = 3 * (lane + 64 * r) + wj;
        v[r] = C2<T>{pro(s, ra, j), pro(s, ra + 1, j)};
    }
    wave3_dit<N, false>(v, grp, wj, tw3, lane);
#pragma unroll
    for (int r = 0; r < E; r++) grp[wj * wave_region<N>() + wpad(lane + 64 * r)] = v[r];
    __syncthreads();
    const T sh = T(0.5) / sqrt(T(N3));
    T ca[HT + 1], sa[HT + 1], cb[HT + 1], sb[HT + 1];
#pragma unroll
    for (int t = 0; t <= HT; t++) {
        const int kk = wj * (N / 2) + lane + 64 * t;
        if (t < HT || (wj == 2 && lane == 0)) {
            const int kc = kk == 0 ? 0 : N3 - kk;
            const bool has_s = kk > 0 && 2 * kk < N3;
            const C2<T> z = grp[g3pos<N>(kk)], w = grp[g3pos<N>(kc)];
            ca[t] = (z.x + w.x) * sh * mul(s, ra, kk);
            cb[t] = (z.y + w.y) * sh * mul(s, ra + 1, kk);
            sa[t] = has_s ? (z.y - w.y) * sh * mul(s, ra, N3 - kk) : T(0);
            sb[t] = has_s ? (w.x - z.x) * sh * mul(s, ra + 1, N3 - kk) : T(0);
            if (spec) {
                T* o = spec + (size_t)s * N3 * N3 + (size_t)ra * N3;
                o[kk] = ca[t];
                o[N3 + kk] = cb[t];
                if (has_s) {
                    o[N3 - kk] = sa[t];
                    o[2 * N3 - kk] = sb[t];
                }
            }
        }
    }
    __syncthreads();  // every wave has read the spectrum
#pragma unroll
    for (int t = 0; t <= HT; t++) {
        const int kk = wj * (N / 2) + lane + 64 * t;
        if (t < HT || (wj == 2 && lane == 0)) {
            grp[g3pos<N>(kk)] = C2<T>{ca[t] - sb[t], sa[t] + cb[t]};
            if (kk > 0 && 2 * kk < N3) grp[g3pos<N>(N3 - kk)] = C2<T>{ca[t] + sb[t], cb[t] - sa[t]};
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; r++) v[r] = grp[g3pos<N>(3 * (lane + 64 * r) + wj)];
    __syncthreads();
    wave3_dit<N, true>(v, grp, wj, tw3, lane);
    const T sc = T(1) / sqrt(T(N3));
    T* stg = reinterpret_cast<T*>(lds_raw);  // [x][row], row stride RS
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int x = wj * N + lane + 64 * r;
        stg[x * RS + 2 * g] = v[r].x * sc;
        stg[x * RS + 2 * g + 1] = v[r].y * sc;
    }
    __syncthreads();
    T* o = tout + (size_t)s * N3 * N3 + r0;
#pragma unroll 4
    for (int e = threadIdx.x; e < RB * N3; e += 192 * NG) {
        const int x = e / RB, rr = e - x * RB;
        o[(size_t)x * N3 + rr] = stg[x * RS + rr];
    }
}

}  // namespace sw
