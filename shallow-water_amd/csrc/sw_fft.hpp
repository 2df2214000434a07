// Batched real 2-D FFT passes for gfx950 (CDNA4), FP64 or FP32.
//
// Semantics follow the reference's stafft/sta2dfft (src/stafft/mod.rs:158-452,
// src/sta2dfft.rs:91-123): a length-n real transform along one axis with the "Hermitian form"
// packing out[k] = Re F_k / sqrt(n) (0 <= k <= n/2), out[n-k] = Im F_k / sqrt(n)
// (1 <= k < n/2), F_k = sum_j x_j exp(-2 pi i j k / n); the reverse transform is its exact inverse.
// The arithmetic is not the reference's (mixed-radix Temperton stages on the CPU) but a
// Stockham autosort complex FFT held in LDS, two real rows packed as one complex sequence.
//
// One pass transforms rows (contiguous runs of n values) of a batch of n x n slabs.  A block owns
// B consecutive rows of one slab; they are packed in pairs as z = a + i b, transformed with radix
// 8/4/2/3/5 passes in place in LDS (data staged through registers between the passes), and
// unpacked.  Output is either straight (row r, frequency f at f + n*r) through an epilogue functor,
// or transposed (at r + n*f) through an LDS-staged coalesced store, which is how the 2-D transform
// gets its transpose for free.  Prologue/epilogue functors let callers fuse pointwise products,
// spectral derivatives and operator multiplies into the loads and stores.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sw {

constexpr int FFT_EPT_MAX = 16;  // complex values per thread held between radix passes
constexpr int FFT_MAX_RADICES = 16;

template <typename T>
struct C2 {
    T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ C2<T> cscale(C2<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename T>
__device__ __forceinline__ C2<T> crot(C2<T> a) {
    return INV ? C2<T>{-a.y, a.x} : C2<T>{a.y, -a.x};
}

template <bool INV, typename T>
__device__ __forceinline__ void dft2(C2<T>* v) {
    C2<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV, typename T>
__device__ __forceinline__ void dft3(C2<T>* v) {
    const T h = T(0.86602540378443864676);  // sin(pi/3)
    C2<T> t1 = cadd(v[1], v[2]);
    C2<T> t2 = {v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
    C2<T> t3 = cscale(crot<INV>(csub(v[1], v[2])), h);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
}

template <bool INV, typename T>
__device__ __forceinline__ void dft4(C2<T>* v) {
    C2<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    C2<T> a2 = cadd(v[1], v[3]), a3 = crot<INV>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

template <bool INV, typename T>
__device__ __forceinline__ void dft5(C2<T>* v) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
    C2<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    C2<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    C2<T> x0 = v[0];
    C2<T> m1 = {x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
    C2<T> m2 = {x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
    C2<T> n1 = crot<INV>(C2<T>{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
    C2<T> n2 = crot<INV>(C2<T>{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
    v[0] = cadd(x0, cadd(a1, a2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
}

template <bool INV, typename T>
__device__ __forceinline__ void dft8(C2<T>* v) {
    const T c = T(0.70710678118654752440);
    C2<T> e[4] = {v[0], v[2], v[4], v[6]};
    C2<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft4<INV>(e);
    dft4<INV>(o);
    // twiddles W8^1, W8^2, W8^3 (conjugated for the inverse)
    C2<T> o1 = INV ? C2<T>{c * (o[1].x - o[1].y), c * (o[1].x + o[1].y)}
                   : C2<T>{c * (o[1].x + o[1].y), c * (o[1].y - o[1].x)};
    C2<T> o2 = crot<INV>(o[2]);
    C2<T> o3 = INV ? C2<T>{-c * (o[3].x + o[3].y), c * (o[3].x - o[3].y)}
                   : C2<T>{c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}

template <int R, bool INV, typename T>
__device__ __forceinline__ void dftR(C2<T>* v) {
    if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 3) dft3<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v);
    else if constexpr (R == 5) dft5<INV>(v);
    else dft8<INV>(v);
}

// Launch description of one transform length (plain data, passed by value to kernels).
// The done flags of device-controlled psolve loops, one per member of a batch: a block of a gated launch
// that works on slab s belongs to member s / spm and does nothing once that member's flag is set (a
// finished iteration; see k_psolve_finalize).  flag == nullptr: not gated.
struct Gate {
    const int* flag = nullptr;  // the first member's flag
    int stride = 0;             // ints between consecutive members' flags
    int spm = 1 << 30;          // slabs per member in the launch (default: every slab is member 0's)
    __device__ __forceinline__ bool member_off(int m) const { return flag && flag[(size_t)m * stride]; }
    __device__ __forceinline__ bool off(int s) const { return flag && flag[(size_t)(s / spm) * stride]; }
};

template <typename T>
struct FftPlan {
    int n;        // transform length
    int rows;     // rows per slab (n for the 2-D passes)
    int nrad;     // number of radix passes
    int rad[FFT_MAX_RADICES];
    int B;        // rows per block (power of two dividing n, or 1)
    int nt;       // threads per block
    int xcd;      // remap block ids so that neighbouring row blocks share an XCD (and its L2)
    const C2<T>* tw;  // tw[k] = exp(-2 pi i k / n), k < n
    Gate gate;        // gated launches: blocks of finished members do nothing
};

// One Stockham radix-R pass over all packed sequences of the block, in place in LDS.
template <int R, bool INV, typename T>
__device__ __forceinline__ void stockham_pass(C2<T>* buf, int n, int npair, int Ns, const C2<T>* tw) {
    constexpr int IT = (FFT_EPT_MAX + R - 1) / R;
    const int m = n / R;
    const int total = npair * m;
    const int stride = n / (Ns * R);
    C2<T> v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        int e = threadIdx.x + it * blockDim.x;
        if (e < total) {
            int p = e / m, j = e - p * m;
            const C2<T>* src = buf + p * n + j;
#pragma unroll
            for (int r = 0; r < R; r++) v[it][r] = src[r * m];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++) {
        int e = threadIdx.x + it * blockDim.x;
        if (e < total) {
            int p = e / m, j = e - p * m;
            int k = j % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C2<T> w = tw[r * k * stride];
                    if (INV) w.y = -w.y;
                    v[it][r] = cmul(v[it][r], w);
                }
            }
            dftR<R, INV>(v[it]);
            C2<T>* dst = buf + p * n + (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) dst[r * Ns] = v[it][r];
        }
    }
    __syncthreads();
}

template <bool INV, typename T>
__device__ __forceinline__ void fft_lds(C2<T>* buf, const FftPlan<T>& pl, int npair) {
    int Ns = 1;
    for (int s = 0; s < pl.nrad; s++) {
        int R = pl.rad[s];
        switch (R) {
            case 8: stockham_pass<8, INV>(buf, pl.n, npair, Ns, pl.tw); break;
            case 4: stockham_pass<4, INV>(buf, pl.n, npair, Ns, pl.tw); break;
            case 2: stockham_pass<2, INV>(buf, pl.n, npair, Ns, pl.tw); break;
            case 3: stockham_pass<3, INV>(buf, pl.n, npair, Ns, pl.tw); break;
            default: stockham_pass<5, INV>(buf, pl.n, npair, Ns, pl.tw); break;
        }
        Ns *= R;
    }
}

__device__ __forceinline__ int fft_block_id(int xcd) {
    int bid = blockIdx.x, nb = gridDim.x;
    if (xcd && (nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);
    return bid;
}

// Plain load/store functors.  A slab s, row r, element j of an array lives at j + n*r + n*n*s.
template <typename T>
struct LoadPlain {
    const T* a;
    size_t slab_stride;  // usually n*n
    int n;
    __device__ T operator()(int s, int r, int j) const { return a[(size_t)s * slab_stride + (size_t)r * n + j]; }
};
template <typename T>
struct StorePlain {
    T* a;
    size_t slab_stride;
    int n;
    __device__ void operator()(int s, int r, int f, T v) const { a[(size_t)s * slab_stride + (size_t)r * n + f] = v; }
};

// The same for the slabs of a batch of members: slab s = m per + l is level lev0 + l of member m,
// whose levels start at m mstride.
template <typename T>
struct LoadML {
    const T* a;
    int per;
    size_t mstride, nn;
    int n;
    __device__ T operator()(int s, int r, int j) const {
        const int m = s / per;
        return a[(size_t)m * mstride + (size_t)(s - m * per) * nn + (size_t)r * n + j];
    }
};
template <typename T>
struct StoreML {
    T* a;
    int per;
    size_t mstride, nn;
    int n;
    __device__ void operator()(int s, int r, int f, T v) const {
        const int m = s / per;
        a[(size_t)m * mstride + (size_t)(s - m * per) * nn + (size_t)r * n + f] = v;
    }
};

// Forward (real -> Hermitian) pass.  Rows come from the prologue pro(s, r, j); the transform runs
// along j.  TRANS: write element (r, f) to out[r + n*f + n*n*s] (coalesced through LDS); otherwise
// hand (s, r, f, value) to the epilogue.
template <typename T, bool TRANS, class Pro, class Epi>
__global__ void __launch_bounds__(512) fft_fwd_pass(FftPlan<T> pl, int nslab, Pro pro, Epi epi, T* tout) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C2<T>* buf = reinterpret_cast<C2<T>*>(lds_raw);
    const int n = pl.n, B = pl.B;
    const int bps = pl.rows / B;
    const int bid = fft_block_id(pl.xcd);
    const int s = bid / bps, r0 = (bid - s * bps) * B;
    if (s >= nslab || pl.gate.off(s)) return;
    const int npair = B > 1 ? B / 2 : 1;
    const int nt = blockDim.x;

    // every load of the thread issued before the first LDS store (npair n <= FFT_EPT_MAX nt), so that
    // their latencies overlap instead of adding up
    T la[FFT_EPT_MAX], lb[FFT_EPT_MAX];
#pragma unroll
    for (int it = 0; it < FFT_EPT_MAX; it++) {
        const int e = threadIdx.x + it * nt;
        if (e < npair * n) {
            const int p = e / n, j = e - p * n;
            la[it] = pro(s, r0 + 2 * p, j);
            lb[it] = (B > 1) ? pro(s, r0 + 2 * p + 1, j) : T(0);
        }
    }
#pragma unroll
    for (int it = 0; it < FFT_EPT_MAX; it++) {
        const int e = threadIdx.x + it * nt;
        if (e < npair * n) buf[e] = C2<T>{la[it], lb[it]};
    }
    __syncthreads();
    fft_lds<false>(buf, pl, npair);

    const T sc = T(1) / sqrt(T(n));
    const int K = n / 2 + 1;
    const int total = npair * K;
    if (!TRANS) {
        for (int e = threadIdx.x; e < total; e += nt) {
            int p = e / K, k = e - p * K;
            int kc = k == 0 ? 0 : n - k;
            C2<T> z = buf[p * n + k], w = buf[p * n + kc];
            T ar = (z.x + w.x) * T(0.5) * sc, ai = (z.y - w.y) * T(0.5) * sc;
            T br = (z.y + w.y) * T(0.5) * sc, bi = (w.x - z.x) * T(0.5) * sc;
            int ra = r0 + 2 * p;
            epi(s, ra, k, ar);
            if (k > 0 && k < kc) epi(s, ra, kc, ai);
            if (B > 1) {
                epi(s, ra + 1, k, br);
                if (k > 0 && k < kc) epi(s, ra + 1, kc, bi);
            }
        }
    } else {
        constexpr int IT = FFT_EPT_MAX / 2 + 1;
        C2<T> zz[IT], ww[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            int e = threadIdx.x + it * nt;
            if (e < total) {
                int p = e / K, k = e - p * K;
                int kc = k == 0 ? 0 : n - k;
                zz[it] = buf[p * n + k];
                ww[it] = buf[p * n + kc];
            }
        }
        __syncthreads();
        T* st = reinterpret_cast<T*>(lds_raw);  // staging [f][row], B rows
#pragma unroll
        for (int it = 0; it < IT; it++) {
            int e = threadIdx.x + it * nt;
            if (e < total) {
                int p = e / K, k = e - p * K;
                int kc = k == 0 ? 0 : n - k;
                C2<T> z = zz[it], w = ww[it];
                int ra = 2 * p;
                st[k * B + ra] = (z.x + w.x) * T(0.5) * sc;
                if (k > 0 && k < kc) st[kc * B + ra] = (z.y - w.y) * T(0.5) * sc;
                if (B > 1) {
                    st[k * B + ra + 1] = (z.y + w.y) * T(0.5) * sc;
                    if (k > 0 && k < kc) st[kc * B + ra + 1] = (w.x - z.x) * T(0.5) * sc;
                }
            }
        }
        __syncthreads();
        T* o = tout + (size_t)s * n * n + r0;
        for (int e = threadIdx.x; e < B * n; e += nt) {
            int f = e / B, rr = e - f * B;
            o[(size_t)f * n + rr] = st[e];
        }
    }
}

// Inverse (Hermitian -> real) pass.  Hermitian rows come from pro(s, r, j); the transform runs
// along j.  TRANS: write element (r, x) to out[r + n*x + n*n*s]; otherwise epi(s, r, x, value).
template <typename T, bool TRANS, class Pro, class Epi>
__global__ void __launch_bounds__(512) fft_inv_pass(FftPlan<T> pl, int nslab, Pro pro, Epi epi, T* tout) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C2<T>* buf = reinterpret_cast<C2<T>*>(lds_raw);
    const int n = pl.n, B = pl.B;
    const int bps = pl.rows / B;
    const int bid = fft_block_id(pl.xcd);
    const int s = bid / bps, r0 = (bid - s * bps) * B;
    if (s >= nslab || pl.gate.off(s)) return;
    const int npair = B > 1 ? B / 2 : 1;
    const int nt = blockDim.x;
    const int K = n / 2 + 1;

    // every load of the thread issued before the first LDS store (npair K <= (FFT_EPT_MAX / 2 + 1) nt)
    constexpr int KIT = FFT_EPT_MAX / 2 + 1;
    T ca[KIT], sa[KIT], cb[KIT], sb[KIT];
#pragma unroll
    for (int it = 0; it < KIT; it++) {
        const int e = threadIdx.x + it * nt;
        if (e < npair * K) {
            const int p = e / K, k = e - p * K, kc = n - k, ra = r0 + 2 * p;
            const bool has_s = k > 0 && k < kc;
            ca[it] = pro(s, ra, k);
            sa[it] = has_s ? pro(s, ra, kc) : T(0);
            cb[it] = T(0), sb[it] = T(0);
            if (B > 1) {
                cb[it] = pro(s, ra + 1, k);
                sb[it] = has_s ? pro(s, ra + 1, kc) : T(0);
            }
        }
    }
#pragma unroll
    for (int it = 0; it < KIT; it++) {
        const int e = threadIdx.x + it * nt;
        if (e < npair * K) {
            const int p = e / K, k = e - p * K, kc = n - k;
            buf[p * n + k] = C2<T>{ca[it] - sb[it], sa[it] + cb[it]};
            if (k > 0 && k < kc) buf[p * n + kc] = C2<T>{ca[it] + sb[it], cb[it] - sa[it]};
        }
    }
    __syncthreads();
    fft_lds<true>(buf, pl, npair);

    const T sc = T(1) / sqrt(T(n));
    if (!TRANS) {
        for (int e = threadIdx.x; e < npair * n; e += nt) {
            int p = e / n, j = e - p * n;
            C2<T> z = buf[p * n + j];
            int ra = r0 + 2 * p;
            epi(s, ra, j, z.x * sc);
            if (B > 1) epi(s, ra + 1, j, z.y * sc);
        }
    } else {
        constexpr int IT = FFT_EPT_MAX;
        C2<T> zz[IT];
        const int total = npair * n;
#pragma unroll
        for (int it = 0; it < IT; it++) {
            int e = threadIdx.x + it * nt;
            if (e < total) zz[it] = buf[e];
        }
        __syncthreads();
        T* st = reinterpret_cast<T*>(lds_raw);
#pragma unroll
        for (int it = 0; it < IT; it++) {
            int e = threadIdx.x + it * nt;
            if (e < total) {
                int p = e / n, j = e - p * n;
                st[j * B + 2 * p] = zz[it].x * sc;
                if (B > 1) st[j * B + 2 * p + 1] = zz[it].y * sc;
            }
        }
        __syncthreads();
        T* o = tout + (size_t)s * n * n + r0;
        for (int e = threadIdx.x; e < B * n; e += nt) {
            int f = e / B, rr = e - f * B;
            o[(size_t)f * n + rr] = st[e];
        }
    }
}

}  // namespace sw
