// The device engine: one per (ng, nz, precision).  Holds the FFT plan, the spectral operators in
// device memory and the work buffers, and implements the reference's spectral and nhswps
// functions as batched, fused HIP launches on one stream.  T is the bulk arithmetic type (double,
// or float in the mixed-precision mode); psolve's tridiagonal solves, its residual and every
// reduction are carried out in double whatever T is.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <climits>

#include "sw_common.hpp"
#include "sw_fft.hpp"
#include "sw_fft_wave.hpp"
#include "sw_host.hpp"

namespace sw {

// spectral derivative along one index: d[k] = dmul(k) * g[mir(k)] (src/sta2dfft.rs:128-184 with
// the wavenumbers of D2FFT::new for a 2 pi domain, which are exact integers)
__device__ __forceinline__ int mir(int k, int n) { return k == 0 ? 0 : n - k; }
template <typename T>
__device__ __forceinline__ T dmul(int k, int n) {
    return (k == 0 || 2 * k == n) ? T(0) : (2 * k < n ? T(-k) : T(n - k));
}

// Base pointers of K arrays, one of which a prologue or epilogue picks with a block-uniform index.
// Kernels must not choose among argument pointers with chains of comparisons (?: on the slab
// group): clang (ROCm 7.2, gfx950) folds such a chain into a switch and, after structurizing it,
// can leave the uniform pointer phi unassigned on one path, which then loads through a stale
// register.  Indexing this table is a plain load from the kernel argument block.
template <typename T, int K>
struct PtrTab {
    T* p[K];
    __device__ __forceinline__ T* operator[](int i) const { return p[i]; }
};

// Slabs of fftw_deriv_lines in K groups of `per` levels for each of M members: output slab
// s = (g M + m) per + l is the derivative of level l of member m of src[g] along x (bit g of xmask set)
// or along y, written to level l of member m of dst[g]; member m's levels start at m mstride.
template <typename T, int K, typename O = T>
struct DerivIo {
    PtrTab<const T, K> src;
    PtrTab<O, K> dst;  // outputs stored as O (rounded from T)
    int per;
    unsigned xmask;
    size_t nn;
    int M = 1;
    size_t mstride = 0;
    __device__ __forceinline__ bool along_x(int s) const { return (xmask >> (s / (per * M))) & 1u; }
    __device__ __forceinline__ size_t off(int s, int& g) const {
        g = s / (per * M);
        const int rr = s - g * per * M, m = rr / per;
        return (size_t)m * mstride + (size_t)(rr - m * per) * nn;
    }
    __device__ __forceinline__ const T* in(int s) const {
        int g;
        const size_t o = off(s, g);
        return src[g] + o;
    }
    __device__ __forceinline__ O* out(int s) const {
        int g;
        const size_t o = off(s, g);
        return dst[g] + o;
    }
};

// Tile shapes of the block-per-rows transform passes: B rows per block (B / 2 complex sequences of two
// packed real rows; B = 1, odd lengths only: one real row), nt threads per block.  Measured per transform length and value
// size by autotune.py (which times the candidates with sw_fft_time_tile) and compiled in from
// sw_fft_tiles.inc; sw_fft_set_tile overrides an entry for the engines created afterwards.  Lengths with
// no entry use the rule of choose_tile (the largest power-of-two B up to 16 within 64 KB of LDS).
struct FftTile {
    int n, bytes, B, nt;
};
#include "sw_fft_tiles.inc"  // kFftTiles[]

constexpr size_t FFT_LDS_MAX = 163840;  // LDS per workgroup on gfx950

inline bool tile_valid(int n, int bytes, int B, int nt) {
    // rows are transformed in pairs (2p, 2p + 1) packed as one complex sequence; the pairing sets the
    // rounding, so even lengths take even B (B = 1, one row alone, only for odd lengths)
    if (B < 1 || n % B != 0 || (n % 2 == 0 ? B % 2 != 0 : B != 1) || nt < 64 || nt > 512 || nt % 64 != 0) return false;
    const int npair = B > 1 ? B / 2 : 1;
    return (size_t)std::max(B, 2) * n * bytes <= FFT_LDS_MAX && (size_t)npair * n <= (size_t)FFT_EPT_MAX * nt;
}

inline std::mutex& tile_mu() {
    static std::mutex m;
    return m;
}
inline std::map<std::pair<int, int>, std::pair<int, int>>& tile_overrides() {
    static std::map<std::pair<int, int>, std::pair<int, int>> m;
    return m;
}

// The tile of the passes of length n over values of `bytes` bytes: forced values if given, else the
// override, else the measured table, else the rule.
inline void choose_tile(int n, int bytes, int force_B, int force_nt, int& B, int& nt) {
    B = nt = 0;
    if (!force_B && !force_nt) {
        std::lock_guard<std::mutex> lk(tile_mu());
        auto it = tile_overrides().find({n, bytes});
        if (it != tile_overrides().end()) {
            B = it->second.first, nt = it->second.second;
        } else {
            for (const FftTile& t : kFftTiles)
                if (t.n == n && t.bytes == bytes && tile_valid(n, bytes, t.B, t.nt)) B = t.B, nt = t.nt;
        }
        if (B) return;
    }
    B = 1;
    for (int b = 16; b >= 2; b /= 2)
        if (n % b == 0 && (size_t)b * n * bytes <= 65536 && (b / 2) * n <= FFT_EPT_MAX * 512) {
            B = b;
            break;
        }
    if (force_B) B = force_B;
    const int npair = B > 1 ? B / 2 : 1;
    nt = ((npair * n + FFT_EPT_MAX - 1) / FFT_EPT_MAX + 63) / 64 * 64;
    nt = std::max(64, std::min(512, nt));
    if (force_nt) nt = force_nt;
}

// Kernels launched with more than 64 KB of dynamic LDS are allowed it once per kernel.
inline void lds_optin(const void* kernel, size_t bytes) {
    if (bytes <= 65536) return;
    static std::mutex mu;
    static std::map<const void*, size_t> done;
    std::lock_guard<std::mutex> lk(mu);
    size_t& d = done[kernel];
    if (d >= bytes) return;
    SW_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    d = bytes;
}

constexpr int RED_BLOCKS = 512;
constexpr int MAX_NZ = 256;

template <int K>
struct Acc {
    double v[K];
};

// Block partials of K quantities over i < n; component c is a sum unless bit c of MAXMASK is set,
// in which case it is a max (Kahan-compensated sums per thread, then a fixed shuffle tree).
// Batched: grid (nb, members); block (b, m) reduces f(m, i) over i < n into part[(m nb + b) K + c]
// and does nothing once member m's gate flag is set.
template <int K, unsigned MAXMASK, class F>
__global__ void __launch_bounds__(256) k_reduce(size_t n, F f, double* part, Gate gate) {
    const int m = blockIdx.y;
    if (gate.member_off(m)) return;  // a finished device-controlled iteration (see k_psolve_finalize)
    __shared__ double sh[256 / 64][K];
    double acc[K], cmp[K];
#pragma unroll
    for (int c = 0; c < K; c++) acc[c] = (MAXMASK >> c) & 1 ? -INFINITY : 0.0, cmp[c] = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        Acc<K> a = f(m, i);
#pragma unroll
        for (int c = 0; c < K; c++) {
            if ((MAXMASK >> c) & 1) {
                acc[c] = fmax(acc[c], a.v[c]);
            } else {  // Kahan-compensated sum
                double y = a.v[c] - cmp[c];
                double t = acc[c] + y;
                cmp[c] = (t - acc[c]) - y;
                acc[c] = t;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < K; c++) {
        double x = acc[c];
        for (int off = 32; off > 0; off >>= 1) {
            double y = __shfl_down(x, off, 64);
            x = ((MAXMASK >> c) & 1) ? fmax(x, y) : x + y;
        }
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][c] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < K; c++) {
            double x = sh[0][c];
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) x = ((MAXMASK >> c) & 1) ? fmax(x, sh[w][c]) : x + sh[w][c];
            part[((size_t)m * gridDim.x + blockIdx.x) * K + c] = x;
        }
    }
}

// spec1d (spectral/mod.rs:558-591) of nslab spectral slabs: block (k, s) sums w v^2 over the members
// of shell k of slab s (a wave, fixed order: deterministic) into out[s (ng + 1) + k].
static __global__ void __launch_bounds__(64) k_shell_sums(const double* x, size_t nn, const int* idx, const int* off,
                                                   const double* w, int ng, double* out) {
    const int k = blockIdx.x, s = blockIdx.y;
    const double* a = x + (size_t)s * nn;
    double acc = 0.0;
    for (int i = off[k] + (int)threadIdx.x; i < off[k + 1]; i += 64) {
        const double v = a[idx[i]];
        acc += w[i] * (v * v);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (threadIdx.x == 0) out[(size_t)s * (ng + 1) + k] = acc;
}

// Per-level partial sums of two products over the points of each level: grid (nb, nlev).
template <class F>
__global__ void __launch_bounds__(256) k_level_sum2(size_t n2, F f, double* part) {
    __shared__ double sh[4][2];
    const int iz = blockIdx.y, nb = gridDim.x;
    double a0 = 0.0, a1 = 0.0;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n2; p += (size_t)nb * blockDim.x) {
        Acc<2> v = f(iz, p);
        a0 += v.v[0];
        a1 += v.v[1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_down(a0, off, 64);
        a1 += __shfl_down(a1, off, 64);
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][0] = a0, sh[threadIdx.x >> 6][1] = a1;
    __syncthreads();
    if (threadIdx.x == 0) {
        double x0 = sh[0][0], x1 = sh[0][1];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) x0 += sh[w][0], x1 += sh[w][1];
        part[((size_t)iz * nb + blockIdx.x) * 2] = x0;
        part[((size_t)iz * nb + blockIdx.x) * 2 + 1] = x1;
    }
}

// Mean flow of main_invert (spectral/mod.rs:409-416): out[c] = -sum_iz w_iz * S_iz,c / ng^2, with
// the level sums S taken in a fixed order from the partials (deterministic): one thread per (level,
// component) sums the nb block partials in block order, then the levels are combined in order.  Block
// m: member m of a batch, whose nlev levels are partial rows m nlev .. m nlev + nlev - 1; out[2 m + c].
static __global__ void __launch_bounds__(256) k_meanflow_finalize(const double* part, int nb, int nlev, const double* w,
                                                           double dsumi, double* out) {
    __shared__ double ls[2 * (MAX_NZ + 1)];
    const int m = blockIdx.x;
    for (int t = threadIdx.x; t < 2 * nlev; t += blockDim.x) {
        const int iz = t >> 1, c = t & 1;
        double s = 0.0;
        for (int b = 0; b < nb; b++) s += part[(((size_t)m * nlev + iz) * nb + b) * 2 + c];
        ls[t] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double uio = 0.0, vio = 0.0;
    for (int iz = 0; iz < nlev; iz++) {
        uio -= w[iz] * ls[2 * iz] * dsumi;
        vio -= w[iz] * ls[2 * iz + 1] * dsumi;
    }
    out[2 * m] = uio;
    out[2 * m + 1] = vio;
}

// The pointwise fold of psolve's fused source launch (fftw_inv_comb_fwd): for level s and physical
// point (y, x), S = sp0 + sigx px + cpt1 pt + sigy py + cpt2 ptt from the level's derivative inputs
// k = 0..3 (x, theta, y, theta-theta; y is the derivative of the theta input along the transform
// axis), or sp0 + cpt2 ptt at the bottom level (one input) (psolve.rs:101-270).
// Slab s of a batch is level s mod lv of member s / lv.
template <typename T>
struct PsolveComb {
    const T* sp0;
    PtrTab<const T, 4> cf;  // coefficient of input k: sigx, cpt1, sigy, cpt2
    int n, lv;
    size_t nn;
    __device__ __forceinline__ int count(int s) const { return s % lv == 0 ? 1 : 4; }
    __device__ __forceinline__ bool deriv_of_prev(int s, int k) const { return s % lv > 0 && k == 2; }
    __device__ __forceinline__ T init(int s, int x, int y) const { return sp0[(size_t)s * nn + y + (size_t)n * x]; }
    __device__ __forceinline__ T coef(int s, int k, int x, int y) const {
        return cf[s % lv == 0 ? 3 : k][(size_t)s * nn + y + (size_t)n * x];
    }
};

// The tiled derivative inputs of psolve's source fold (fftw_inv_comb_fwd with TILED): input k of level
// l = s mod lv of member s / lv is slab 0 at the bottom and slab 3 l - 2 + (0, 1, 2) for k = (0, 1, 3)
// (x derivative, theta, theta-theta; input 2, the y derivative, is formed from input 1), nd slabs per member,
// in the layout tile_pos<N, RB>; tile() points at row r0, position 0.
template <typename T, int N, int RB_>
struct PsolveTiles {
    static constexpr int RB = RB_;
    const T* D;
    int lv, nd;
    size_t nn;
    __device__ __forceinline__ const T* tile(int s, int k, int r0) const {
        const int m = s / lv, l = s - m * lv;
        const int sl = l == 0 ? 0 : 3 * l - 2 + (k == 0 ? 0 : (k == 1 ? 1 : 2));
        return D + ((size_t)m * nd + sl) * nn + tile_pos<N, RB>(r0, 0);
    }
};

// The LU factors of the pressure tridiagonal at one (kx, ky) (Spectral::new, spectral/mod.rs:214-256):
// coefficients a0, a0b, ap, apb of rks = kx^2 + ky^2 (integer wavenumbers), then htdv(iz) = filt / (a0 +
// ap etdv(iz-1)) and etdv(iz) = -ap htdv(iz) level by level.  Contraction is off so that the factors are
// bitwise those of the host's IEEE arithmetic.
struct PressureLU {
    double a0, a0b, ap, apb, filt;
    __device__ __forceinline__ PressureLU(double rks, double dzisq, double f) {
#pragma clang fp contract(off)
        a0 = -2.0 * dzisq - (5.0 / 6.0) * rks;
        a0b = -dzisq - (1.0 / 3.0) * rks;
        ap = dzisq - (1.0 / 12.0) * rks;
        apb = dzisq - (1.0 / 6.0) * rks;
        filt = f;
    }
    __device__ __forceinline__ double h0() const { return filt / a0b; }
    __device__ __forceinline__ double e0(double h) const {
#pragma clang fp contract(off)
        return -apb * h;
    }
    __device__ __forceinline__ double h(double eprev) const {
#pragma clang fp contract(off)
        return filt / (a0 + ap * eprev);
    }
    __device__ __forceinline__ double e(double h) const {
#pragma clang fp contract(off)
        return -ap * h;
    }
};

// psolve's column solve (psolve.rs:285-346) for one (kx, ky): the fourth-order compact tridiagonal in
// theta, in FP64, on the column c[iz n] = S(kx, ky, iz) of the source spectrum, with the LU factors
// formed on the way down (PressureLU: no per-level factor arrays read from HBM); writes ps and the top
// level of S (used after the iteration), and returns the Parseval-weighted squared change and squared
// old value of ps for the residual.  NZC = nz.
// Member mb of a batch: its ps and sptop at mb * mstride.
template <typename T, int NZC>
struct PsolveCol {
    const T* filt;
    double dzisq;
    T* ps0;
    T* sptop0;
    int n;
    size_t nn, mstride;
    __device__ __forceinline__ Acc<2> operator()(int mb, int kx, int ky, const T* c) const {
        T pc[NZC + 1];
        return (*this)(mb, kx, ky, c, pc);
    }
    // the same, leaving the column's new ps (levels 0 .. NZC - 1) and its top level NZC (which the
    // solve does not change) in pc
    __device__ __forceinline__ Acc<2> operator()(int mb, int kx, int ky, const T* c, T (&pc)[NZC + 1]) const {
        const Acc<2> r = solve(mb, kx, ky, [&](int iz) { return (double)c[iz * n]; }, pc);
        sptop0[(size_t)mb * mstride + kx + (size_t)n * ky] = c[NZC * n];
        return r;
    }
    // the solve alone, S(iz) giving the source at level iz (in LDS or in HBM)
    template <class SF>
    __device__ __forceinline__ Acc<2> solve(int mb, int kx, int ky, SF S, T (&pc)[NZC + 1]) const {
        T* ps = ps0 + (size_t)mb * mstride;
        const size_t p = kx + (size_t)n * ky;
        const double wgt = ((kx == 0 || 2 * kx == n) ? 1.0 : 2.0) * ((ky == 0 || 2 * ky == n) ? 1.0 : 2.0);
        const double rx = 2 * kx <= n ? kx : n - kx, ry = 2 * ky <= n ? ky : n - ky;
        const PressureLU lu(rx * rx + ry * ry, dzisq, (double)filt[p]);
        double colr[NZC], er[NZC];
        double h = lu.h0();
        er[0] = lu.e0(h);
        double prev = ((1.0 / 3.0) * S(0) + (1.0 / 6.0) * S(1)) * h;
        colr[0] = prev;
#pragma unroll
        for (int iz = 1; iz < NZC; iz++) {
            h = lu.h(er[iz - 1]);
            if (iz < NZC - 1) er[iz] = lu.e(h);
            const double gg = (1.0 / 12.0) * (S(iz - 1) + S(iz + 1)) + (5.0 / 6.0) * S(iz);
            prev = (gg - lu.ap * prev) * h;
            colr[iz] = prev;
        }
        double nxt = prev, dsum = 0.0, osum = 0.0;
#pragma unroll
        for (int iz = NZC - 1; iz >= 0; iz--) {
            const size_t o = (size_t)iz * nn + p;
            if (iz < NZC - 1) nxt = colr[iz] + er[iz] * nxt;
            const T neu = T(nxt);
            const double old = (double)ps[o];
            const double d = (double)neu - old;
            dsum += d * d;
            osum += old * old;
            ps[o] = neu;
            pc[iz] = neu;
        }
        pc[NZC] = ps[(size_t)NZC * nn + p];
        Acc<2> r;
        r.v[0] = wgt * dsum;
        r.v[1] = wgt * osum;
        return r;
    }
};

// The theta-derivative spectra of psolve's source read from ps in HBM (psolve.rs:101-124, 127-227,
// 229-270), one value per call: slab s of member m (sg = m nd + s, nd = 1 + 3 nz) at (kx, ky), with the
// rows of PsolveDer (the x derivative taken here: dmul(kx) times the value at mir(kx)).
template <typename T>
struct PsolveDerivSpec {
    const T* ps;
    int n, nzv, nd;
    size_t nn, NN;
    T hdzi, dzisq, dz2;
    __device__ __forceinline__ T operator()(int sg, int ky, int kx) const {
        const int m = sg / nd, s = sg - m * nd;
        const T* pm = ps + (size_t)m * NN;
        size_t p = kx + (size_t)n * ky;
        auto wka = [&](int iz, size_t q) { return (pm[(size_t)(iz + 1) * nn + q] - pm[(size_t)(iz - 1) * nn + q]) * hdzi; };
        auto wkd = [&](int iz, size_t q) {
            size_t o = (size_t)iz * nn + q;
            return (pm[o + nn] - T(2) * pm[o] + pm[o - nn]) * dzisq;
        };
        if (s == 0) return (T(2) * pm[p] - T(5) * pm[nn + p] + T(4) * pm[2 * nn + p] - pm[3 * nn + p]) * dzisq;
        const int iz = (s + 2) / 3, t = s + 2 - 3 * iz;
        const size_t qx = mir(kx, n) + (size_t)n * ky;
        if (iz < nzv) {
            if (t == 0) return dmul<T>(kx, n) * wka(iz, qx);
            if (t == 1) return wka(iz, p);
            return wkd(iz, p);
        }
        // top: dpdt(nz) = dpdt(nz-1) + dz2 (3 d2pdt2(nz-1) - d2pdt2(nz-2)), d2pdt2(nz) = 2 d2pdt2(nz-1) - d2pdt2(nz-2)
        auto w2 = [&](size_t q) { return nzv >= 3 ? wkd(nzv - 2, q) : T(0); };
        auto top = [&](size_t q) { return wka(nzv - 1, q) + dz2 * (T(3) * wkd(nzv - 1, q) - w2(q)); };
        if (t == 0) return dmul<T>(kx, n) * top(qx);
        if (t == 1) return top(p);
        return T(2) * wkd(nzv - 1, p) - w2(p);
    }
};

// The columns of ps (levels 0 .. NZC) of member mb at (kx, ky), loaded from HBM, for the derivative rows of
// psolve's first guess (fftw_rows_cols_inv without the solve).
template <typename T, int NZC>
struct PsolveColLoad {
    const T* ps0;
    int n;
    size_t nn, mstride;
    __device__ __forceinline__ void load(int mb, int kx, int ky, T (&pc)[NZC + 1]) const {
        const T* ps = ps0 + (size_t)mb * mstride + kx + (size_t)n * ky;
#pragma unroll
        for (int iz = 0; iz <= NZC; iz++) pc[iz] = ps[(size_t)iz * nn];
    }
};

// The 1 + 3 NZC theta-derivative rows of psolve's source (psolve.rs:101-124, 127-227, 229-270) from a
// column of ps (levels 0 .. NZC) held in registers: row 0 = d2p/dtheta2 at the bottom; row 3 iz - 2 + t for
// iz = 1 .. NZC: t = 0 its x derivative (xd: the derivative along kx is taken when the rows are
// packed), t = 1 dp/dtheta, t = 2 d2p/dtheta2; at the top (iz = NZC) the extrapolated dp/dtheta =
// dpdt(nz-1) + dz2 (3 d2pdt2(nz-1) - d2pdt2(nz-2)) and d2p/dtheta2 = 2 d2pdt2(nz-1) - d2pdt2(nz-2), formed
// spectrally (the extrapolation is linear).
template <typename T, int NZC>
struct PsolveDer {
    T hdzi, dzisq, dz2;
    static constexpr int ND = 1 + 3 * NZC;
    __device__ __forceinline__ bool xd(int s) const { return s >= 1 && (s + 2) % 3 == 0; }
    template <int S>
    __device__ __forceinline__ T raw(const T (&p)[NZC + 1]) const {
        static_assert(NZC >= 3 && S >= 0 && S < ND, "row index");
        auto wka = [&](auto izc) {
            constexpr int iz = decltype(izc)::value;
            return (p[iz + 1] - p[iz - 1]) * hdzi;
        };
        auto wkd = [&](auto izc) {
            constexpr int iz = decltype(izc)::value;
            return (p[iz + 1] - T(2) * p[iz] + p[iz - 1]) * dzisq;
        };
        if constexpr (S == 0) {
            return (T(2) * p[0] - T(5) * p[1] + T(4) * p[2] - p[3]) * dzisq;
        } else {
            constexpr int iz = (S + 2) / 3, t = S + 2 - 3 * iz;
            if constexpr (iz < NZC) {
                if constexpr (t == 2) return wkd(std::integral_constant<int, iz>{});
                else return wka(std::integral_constant<int, iz>{});
            } else {
                const T w2 = wkd(std::integral_constant<int, NZC - 2>{});
                const T d1 = wkd(std::integral_constant<int, NZC - 1>{});
                if constexpr (t == 2) return T(2) * d1 - w2;
                else return wka(std::integral_constant<int, NZC - 1>{}) + dz2 * (T(3) * d1 - w2);
            }
        }
    }
};

// Convergence control of psolve's iteration when it runs without a host round trip per iteration:
// the state of the host loop psolve.rs:339-366 (as restated on the host in the block-per-rows path)
// lives in device memory and k_psolve_finalize advances it after each iteration's column solve.
struct PsolveCtrl {
    int done;    // no further iteration: every gated launch of later iterations is a no-op
    int iters;   // iterations completed
    int status;  // 0, SW_ERR_NONFINITE or SW_ERR_PSOLVE_DIVERGED
    int stall;   // mixed precision: consecutive iterations whose residual failed to halve
    double errp, errp_prev;
    double top2;  // squared norm of the caller's top level of pn (enters the first residual)
};

// Block m (member m of a batch; 256 threads): resets its control state, with top2 = the nb block
// partials of its reduction summed as k_psolve_finalize sums the residual partials (Kahan per thread over
// a fixed set of blocks, then a fixed shuffle tree).
static __global__ void __launch_bounds__(256) k_psolve_ctrl_init(PsolveCtrl* c, const double* top2_part, int nb) {
    __shared__ double sh[4];
    c += blockIdx.x;
    top2_part += (size_t)blockIdx.x * nb;
    double acc = 0.0, cmp = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        double y = top2_part[b] - cmp, t = acc + y;
        cmp = (t - acc) - y;
        acc = t;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x != 0) return;
    *c = PsolveCtrl{0, 0, 0, 0, 1.0, 1.0, sh[0] + sh[1] + sh[2] + sh[3]};
}

// Block m (member m of a batch; 256 threads): the residual sums from its nb block partials of the column
// solve (part + 2 m nb; Kahan per thread over a fixed set of blocks, then a fixed shuffle tree), then the
// stopping rule: errp =
// sqrt(dsum / (osum + 1e-20)) with the top level of the caller's pn added at the first iteration
// (c->top2); non-finite or (after the first iteration) errp > 1 stop with an error; errp <= toler or
// nitmax iterations stop; in the mixed mode also two consecutive iterations from the third on whose
// residual failed to halve while below 1e-4.
static __global__ void __launch_bounds__(256) k_psolve_finalize(const double* part, int nb, double toler, int nitmax,
                                                         int mixed, PsolveCtrl* c) {
    __shared__ double sh[4][2];
    c += blockIdx.x;
    part += (size_t)2 * blockIdx.x * nb;
    if (c->done) return;
    double acc[2] = {0.0, 0.0}, cmp[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int k = 0; k < 2; k++) {
            double y = part[2 * b + k] - cmp[k], t = acc[k] + y;
            cmp[k] = (t - acc[k]) - y;
            acc[k] = t;
        }
    for (int k = 0; k < 2; k++) {
        double x = acc[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][k] = x;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double s0 = sh[0][0] + sh[1][0] + sh[2][0] + sh[3][0];
    double s1 = sh[0][1] + sh[1][1] + sh[2][1] + sh[3][1];
    if (c->iters == 0) s0 += c->top2, s1 += c->top2;
    const double errp = sqrt(s0 / (s1 + 1.0E-20));
    c->errp = errp;
    if (!isfinite(errp)) {
        c->iters += 1, c->status = SW_ERR_NONFINITE, c->done = 1;
        return;
    }
    if (c->iters > 0 && errp > 1.0) {
        c->iters += 1, c->status = SW_ERR_PSOLVE_DIVERGED, c->done = 1;
        return;
    }
    c->iters += 1;
    if (mixed && c->iters >= 3 && errp > toler) {
        c->stall = errp > 0.5 * c->errp_prev ? c->stall + 1 : 0;
        if (c->stall >= 2 && errp < 1.0e-4) {
            c->done = 1;
            return;
        }
    }
    c->errp_prev = errp;
    if (errp <= toler || c->iters >= nitmax) c->done = 1;
}

template <typename T>
struct Engine {
    int ng = 0, nz = 0, L = 0;
    size_t n2 = 0, N = 0;
    // members of a batch: every field holds M states, member m's (ng, ng, nz+1) block at m N (NM = M N);
    // each stage runs on all members in the same launches, psolve with per-member control
    int M = 1;
    size_t NM = 0;
    bool mixed = false;
    hipStream_t st = nullptr;
    HostOps host;
    FftPlan<T> plan{};
    size_t lds_bytes = 0;
    DevBuf<C2<T>> tw;
    DevBuf<T> hlap, glap, rlap, c2g2, simp, rope, fope, filt, diss, rdis;
    // Rows of the de-aliasing multipliers (filt, glap) are zero at every packed ky whose wavenumber
    // min(ky, ng - ky) exceeds band_ (the 2/3 rule's circle, |k| > ng/3, contains those whole rows):
    // the de-aliasing transform pairs neither store nor read them.  ng / 2 (no such rows) unless the
    // operators are seen to vanish there.
    int band_ = 0;
    DevBuf<double> ap, htdv, etdv;  // FP64 always: the pressure tridiagonal
    double weight[MAX_NZ + 1], theta[MAX_NZ + 1], htd1[MAX_NZ], etd1[MAX_NZ];
    DevBuf<double> vc;  // device copy: weight[0..L), theta[L..2L), htd1[2L..2L+nz), etd1[2L+nz..2L+2nz)
    DevBuf<T> tmp;  // intermediate of the two FFT passes
    DevBuf<double> part, scal, dwork;
    std::map<std::string, DevBuf<T>> work;
    // live timing of one kind of FFT pass (sw_profile modes): events on this engine's stream around
    // each launch of it, algorithmic bytes = read + write of nslab * ng^2 values
    int prof = SW_PROF_OFF;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<double> ev_bytes;
    size_t ev_used = 0;
    double prof_ms = 0.0, prof_bytes = 0.0;
    long prof_launches = 0;
    // tag: the psolve iteration a launch belongs to in a device-controlled loop (-1: none); launches
    // of iterations at or beyond the loop's final count were no-ops and are not counted
    std::vector<int> ev_tag;
    int prof_tag = -1;
    void prof_start(double bytes) {
        if (ev_used == ev.size()) {
            hipEvent_t a, b;
            SW_HIP(hipEventCreate(&a));
            SW_HIP(hipEventCreate(&b));
            ev.push_back({a, b});
            ev_bytes.push_back(0.0);
            ev_tag.push_back(-1);
        }
        ev_bytes[ev_used] = bytes;
        ev_tag[ev_used] = prof_tag;
        SW_HIP(hipEventRecord(ev[ev_used].first, st));
    }
    void prof_stop() {
        SW_HIP(hipEventRecord(ev[ev_used].second, st));
        ev_used++;
    }
    void prof_collect(int drop_from = INT_MAX) {
        if (!ev_used) return;
        SW_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < ev_used; i++) {
            if (ev_tag[i] >= drop_from) continue;
            float ms = 0.f;
            SW_HIP(hipEventElapsedTime(&ms, ev[i].first, ev[i].second));
            prof_ms += ms;
            prof_bytes += ev_bytes[i];
            prof_launches++;
        }
        ev_used = 0;
    }
    // stage timing (SW_PROF_STAGES): HIP events on this engine's stream around the stages of advance;
    // stage(k) closes the open stage and opens k (k < 0: closes only)
    enum { ST_INVERT = 0, ST_PSETUP, ST_PITER, ST_PEND, ST_SOURCE, ST_UPDATE, ST_DIAG, NSTAGE };
    std::vector<std::pair<hipEvent_t, hipEvent_t>> sev;
    std::vector<int> sev_kind;
    size_t sev_used = 0;
    int stage_open = -1;
    double stage_ms[NSTAGE] = {};
    long stage_iters = 0, stage_steps = 0;
    void stage(int k) {
        if (prof != SW_PROF_STAGES) return;
        if (stage_open >= 0) {
            SW_HIP(hipEventRecord(sev[sev_used].second, st));
            sev_used++;
        }
        stage_open = k;
        if (k < 0) return;
        if (sev_used == sev.size()) {
            hipEvent_t a, b;
            SW_HIP(hipEventCreate(&a));
            SW_HIP(hipEventCreate(&b));
            sev.push_back({a, b});
            sev_kind.push_back(0);
        }
        sev_kind[sev_used] = k;
        SW_HIP(hipEventRecord(sev[sev_used].first, st));
    }
    void stage_collect() {
        stage(-1);
        if (!sev_used) return;
        SW_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < sev_used; i++) {
            float ms = 0.f;
            SW_HIP(hipEventElapsedTime(&ms, sev[i].first, sev[i].second));
            stage_ms[sev_kind[i]] += ms;
        }
        sev_used = 0;
    }
    ~Engine() {
        for (auto& e : ev) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        for (auto& e : sev) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
    // per-member flags that turn the gated launches into no-ops (set while a device-controlled psolve
    // loop is being enqueued; flag null otherwise)
    Gate gate_{};
    DevBuf<PsolveCtrl> ctrl;
    // first batch of a device-controlled psolve loop: the iteration count of the previous call in the
    // same position (advance's predictor and two correctors; slot 3 for any other call), which
    // changes slowly from step to step
    int iter_hint[4] = {8, 8, 8, 8};
    int psolve_slot = 3;
    // graph capture of a step (the ensemble): while capturing, each device-controlled psolve enqueues
    // exactly cap_iters[slot] gated iterations and its control state is read after the graph has run;
    // diagnose leaves its partial sums in dpart
    bool capturing = false;
    int cap_iters[4] = {0, 0, 0, 0};
    // psolve controls
    double toler = 1.0e-9;
    int last_iters = 0;
    double last_errp = 0.0;
    // the last psolve call's results per member of the batch (iterations, errp, 0 or an error status)
    std::vector<int> m_iters, m_status;
    std::vector<double> m_errp;

    Engine(int ng_, int nz_, hipStream_t s, int members = 1) : ng(ng_), nz(nz_), L(nz_ + 1), M(members), st(s) {
        mixed = sizeof(T) == 4;
        n2 = (size_t)ng * ng;
        N = n2 * L;
        NM = N * (size_t)M;
        SW_CHECK(M >= 1 && (M == 1 || nz_spec()), SW_ERR_ARG, "a batch of members needs nz in {4, 8, 16, 24, 32}");
        m_iters.assign(M, 0);
        m_status.assign(M, 0);
        m_errp.assign(M, 0.0);
        SW_CHECK(ng >= 4 && nz >= 2 && nz <= MAX_NZ, SW_ERR_ARG, "need ng >= 4 and 2 <= nz <= 256");
        SW_CHECK(host.build(ng, nz), SW_ERR_FACTOR, "ng is not a product of 2, 3, 4, 5, 6");
        SW_CHECK(reference_transform_is_dft(host.xf), SW_ERR_UNSUPPORTED,
                 "ng factorises into stages the reference cannot transform (radix 2 after 5, or 25 | ng)");
        make_plan(ng, plan, tw, lds_bytes);
        auto up = [&](DevBuf<T>& d, const std::vector<double>& h) {
            std::vector<T> c(h.begin(), h.end());
            d.alloc(c.size());
            copy_h2d(d.p, c.data(), c.size() * sizeof(T), st);
        };
        auto upd = [&](DevBuf<double>& d, const std::vector<double>& h) {
            d.alloc(h.size());
            copy_h2d(d.p, h.data(), h.size() * sizeof(double), st);
        };
        up(hlap, host.hlap); up(glap, host.glap); up(rlap, host.rlap); up(c2g2, host.c2g2);
        up(simp, host.simp); up(rope, host.rope); up(fope, host.fope); up(filt, host.filt);
        up(diss, host.diss); up(rdis, host.rdis);
        {
            const int kc = ng / 3;
            bool zero = true;
            for (int ky = 0; ky < ng && zero; ky++)
                if (std::min(ky, ng - ky) > kc)
                    for (int kx = 0; kx < ng; kx++)
                        if (host.filt[(size_t)kx + (size_t)ng * ky] != 0.0 || host.glap[(size_t)kx + (size_t)ng * ky] != 0.0) {
                            zero = false;
                            break;
                        }
            band_ = zero ? kc : ng / 2;
        }
        upd(ap, host.ap); upd(htdv, host.htdv); upd(etdv, host.etdv);
        for (int i = 0; i <= nz; i++) weight[i] = host.weight[i], theta[i] = host.theta[i];
        for (int i = 0; i < nz; i++) htd1[i] = host.htd1[i], etd1[i] = host.etd1[i];
        {
            std::vector<double> h(2 * L + 2 * nz);
            for (int i = 0; i < L; i++) h[i] = weight[i], h[L + i] = theta[i];
            for (int i = 0; i < nz; i++) h[2 * L + i] = htd1[i], h[2 * L + nz + i] = etd1[i];
            upd(vc, h);
        }
        part.alloc((size_t)RED_BLOCKS * 4 * (MAX_NZ + 1) * M);
        tmp.alloc((size_t)9 * L * n2 * M);  // the largest batch (source: 9 levels' worth of slabs)
        scal.alloc(std::max(16, 2 * M));
        if (mixed) toler = 1.0e-6;
    }

    // ---------------------------------------------------------------- FFT plumbing
    static void make_plan(int n, FftPlan<T>& pl, DevBuf<C2<T>>& twb, size_t& lds, int force_B = 0, int force_nt = 0) {
        std::vector<int> r = gpu_radices(n);
        SW_CHECK(!r.empty() && (int)r.size() <= FFT_MAX_RADICES, SW_ERR_FACTOR, "unsupported transform length");
        pl.n = n;
        pl.rows = n;
        pl.nrad = (int)r.size();
        for (int i = 0; i < pl.nrad; i++) pl.rad[i] = r[i];
        int B, nt;
        choose_tile(n, (int)sizeof(T), force_B, force_nt, B, nt);
        SW_CHECK(tile_valid(n, (int)sizeof(T), B, nt) || (force_B && !force_nt), SW_ERR_ARG, "invalid FFT tile");
        const int npair = B > 1 ? B / 2 : 1;
        SW_CHECK(npair * n <= FFT_EPT_MAX * nt, SW_ERR_ARG, "FFT tile too large for the thread count");
        pl.B = B;
        pl.nt = nt;
        pl.xcd = 1;
        lds = (size_t)std::max(B, 2) * n * sizeof(T);
        std::vector<C2<T>> h(n);
        for (int k = 0; k < n; k++) {
            double a = -2.0 * PI * (double)k / (double)n;
            h[k] = C2<T>{T(std::cos(a)), T(std::sin(a))};
        }
        twb.alloc(n);
        copy_h2d(twb.p, h.data(), n * sizeof(C2<T>), nullptr);
        pl.tw = twb.p;
    }

    T* W(const std::string& name, size_t count) {
        auto& b = work[name];
        if (b.n < count) b.alloc(count);
        return b.p;
    }
    T* tmp_for(int nslab) {
        size_t need = (size_t)nslab * n2;
        if (tmp.n < need) tmp.alloc(need);
        return tmp.p;
    }

    // call sites a profiling mode can single out
    enum { SITE_OTHER = 0, SITE_PSOLVE_DERIV = 1, SITE_PSOLVE_COMB = 2, SITE_PSOLVE_TRI = 3 };

    struct NoEpi {
        __device__ void operator()(int, int, int, T) const {}
    };

    // Transform lengths with wave-per-sequence kernels (sw_fft_wave.hpp); 1536 = 3 x 512 with
    // three-wave groups for its 2-D transforms (fwd2d, inv2d); others use the block-per-rows kernels of
    // sw_fft.hpp.
    bool wave_n() const { return ng == 128 || ng == 256 || ng == 512; }
    bool wave3_n() const { return ng == 1536; }
    // groups per block of the three-wave transposing passes: 8 rows, runs of 8 values per output row
    static constexpr int NG3 = 4;
    template <class Pro, class Epi>
    void fwd2d_wave3(int nslab, Pro pro, Epi epi, T* t) {
        constexpr int N = 512, N3 = 1536;
        const size_t l1 = wave3_lds_bytes<N, NG3, T>(), l2 = wave3_lds_bytes<N, 1, T>();
        lds_optin((const void*)fftw3_fwd_pass<T, N, NG3, true, Pro, NoEpi>, l1);
        lds_optin((const void*)fftw3_fwd_pass<T, N, 1, false, LoadPlain<T>, Epi>, l2);
        hipLaunchKernelGGL((fftw3_fwd_pass<T, N, NG3, true, Pro, NoEpi>), dim3(nslab * (N3 / (2 * NG3))), dim3(192 * NG3), l1, st,
                           plan.tw, nslab, pro, NoEpi{}, t, gate_);
        SW_HIP(hipGetLastError());
        LoadPlain<T> lp{t, n2, ng};
        hipLaunchKernelGGL((fftw3_fwd_pass<T, N, 1, false, LoadPlain<T>, Epi>), dim3(nslab * (N3 / 2)), dim3(192), l2, st, plan.tw,
                           nslab, lp, epi, (T*)nullptr, gate_);
        SW_HIP(hipGetLastError());
    }
    template <class Pro, class Epi>
    void inv2d_wave3(int nslab, Pro pro, Epi epi, T* t) {
        constexpr int N = 512, N3 = 1536;
        const size_t l1 = wave3_lds_bytes<N, NG3, T>(), l2 = wave3_lds_bytes<N, 1, T>();
        lds_optin((const void*)fftw3_inv_pass<T, N, NG3, true, Pro, NoEpi>, l1);
        lds_optin((const void*)fftw3_inv_pass<T, N, 1, false, LoadPlain<T>, Epi>, l2);
        hipLaunchKernelGGL((fftw3_inv_pass<T, N, NG3, true, Pro, NoEpi>), dim3(nslab * (N3 / (2 * NG3))), dim3(192 * NG3), l1, st,
                           plan.tw, nslab, pro, NoEpi{}, t, gate_);
        SW_HIP(hipGetLastError());
        LoadPlain<T> lp{t, n2, ng};
        hipLaunchKernelGGL((fftw3_inv_pass<T, N, 1, false, LoadPlain<T>, Epi>), dim3(nslab * (N3 / 2)), dim3(192), l2, st, plan.tw,
                           nslab, lp, epi, (T*)nullptr, gate_);
        SW_HIP(hipGetLastError());
    }
    template <class F>
    void wave_dispatch(F f) {
        if (ng == 512) f(std::integral_constant<int, 512>{});
        else if (ng == 256) f(std::integral_constant<int, 256>{});
        else f(std::integral_constant<int, 128>{});
    }

    // f(integral_constant<int, nz>) for the level counts with a compile-time specialisation, else
    // f(integral_constant<int, 0>).
    bool nz_spec() const { return nz == 4 || nz == 8 || nz == 16 || nz == 24 || nz == 32; }
    template <class F>
    void nz_dispatch(F f) {
        switch (nz) {
            case 4: f(std::integral_constant<int, 4>{}); break;
            case 8: f(std::integral_constant<int, 8>{}); break;
            case 16: f(std::integral_constant<int, 16>{}); break;
            case 24: f(std::integral_constant<int, 24>{}); break;
            case 32: f(std::integral_constant<int, 32>{}); break;
            default: f(std::integral_constant<int, 0>{}); break;
        }
    }

    // f(integral_constant<int, nz>) when nz has a compile-time specialisation; false (f not called)
    // otherwise.  The per-point level loops use it to hold a column in registers and to issue every
    // level's loads before the first store (the compiler may not move loads past stores to arrays that
    // might alias, which otherwise costs one memory latency per level).
    template <class F>
    bool with_nz(F f) {
        bool done = false;
        nz_dispatch([&](auto nzc) {
            if constexpr (decltype(nzc)::value > 0) {
                f(nzc);
                done = true;
            }
        });
        return done;
    }

    // Forward 2-D transform of nslab slabs: pro(s, x, y) gives the physical value at (y, x);
    // epi(s, ky, kx, v) receives the spectral coefficient at (kx, ky).
    // NW: waves per block of the transposing pass (0: NWT).  inter: the buffer between the two passes
    // (nslab slabs; null: the engine's own, grown as needed).
    template <int NW = 0, class Pro, class Epi>
    void fwd2d(int nslab, Pro pro, Epi epi, int site = SITE_OTHER, T* inter = nullptr) {
        if (nslab <= 0) return;
        if (wave3_n()) {
            T* t = inter ? inter : tmp_for(nslab);
            fwd2d_wave3(nslab, pro, epi, t);
            return;
        }
        if (wave_n()) {
            T* t = inter ? inter : tmp_for(nslab);
            wave_dispatch([&](auto nn) {
                constexpr int NN = decltype(nn)::value;
                fwd_pass1_wave<NN, NW ? NW : NWT>(nslab, pro, t, site);
                fwd_pass2_wave<NN>(nslab, t, epi);
            });
            return;
        }
        T* t = inter ? inter : tmp_for(nslab);
        dim3 grid(nslab * (ng / plan.B)), block(plan.nt);
        const double bytes = 2.0 * sizeof(T) * (double)n2 * nslab;
        const bool time1 = prof == SW_PROF_PSOLVE_DERIV && site == SITE_PSOLVE_DERIV, time2 = prof == SW_PROF_STRAIGHT;
        FftPlan<T> pl = plan;
        pl.gate = gate_;
        lds_optin((const void*)fft_fwd_pass<T, true, Pro, NoEpi>, lds_bytes);
        lds_optin((const void*)fft_fwd_pass<T, false, LoadPlain<T>, Epi>, lds_bytes);
        if (time1) prof_start(bytes);
        hipLaunchKernelGGL((fft_fwd_pass<T, true, Pro, NoEpi>), grid, block, lds_bytes, st, pl, nslab, pro, NoEpi{}, t);
        if (time1) prof_stop();
        SW_HIP(hipGetLastError());
        LoadPlain<T> lp{t, n2, ng};
        if (time2) prof_start(bytes);
        hipLaunchKernelGGL((fft_fwd_pass<T, false, LoadPlain<T>, Epi>), grid, block, lds_bytes, st, pl, nslab, lp, epi,
                           (T*)nullptr);
        if (time2) prof_stop();
        SW_HIP(hipGetLastError());
    }

    // Inverse 2-D transform: pro(s, ky, kx) gives the spectral coefficient at (kx, ky);
    // epi(s, x, y, v) receives the physical value at (y, x).
    template <int NW = 0, class Pro, class Epi>
    void inv2d(int nslab, Pro pro, Epi epi, int site = SITE_OTHER) {
        if (nslab <= 0) return;
        if (wave3_n()) {
            inv2d_wave3(nslab, pro, epi, tmp_for(nslab));
            return;
        }
        if (wave_n()) {
            T* t = tmp_for(nslab);
            wave_dispatch([&](auto nn) {
                constexpr int NN = decltype(nn)::value;
                inv_pass1_wave<NN, NW ? NW : NWT>(nslab, pro, t, site);
                inv_pass2_wave<NN>(nslab, t, epi);
            });
            return;
        }
        T* t = tmp_for(nslab);
        dim3 grid(nslab * (ng / plan.B)), block(plan.nt);
        const double bytes = 2.0 * sizeof(T) * (double)n2 * nslab;
        const bool time1 = prof == SW_PROF_PSOLVE_DERIV && site == SITE_PSOLVE_DERIV, time2 = prof == SW_PROF_STRAIGHT;
        FftPlan<T> pl = plan;
        pl.gate = gate_;
        lds_optin((const void*)fft_inv_pass<T, true, Pro, NoEpi>, lds_bytes);
        lds_optin((const void*)fft_inv_pass<T, false, LoadPlain<T>, Epi>, lds_bytes);
        if (time1) prof_start(bytes);
        hipLaunchKernelGGL((fft_inv_pass<T, true, Pro, NoEpi>), grid, block, lds_bytes, st, pl, nslab, pro, NoEpi{}, t);
        if (time1) prof_stop();
        SW_HIP(hipGetLastError());
        LoadPlain<T> lp{t, n2, ng};
        if (time2) prof_start(bytes);
        hipLaunchKernelGGL((fft_inv_pass<T, false, LoadPlain<T>, Epi>), grid, block, lds_bytes, st, pl, nslab, lp, epi,
                           (T*)nullptr);
        if (time2) prof_stop();
        SW_HIP(hipGetLastError());
    }

    // Single passes of the wave-per-sequence kernels for n = NN.  A transposing pass groups NW waves
    // (NWT by default) so that its store writes 2 NW consecutive values per output row: 4 waves in FP64
    // (64-byte runs; with the padded stage these blocks outrun 8-wave ones, 5.9 against 5.2 TB/s for a
    // plain 49-slab pass at n = 512), 8 in FP32 (64-byte runs again; 4-wave blocks drop to 3.5 TB/s);
    // a straight pass runs one wave per block.  Pass 1 of a forward transform runs along y (physical
    // rows x) and leaves the spectrum in y transposed: out[x + n ky]; pass 2 runs along x.  Pass 1 of an
    // inverse transform runs along kx (rows ky) and leaves out[ky + n x]; pass 2 runs along ky.
    static constexpr int NWT = sizeof(T) == 8 ? 4 : 8;
    bool timed(int site) const {
        return (prof == SW_PROF_PSOLVE_DERIV && site == SITE_PSOLVE_DERIV) ||
               (prof == SW_PROF_PSOLVE_COMB && site == SITE_PSOLVE_COMB) ||
               (prof == SW_PROF_PSOLVE_TRI && site == SITE_PSOLVE_TRI);
    }
    // CUT: output rows whose wavenumber exceeds kcut are not stored
    template <int NN, int NW, bool CUT = false, class Pro, typename TO>
    void fwd_pass1_wave(int nslab, Pro pro, TO* out, int site, int kcut = NN / 2) {
        const size_t lds_t = wave_lds_bytes<NN, NW, T>();
        const bool tm = timed(site);
        if (tm) prof_start(2.0 * sizeof(T) * (double)n2 * nslab);
        hipLaunchKernelGGL((fftw_fwd_pass<T, NN, NW, true, Pro, NoEpi, CUT, TO>), dim3(nslab * (NN / (2 * NW))), dim3(64 * NW),
                           lds_t, st, plan.tw, nslab, pro, NoEpi{}, out, kcut);
        if (tm) prof_stop();
        SW_HIP(hipGetLastError());
    }
    template <int NN, class Epi>
    void fwd_pass2_wave(int nslab, const T* in, Epi epi) {
        const size_t lds_s = wave_lds_bytes<NN, 1, T>();
        const bool tm = prof == SW_PROF_STRAIGHT;
        if (tm) prof_start(2.0 * sizeof(T) * (double)n2 * nslab);
        LoadPlain<T> lp{in, n2, ng};
        hipLaunchKernelGGL((fftw_fwd_pass<T, NN, 1, false, LoadPlain<T>, Epi>), dim3(nslab * (NN / 2)), dim3(64),
                           lds_s, st, plan.tw, nslab, lp, epi, (T*)nullptr, NN / 2);
        if (tm) prof_stop();
        SW_HIP(hipGetLastError());
    }
    // bytes: algorithmic bytes of the launch for the profiling mode that times this site (0: read +
    // write of every transformed value)
    template <int NN, int NW, class Pro>
    void inv_pass1_wave(int nslab, Pro pro, T* out, int site, double bytes = 0.0) {
        const size_t lds_t = wave_lds_bytes<NN, NW, T>();
        const bool tm = timed(site);
        if (tm) prof_start(bytes > 0.0 ? bytes : 2.0 * sizeof(T) * (double)n2 * nslab);
        hipLaunchKernelGGL((fftw_inv_pass<T, NN, NW, true, Pro, NoEpi>), dim3(nslab * (NN / (2 * NW))), dim3(64 * NW),
                           lds_t, st, plan.tw, nslab, pro, NoEpi{}, out, gate_, NN / 2);
        if (tm) prof_stop();
        SW_HIP(hipGetLastError());
    }
    // CUT: input positions whose wavenumber exceeds kcut are taken as zero and not read
    template <int NN, bool CUT = false, class Epi>
    void inv_pass2_wave(int nslab, const T* in, Epi epi, int kcut = NN / 2) {
        const size_t lds_s = wave_lds_bytes<NN, 1, T>();
        const bool tm = prof == SW_PROF_STRAIGHT;
        if (tm) prof_start(2.0 * sizeof(T) * (double)n2 * nslab);
        LoadPlain<T> lp{in, n2, ng};
        hipLaunchKernelGGL((fftw_inv_pass<T, NN, 1, false, LoadPlain<T>, Epi, CUT>), dim3(nslab * (NN / 2)), dim3(64),
                           lds_s, st, plan.tw, nslab, lp, epi, (T*)nullptr, gate_, kcut);
        if (tm) prof_stop();
        SW_HIP(hipGetLastError());
    }
    // fftw_inv_comb_fwd over nslab slab groups (transposed output); bytes = the algorithmic bytes of
    // the launch, for the profiling mode that times this site.
    // FP64: 4-wave blocks (the kernel runs three waves per SIMD, so three blocks share a CU); FP32:
    // 8-wave blocks, so that a block's 16 rows are one 16-row tile of the tiled derivative layout below
    static constexpr int NWC = sizeof(T) == 4 ? 8 : 4;
    template <int NN, class Pro, class Comb>
    void inv_comb_fwd_wave(int nslab, Pro pro, Comb comb, T* out, double bytes, int site) {
        const size_t lds_t = wave_lds_bytes<NN, NWC, T>();
        const bool tm = timed(site);
        if (tm) prof_start(bytes);
        hipLaunchKernelGGL((fftw_inv_comb_fwd<T, NN, NWC, Pro, Comb>), dim3(nslab * (NN / (2 * NWC))), dim3(64 * NWC),
                           lds_t, st, plan.tw, nslab, pro, comb, out, gate_);
        if (tm) prof_stop();
        SW_HIP(hipGetLastError());
    }

    // A de-aliasing transform pair: the forward 2-D transform of pro (pro(s, x, y): physical value at
    // (y, x)), each spectral coefficient multiplied by mul(s, ky, kx) (and stored at
    // spec[s n^2 + kx + n ky] when spec is not null), then the inverse 2-D transform into epi
    // (epi(s, x, y, v)).  On the wave path the forward pass along x, the multiply and the inverse pass
    // along kx are one launch (fftw_fwd_mul_inv).  NW1: waves per block of the first pass (0: NWT).
    // kcut: mul is zero on every row ky whose wavenumber exceeds it (band_ for the filter and glap;
    // < 0: no such rows); those rows are neither stored nor read between the three launches.
    template <int NW1 = 0, class Pro, class Mul, class Epi>
    void fwd_mul_inv(int nslab, Pro pro, Mul mul, Epi epi, T* spec = nullptr, int kcut = -1) {
        if (nslab <= 0) return;
        if (wave3_n()) {
            // pass 1 along y (transposing), then forward along x, multiply and inverse along kx in one
            // launch, then the straight pass along ky
            constexpr int N = 512, N3 = 1536;
            T* t1 = tmp_for(nslab);
            T* t2 = W("fmi", (size_t)nslab * n2);
            const size_t l1 = wave3_lds_bytes<N, NG3, T>(), l2 = wave3_lds_bytes<N, 1, T>();
            lds_optin((const void*)fftw3_fwd_pass<T, N, NG3, true, Pro, NoEpi>, l1);
            lds_optin((const void*)fftw3_fwd_mul_inv<T, N, NG3, LoadPlain<T>, Mul>, l1);
            lds_optin((const void*)fftw3_inv_pass<T, N, 1, false, LoadPlain<T>, Epi>, l2);
            hipLaunchKernelGGL((fftw3_fwd_pass<T, N, NG3, true, Pro, NoEpi>), dim3(nslab * (N3 / (2 * NG3))), dim3(192 * NG3), l1,
                               st, plan.tw, nslab, pro, NoEpi{}, t1, Gate{});
            SW_HIP(hipGetLastError());
            LoadPlain<T> lp1{t1, n2, ng};
            hipLaunchKernelGGL((fftw3_fwd_mul_inv<T, N, NG3, LoadPlain<T>, Mul>), dim3(nslab * (N3 / (2 * NG3))), dim3(192 * NG3),
                               l1, st, plan.tw, nslab, lp1, mul, spec, t2);
            SW_HIP(hipGetLastError());
            LoadPlain<T> lp2{t2, n2, ng};
            hipLaunchKernelGGL((fftw3_inv_pass<T, N, 1, false, LoadPlain<T>, Epi>), dim3(nslab * (N3 / 2)), dim3(192), l2, st,
                               plan.tw, nslab, lp2, epi, (T*)nullptr, Gate{});
            SW_HIP(hipGetLastError());
            return;
        }
        if (wave_n()) {
            T* t1 = tmp_for(nslab);
            T* t2 = W("fmi", (size_t)nslab * n2);
            wave_dispatch([&](auto nn_) {
                constexpr int NN = decltype(nn_)::value;
                const int kc = kcut < 0 ? NN / 2 : kcut;
                fwd_pass1_wave<NN, NW1 ? NW1 : NWT, true>(nslab, pro, t1, SITE_OTHER, kc);
                const size_t lds = wave_lds_bytes<NN, NWT, T>();
                LoadPlain<T> lp{t1, n2, ng};
                hipLaunchKernelGGL((fftw_fwd_mul_inv<T, NN, NWT, LoadPlain<T>, Mul>), dim3(nslab * (NN / (2 * NWT))),
                                   dim3(64 * NWT), lds, st, plan.tw, nslab, lp, mul, spec, t2, kc);
                SW_HIP(hipGetLastError());
                inv_pass2_wave<NN, true>(nslab, t2, epi, kc);
            });
            return;
        }
        T* S = spec ? spec : W("fmis", (size_t)nslab * n2);
        const int n = ng;
        const size_t nn = n2;
        fwd2d<NW1>(nslab, pro, [=] __device__(int s, int ky, int kx, T v) { S[(size_t)s * nn + kx + (size_t)n * ky] = v * mul(s, ky, kx); });
        inv2d(nslab, ld(S), epi);
    }

    // Horizontal derivatives of physical slabs (DerivIo) as 1-D transform pairs along their axis
    // (fftw_deriv_lines, one launch for all nslab outputs).  Returns false, launching nothing, when
    // ng has no wave kernels: the caller then forms them with a 2-D transform pair.
    template <int K, typename O>
    bool deriv_lines(int nslab, const DerivIo<T, K, O>& io) {
        if (wave3_n()) {
            if (nslab <= 0) return true;
            constexpr int N = 512, N3 = 1536;
            const size_t lds = wave3_lds_bytes<N, NG3, T>();
            lds_optin((const void*)fftw3_deriv_lines<T, N, NG3, DerivIo<T, K, O>>, lds);
            hipLaunchKernelGGL((fftw3_deriv_lines<T, N, NG3, DerivIo<T, K, O>>), dim3(nslab * (N3 / (2 * NG3))), dim3(192 * NG3), lds,
                               st, plan.tw, nslab, io);
            SW_HIP(hipGetLastError());
            return true;
        }
        if (!wave_n()) return false;
        if (nslab <= 0) return true;
        wave_dispatch([&](auto nn_) {
            constexpr int NN = decltype(nn_)::value;
            const size_t lds = wave_lds_bytes<NN, NWT, T>();
            hipLaunchKernelGGL((fftw_deriv_lines<T, NN, NWT, DerivIo<T, K, O>>), dim3(nslab * (NN / (2 * NWT))),
                               dim3(64 * NWT), lds, st, plan.tw, nslab, io);
            SW_HIP(hipGetLastError());
        });
        return true;
    }

    // plain slab accessors (level offset lev0)
    LoadPlain<T> ld(const T* a, int lev0 = 0) const { return LoadPlain<T>{a + (size_t)lev0 * n2, n2, ng}; }
    StorePlain<T> stp(T* a, int lev0 = 0) const { return StorePlain<T>{a + (size_t)lev0 * n2, n2, ng}; }
    // per consecutive levels lev0 .. lev0 + per - 1 of every member of the batch (slab m per + l)
    LoadML<T> ldm(const T* a, int lev0, int per) const { return LoadML<T>{a + (size_t)lev0 * n2, per, N, n2, ng}; }
    StoreML<T> stm(T* a, int lev0, int per) const { return StoreML<T>{a + (size_t)lev0 * n2, per, N, n2, ng}; }

    // per member m of the batch, the block partials of f(m, i) over i < n left in out[] (gated by gate_);
    // returns the block count per member
    template <int K, unsigned MAXMASK, class F>
    int reduce_dev(size_t n, F f, double* out) {
        int blocks = (int)std::min<size_t>(RED_BLOCKS, (n + 255) / 256);
        hipLaunchKernelGGL((k_reduce<K, MAXMASK, F>), dim3(blocks, M), dim3(256), 0, st, n, f, out, gate_);
        SW_HIP(hipGetLastError());
        return blocks;
    }
    // a reduction of f(i) over i < n, finished on the host (one member)
    template <int K, unsigned MAXMASK, class F>
    void reduce(size_t n, F f0, double* out_host) {
        int blocks = (int)std::min<size_t>(RED_BLOCKS, (n + 255) / 256);
        auto f = [=] __device__(int, size_t i) { return f0(i); };
        hipLaunchKernelGGL((k_reduce<K, MAXMASK, decltype(f)>), dim3(blocks), dim3(256), 0, st, n, f, part.p, Gate{});
        SW_HIP(hipGetLastError());
        std::vector<double> h((size_t)blocks * K);
        copy_d2h(h.data(), part.p, h.size() * sizeof(double), st);
        for (int c = 0; c < K; c++) {
            bool mx = (MAXMASK >> c) & 1;
            double acc = mx ? -INFINITY : 0.0, cmp = 0.0;
            for (int b = 0; b < blocks; b++) {
                double x = h[(size_t)b * K + c];
                if (mx) {
                    acc = std::fmax(acc, x);
                } else {
                    double y = x - cmp, t = acc + y;
                    cmp = (t - acc) - y;
                    acc = t;
                }
            }
            out_host[c] = acc;
        }
    }

    // ---------------------------------------------------------------- spectral module
    void ptospc3d(const T* fp, T* fs, int izbeg, int izend) { fwd2d(izend - izbeg + 1, ld(fp, izbeg), stp(fs, izbeg)); }
    void spctop3d(const T* fs, T* fp, int izbeg, int izend) { inv2d(izend - izbeg + 1, ld(fs, izbeg), stp(fp, izbeg)); }

    // deal3d / deal2d: ptospc, multiply by the 2/3-rule filter, spctop (spectral/mod.rs:524-554)
    void deal(T* fp, int nlev) {
        const T* F = filt.p;
        const int n = ng;
        fwd_mul_inv(nlev, ld(fp), [=] __device__(int, int ky, int kx) { return F[kx + (size_t)n * ky]; }, stp(fp), nullptr, band_);
    }

    void divs(const T* a, const T* b, T* cs) {  // spectral/mod.rs:461-475
        T* S = W("divs", 2 * n2);
        const int n = ng;
        const size_t nn = n2;
        const PtrTab<const T, 2> ab{{a, b}};
        fwd2d(2, [=] __device__(int s, int x, int y) { return ab[s][y + (size_t)n * x]; }, stp(S));
        each(st, n2, [=] __device__(size_t p) {
            int kx = p % n, ky = p / n;
            T xd = dmul<T>(kx, n) * S[mir(kx, n) + (size_t)n * ky];
            T yd = dmul<T>(ky, n) * S[nn + kx + (size_t)n * mir(ky, n)];
            cs[p] = yd + xd;
        });
    }

    void jacob(const T* a, const T* b, T* cs) {  // spectral/mod.rs:423-457
        T* S = W("jacS", 2 * n2);
        T* D = W("jacD", 4 * n2);
        const int n = ng;
        const size_t nn = n2;
        const PtrTab<const T, 2> ab{{a, b}};
        fwd2d(2, [=] __device__(int s, int x, int y) { return ab[s][y + (size_t)n * x]; }, stp(S));
        inv2d(4, [=] __device__(int s, int ky, int kx) {
            const T* g = S + (s >> 1) * nn;
            return (s & 1) ? dmul<T>(ky, n) * g[kx + (size_t)n * mir(ky, n)] : dmul<T>(kx, n) * g[mir(kx, n) + (size_t)n * ky];
        }, stp(D));
        fwd2d(1, [=] __device__(int, int x, int y) {
            size_t q = y + (size_t)n * x;
            return D[q] * D[3 * nn + q] - D[nn + q] * D[2 * nn + q];
        }, stp(cs));
    }

    // ---------------------------------------------------------------- mixed precision
    // SW_MIXED: the FP32 engine (same ng, nz and stream) that computes psolve's coefficient fields and
    // runs its iteration; null in FP64.  Every other stage, and the state, stay in FP64: the FP32
    // rounding of a physical field is white in wavenumber, and the Laplacians and derivative products
    // downstream of the inversion, vertical, cpsource and source amplify it by up to (ng/3)^2, whereas
    // the pressure iteration solves an elliptic problem, which damps it.
    Engine<float>* sub = nullptr;
    // within psolve, after the FP32 iteration (nz with a compile-time specialisation): its ps and S,
    // read by the launches that follow instead of FP64 copies (null otherwise)
    const float* ps_f32 = nullptr;
    const float* sptop_f32 = nullptr;
    // J(u, v) = ux vy - uy vx on every level, formed by cpsource when it stores the derivatives of u, v in
    // FP32 (the mixed mode), for source's reuse of them (null otherwise)
    const T* juv_src = nullptr;
    void to_f32(const T* a, float* b, size_t n) {
        each(st, n, [=] __device__(size_t i) { b[i] = (float)a[i]; });
    }
    void to_f64(const float* a, T* b, size_t n) {
        each(st, n, [=] __device__(size_t i) { b[i] = (T)a[i]; });
    }
    void set_stream(hipStream_t s) {
        st = s;
        if (sub) sub->st = s;
    }
    // the device control state of psolve call slot k (0: advance's predictor, 1, 2: its correctors), in
    // the engine that runs the iteration
    PsolveCtrl* ctrl_slot(int k) { return (sub ? sub->ctrl.p : ctrl.p) + k; }
    int* hint_slot(int k) { return sub ? &sub->iter_hint[k] : &iter_hint[k]; }

    // ---------------------------------------------------------------- nhswps and the time loop
    // Defined in sw_engine_ops.inc and instantiated once per precision (sw_engine_f64.hip,
    // sw_engine_f32.hip).
    struct Fields {
        T* a[SW_NFIELDS];
    };
    void main_invert(const T* qs, const T* ds, const T* gs, T* r, T* u, T* v, T* zeta);
    void vertical(Fields f);
    void coeffs(Fields f, T* sigx, T* sigy, T* cpt1, T* cpt2);
    template <typename U>
    void coeffs_from(const U* ri, const U* zx, const U* zy, T* sigx, T* sigy, T* cpt1, T* cpt2);
    // sp0f: when not null, an FP32 copy of sp0 written beside it (the mixed mode's FP32 engine reads it)
    void cpsource(Fields f, T* sp0, float* sp0f = nullptr);
    T* psolve_source_wave(const T* ps, const T* sp0, const T* sigx, const T* sigy, const T* cpt1, const T* cpt2);
    T* psolve_source_wave3(const T* ps, const T* sp0, const T* sigx, const T* sigy, const T* cpt1, const T* cpt2);
    // the fused iteration of the wave path with a compile-time nz: derivative rows in the tiled layout
    // (tile_pos, TILE_RB = the rows of one source-fold block: 8 in FP64, 16 in FP32, so that the
    // column-solve launch stores runs of 64 bytes in either precision), written by the column-solve launch
    // and read by the source fold.  In FP32 the 16-row tiles against 8-row ones: the column solve 41.7 ->
    // 36.0 us and the first-guess rows 39.6 -> 25.7 us per launch, the fold (whose two-row loads are then
    // 64 bytes apart) 39.6 -> 45.7 us, at 512x16: psolve's launches 3% faster per step.
    static constexpr int TILE_RB = 2 * NWC;
    bool psolve_fused() const { return wave_n() && nz_spec(); }
    void psolve_deriv_tiled(const T* ps, T* D);
    T* psolve_fold_tiled(const T* D, const T* sp0, const T* sigx, const T* sigy, const T* cpt1, const T* cpt2);
    void psolve_solve_fused(const T* SPt, T* SP, T* ps, T* D);
    void psolve_iterate_device(const T* sp0, const T* sigx, const T* sigy, const T* cpt1, const T* cpt2, T* SP, T* ps,
                               const T* pn, int nitmax, int* iters, double* errp, int* status);
    int psolve_step(const T* sp0, const T* sigx, const T* sigy, const T* cpt1, const T* cpt2, T* SP, T* ps, double* sums);
    int iterate(const T* sp0, const T* sigx, const T* sigy, const T* cpt1, const T* cpt2, T* SP, T* ps, const T* pn,
                int nitmax, int* iters_out, double* errp_out);
    int psolve(Fields f, int nitmax, int* iters_out, double* errp_out);
    void source(Fields f, T* sqs, T* sds, T* sgs, bool reuse_derivs = false);
    void diagnose(Fields f, double t, double out[5]);
    void diagnose_launch(Fields f);
    void diagnose_finish(double t, double out[5]);
    DevBuf<double> dpart;
    int diag_blocks = 0;
    // savegrid's output (src/nhswps/mod.rs:438-652): d3 = ql, d, g, r, w, pn as f32 (6 fields), d2 = the
    // vertically integrated d, zeta, q, g, h as f32 (5 slabs), spec3 = the zeta, delta, gamma spectra
    // (3 x (ng + 1)); host arrays
    void savegrid(Fields f, float* d3, float* d2, double* spec3);
    // spec1d's shells (spectral/mod.rs:558-591): the packed indices of each shell k = kmag in the
    // order spec1d visits them, with its weights (1/4 at (0, 0), 1/2 on the axes, 1 elsewhere)
    DevBuf<int> shell_idx, shell_off;
    DevBuf<double> shell_w;
    void energy(Fields f, double t, double out[6]);
    // balinit (src/balinit.rs:21-436) on one ng x ng slab: qq0 in, the balanced qq, dd, gg out (physical,
    // device arrays); returns the iteration count
    int balinit(const T* qq0, T* qq, T* dd, T* gg);
    // observer of the time loop: a monitor row per step, and each save (after its energies, with the
    // state's fields inverted and the pressure solved)
    struct RunHooks {
        virtual ~RunHooks() = default;
        virtual void monitor(const double d[5]) { (void)d; }
        virtual void save(const double e[6]) { (void)e; }
    };
    int run(Fields f, double& t, double tsim, double tgsave, double tstep, double* mon, int maxmon, int* nmon,
            double* eco, int maxeco, int* neco, int* iters_total, RunHooks* hooks = nullptr);
    int advance(Fields f, double& t, bool ggen, double diag[5], int iters[3]);
    // one step of every member of the batch: t[M], diag[5 M], iters[3 M], status[M]; returns the first
    // member's non-zero status
    int advance_batch(Fields f, double* t, bool ggen, double* diag, int* iters, int* status);
    // the monitor values of every member from diagnose's partials: t[M] in, out[5 M]
    void diagnose_finish_batch(const double* t, double* out);
};

// The two precisions are instantiated in their own translation units.
extern template struct Engine<double>;
extern template struct Engine<float>;

}  // namespace sw
