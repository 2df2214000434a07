// extern "C" entry points of libsw_hip.so (declared in include/shallow_water.h).  Each entry point
// validates its arguments, stages host arrays through device memory when needed, dispatches on the
// precision of the handle and turns exceptions into sw_status codes.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <atomic>
#include <vector>

#include "shallow_water.h"
#include "sw_engine.hpp"

using namespace sw;

namespace {

thread_local std::string g_err;
int g_device = 0;
hipStream_t g_stream = nullptr;
std::mutex g_mu;

hipStream_t stream() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_stream) {
        SW_HIP(hipSetDevice(g_device));
        SW_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    }
    return g_stream;
}

template <class F>
int guard(F f) {
    try {
        return f();
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return SW_ERR_ARG;
    }
}

__global__ void k_cvt_d2f(const double* a, float* b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_cvt_f2d(const float* a, double* b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

// One array argument at the boundary: a double* that may be host or device memory, seen by the
// kernels as a device T*.
template <typename T>
struct Arg {
    double* user;
    size_t n;
    bool out;
    bool dev;
    DevBuf<double> stage;
    DevBuf<T> conv;
    T* d = nullptr;
    hipStream_t st;
    Arg(const double* u, size_t count, bool in, bool out_, hipStream_t s)
        : user(const_cast<double*>(u)), n(count), out(out_), st(s) {
        SW_CHECK(u != nullptr, SW_ERR_ARG, "null array argument");
        dev = is_device_ptr(u);
        double* dd = user;
        if (!dev) {
            stage.alloc(n);
            dd = stage.p;
            if (in) copy_h2d(dd, user, n * sizeof(double), st);
        }
        if constexpr (sizeof(T) == sizeof(double)) {
            d = reinterpret_cast<T*>(dd);
        } else {
            conv.alloc(n);
            d = conv.p;
            if (in) hipLaunchKernelGGL(k_cvt_d2f, dim3((n + 255) / 256), dim3(256), 0, st, dd, conv.p, n);
        }
    }
    void finish() {
        if (!out) return;
        double* dd = dev ? user : stage.p;
        if constexpr (sizeof(T) != sizeof(double))
            hipLaunchKernelGGL(k_cvt_f2d, dim3((n + 255) / 256), dim3(256), 0, st, conv.p, dd, n);
        SW_HIP(hipStreamSynchronize(st));
        if (!dev) copy_d2h(user, dd, n * sizeof(double), st);
    }
};

// Scratch of sw_ptospc_noalloc: the intermediate slab between the two passes of its 2-D transform,
// for grids up to NOALLOC_MAX_NG x NOALLOC_MAX_NG (2 MiB, held by every device the library runs on
// until the process ends).  Each device has its own copy at its own address: host code never uses
// the array's name as an address, only the pointer hipGetSymbolAddress gives for the current device
// (scratch_on_device()).  All calls on one device share it, so such calls must not overlap.
constexpr size_t NOALLOC_MAX_NG = 512;
static __device__ __align__(256) double scratch[NOALLOC_MAX_NG * NOALLOC_MAX_NG];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// The address of this device's scratch, looked up once per device with that device current.
double* scratch_on_device() {
    int dev = 0;
    SW_HIP(hipGetDevice(&dev));
    SW_CHECK(dev >= 0 && dev < MAX_DEVICES, SW_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SW_CHECK(p != nullptr, SW_ERR_HIP, "no device address for the transform scratch");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

struct sw_d2fft {
    size_t nx, ny;
    std::unique_ptr<Engine<double>> e;
};

// A spectral handle: the FP64 engine, and for SW_MIXED the FP32 engine that runs psolve's coefficient
// fields and iteration inside it (Engine::sub).
struct sw_spectral {
    int precision;
    size_t ng, nz;
    std::unique_ptr<Engine<double>> e64;
    std::unique_ptr<Engine<float>> e32;
};

// The 16 fields of the reference State, FP64 in either precision mode.
struct sw_state {
    sw_spectral* s;
    double t = 0.0;
    int ggen = 0;
    DevBuf<double> f64[SW_NFIELDS];
    Engine<double>::Fields fields() {
        Engine<double>::Fields f;
        for (int i = 0; i < SW_NFIELDS; i++) f.a[i] = f64[i].p;
        return f;
    }
};

// A member's captured step: one advance (with ggen set) recorded as a hipGraph on its stream, psolve's
// three iterations each given a fixed number of gated iterations (cap); the 16 fields are first copied
// to snap, so that a step whose psolve needed more iterations than captured can be undone and redone
// eagerly.
struct MemberGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool valid = false;
    int cap[3] = {0, 0, 0};
    DevBuf<double> snap;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        valid = false;
    }
    ~MemberGraph() { reset(); }
};

// Members of one ng, nz and precision advanced as one batch (an engine with M members, which runs every
// stage of the step for all of them in the same launches): their 16 fields are gathered into
// member-contiguous batch fields before the step and scattered back after it.
struct MemberBatch {
    std::vector<int> idx;  // member indices
    int precision = SW_FP64;
    std::unique_ptr<Engine<double>> e64;
    std::unique_ptr<Engine<float>> e32;
    DevBuf<double> f[SW_NFIELDS];  // M N values each
    DevBuf<double*> tab;           // tab[m SW_NFIELDS + k]: member m's field k
    Engine<double>::Fields fields() {
        Engine<double>::Fields a;
        for (int k = 0; k < SW_NFIELDS; k++) a.a[k] = f[k].p;
        return a;
    }
};

// An ensemble: member states (each with its own spectral handle, hence its own engine and work buffers)
// advanced together, spread over a few HIP streams each driven by one host thread.  Members of the same
// shape and precision (nz with a compile-time specialisation) run as batches of at most batch_max; the
// others one by one, and with graphs on each such member's step after the first is a replay of its
// captured step.
struct sw_ensemble {
    std::vector<sw_state*> m;
    std::vector<hipStream_t> streams;
    std::vector<std::vector<int>> groups;   // unbatched member indices per stream
    std::vector<std::vector<int>> bgroups;  // batch indices per stream
    std::vector<std::unique_ptr<MemberBatch>> batches;
    std::vector<std::unique_ptr<MemberGraph>> graphs;
    int batch_max = 32;
    int nstreams_req = 0;  // 0: one stream per unit of work (batch or single member), at most 8
    int use_graphs = 0;
    int margin = 4;  // iterations captured beyond the last count, per psolve call
    long replays = 0, redos = 0, captures = 0, batched_steps = 0;
    ~sw_ensemble() {
        for (auto s : streams) (void)hipStreamSynchronize(s);
        batches.clear();
        graphs.clear();
        for (auto s : streams) (void)hipStreamDestroy(s);
    }
};

// Run fn(engine) with the handle's engine (the FP64 one; in SW_MIXED it carries the FP32 one).
template <class F>
static int with_engine(sw_spectral* s, F fn) {
    SW_CHECK(s != nullptr, SW_ERR_ARG, "null spectral handle");
    return fn(*s->e64);
}

template <class F>
static int with_state(sw_state* st, F fn) {
    SW_CHECK(st && st->s, SW_ERR_ARG, "null state");
    return with_engine(st->s, [&](auto& e) { return fn(e, st->fields()); });
}

extern "C" {

const char* sw_last_error(void) { return g_err.c_str(); }

int sw_init(int device) {
    return guard([&] {
        std::lock_guard<std::mutex> lk(g_mu);
        int cnt = 0;
        SW_HIP(hipGetDeviceCount(&cnt));
        SW_CHECK(device >= 0 && device < cnt, SW_ERR_ARG, "no such device");
        g_device = device;
        SW_HIP(hipSetDevice(device));
        if (g_stream) {
            SW_HIP(hipStreamDestroy(g_stream));
            g_stream = nullptr;
        }
        (void)scratch_on_device();
        return 0;
    });
}

int sw_sync(void) {
    return guard([&] {
        SW_HIP(hipStreamSynchronize(stream()));
        return 0;
    });
}

int sw_factorisen(size_t n, size_t factors[5]) {
    return guard([&] {
        SW_CHECK(factors, SW_ERR_ARG, "null factors");
        return factorise_ref(n, factors) ? 0 : (int)SW_ERR_FACTOR;
    });
}

int sw_initfft(size_t n, size_t factors[5], double* trig) {
    return guard([&] {
        SW_CHECK(factors && trig, SW_ERR_ARG, "null argument");
        SW_CHECK(factorise_ref(n, factors), SW_ERR_FACTOR, "factorisation failed");
        initfft_ref(n, factors, trig);
        return 0;
    });
}

// forfft/revfft take the caller's trig table (stafft/mod.rs:158-162 asserts only its length); the kernels
// use twiddles of their own, exp(-2 pi i k / n), so the caller's table must be initfft(n)'s, which is
// checked value for value (host or device memory) to 1e-12: tables from the reference's initfft, whose
// libm sin/cos may round differently, pass; a table of another length or content is an error.
static void check_trig(size_t n, const double* trig) {
    SW_CHECK(trig != nullptr, SW_ERR_ARG, "null trig");
    size_t f[5];
    std::vector<double> want(2 * n), got(2 * n);
    SW_CHECK(factorise_ref(n, f), SW_ERR_FACTOR, "factorisation failed");
    initfft_ref(n, f, want.data());
    if (is_device_ptr(trig)) copy_d2h(got.data(), trig, 2 * n * sizeof(double), stream());
    else std::memcpy(got.data(), trig, 2 * n * sizeof(double));
    bool same = true;
    for (size_t i = 0; i < 2 * n; i++) same = same && std::fabs(want[i] - got[i]) <= 1e-12;
    SW_CHECK(same, SW_ERR_ARG, "trig is not initfft(n)'s table (the transforms of length n use that table)");
}

static int fft1d(size_t m, size_t n, double* x, const double* trig, const size_t factors[5], bool inv) {
    return guard([&] {
        SW_CHECK(x && m > 0 && n >= 2 && m <= (1u << 30) && n <= (1u << 16), SW_ERR_ARG, "bad forfft/revfft shape");
        size_t f[5];
        SW_CHECK(factorise_ref(n, f), SW_ERR_FACTOR, "factorisation failed");
        if (factors) SW_CHECK(std::memcmp(f, factors, sizeof(f)) == 0, SW_ERR_ARG, "factors do not match n");
        check_trig(n, trig);
        SW_CHECK(reference_transform_is_dft(f), SW_ERR_UNSUPPORTED, "length not supported");
        hipStream_t st = stream();
        FftPlan<double> pl{};
        DevBuf<C2<double>> tw;
        size_t lds;
        // rows (the batch index i of x(m,n)) per block must divide m
        int B = 1;
        for (int b = 16; b >= 2; b /= 2)
            if (m % b == 0 && (size_t)b * n * sizeof(double) <= 65536 && (size_t)(b / 2) * n <= (size_t)FFT_EPT_MAX * 512) {
                B = b;
                break;
            }
        Engine<double>::make_plan((int)n, pl, tw, lds, B, 0);
        pl.rows = (int)m;
        pl.xcd = 0;
        Arg<double> a(x, m * n, true, true, st);
        double* d = a.d;
        const size_t mm = m;
        auto pro = [=] __device__(int, int r, int j) { return d[(size_t)r + mm * j]; };
        auto epi = [=] __device__(int, int r, int f, double v) { d[(size_t)r + mm * f] = v; };
        dim3 grid((unsigned)(m / B)), block(pl.nt);
        if (!inv)
            hipLaunchKernelGGL((fft_fwd_pass<double, false, decltype(pro), decltype(epi)>), grid, block, lds, st, pl, 1, pro,
                               epi, (double*)nullptr);
        else
            hipLaunchKernelGGL((fft_inv_pass<double, false, decltype(pro), decltype(epi)>), grid, block, lds, st, pl, 1, pro,
                               epi, (double*)nullptr);
        SW_HIP(hipGetLastError());
        a.finish();
        return 0;
    });
}

// ---- tiles of the block-per-rows transform passes

int sw_fft_set_tile(size_t n, int bytes, int B, int nt) {
    return guard([&] {
        SW_CHECK((bytes == 4 || bytes == 8) && n >= 2 && n <= 65536, SW_ERR_ARG, "bad length or value size");
        std::lock_guard<std::mutex> lk(tile_mu());
        if (B == 0) {
            tile_overrides().erase({(int)n, bytes});
            return 0;
        }
        SW_CHECK(tile_valid((int)n, bytes, B, nt), SW_ERR_ARG, "invalid tile for this length");
        tile_overrides()[{(int)n, bytes}] = {B, nt};
        return 0;
    });
}

int sw_fft_get_tile(size_t n, int bytes, int* B, int* nt) {
    return guard([&] {
        SW_CHECK(B && nt && (bytes == 4 || bytes == 8) && n >= 2 && n <= 65536, SW_ERR_ARG, "bad argument");
        choose_tile((int)n, bytes, 0, 0, *B, *nt);
        return 0;
    });
}

}  // extern "C"

namespace {

// fwd2d then inv2d of nslab plain slabs with the passes of tile (B, nt), as the engine launches them
// (transposing first pass, straight second pass); returns microseconds per transform pair
template <typename T>
double time_tile(int n, int B, int nt, int nslab, int reps, hipStream_t st) {
    FftPlan<T> pl{};
    DevBuf<C2<T>> tw;
    size_t lds = 0;
    Engine<T>::make_plan(n, pl, tw, lds, B, nt);
    const size_t n2 = (size_t)n * n, total = n2 * (size_t)nslab;
    DevBuf<T> a(total), t(total);
    T* ap = a.p;
    each(st, total, [=] __device__(size_t i) { ap[i] = T((double)((i * 2654435761u) % 1000003u) / 1000003.0 - 0.5); });
    using NoEpi = typename Engine<T>::NoEpi;
    const void* k[4] = {(const void*)fft_fwd_pass<T, true, LoadPlain<T>, NoEpi>,
                        (const void*)fft_fwd_pass<T, false, LoadPlain<T>, StorePlain<T>>,
                        (const void*)fft_inv_pass<T, true, LoadPlain<T>, NoEpi>,
                        (const void*)fft_inv_pass<T, false, LoadPlain<T>, StorePlain<T>>};
    for (const void* f : k) lds_optin(f, lds);
    dim3 grid((unsigned)(nslab * (n / B))), block((unsigned)nt);
    const LoadPlain<T> la{a.p, n2, n}, lt{t.p, n2, n};
    const StorePlain<T> sa{a.p, n2, n};
    auto run = [&]() {
        hipLaunchKernelGGL((fft_fwd_pass<T, true, LoadPlain<T>, NoEpi>), grid, block, lds, st, pl, nslab, la, NoEpi{}, t.p);
        hipLaunchKernelGGL((fft_fwd_pass<T, false, LoadPlain<T>, StorePlain<T>>), grid, block, lds, st, pl, nslab, lt, sa,
                           (T*)nullptr);
        hipLaunchKernelGGL((fft_inv_pass<T, true, LoadPlain<T>, NoEpi>), grid, block, lds, st, pl, nslab, la, NoEpi{}, t.p);
        hipLaunchKernelGGL((fft_inv_pass<T, false, LoadPlain<T>, StorePlain<T>>), grid, block, lds, st, pl, nslab, lt, sa,
                           (T*)nullptr);
        SW_HIP(hipGetLastError());
    };
    run();
    SW_HIP(hipStreamSynchronize(st));
    hipEvent_t e0, e1;
    SW_HIP(hipEventCreate(&e0));
    SW_HIP(hipEventCreate(&e1));
    SW_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) run();
    SW_HIP(hipEventRecord(e1, st));
    SW_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SW_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1000.0 * (double)ms / (double)reps;
}

}  // namespace

extern "C" {

int sw_fft_time_tile(size_t n, int bytes, int B, int nt, int nslab, int reps, double* us) {
    return guard([&] {
        SW_CHECK(us && nslab > 0 && reps > 0 && (bytes == 4 || bytes == 8), SW_ERR_ARG, "bad argument");
        SW_CHECK(n >= 4 && n <= 65536 && tile_valid((int)n, bytes, B, nt), SW_ERR_ARG, "invalid tile for this length");
        std::vector<int> r = gpu_radices((int)n);
        SW_CHECK(!r.empty(), SW_ERR_FACTOR, "unsupported transform length");
        hipStream_t st = stream();
        *us = bytes == 8 ? time_tile<double>((int)n, B, nt, nslab, reps, st) : time_tile<float>((int)n, B, nt, nslab, reps, st);
        return 0;
    });
}

int sw_forfft(size_t m, size_t n, double* x, const double* trig, const size_t factors[5]) {
    return fft1d(m, n, x, trig, factors, false);
}
int sw_revfft(size_t m, size_t n, double* x, const double* trig, const size_t factors[5]) {
    return fft1d(m, n, x, trig, factors, true);
}

int sw_d2fft_create(size_t nx, size_t ny, double lx, double ly, double* kx, double* ky, sw_d2fft** out) {
    return guard([&] {
        SW_CHECK(out, SW_ERR_ARG, "null output handle");
        SW_CHECK(nx == ny, SW_ERR_UNSUPPORTED, "only square grids (nx == ny) are supported");
        SW_CHECK(lx != 0.0, SW_ERR_ARG, "domain length in x equal to zero not allowed");
        auto d = std::make_unique<sw_d2fft>();
        d->nx = nx;
        d->ny = ny;
        d->e = std::make_unique<Engine<double>>((int)nx, 2, stream());
        if (kx) for (size_t k = 1; k <= nx; k++) kx[k - 1] = (PI / lx) * (double)k;
        if (ky && ly != 0.0) for (size_t k = 1; k <= ny; k++) ky[k - 1] = (PI / ly) * (double)k;
        *out = d.release();
        return 0;
    });
}

int sw_d2fft_destroy(sw_d2fft* d) {
    delete d;
    return 0;
}

int sw_ptospc(const sw_d2fft* d, double* rvar, double* svar) {
    return guard([&] {
        SW_CHECK(d, SW_ERR_ARG, "null handle");
        auto& e = *d->e;
        Arg<double> a(rvar, e.n2, true, false, e.st), b(svar, e.n2, false, true, e.st);
        e.fwd2d(1, e.ld(a.d), e.stp(b.d));
        b.finish();
        return 0;
    });
}

int sw_ptospc_noalloc(const sw_d2fft* d, const double* rvar, double* svar) {
    return guard([&] {
        SW_CHECK(d && rvar && svar, SW_ERR_ARG, "null argument");
        auto& e = *d->e;
        SW_CHECK(d->nx <= NOALLOC_MAX_NG, SW_ERR_ARG,
                 "sw_ptospc_noalloc: ng = " + std::to_string(d->nx) + " exceeds its scratch (ng <= " +
                     std::to_string(NOALLOC_MAX_NG) + ")");
        SW_CHECK(is_device_ptr(rvar) && is_device_ptr(svar), SW_ERR_ARG,
                 "sw_ptospc_noalloc takes device arrays only (a host array would need a staging buffer)");
        double* ws = scratch_on_device();
        e.fwd2d(1, e.ld(rvar), e.stp(svar), Engine<double>::SITE_OTHER, ws);
        SW_HIP(hipStreamSynchronize(e.st));
        return 0;
    });
}

int sw_spctop(const sw_d2fft* d, double* svar, double* rvar) {
    return guard([&] {
        SW_CHECK(d, SW_ERR_ARG, "null handle");
        auto& e = *d->e;
        Arg<double> a(svar, e.n2, true, false, e.st), b(rvar, e.n2, false, true, e.st);
        e.inv2d(1, e.ld(a.d), e.stp(b.d));
        b.finish();
        return 0;
    });
}

// xderiv/yderiv with the caller's wavenumber array, exactly as sta2dfft.rs:128-184
static int deriv(const sw_d2fft* d, const double* rk, const double* var, double* der, bool ydir) {
    return guard([&] {
        SW_CHECK(d && rk, SW_ERR_ARG, "null argument");
        auto& e = *d->e;
        const int n = (int)d->nx;
        Arg<double> k(rk, (size_t)n, true, false, e.st), a(var, e.n2, true, false, e.st), b(der, e.n2, false, true, e.st);
        const double *K = k.d, *A = a.d;
        double* Dd = b.d;
        each(e.st, e.n2, [=] __device__(size_t p) {
            int kx = p % n, ky = p / n;
            int i = ydir ? ky : kx;
            double val = 0.0;
            if (i != 0 && 2 * i != n) {
                size_t q = ydir ? kx + (size_t)n * (n - ky) : (n - kx) + (size_t)n * ky;
                val = 2 * i < n ? -K[2 * i - 1] * A[q] : K[2 * (n - i) - 1] * A[q];
            }
            Dd[p] = val;
        });
        b.finish();
        return 0;
    });
}
int sw_xderiv(const sw_d2fft* d, const double* rkx, const double* var, double* der) { return deriv(d, rkx, var, der, false); }
int sw_yderiv(const sw_d2fft* d, const double* rky, const double* var, double* der) { return deriv(d, rky, var, der, true); }

int sw_spectral_create(size_t ng, size_t nz, int precision, sw_spectral** out) {
    return guard([&] {
        SW_CHECK(out, SW_ERR_ARG, "null output handle");
        SW_CHECK(precision == SW_FP64 || precision == SW_MIXED, SW_ERR_ARG, "bad precision");
        auto s = std::make_unique<sw_spectral>();
        s->precision = precision;
        s->ng = ng;
        s->nz = nz;
        s->e64 = std::make_unique<Engine<double>>((int)ng, (int)nz, stream());
        if (precision == SW_MIXED) {
            s->e32 = std::make_unique<Engine<float>>((int)ng, (int)nz, stream());
            s->e64->sub = s->e32.get();
        }
        *out = s.release();
        return 0;
    });
}

int sw_spectral_destroy(sw_spectral* s) {
    delete s;
    return 0;
}

long sw_spectral_get(const sw_spectral* s, const char* name, double* out) {
    try {
        SW_CHECK(s && name, SW_ERR_ARG, "null argument");
        const HostOps& h = s->e64->host;
        std::string n(name);
        const std::vector<double>* v = nullptr;
        if (n == "hlap") v = &h.hlap; else if (n == "glap") v = &h.glap; else if (n == "rlap") v = &h.rlap;
        else if (n == "helm") v = &h.helm; else if (n == "c2g2") v = &h.c2g2; else if (n == "simp") v = &h.simp;
        else if (n == "rope") v = &h.rope; else if (n == "fope") v = &h.fope; else if (n == "filt") v = &h.filt;
        else if (n == "diss") v = &h.diss; else if (n == "opak") v = &h.opak; else if (n == "rdis") v = &h.rdis;
        else if (n == "etdv") v = &h.etdv; else if (n == "htdv") v = &h.htdv; else if (n == "ap") v = &h.ap;
        else if (n == "etd1") v = &h.etd1; else if (n == "htd1") v = &h.htd1; else if (n == "theta") v = &h.theta;
        else if (n == "weight") v = &h.weight; else if (n == "hrkx") v = &h.hrkx; else if (n == "hrky") v = &h.hrky;
        else if (n == "rk") v = &h.rk; else if (n == "spmf") v = &h.spmf; else if (n == "alk") v = &h.alk;
        else if (n == "kmag") v = &h.kmag; else if (n == "xtrig" || n == "ytrig") v = &h.xtrig;
        else if (n == "kmax" || n == "kmaxred") {
            if (out) out[0] = (double)(n == "kmax" ? h.kmax : h.kmaxred);
            return 1;
        } else if (n == "xfactors" || n == "yfactors") {
            if (out) for (int i = 0; i < 5; i++) out[i] = (double)h.xf[i];
            return 5;
        }
        SW_CHECK(v != nullptr, SW_ERR_ARG, "unknown operator name");
        if (out) std::memcpy(out, v->data(), v->size() * sizeof(double));
        return (long)v->size();
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    }
}

int sw_main_invert(sw_spectral* s, const double* qs, const double* ds, const double* gs, double* r, double* u, double* v,
                   double* zeta) {
    return guard([&] {
        return with_engine(s, [&](auto& e) {
            using T = std::remove_reference_t<decltype(*e.tmp.p)>;
            Arg<T> aq(qs, e.N, true, false, e.st), ad(ds, e.N, true, false, e.st), ag(gs, e.N, true, false, e.st);
            Arg<T> ar(r, e.N, false, true, e.st), au(u, e.N, false, true, e.st), av(v, e.N, false, true, e.st),
                az(zeta, e.N, false, true, e.st);
            e.main_invert(aq.d, ad.d, ag.d, ar.d, au.d, av.d, az.d);
            ar.finish(); au.finish(); av.finish(); az.finish();
            return 0;
        });
    });
}

int sw_jacob(sw_spectral* s, const double* aa, const double* bb, double* cs) {
    return guard([&] {
        return with_engine(s, [&](auto& e) {
            using T = std::remove_reference_t<decltype(*e.tmp.p)>;
            Arg<T> a(aa, e.n2, true, false, e.st), b(bb, e.n2, true, false, e.st), c(cs, e.n2, false, true, e.st);
            e.jacob(a.d, b.d, c.d);
            c.finish();
            return 0;
        });
    });
}

int sw_divs(sw_spectral* s, const double* aa, const double* bb, double* cs) {
    return guard([&] {
        return with_engine(s, [&](auto& e) {
            using T = std::remove_reference_t<decltype(*e.tmp.p)>;
            Arg<T> a(aa, e.n2, true, false, e.st), b(bb, e.n2, true, false, e.st), c(cs, e.n2, false, true, e.st);
            e.divs(a.d, b.d, c.d);
            c.finish();
            return 0;
        });
    });
}

static int p3d(sw_spectral* s, const double* in, double* out, size_t izbeg, size_t izend, bool fwd) {
    return guard([&] {
        SW_CHECK(s && izbeg <= izend && izend <= s->nz, SW_ERR_ARG, "bad level range");
        return with_engine(s, [&](auto& e) {
            using T = std::remove_reference_t<decltype(*e.tmp.p)>;
            // the output's untouched levels must survive: read it in too
            Arg<T> a(in, e.N, true, false, e.st), b(out, e.N, true, true, e.st);
            if (fwd) e.ptospc3d(a.d, b.d, (int)izbeg, (int)izend);
            else e.spctop3d(a.d, b.d, (int)izbeg, (int)izend);
            b.finish();
            return 0;
        });
    });
}
int sw_ptospc3d(sw_spectral* s, const double* fp, double* fs, size_t izbeg, size_t izend) { return p3d(s, fp, fs, izbeg, izend, true); }
int sw_spctop3d(sw_spectral* s, const double* fs, double* fp, size_t izbeg, size_t izend) { return p3d(s, fs, fp, izbeg, izend, false); }

static int deal_n(sw_spectral* s, double* fp, bool three) {
    return guard([&] {
        return with_engine(s, [&](auto& e) {
            using T = std::remove_reference_t<decltype(*e.tmp.p)>;
            size_t cnt = three ? e.N : e.n2;
            Arg<T> a(fp, cnt, true, true, e.st);
            e.deal(a.d, three ? e.L : 1);
            a.finish();
            return 0;
        });
    });
}
int sw_deal3d(sw_spectral* s, double* fp) { return deal_n(s, fp, true); }
int sw_deal2d(sw_spectral* s, double* fp) { return deal_n(s, fp, false); }

int sw_state_create(sw_spectral* s, sw_state** out) {
    return guard([&] {
        SW_CHECK(s && out, SW_ERR_ARG, "null argument");
        auto st = std::make_unique<sw_state>();
        st->s = s;
        size_t N = s->ng * s->ng * (s->nz + 1);
        hipStream_t hs = stream();
        for (int i = 0; i < SW_NFIELDS; i++) {
            st->f64[i].alloc(N);
            SW_HIP(hipMemsetAsync(st->f64[i].p, 0, N * sizeof(double), hs));
        }
        SW_HIP(hipStreamSynchronize(hs));
        *out = st.release();
        return 0;
    });
}

int sw_state_destroy(sw_state* st) {
    delete st;
    return 0;
}

int sw_state_set(sw_state* st, int field, const double* src) {
    return guard([&] {
        SW_CHECK(field >= 0 && field < SW_NFIELDS, SW_ERR_ARG, "bad field id");
        return with_state(st, [&](auto& e, auto f) {
            using T = std::remove_reference_t<decltype(*e.tmp.p)>;
            Arg<T> a(src, e.N, true, false, e.st);
            SW_HIP(hipMemcpyAsync(f.a[field], a.d, e.N * sizeof(T), hipMemcpyDeviceToDevice, e.st));
            SW_HIP(hipStreamSynchronize(e.st));
            return 0;
        });
    });
}

int sw_state_get(sw_state* st, int field, double* dst) {
    return guard([&] {
        SW_CHECK(field >= 0 && field < SW_NFIELDS, SW_ERR_ARG, "bad field id");
        return with_state(st, [&](auto& e, auto f) {
            using T = std::remove_reference_t<decltype(*e.tmp.p)>;
            Arg<T> a(dst, e.N, false, true, e.st);
            SW_HIP(hipMemcpyAsync(a.d, f.a[field], e.N * sizeof(T), hipMemcpyDeviceToDevice, e.st));
            a.finish();
            return 0;
        });
    });
}

int sw_state_set_time(sw_state* st, double t, int ggen) {
    return guard([&] {
        SW_CHECK(st, SW_ERR_ARG, "null state");
        st->t = t;
        st->ggen = ggen;
        return 0;
    });
}

int sw_state_get_time(const sw_state* st, double* t) {
    return guard([&] {
        SW_CHECK(st && t, SW_ERR_ARG, "null argument");
        *t = st->t;
        return 0;
    });
}

int sw_vertical(sw_state* st) {
    return guard([&] {
        return with_state(st, [&](auto& e, auto f) {
            e.vertical(f);
            SW_HIP(hipStreamSynchronize(e.st));
            return 0;
        });
    });
}

int sw_coeffs(sw_state* st, double* sigx, double* sigy, double* cpt1, double* cpt2) {
    return guard([&] {
        return with_state(st, [&](auto& e, auto f) {
            using T = std::remove_reference_t<decltype(*e.tmp.p)>;
            Arg<T> a(sigx, e.N, false, true, e.st), b(sigy, e.N, false, true, e.st), c(cpt1, e.N, false, true, e.st),
                d(cpt2, e.N, false, true, e.st);
            e.coeffs(f, a.d, b.d, c.d, d.d);
            a.finish(); b.finish(); c.finish(); d.finish();
            return 0;
        });
    });
}

int sw_cpsource(sw_state* st, double* sp0) {
    return guard([&] {
        return with_state(st, [&](auto& e, auto f) {
            using T = std::remove_reference_t<decltype(*e.tmp.p)>;
            Arg<T> a(sp0, e.N, false, true, e.st);
            e.cpsource(f, a.d);
            a.finish();
            return 0;
        });
    });
}

int sw_psolve(sw_state* st, int nitmax, int* iters, double* errp) {
    return guard([&] {
        SW_CHECK(nitmax >= 1, SW_ERR_ARG, "nitmax must be >= 1");
        SW_CHECK(st && st->s && st->s->nz >= 3, SW_ERR_ARG, "psolve needs nz >= 3");
        return with_state(st, [&](auto& e, auto f) {
            int rc = e.psolve(f, nitmax, iters, errp);
            SW_HIP(hipStreamSynchronize(e.st));
            return rc;
        });
    });
}

int sw_source(sw_state* st, double* sqs, double* sds, double* sgs) {
    return guard([&] {
        return with_state(st, [&](auto& e, auto f) {
            using T = std::remove_reference_t<decltype(*e.tmp.p)>;
            Arg<T> a(sqs, e.N, false, true, e.st), b(sds, e.N, false, true, e.st), c(sgs, e.N, false, true, e.st);
            e.source(f, a.d, b.d, c.d);
            a.finish(); b.finish(); c.finish();
            return 0;
        });
    });
}

int sw_diagnose(sw_state* st, double diag[5]) {
    return guard([&] {
        SW_CHECK(diag, SW_ERR_ARG, "null diag");
        return with_state(st, [&](auto& e, auto f) {
            e.diagnose(f, st->t, diag);
            return 0;
        });
    });
}

int sw_advance(sw_state* st, double diag[5], int iters[3]) {
    return guard([&] {
        SW_CHECK(st && st->s && st->s->nz >= 3, SW_ERR_ARG, "advance needs nz >= 3");
        double dg[5];
        int it[3];
        int rc = with_state(st, [&](auto& e, auto f) {
            int r = e.advance(f, st->t, st->ggen != 0, dg, it);
            SW_HIP(hipStreamSynchronize(e.st));
            return r;
        });
        if (diag) std::memcpy(diag, dg, sizeof(dg));
        if (iters) std::memcpy(iters, it, sizeof(it));
        return rc;
    });
}

int sw_energy(sw_state* st, double out[6]) {
    return guard([&] {
        SW_CHECK(out, SW_ERR_ARG, "null output");
        return with_state(st, [&](auto& e, auto f) {
            e.energy(f, st->t, out);
            return 0;
        });
    });
}

int sw_run(sw_state* st, double tsim, double tgsave, double* monitor, int max_monitor, int* nmonitor, double* ecomp,
           int max_ecomp, int* necomp, int* iters_total) {
    return guard([&] {
        SW_CHECK(st && st->s && st->s->nz >= 3, SW_ERR_ARG, "run needs nz >= 3");
        SW_CHECK(tgsave > 0.0 && std::round(tgsave * (double)st->s->ng) >= 1.0, SW_ERR_ARG,
                 "save interval below one time step");
        SW_CHECK((monitor || max_monitor == 0) && (ecomp || max_ecomp == 0), SW_ERR_ARG, "null row buffer");
        int nm = 0, ne = 0, itot = 0;
        int rc = with_state(st, [&](auto& e, auto f) {
            int r = e.run(f, st->t, tsim, tgsave, 1.0 / (double)st->s->ng, monitor, max_monitor, &nm, ecomp, max_ecomp,
                          &ne, &itot);
            SW_HIP(hipStreamSynchronize(e.st));
            return r;
        });
        st->ggen = 0;
        if (nmonitor) *nmonitor = nm;
        if (necomp) *necomp = ne;
        if (iters_total) *iters_total = itot;
        return rc;
    });
}

// ---------------------------------------------------------------- on-disk formats (src/nhswps/mod.rs)
namespace {

struct File {
    FILE* f = nullptr;
    File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {
        SW_CHECK(f != nullptr, SW_ERR_ARG, "cannot open " + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
    void write(const void* p, size_t bytes) {
        SW_CHECK(std::fwrite(p, 1, bytes, f) == bytes, SW_ERR_ARG, "short write");
    }
};

// n little-endian f64 values after an 8-byte header: the *_init.r8 records of the nhswps loader
// (mod.rs:159-188) and the checkpoint's layout.  The host is little-endian (x86-64).
void read_r8(const std::string& path, double* out, size_t n, double* header = nullptr) {
    File f(path, "rb");
    SW_CHECK(std::fseek(f.f, 0, SEEK_END) == 0, SW_ERR_ARG, "cannot seek " + path);
    const long size = std::ftell(f.f);
    SW_CHECK(size == (long)(8 + 8 * n), SW_ERR_ARG,
             path + ": " + std::to_string(size) + " bytes, expected an 8-byte header and " + std::to_string(n) +
                 " f64 values (" + std::to_string(8 + 8 * n) + " bytes)");
    std::rewind(f.f);
    double h = 0.0;
    SW_CHECK(std::fread(&h, 8, 1, f.f) == 1 && std::fread(out, 8, n, f.f) == n, SW_ERR_ARG, "short read of " + path);
    if (header) *header = h;
}

// append_output (mod.rs:654-663): a record of t then the values, all little-endian f32
void append_r4(File& f, double t, const float* data, size_t n) {
    const float tf = (float)t;
    f.write(&tf, 4);
    f.write(data, 4 * n);
}

}  // namespace

int sw_read_r8(const char* path, double* out, size_t n) {
    return guard([&] {
        SW_CHECK(path && out, SW_ERR_ARG, "null argument");
        read_r8(path, out, n);
        return 0;
    });
}

int sw_savegrid(sw_state* st, float* d3, float* d2, double* spec3) {
    return guard([&] {
        SW_CHECK(d3 && d2 && spec3, SW_ERR_ARG, "null output");
        return with_state(st, [&](auto& e, auto f) {
            e.savegrid(f, d3, d2, spec3);
            return 0;
        });
    });
}

int sw_state_save(sw_state* st, const char* path) {
    return guard([&] {
        SW_CHECK(st && st->s && path, SW_ERR_ARG, "null argument");
        const size_t N = st->s->ng * st->s->ng * (st->s->nz + 1);
        std::vector<double> h(N);
        File f(path, "wb");
        f.write(&st->t, 8);
        return with_state(st, [&](auto& e, auto fl) {
            for (int k = 0; k < SW_NFIELDS; k++) {
                copy_d2h(h.data(), fl.a[k], N * sizeof(double), e.st);
                f.write(h.data(), N * sizeof(double));
            }
            return 0;
        });
    });
}

int sw_state_load(sw_state* st, const char* path) {
    return guard([&] {
        SW_CHECK(st && st->s && path, SW_ERR_ARG, "null argument");
        const size_t N = st->s->ng * st->s->ng * (st->s->nz + 1);
        std::vector<double> h((size_t)SW_NFIELDS * N);
        double t = 0.0;
        read_r8(path, h.data(), h.size(), &t);
        return with_state(st, [&](auto& e, auto fl) {
            for (int k = 0; k < SW_NFIELDS; k++) copy_h2d(fl.a[k], h.data() + (size_t)k * N, N * sizeof(double), e.st);
            st->t = t;
            return 0;
        });
    });
}

int sw_balinit(sw_spectral* s, const double* qq_init, double* qq, double* dd, double* gg, int* iters) {
    return guard([&] {
        return with_engine(s, [&](auto& e) {
            Arg<double> a(qq_init, e.n2, true, false, e.st), q(qq, e.n2, false, true, e.st), d(dd, e.n2, false, true, e.st),
                g(gg, e.n2, false, true, e.st);
            const int it = e.balinit(a.d, q.d, d.d, g.d);
            q.finish();
            d.finish();
            g.finish();
            if (iters) *iters = it;
            return 0;
        });
    });
}

int sw_balinit_file(const char* dir, size_t ng, size_t nz, int* iters) {
    return guard([&] {
        SW_CHECK(dir != nullptr, SW_ERR_ARG, "null directory");
        const std::string d(dir);
        const size_t n2 = ng * ng;
        std::vector<double> q0(n2), out(3 * (n2 + 1), 0.0);
        read_r8(d + "/qq_init.r8", q0.data(), n2);
        sw_spectral* sp = nullptr;
        int rc = sw_spectral_create(ng, nz, SW_FP64, &sp);
        if (rc) throw Error(rc, g_err);
        std::unique_ptr<sw_spectral, int (*)(sw_spectral*)> sp_guard(sp, sw_spectral_destroy);
        // sw_init.r8 (balinit.rs:419-433): three records, each a zero header and ng^2 values: q, delta, gamma
        rc = sw_balinit(sp, q0.data(), out.data() + 1, out.data() + n2 + 2, out.data() + 2 * n2 + 3, iters);
        if (rc) throw Error(rc, g_err);
        File f(d + "/sw_init.r8", "wb");
        f.write(out.data(), out.size() * sizeof(double));
        return 0;
    });
}

int sw_nhswps(const char* dir, size_t ng, size_t nz, double time_step, double duration, double save_interval, int precision,
              int* iters_total) {
    return guard([&] {
        SW_CHECK(dir && time_step > 0.0 && save_interval > 0.0 && std::round(save_interval / time_step) >= 1.0, SW_ERR_ARG,
                 "bad nhswps parameters");
        const std::string d(dir);
        const size_t n2 = ng * ng, N = n2 * (nz + 1);
        std::vector<double> init[3];
        const char* names[3] = {"qq_init.r8", "dd_init.r8", "gg_init.r8"};
        for (int k = 0; k < 3; k++) {
            init[k].resize(N);
            read_r8(d + "/" + names[k], init[k].data(), N);
        }
        sw_spectral* sp = nullptr;
        int rc = sw_spectral_create(ng, nz, precision, &sp);
        if (rc) throw Error(rc, g_err);
        std::unique_ptr<sw_spectral, int (*)(sw_spectral*)> sp_guard(sp, sw_spectral_destroy);
        sw_state* stp = nullptr;
        rc = sw_state_create(sp, &stp);
        if (rc) throw Error(rc, g_err);
        std::unique_ptr<sw_state, int (*)(sw_state*)> st_guard(stp, sw_state_destroy);
        auto& e = *sp->e64;
        auto f = stp->fields();
        // the loader (mod.rs:238-275): to spectral space, the (0, 0) mode zeroed, times the filter
        const int fid[3] = {SW_QS, SW_DS, SW_GS};
        for (int k = 0; k < 3; k++) {
            Arg<double> a(init[k].data(), N, true, false, e.st);
            e.ptospc3d(a.d, f.a[fid[k]], 0, (int)nz);
            double* x = f.a[fid[k]];
            const double* F = e.filt.p;
            const size_t nn = n2;
            each(e.st, N, [=] __device__(size_t q) {
                const size_t p = q % nn;
                x[q] = p == 0 ? 0.0 : x[q] * F[p];
            });
            SW_HIP(hipStreamSynchronize(e.st));
        }
        mkdir((d + "/2d").c_str(), 0777);
        mkdir((d + "/3d").c_str(), 0777);
        struct Out : Engine<double>::RunHooks {
            Engine<double>* e = nullptr;
            Engine<double>::Fields f{};
            double* t = nullptr;
            size_t n2 = 0, N = 0;
            File mon, eco, spc, ql, d3, g3, r3, w3, p3, q2, d2, g2, h2, z2;
            explicit Out(const std::string& d)
                : mon(d + "/monitor.asc", "w"), eco(d + "/ecomp.asc", "w"), spc(d + "/spectra.asc", "w"),
                  ql(d + "/3d/ql.r4", "wb"), d3(d + "/3d/d.r4", "wb"), g3(d + "/3d/g.r4", "wb"), r3(d + "/3d/r.r4", "wb"),
                  w3(d + "/3d/w.r4", "wb"), p3(d + "/3d/pn.r4", "wb"), q2(d + "/2d/q.r4", "wb"), d2(d + "/2d/d.r4", "wb"),
                  g2(d + "/2d/g.r4", "wb"), h2(d + "/2d/h.r4", "wb"), z2(d + "/2d/zeta.r4", "wb") {}
            // diagnose's line (mod.rs:700-708)
            void monitor(const double r[5]) override {
                std::fprintf(mon.f, "%.5f %.6f %.6f %.6f %.6f\n", r[0], r[1], r[2], r[3], r[4]);
            }
            // savegrid (mod.rs:416-420, 474-489, 491-652)
            void save(const double en[6]) override {
                std::fprintf(eco.f, "%.6f %.9f %.9f %.9f %.9f %.9f\n", en[0], en[1], en[2], en[3], en[4], en[5]);
                std::vector<float> f3(6 * N), f2(5 * n2);
                std::vector<double> sp(3 * (e->ng + 1));
                e->savegrid(f, f3.data(), f2.data(), sp.data());
                const size_t nk = e->ng + 1, kr = e->host.kmaxred;
                std::fprintf(spc.f, "%.6f %zu\n", *t, kr);
                for (size_t k = 1; k <= kr; k++)
                    std::fprintf(spc.f, "%.8f %.8f %.8f %.8f\n", e->host.alk[k - 1], std::log10(sp[k]),
                                 std::log10(sp[nk + k] + 1.0E-32), std::log10(sp[2 * nk + k] + 1.0E-32));
                File* o3[6] = {&ql, &d3, &g3, &r3, &w3, &p3};
                for (int q = 0; q < 6; q++) append_r4(*o3[q], *t, f3.data() + (size_t)q * N, N);
                File* o2[5] = {&d2, &z2, &q2, &g2, &h2};
                for (int q = 0; q < 5; q++) append_r4(*o2[q], *t, f2.data() + (size_t)q * n2, n2);
            }
        } out(d);
        out.e = &e;
        out.f = f;
        out.t = &stp->t;
        out.n2 = n2;
        out.N = N;
        int nm = 0, ne = 0, itot = 0;
        rc = e.run(f, stp->t, duration, save_interval, time_step, nullptr, 0, &nm, nullptr, 0, &ne, &itot, &out);
        SW_HIP(hipStreamSynchronize(e.st));
        if (iters_total) *iters_total = itot;
        return rc;
    });
}

namespace {

bool batchable_nz(size_t nz) { return nz == 4 || nz == 8 || nz == 16 || nz == 24 || nz == 32; }

// Units of work (batches of same-shape members, single members) to streams: largest first, each to the
// stream with the least work so far (work ~ points, FP64 counted twice).  Batch engines are created on
// their stream.
void partition(sw_ensemble* e) {
    for (auto s : e->streams) SW_HIP(hipStreamSynchronize(s));
    e->batches.clear();
    const int n = (int)e->m.size();
    std::map<std::tuple<size_t, size_t, int>, std::vector<int>> by;
    std::vector<int> singles;
    for (int i = 0; i < n; i++) {
        const sw_spectral* sp = e->m[i]->s;
        if (e->batch_max >= 2 && batchable_nz(sp->nz)) by[{sp->ng, sp->nz, sp->precision}].push_back(i);
        else singles.push_back(i);
    }
    std::vector<std::vector<int>> bl;  // members of each batch
    for (auto& kv : by) {
        const auto& v = kv.second;
        if (v.size() < 2) {
            singles.push_back(v[0]);
            continue;
        }
        const size_t nb = (v.size() + e->batch_max - 1) / e->batch_max, base = v.size() / nb, extra = v.size() % nb;
        size_t at = 0;
        for (size_t b = 0; b < nb; b++) {
            const size_t cnt = base + (b < extra ? 1 : 0);
            bl.emplace_back(v.begin() + at, v.begin() + at + cnt);
            at += cnt;
        }
    }
    // streams: as requested, else one per unit of work (a batch's launches and its host synchronisations
    // then overlap those of the others), at most 8
    const int units_n = (int)(bl.size() + singles.size());
    const int ns = std::max(1, std::min(e->nstreams_req > 0 ? e->nstreams_req : std::min(units_n, 8), n));
    while ((int)e->streams.size() > ns) {
        (void)hipStreamDestroy(e->streams.back());
        e->streams.pop_back();
    }
    while ((int)e->streams.size() < ns) {
        hipStream_t h;
        SW_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
        e->streams.push_back(h);
    }
    auto work1 = [&](int i) {
        const sw_spectral* sp = e->m[i]->s;
        return (double)(sp->ng * sp->ng * (sp->nz + 1)) * (sp->precision == SW_FP64 ? 2.0 : 1.0);
    };
    struct Unit {
        double work;
        int batch, member;
    };
    std::vector<Unit> units;
    for (int b = 0; b < (int)bl.size(); b++) {
        double w = 0.0;
        for (int i : bl[b]) w += work1(i);
        units.push_back({w, b, -1});
    }
    for (int i : singles) units.push_back({work1(i), -1, i});
    std::stable_sort(units.begin(), units.end(), [](const Unit& a, const Unit& b) { return a.work > b.work; });
    e->groups.assign(ns, {});
    e->bgroups.assign(ns, {});
    std::vector<double> load(ns, 0.0);
    std::vector<int> bstream(bl.size(), 0);
    for (const Unit& u : units) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        if (u.batch >= 0) e->bgroups[g].push_back(u.batch), bstream[u.batch] = g;
        else e->groups[g].push_back(u.member);
        load[g] += u.work;
    }
    for (int b = 0; b < (int)bl.size(); b++) {
        auto B = std::make_unique<MemberBatch>();
        B->idx = bl[b];
        const int M = (int)B->idx.size();
        const sw_spectral* sp = e->m[B->idx[0]]->s;
        hipStream_t S = e->streams[bstream[b]];
        B->precision = sp->precision;
        B->e64 = std::make_unique<Engine<double>>((int)sp->ng, (int)sp->nz, S, M);
        if (sp->precision == SW_MIXED) {
            B->e32 = std::make_unique<Engine<float>>((int)sp->ng, (int)sp->nz, S, M);
            B->e64->sub = B->e32.get();
        }
        for (int k = 0; k < SW_NFIELDS; k++) B->f[k].alloc(B->e64->NM);
        std::vector<double*> h((size_t)M * SW_NFIELDS);
        for (int m = 0; m < M; m++)
            for (int k = 0; k < SW_NFIELDS; k++) h[(size_t)m * SW_NFIELDS + k] = e->m[B->idx[m]]->f64[k].p;
        B->tab.alloc(h.size());
        copy_h2d(B->tab.p, h.data(), h.size() * sizeof(double*), S);
        e->batches.push_back(std::move(B));
    }
}

// the members' 16 fields into the batch fields (in) or back (out), one launch
void batch_copy(hipStream_t st, MemberBatch& B, bool in) {
    PtrTab<double, SW_NFIELDS> bf;
    for (int k = 0; k < SW_NFIELDS; k++) bf.p[k] = B.f[k].p;
    double* const* tab = B.tab.p;
    const size_t N = B.e64->N, MN = B.e64->NM;
    each(st, (size_t)SW_NFIELDS * MN, [=] __device__(size_t i) {
        const int k = (int)(i / MN);
        const size_t r = i - (size_t)k * MN, m = r / N, q = r - m * N;
        double* a = tab[m * SW_NFIELDS + k];
        if (in) bf[k][r] = a[q];
        else a[q] = bf[k][r];
    });
}

// psolve's tolerance of a member (the engine that runs its iteration)
double member_toler(const sw_state* st) { return st->s->e32 ? st->s->e32->toler : st->s->e64->toler; }

}  // namespace

int sw_ensemble_create(int nmember, sw_state* const* members, int nstreams, sw_ensemble** out) {
    return guard([&] {
        SW_CHECK(out && members && nmember > 0, SW_ERR_ARG, "bad ensemble arguments");
        auto e = std::make_unique<sw_ensemble>();
        for (int i = 0; i < nmember; i++) {
            sw_state* st = members[i];
            SW_CHECK(st && st->s && st->s->nz >= 3, SW_ERR_ARG, "ensemble members need a state with nz >= 3");
            for (int j = 0; j < i; j++)
                SW_CHECK(members[j]->s != st->s, SW_ERR_ARG, "ensemble members must not share a spectral handle");
            e->m.push_back(st);
        }
        e->nstreams_req = std::max(0, nstreams);
        stream();  // the library device is selected
        for (int i = 0; i < nmember; i++) e->graphs.push_back(std::make_unique<MemberGraph>());
        partition(e.get());
        *out = e.release();
        return 0;
    });
}

int sw_ensemble_set_batching(sw_ensemble* e, int max_members) {
    return guard([&] {
        SW_CHECK(e && max_members >= 0, SW_ERR_ARG, "bad argument");
        e->batch_max = max_members;
        for (auto& g : e->graphs) g->reset();
        partition(e);
        return 0;
    });
}

int sw_ensemble_destroy(sw_ensemble* e) {
    return guard([&] {
        if (e) {
            for (auto s : e->streams) (void)hipStreamSynchronize(s);
            delete e;
        }
        return 0;
    });
}

namespace {

// the 16 fields of a member copied between the state and its snapshot (one launch)
void copy_fields(hipStream_t st, const Engine<double>::Fields& f, double* snap, size_t N, bool to_snap) {
    PtrTab<double, SW_NFIELDS> a;
    for (int k = 0; k < SW_NFIELDS; k++) a.p[k] = f.a[k];
    each(st, (size_t)SW_NFIELDS * N, [=] __device__(size_t i) {
        const int k = (int)(i / N);
        const size_t q = i - (size_t)k * N;
        if (to_snap) snap[i] = a[k][q];
        else a[k][q] = snap[i];
    });
}

// Record member st's next step (advance with ggen set) as a graph on its engine's stream: the snapshot
// copy, then the step with each psolve iteration capped at the last count plus margin.
void capture_step(sw_state* st, Engine<double>& eng, MemberGraph& G, int margin) {
    G.reset();
    auto f = st->fields();
    if (G.snap.n < (size_t)SW_NFIELDS * eng.N) G.snap.alloc((size_t)SW_NFIELDS * eng.N);
    for (int k = 0; k < 3; k++) G.cap[k] = std::min(100, *eng.hint_slot(k) + margin);
    // relaxed: other host threads keep launching, allocating and synchronising their own streams
    SW_HIP(hipStreamBeginCapture(eng.st, hipStreamCaptureModeRelaxed));
    double tcap = st->t, dg[5];
    int it[3];
    int rc = 0;
    try {
        copy_fields(eng.st, f, G.snap.p, eng.N, true);
        eng.capturing = true;
        for (int k = 0; k < 3; k++) eng.cap_iters[k] = G.cap[k];
        rc = eng.advance(f, tcap, true, dg, it);
    } catch (...) {
        eng.capturing = false;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(eng.st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    eng.capturing = false;
    SW_HIP(hipStreamEndCapture(eng.st, &G.graph));
    SW_CHECK(rc == 0, rc, "capture of a step failed");
    SW_HIP(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
    G.valid = true;
}

}  // namespace

int sw_ensemble_set_graphs(sw_ensemble* e, int enable, int margin) {
    return guard([&] {
        SW_CHECK(e && margin >= 0, SW_ERR_ARG, "bad argument");
        e->use_graphs = enable ? 1 : 0;
        e->margin = margin;
        for (auto& g : e->graphs) g->reset();
        return 0;
    });
}

int sw_ensemble_stats(sw_ensemble* e, long* replays, long* redos, long* captures) {
    return guard([&] {
        SW_CHECK(e && replays && redos && captures, SW_ERR_ARG, "null argument");
        *replays = e->replays;
        *redos = e->redos;
        *captures = e->captures;
        return 0;
    });
}

int sw_ensemble_batches(sw_ensemble* e, int* nbatch, int* members, long* batched_steps, int* nstreams) {
    return guard([&] {
        SW_CHECK(e && nbatch && members && batched_steps && nstreams, SW_ERR_ARG, "null argument");
        *nstreams = (int)e->streams.size();
        *nbatch = (int)e->batches.size();
        int c = 0;
        for (auto& b : e->batches) c += (int)b->idx.size();
        *members = c;
        *batched_steps = e->batched_steps;
        return 0;
    });
}

int sw_ensemble_advance(sw_ensemble* e, double* diag, int* iters, int* status) {
    return guard([&] {
        SW_CHECK(e, SW_ERR_ARG, "null ensemble");
        const int n = (int)e->m.size();
        std::vector<int> rc(n, 0);
        std::vector<double> dg((size_t)5 * n, 0.0);
        std::vector<int> it((size_t)3 * n, 0);
        std::vector<std::string> msg(n);
        std::atomic<long> replays{0}, redos{0}, batched{0};
        // members whose step is recorded after the worker threads have joined: a capture runs while no
        // other host thread makes HIP calls (concurrent calls have invalidated captures on this runtime)
        std::vector<int> to_capture;
        std::mutex cap_mu;
        auto want_capture = [&](int i) {
            std::lock_guard<std::mutex> lk(cap_mu);
            to_capture.push_back(i);
        };
        auto run_group = [&](int g) {
            (void)hipSetDevice(g_device);
            hipStream_t S = e->streams[g];
            auto fail = [&](int i, const std::exception& x) {
                const Error* er = dynamic_cast<const Error*>(&x);
                rc[i] = er ? er->code : SW_ERR_ARG;
                msg[i] = x.what();
            };
            std::vector<int> replayed, eager;
            // batches: one batched step when the members agree on ggen and psolve's tolerance, else their
            // members one by one
            for (int b : e->bgroups[g]) {
                MemberBatch& B = *e->batches[b];
                const int M = (int)B.idx.size();
                const sw_state* s0 = e->m[B.idx[0]];
                bool same = true;
                for (int i : B.idx) same = same && e->m[i]->ggen == s0->ggen && member_toler(e->m[i]) == member_toler(s0);
                if (!same) {
                    for (int i : B.idx) eager.push_back(i);
                    continue;
                }
                try {
                    auto& eng = *B.e64;
                    eng.set_stream(S);
                    eng.toler = s0->s->e64->toler;
                    if (B.e32) B.e32->toler = member_toler(s0);
                    std::vector<double> tb(M), db((size_t)5 * M);
                    std::vector<int> ib((size_t)3 * M), sb(M);
                    for (int m = 0; m < M; m++) tb[m] = e->m[B.idx[m]]->t;
                    batch_copy(S, B, true);
                    eng.advance_batch(B.fields(), tb.data(), s0->ggen != 0, db.data(), ib.data(), sb.data());
                    batch_copy(S, B, false);
                    SW_HIP(hipStreamSynchronize(S));
                    for (int m = 0; m < M; m++) {
                        const int i = B.idx[m];
                        std::memcpy(&dg[(size_t)5 * i], &db[(size_t)5 * m], 5 * sizeof(double));
                        std::memcpy(&it[(size_t)3 * i], &ib[(size_t)3 * m], 3 * sizeof(int));
                        rc[i] = sb[m];
                        if (sb[m]) msg[i] = "psolve failed in a batched step";
                        else e->m[i]->t = tb[m];
                    }
                    batched++;
                } catch (const std::exception& x) {
                    (void)hipStreamSynchronize(S);
                    for (int i : B.idx) fail(i, x);
                }
            }
            // replays first: every member whose captured step applies, launched back to back
            for (int i : e->groups[g]) {
                sw_state* st = e->m[i];
                MemberGraph& G = *e->graphs[i];
                if (e->use_graphs && G.valid && st->ggen) {
                    try {
                        SW_HIP(hipGraphLaunch(G.exec, S));
                        replayed.push_back(i);
                    } catch (const std::exception& x) {
                        fail(i, x);
                    }
                } else {
                    eager.push_back(i);
                }
            }
            (void)hipStreamSynchronize(S);
            for (int i : replayed) {
                sw_state* st = e->m[i];
                MemberGraph& G = *e->graphs[i];
                auto& eng = *st->s->e64;
                hipStream_t old = eng.st;
                eng.set_stream(S);
                try {
                    PsolveCtrl c[3];
                    bool ok = true;
                    for (int k = 0; k < 3; k++) {
                        copy_d2h(&c[k], eng.ctrl_slot(k), sizeof(PsolveCtrl), S);
                        ok = ok && c[k].done && c[k].status == 0 && c[k].iters < 100;
                    }
                    if (ok) {
                        eng.diagnose_finish(st->t, &dg[(size_t)5 * i]);
                        for (int k = 0; k < 3; k++) {
                            it[(size_t)3 * i + k] = c[k].iters;
                            *eng.hint_slot(k) = c[k].iters;
                        }
                        st->t += 1.0 / (double)st->s->ng;
                        replays++;
                        // iteration counts creep up: capture again before they reach the cap
                        bool near = false;
                        for (int k = 0; k < 3; k++) near = near || c[k].iters >= G.cap[k];
                        if (near) want_capture(i);
                    } else {
                        // a psolve needed more iterations than captured, or failed: undo the step and redo
                        // it eagerly (which also reports any error as sw_advance would)
                        copy_fields(S, st->fields(), G.snap.p, eng.N, false);
                        G.reset();
                        redos++;
                        eager.push_back(i);
                    }
                } catch (const std::exception& x) {
                    fail(i, x);
                }
                eng.set_stream(old);
            }
            for (int i : eager) {
                sw_state* st = e->m[i];
                try {
                    rc[i] = with_state(st, [&](auto& eng, auto f) {
                        hipStream_t old = eng.st;
                        eng.set_stream(S);
                        int r = 0;
                        try {
                            r = eng.advance(f, st->t, st->ggen != 0, &dg[(size_t)5 * i], &it[(size_t)3 * i]);
                            SW_HIP(hipStreamSynchronize(eng.st));
                            if (r == 0 && e->use_graphs && eng.nz_spec()) want_capture(i);
                        } catch (...) {
                            (void)hipStreamSynchronize(eng.st);
                            eng.set_stream(old);
                            throw;
                        }
                        eng.set_stream(old);
                        return r;
                    });
                } catch (const std::exception& x) {
                    fail(i, x);
                }
            }
        };
        std::vector<std::thread> th;
        for (int g = 1; g < (int)e->groups.size(); g++) th.emplace_back(run_group, g);
        run_group(0);
        for (auto& t : th) t.join();
        for (int i : to_capture) {
            if (rc[i]) continue;
            sw_state* st = e->m[i];
            auto& eng = *st->s->e64;
            int g = 0;
            for (int k = 0; k < (int)e->groups.size(); k++)
                if (std::find(e->groups[k].begin(), e->groups[k].end(), i) != e->groups[k].end()) g = k;
            hipStream_t old = eng.st;
            eng.set_stream(e->streams[g]);
            try {
                capture_step(st, eng, *e->graphs[i], e->margin);
                e->captures++;
            } catch (const std::exception& x) {
                const Error* er = dynamic_cast<const Error*>(&x);
                rc[i] = er ? er->code : SW_ERR_ARG;
                msg[i] = x.what();
            }
            eng.set_stream(old);
        }
        e->replays += replays;
        e->redos += redos;
        e->batched_steps += batched;
        if (diag) std::memcpy(diag, dg.data(), dg.size() * sizeof(double));
        if (iters) std::memcpy(iters, it.data(), it.size() * sizeof(int));
        if (status) std::memcpy(status, rc.data(), rc.size() * sizeof(int));
        for (int i = 0; i < n; i++)
            if (rc[i]) {
                g_err = "ensemble member " + std::to_string(i) + ": " + msg[i];
                return rc[i];
            }
        return 0;
    });
}

int sw_profile(sw_spectral* s, int mode) {
    return guard([&] {
        SW_CHECK(mode >= SW_PROF_OFF && mode <= SW_PROF_STAGES, SW_ERR_ARG, "unknown profiling mode");
        return with_engine(s, [&](auto& e) {
            auto reset = [&](auto& x) {
                x.prof_collect();
                x.stage_collect();
                x.prof = mode;
                x.prof_ms = x.prof_bytes = 0.0;
                x.prof_launches = 0;
                for (double& m : x.stage_ms) m = 0.0;
                x.stage_steps = x.stage_iters = 0;
            };
            reset(e);
            if (e.sub) reset(*e.sub);
            return 0;
        });
    });
}

int sw_profile_read(sw_spectral* s, double* ms, double* bytes, long* launches) {
    return guard([&] {
        SW_CHECK(ms && bytes && launches, SW_ERR_ARG, "null argument");
        return with_engine(s, [&](auto& e) {
            e.prof_collect();
            *ms = e.prof_ms;
            *bytes = e.prof_bytes;
            *launches = e.prof_launches;
            if (e.sub) {
                e.sub->prof_collect();
                *ms += e.sub->prof_ms;
                *bytes += e.sub->prof_bytes;
                *launches += e.sub->prof_launches;
            }
            return 0;
        });
    });
}

int sw_profile_stages(sw_spectral* s, double ms[7], long* steps, long* iterations) {
    return guard([&] {
        SW_CHECK(ms && steps && iterations, SW_ERR_ARG, "null argument");
        return with_engine(s, [&](auto& e) {
            e.stage_collect();
            static_assert(std::remove_reference_t<decltype(e)>::NSTAGE == 7, "seven stages");
            for (int k = 0; k < 7; k++) ms[k] = e.stage_ms[k];
            *steps = e.stage_steps;
            *iterations = e.stage_iters;
            return 0;
        });
    });
}

int sw_set_tolerance(sw_spectral* s, double toler) {
    return guard([&] {
        SW_CHECK(toler > 0.0, SW_ERR_ARG, "tolerance must be positive");
        return with_engine(s, [&](auto& e) {
            if (e.sub) e.sub->toler = toler;  // the mixed mode's iteration runs in the FP32 engine
            else e.toler = toler;
            return 0;
        });
    });
}

// ---------------------------------------------------------------- synthetic initial state
// splitmix64 seeding xoshiro256**, Box-Muller normals.
namespace {
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; i++) {
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
    }
};

// one spectral slab with an isotropic k^3/(k0^6+k^6)/k power law, de-aliased, zero mean, scaled so
// that the physical rms of op(k)*field equals rms
void synth_slab(const HostOps& h, Rng& g, double k0, double rms, const std::vector<double>* op, std::vector<double>& out) {
    size_t ng = h.ng, n2 = ng * ng;
    out.assign(n2, 0.0);
    double ss = 0.0;
    for (size_t ky = 0; ky < ng; ky++)
        for (size_t kx = 0; kx < ng; kx++) {
            size_t p = kx + ng * ky;
            double k = std::sqrt(h.rk[kx] * h.rk[kx] + h.rk[ky] * h.rk[ky]);
            double x = g.normal();
            if (k == 0.0 || h.filt[p] == 0.0) continue;
            double spec = k * k * k / (std::pow(k0, 6.0) + std::pow(k, 6.0));
            out[p] = std::sqrt(spec / k) * x;
            double wx = (kx == 0 || 2 * kx == ng) ? 1.0 : 2.0, wy = (ky == 0 || 2 * ky == ng) ? 1.0 : 2.0;
            double y = op ? (*op)[p] * out[p] : out[p];
            ss += wx * wy * y * y;
        }
    double cur = std::sqrt(ss / (double)n2);
    double a = cur > 0.0 ? rms / cur : 0.0;
    for (auto& x : out) x *= a;
}
}  // namespace

int sw_synth_kminus3(size_t ng, size_t nz, unsigned long long seed, double k0, double zeta_rms, double delta_rms, double* qq,
                     double* dd) {
    return guard([&] {
        SW_CHECK(qq && dd && nz >= 1 && ng >= 4, SW_ERR_ARG, "bad argument");
        HostOps h;
        SW_CHECK(h.build(ng, nz < 2 ? 2 : nz), SW_ERR_FACTOR, "ng not factorisable");
        Rng g(seed);
        std::vector<double> q, d;
        // zeta_hat = rope * q_hat for identical layers with delta = gamma = 0
        synth_slab(h, g, k0, zeta_rms, &h.rope, q);
        if (delta_rms > 0.0) synth_slab(h, g, k0, delta_rms, nullptr, d);
        else d.assign(ng * ng, 0.0);
        size_t n2 = ng * ng;
        for (size_t iz = 0; iz <= nz; iz++) {
            std::memcpy(qq + iz * n2, q.data(), n2 * sizeof(double));
            std::memcpy(dd + iz * n2, d.data(), n2 * sizeof(double));
        }
        return 0;
    });
}

}  // extern "C"
