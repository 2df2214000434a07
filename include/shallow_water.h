/* shallow_water.h -- C ABI of the MI355X (gfx950) implementation of the nhswps per-timestep hot
 * path.  Each entry point replaces one public function of the reference crate
 * (rse-standrewscs/shallow-water); the citation next to it names the Rust function it stands in
 * for.  All arrays use the reference's layout: element (i,j,iz) of an (ng,ng,nz+1) field lives at
 * i + ng*j + ng*ng*iz (src/utils.rs:26-31); physical fields have i = y, j = x, spectral fields
 * i = kx, j = ky with the stafft Hermitian packing.
 *
 * Pointers: every double* array argument may be a host pointer or a device pointer (HIP memory
 * on the library's device).  Host arrays are staged through device memory; device arrays are used
 * in place.  Unlike the reference, inputs documented as "destroyed" there are left intact here.
 *
 * Errors: functions return 0 or a negative sw_status instead of the reference's exit(1)/panic.
 * Thread safety: per handle; calls on one handle must not overlap.
 */
#ifndef SHALLOW_WATER_H
#define SHALLOW_WATER_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SW_OK = 0,
    SW_ERR_ARG = -1,              /* shape mismatch / bad argument (reference: assert_eq! panic) */
    SW_ERR_PSOLVE_MAXITER = -2,   /* psolve iteration cap reached (reference: psolve.rs:360-366 exit(1)) */
    SW_ERR_PSOLVE_DIVERGED = -3,  /* psolve error > 1 after the first iteration (psolve.rs:349-352 exit(1)) */
    SW_ERR_FACTOR = -4,           /* length not a product of 2,3,4,5,6 (stafft/mod.rs:148-149 exit(1)) */
    SW_ERR_UNSUPPORTED = -5,      /* length the reference itself cannot transform correctly */
    SW_ERR_HIP = -6,              /* HIP runtime error */
    SW_ERR_NONFINITE = -7         /* NaN/Inf reached the psolve residual */
} sw_status;

/* SW_FP64: every stage in FP64.  SW_MIXED: psolve's coefficient fields (coeffs) and its fixed-point
 * iteration in FP32, the pressure tridiagonal and residual in FP64, with the iteration's tolerance 1e-6
 * (FP32 transforms put a floor near 1e-7 under the residual); the state and every other stage in FP64. */
typedef enum { SW_FP64 = 0, SW_MIXED = 1 } sw_precision;

/* Field ids of a state, in the order of the reference State (src/nhswps/mod.rs:87-133). */
typedef enum {
    SW_U = 0, SW_V, SW_W, SW_Z, SW_ZX, SW_ZY, SW_R, SW_RI, SW_AA, SW_ZETA,
    SW_PN, SW_DPN, SW_PS, SW_QS, SW_DS, SW_GS, SW_NFIELDS
} sw_field;

typedef struct sw_d2fft sw_d2fft;
typedef struct sw_spectral sw_spectral;
typedef struct sw_state sw_state;

/* ---- library ---- */
int sw_init(int device);            /* select the HIP device used by later calls (default 0) */
int sw_sync(void);                  /* wait for all work queued by the library */
const char* sw_last_error(void);    /* text of the last error on this thread */

/* ---- 1-D FFT: src/stafft/mod.rs ---- */
int sw_factorisen(size_t n, size_t factors[5]);                       /* factorisen  mod.rs:97  */
int sw_initfft(size_t n, size_t factors[5], double* trig);            /* initfft     mod.rs:67  (host) */
/* forfft/revfft: trig must be initfft(n)'s table (compared value for value to 1e-12, host or device
   memory; SW_ERR_ARG otherwise); factors may be null */
int sw_forfft(size_t m, size_t n, double* x, const double* trig,      /* forfft      mod.rs:158 */
              const size_t factors[5]);
int sw_revfft(size_t m, size_t n, double* x, const double* trig,      /* revfft      mod.rs:301 */
              const size_t factors[5]);

/* ---- 2-D FFT and spectral derivatives: src/sta2dfft.rs ---- */
int sw_d2fft_create(size_t nx, size_t ny, double lx, double ly, double* kx, double* ky,
                    sw_d2fft** out);                                  /* D2FFT::new   :37 */
int sw_d2fft_destroy(sw_d2fft* d);
int sw_ptospc(const sw_d2fft* d, double* rvar, double* svar);         /* ptospc  :91  */
int sw_spctop(const sw_d2fft* d, double* svar, double* rvar);         /* spctop  :113 */
/* ptospc without any allocation: no counterpart in the reference.  For callers that cannot allocate
 * device or host memory during the call (code between graph captures, tight serving loops, harnesses
 * that forbid hipMalloc): rvar and svar must be device arrays; the buffer between the two passes is a
 * fixed scratch array that the library holds on each device (one slab of 512 x 512 doubles, 2 MiB), so
 * nx <= 512 (SW_ERR_ARG otherwise).  Every call on a device shares that scratch: calls that overlap on
 * one device race.  The result equals sw_ptospc's bit for bit. */
int sw_ptospc_noalloc(const sw_d2fft* d, const double* rvar, double* svar);
int sw_xderiv(const sw_d2fft* d, const double* rkx, const double* var, double* der); /* :128 */
int sw_yderiv(const sw_d2fft* d, const double* rky, const double* var, double* der); /* :157 */

/* ---- spectral operators and inversion: src/spectral/mod.rs ---- */
int sw_spectral_create(size_t ng, size_t nz, int precision, sw_spectral** out); /* Spectral::new :65 */
int sw_spectral_destroy(sw_spectral* s);
long sw_spectral_get(const sw_spectral* s, const char* name, double* out); /* operator export (host) */
int sw_main_invert(sw_spectral* s, const double* qs, const double* ds, const double* gs,
                   double* r, double* u, double* v, double* zeta);    /* main_invert :305 */
int sw_jacob(sw_spectral* s, const double* aa, const double* bb, double* cs);   /* jacob :423 */
int sw_divs(sw_spectral* s, const double* aa, const double* bb, double* cs);    /* divs  :461 */
int sw_ptospc3d(sw_spectral* s, const double* fp, double* fs, size_t izbeg, size_t izend); /* :479 */
int sw_spctop3d(sw_spectral* s, const double* fs, double* fp, size_t izbeg, size_t izend); /* :502 */
int sw_deal3d(sw_spectral* s, double* fp);                            /* deal3d :524 */
int sw_deal2d(sw_spectral* s, double* fp);                            /* deal2d :544 */

/* ---- state and the nhswps kernels: src/nhswps ---- */
int sw_state_create(sw_spectral* s, sw_state** out);                  /* State (mod.rs:87) */
int sw_state_destroy(sw_state* st);
int sw_state_set(sw_state* st, int field, const double* src);
int sw_state_get(sw_state* st, int field, double* dst);
int sw_state_set_time(sw_state* st, double t, int ggen);
int sw_state_get_time(const sw_state* st, double* t);
int sw_vertical(sw_state* st);                                        /* vertical.rs:8  */
int sw_coeffs(sw_state* st, double* sigx, double* sigy, double* cpt1, double* cpt2); /* coeffs.rs:8 */
int sw_cpsource(sw_state* st, double* sp0);                           /* cpsource.rs:13 */
int sw_psolve(sw_state* st, int nitmax, int* iters, double* errp);    /* psolve.rs:16   */
int sw_source(sw_state* st, double* sqs, double* sds, double* sgs);   /* source.rs:19   */
int sw_diagnose(sw_state* st, double diag[5]);                        /* diagnose mod.rs:666 */
int sw_advance(sw_state* st, double diag[5], int iters[3]);           /* advance.rs:13  */

/* ---- time loop and diagnostics: src/nhswps/mod.rs ---- */
/* The energies of savegrid (mod.rs:365-420): out = {t, 0, ekin, ekin, epot, etot}, the ecomp.asc row. */
int sw_energy(sw_state* st, double out[6]);
/* The time loop of nhswps (mod.rs:285-345) from the state's prognostic fields (qs, ds, gs) and time,
 * up to t > tsim, with saves every tgsave: monitor rows (t, zrms^2/2, zrms, zmax, umax; one per step,
 * the monitor.asc line) into monitor[5 i ..] and ecomp rows (one per save) into ecomp[6 i ..], at most
 * max_monitor / max_ecomp of them; *nmonitor / *necomp = the rows produced, *iters_total = the psolve
 * iterations.  File output (savegrid's .r4 fields and spectra) is the caller's. */
int sw_run(sw_state* st, double tsim, double tgsave, double* monitor, int max_monitor, int* nmonitor, double* ecomp,
           int max_ecomp, int* necomp, int* iters_total);

/* ---- on-disk formats and the whole program: src/nhswps/mod.rs ---- */
/* n little-endian f64 values after an 8-byte header (the *_init.r8 files the loader reads, mod.rs:159-188);
 * SW_ERR_ARG when the file is not exactly that size. */
int sw_read_r8(const char* path, double* out, size_t n);
/* savegrid's output (mod.rs:347-652) as numbers, into host arrays: d3 = the 3-D fields ql, d, g, r, w, pn
 * as f32 (6 x ng*ng*(nz+1)), d2 = the vertically integrated d, zeta, q, g, h as f32 (5 x ng*ng), spec3 =
 * the zeta, delta and gamma spectra (3 x (ng+1); spectra.asc prints log10 of entries 1..ng/2).  Its
 * energies are sw_energy. */
int sw_savegrid(sw_state* st, float* d3, float* d2, double* spec3);
/* nhswps(&Parameters) (mod.rs:150-345): reads dir/{qq,dd,gg}_init.r8 (ng*ng*(nz+1) values each), forms the
 * spectral state as the loader does, runs the time loop to t > duration with saves every save_interval
 * (the cadence counted in time_step units; the solver's step is 1/ng), and writes dir/monitor.asc,
 * dir/ecomp.asc, dir/spectra.asc, dir/3d/{ql,d,g,r,w,pn}.r4 and dir/2d/{d,zeta,q,g,h}.r4 in the reference's
 * formats.  *iters_total = the psolve iterations of the run. */
int sw_nhswps(const char* dir, size_t ng, size_t nz, double time_step, double duration, double save_interval,
              int precision, int* iters_total);
/* balinit (src/balinit.rs:21-436): the balanced initial state of a 2-D PV anomaly.  qq_init: ng*ng values
   (the record of qq_init.r8); out: the balanced q, delta and gamma (the three records of sw_init.r8),
   physical, ng*ng each; iters: the iterations run (may be null).  Any precision handle: it runs in FP64. */
int sw_balinit(sw_spectral* s, const double* qq_init, double* qq, double* dd, double* gg, int* iters);
/* balinit(&Parameters) as the program runs it: reads dir/qq_init.r8, writes dir/sw_init.r8 (balinit.rs:28-36,
   419-433) */
int sw_balinit_file(const char* dir, size_t ng, size_t nz, int* iters);
/* FP64 checkpoint of a state in the r8 layout: an 8-byte header holding t, then the 16 fields in the order
 * of the State (no reference counterpart: the reference keeps only the t = 0 *_init.r8 files). */
int sw_state_save(sw_state* st, const char* path);
int sw_state_load(sw_state* st, const char* path);

/* ---- ensemble: members advanced together (no reference function; the ensemble configuration of
 * the batched workload) ---- */
typedef struct sw_ensemble sw_ensemble;
/* members: nmember states, each with its own spectral handle (any ng, nz >= 3, precision); the
 * ensemble does not own them.  nstreams: HIP streams the units of work (batches, below, and single
 * members) are spread over, each driven by one host thread (0: one per unit, at most 8). */
int sw_ensemble_create(int nmember, sw_state* const* members, int nstreams, sw_ensemble** out);
int sw_ensemble_destroy(sw_ensemble* e);
/* One advance step (advance.rs:13) of every member, as sw_advance does for one state: diag[5 i ..],
 * iters[3 i ..] and status[i] (0 or a negative sw_status) per member i; returns the first member's
 * non-zero status, if any. */
int sw_ensemble_advance(sw_ensemble* e, double* diag, int* iters, int* status);
/* Graph replay: after a member's first step with ggen set, its next step is
 * recorded as a hipGraph (a snapshot copy, then advance with each psolve given its last iteration count
 * plus margin gated iterations) and replayed on later steps; a replay whose psolve needed more iterations
 * is undone from the snapshot and redone eagerly, and the step is recorded again.  Results are
 * bit-identical either way.  Setting drops the recorded graphs.  Default: off (on MI355X with ROCm 7.2
 * the replays measure no faster than eager launches), margin 4. */
int sw_ensemble_set_graphs(sw_ensemble* e, int enable, int margin);
/* graph replays, eager redos and captures since the ensemble was created */
int sw_ensemble_stats(sw_ensemble* e, long* replays, long* redos, long* captures);
/* Batching: members of the same ng, nz and precision, with nz in {4, 8, 16, 24, 32}, are advanced as
 * batches of at most max_members (every stage of the step runs for all of them in the same launches, on
 * member-contiguous copies of their fields; results bit-identical to advancing them alone).  A batch whose
 * members differ in ggen or psolve tolerance at a step runs them one by one.  A member whose psolve fails
 * in a batched step has its status set and its fields unspecified, as after a failed sw_advance; the others
 * complete the step.  0 or 1: off.  Default: 32.  Graph replay (above) applies to unbatched members. */
int sw_ensemble_set_batching(sw_ensemble* e, int max_members);
/* the current batches: their count, the members they hold, the batched steps run so far, and the
 * number of streams */
int sw_ensemble_batches(sw_ensemble* e, int* nbatch, int* members, long* batched_steps, int* nstreams);

/* ---- tiles of the block-per-rows transform passes (no reference counterpart; the transform lengths
 * without wave-per-sequence kernels, i.e. other than 128, 256, 512) ---- */
/* Rows per block B (B / 2 complex sequences of two packed real rows) and threads per block nt of the
 * passes of length n over values of `bytes` bytes (8: FP64, 4: FP32), for the engines created after the
 * call; B = 0 drops the override.  Valid: B divides n, B even (1 for odd n), nt a multiple of 64 up to 512, the
 * block's B rows within the 160 KB of LDS per workgroup, at most 16 complex values per thread.  Results
 * do not depend on the tile.  Default: the measured table compiled into the library (autotune.py), else
 * the largest power-of-two B up to 16 within 64 KB of LDS. */
int sw_fft_set_tile(size_t n, int bytes, int B, int nt);
/* the tile an engine created now uses */
int sw_fft_get_tile(size_t n, int bytes, int* B, int* nt);
/* Time of one forward and one inverse 2-D transform of nslab plain slabs with tile (B, nt), averaged
 * over reps calls (microseconds per transform pair); the autotuner's measurement. */
int sw_fft_time_tile(size_t n, int bytes, int B, int nt, int nslab, int reps, double* us);

/* ---- instrumentation and controls (no reference counterpart) ---- */
/* Time one kind of launch of this handle with HIP events on the library stream:
 * SW_PROF_STRAIGHT = the straight (second) pass of every 2-D transform, SW_PROF_PSOLVE_DERIV = the
 * transposing pass of psolve's theta-derivative transform, SW_PROF_PSOLVE_COMB = psolve's fused
 * inverse-transform / source / forward-transform launch, SW_PROF_PSOLVE_TRI = psolve's fused forward
 * pass and column solve (the last two for ng = 128, 256, 512). */
typedef enum {
    SW_PROF_OFF = 0, SW_PROF_STRAIGHT = 1, SW_PROF_PSOLVE_DERIV = 2, SW_PROF_PSOLVE_COMB = 3, SW_PROF_PSOLVE_TRI = 4,
    SW_PROF_STAGES = 5
} sw_prof_mode;
int sw_profile(sw_spectral* s, int mode);
int sw_profile_read(sw_spectral* s, double* ms, double* bytes, long* launches);
/* SW_PROF_STAGES: time spent in each stage of the advance steps since sw_profile (HIP events on the
 * library stream): ms[0] main_invert, [1] psolve before its iteration (ri, vertical, coeffs, cpsource),
 * [2] psolve's iteration, [3] psolve after it (pn, dpn), [4] source, [5] the pointwise updates,
 * [6] diagnose; *steps the advance steps and *iterations their psolve iterations. */
int sw_profile_stages(sw_spectral* s, double ms[7], long* steps, long* iterations);
/* psolve stopping tolerance; default 1e-9 (psolve.rs:17) in FP64, 1e-6 for the mixed mode's FP32
 * iteration, where the FP32 transforms put a floor under the residual. */
int sw_set_tolerance(sw_spectral* s, double toler);

/* ---- synthetic initial state (host, no reference counterpart: the bench/test generator) ---- */
int sw_synth_kminus3(size_t ng, size_t nz, unsigned long long seed, double k0, double zeta_rms,
                     double delta_rms, double* qq, double* dd);

#ifdef __cplusplus
}
#endif
#endif
