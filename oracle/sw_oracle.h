/* C ABI of the CPU oracle: a restatement of the reference crate's nhswps hot path
 * (src/stafft/, src/sta2dfft.rs, src/spectral/mod.rs and the src/nhswps/ kernels).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load this library, and only as the parity checker / CPU baseline.  The GPU product path
 * never links or calls it.
 *
 * Arrays use the reference's layout: element (i,j,iz) at i + ng*j + ng*ng*iz.  A "fields" argument
 * is an array of 16 pointers to (ng,ng,nz+1) arrays in the order of the reference State
 * (src/nhswps/mod.rs:87-133): u v w z zx zy r ri aa zeta pn dpn ps qs ds gs.
 * Return codes: 0 ok, -1 bad argument/factorisation, -2 psolve iteration cap, -3 psolve diverged.
 */
#pragma once
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif
int orc_set_threads(int n);
int orc_factorisen(size_t n, size_t* f5);
int orc_initfft(size_t n, size_t* f5, double* trig);
int orc_forfft(size_t m, size_t n, double* x, const double* trig, const size_t* f5);
int orc_revfft(size_t m, size_t n, double* x, const double* trig, const size_t* f5);
int orc_init2dfft(size_t nx, size_t ny, double lx, double ly, size_t* xf5, size_t* yf5, double* xtrig,
                  double* ytrig, double* kx, double* ky);
int orc_ptospc(size_t nx, size_t ny, const size_t* xf5, const size_t* yf5, const double* xtrig,
               const double* ytrig, double* rvar, double* svar);
int orc_spctop(size_t nx, size_t ny, const size_t* xf5, const size_t* yf5, const double* xtrig,
               const double* ytrig, double* svar, double* rvar);
int orc_xderiv(size_t nx, size_t ny, const double* rkx, const double* var, double* der);
int orc_yderiv(size_t nx, size_t ny, const double* rky, const double* var, double* der);
void* orc_spectral_new(size_t ng, size_t nz);
void orc_spectral_free(void* s);
long orc_spectral_get(void* s, const char* name, double* out);
int orc_main_invert(void* s, const double* qs, const double* ds, const double* gs, double* r, double* u,
                    double* v, double* zeta);
int orc_jacob(void* s, const double* aa, const double* bb, double* cs);
int orc_divs(void* s, const double* aa, const double* bb, double* cs);
int orc_ptospc3d(void* s, const double* fp, double* fs, size_t izbeg, size_t izend);
int orc_spctop3d(void* s, const double* fs, double* fp, size_t izbeg, size_t izend);
int orc_deal3d(void* s, double* fp);
int orc_deal2d(void* s, double* fp);
int orc_spec1d(void* s, const double* ss, double* spec);
int orc_vertical(void* s, double** fields);
int orc_coeffs(void* s, double** fields, double* sigx, double* sigy, double* cpt1, double* cpt2);
int orc_cpsource(void* s, double** fields, double* sp0);
int orc_psolve(void* s, double** fields, int nitmax, int* iters, double* errp);
int orc_source(void* s, double** fields, double* sqs, double* sds, double* sgs);
int orc_diagnose(void* s, double** fields, double t, double* out5);
int orc_advance(void* s, double** fields, double* t, int ggen, int nitmax, double* diag5, int* iters3);
int orc_energy(void* s, double** fields, double t, double* out6);
/* savegrid's output as numbers (src/nhswps/mod.rs:438-652): d3 = 6 (ng,ng,nz+1) f32 fields ql d g r w
 * pn; d2 = 5 (ng,ng) f32 fields d zeta q g h; spec3 = 3 arrays of ng+1 (zeta, delta, gamma spectra) */
int orc_savegrid(void* s, double** fields, float* d3, float* d2, double* spec3);
int orc_run(void* s, double** fields, double* t, double tsim, double tgsave, int nitmax, double* mon, int maxmon,
            int* nmon, double* eco, int maxeco, int* neco, int* iters_total);
/* balinit (src/balinit.rs:21-436): the balanced q, delta, gamma (the three records of sw_init.r8) of the
 * 2-D PV anomaly qq_init (the record of qq_init.r8), ng*ng each; returns the iteration count (>= 1) */
int orc_balinit(void* s, const double* qq_init, double* qq, double* dd, double* gg);
#ifdef __cplusplus
}
#endif
