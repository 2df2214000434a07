// Internal declarations of the CPU oracle (test infrastructure: parity checker and CPU baseline).
#pragma once
#include <cstddef>
#include <vector>

namespace orc {

const double PI = 3.14159265358979323846;
extern const double SINFPI3, RTF516, SINF2PI5, SINRAT, FRAC_1_SQRT_2;

// src/constants.rs:8-17
const double COF = 4.0 * PI;
const double CGW = 2.0 * PI;
const double HBAR = 0.4;
const double CDAMP = 10.0;
const double NNU = 3.0;
const double CSQ = CGW * CGW;
const double FSQ = COF * COF;
const double COFI = 1.0 / COF;
const double GRAVITY = CSQ / HBAR;

int factorisen(size_t n, size_t f[5]);
int initfft(size_t n, size_t f[5], double* trig);
int forfft(size_t m, size_t n, double* xs, const double* trig, const size_t f[5]);
int revfft(size_t m, size_t n, double* xs, const double* trig, const size_t f[5]);

struct D2FFT {
    size_t nx = 0, ny = 0;
    size_t xf[5] = {0, 0, 0, 0, 0}, yf[5] = {0, 0, 0, 0, 0};
    std::vector<double> xtrig, ytrig;
    int init(size_t nx, size_t ny, double lx, double ly, double* kx, double* ky);
    void ptospc(double* rvar, double* svar) const;
    void spctop(double* svar, double* rvar) const;
    void xderiv(const double* rkx, const double* var, double* der) const;
    void yderiv(const double* rky, const double* var, double* der) const;
};

// src/spectral/mod.rs:17-62
struct Spectral {
    size_t ng = 0, nz = 0;
    std::vector<double> hlap, glap, rlap, helm, c2g2, simp, rope, fope, filt, diss, opak, rdis;
    std::vector<double> etdv, htdv, ap;  // etdv/htdv: (ng,ng,nz) element (kx,ky,iz) at kx+ng*ky+ng*ng*iz
    std::vector<double> etd1, htd1, theta, weight, hrkx, hrky, rk, spmf, alk;
    std::vector<double> kmag;  // stored as double for export
    size_t kmax = 0, kmaxred = 0;
    D2FFT d2fft;
    int init(size_t ng, size_t nz);

    void main_invert(const double* qs, const double* ds, const double* gs, double* r, double* u,
                     double* v, double* zeta) const;
    void jacob(const double* aa, const double* bb, double* cs) const;
    void divs(const double* aa, const double* bb, double* cs) const;
    void ptospc3d(const double* fp, double* fs, size_t izbeg, size_t izend) const;
    void spctop3d(const double* fs, double* fp, size_t izbeg, size_t izend) const;
    void deal3d(double* fp) const;
    void deal2d(double* fp) const;
    void spec1d(const double* ss, double* spec) const;
};

// Field order of a State (src/nhswps/mod.rs:87-133).
enum Field { U = 0, V, W, Z, ZX, ZY, R, RI, AA, ZETA, PN, DPN, PS, QS, DS, GS, NFIELDS };

struct State {
    const Spectral* sp;
    double* f[NFIELDS];
    double t;
    int ggen;
};

void vertical(State& st);
void coeffs(const State& st, double* sigx, double* sigy, double* cpt1, double* cpt2);
void cpsource(const State& st, double* sp0);
int psolve(State& st, int nitmax, int* iters, double* errp_out);
void source(const State& st, double* sqs, double* sds, double* sgs);
void diagnose(const State& st, double out[5]);
int advance(State& st, int nitmax, double diag[5], int iters[3]);
void energy(const State& st, double out[6]);
void savegrid(const State& st, float* d3, float* d2, double* spec3);
int run(State& st, double tsim, double tgsave, int nitmax, double* mon, int maxmon, int* nmon, double* eco, int maxeco,
        int* neco, int* iters_total);

double unrolled_sum(const double* x, size_t n);

// src/balinit.rs:21-436 (2-D; returns the iteration count)
int balinit(const Spectral& S, const double* qq_init, double* qq_out, double* dd_out, double* gg_out);

}  // namespace orc
