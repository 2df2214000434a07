// extern "C" wrapper of the CPU oracle (test infrastructure only).
#include "sw_oracle.h"

#include <omp.h>

#include <cstring>
#include <string>

#include "sw_oracle_internal.h"

using namespace orc;

static State make_state(void* s, double** fields, double t, int ggen) {
    State st;
    st.sp = static_cast<const Spectral*>(s);
    for (int i = 0; i < NFIELDS; i++) st.f[i] = fields[i];
    st.t = t;
    st.ggen = ggen;
    return st;
}

extern "C" {

int orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}
int orc_factorisen(size_t n, size_t* f5) { return factorisen(n, f5); }
int orc_initfft(size_t n, size_t* f5, double* trig) { return initfft(n, f5, trig); }
int orc_forfft(size_t m, size_t n, double* x, const double* trig, const size_t* f5) {
    return forfft(m, n, x, trig, f5);
}
int orc_revfft(size_t m, size_t n, double* x, const double* trig, const size_t* f5) {
    return revfft(m, n, x, trig, f5);
}
int orc_init2dfft(size_t nx, size_t ny, double lx, double ly, size_t* xf5, size_t* yf5, double* xtrig,
                  double* ytrig, double* kx, double* ky) {
    D2FFT d;
    if (d.init(nx, ny, lx, ly, kx, ky)) return -1;
    std::memcpy(xf5, d.xf, sizeof(d.xf));
    std::memcpy(yf5, d.yf, sizeof(d.yf));
    std::memcpy(xtrig, d.xtrig.data(), sizeof(double) * 2 * nx);
    std::memcpy(ytrig, d.ytrig.data(), sizeof(double) * 2 * ny);
    return 0;
}
static D2FFT make_d2(size_t nx, size_t ny, const size_t* xf5, const size_t* yf5, const double* xtrig,
                     const double* ytrig) {
    D2FFT d;
    d.nx = nx;
    d.ny = ny;
    std::memcpy(d.xf, xf5, sizeof(d.xf));
    std::memcpy(d.yf, yf5, sizeof(d.yf));
    d.xtrig.assign(xtrig, xtrig + 2 * nx);
    d.ytrig.assign(ytrig, ytrig + 2 * ny);
    return d;
}
int orc_ptospc(size_t nx, size_t ny, const size_t* xf5, const size_t* yf5, const double* xtrig,
               const double* ytrig, double* rvar, double* svar) {
    make_d2(nx, ny, xf5, yf5, xtrig, ytrig).ptospc(rvar, svar);
    return 0;
}
int orc_spctop(size_t nx, size_t ny, const size_t* xf5, const size_t* yf5, const double* xtrig,
               const double* ytrig, double* svar, double* rvar) {
    make_d2(nx, ny, xf5, yf5, xtrig, ytrig).spctop(svar, rvar);
    return 0;
}
int orc_xderiv(size_t nx, size_t ny, const double* rkx, const double* var, double* der) {
    D2FFT d;
    d.nx = nx;
    d.ny = ny;
    d.xderiv(rkx, var, der);
    return 0;
}
int orc_yderiv(size_t nx, size_t ny, const double* rky, const double* var, double* der) {
    D2FFT d;
    d.nx = nx;
    d.ny = ny;
    d.yderiv(rky, var, der);
    return 0;
}
void* orc_spectral_new(size_t ng, size_t nz) {
    Spectral* s = new Spectral();
    if (s->init(ng, nz)) {
        delete s;
        return nullptr;
    }
    return s;
}
void orc_spectral_free(void* s) { delete static_cast<Spectral*>(s); }
long orc_spectral_get(void* sv, const char* name, double* out) {
    const Spectral* s = static_cast<const Spectral*>(sv);
    std::string n(name);
    const std::vector<double>* v = nullptr;
    if (n == "hlap") v = &s->hlap;
    else if (n == "glap") v = &s->glap;
    else if (n == "rlap") v = &s->rlap;
    else if (n == "helm") v = &s->helm;
    else if (n == "c2g2") v = &s->c2g2;
    else if (n == "simp") v = &s->simp;
    else if (n == "rope") v = &s->rope;
    else if (n == "fope") v = &s->fope;
    else if (n == "filt") v = &s->filt;
    else if (n == "diss") v = &s->diss;
    else if (n == "opak") v = &s->opak;
    else if (n == "rdis") v = &s->rdis;
    else if (n == "etdv") v = &s->etdv;
    else if (n == "htdv") v = &s->htdv;
    else if (n == "ap") v = &s->ap;
    else if (n == "etd1") v = &s->etd1;
    else if (n == "htd1") v = &s->htd1;
    else if (n == "theta") v = &s->theta;
    else if (n == "weight") v = &s->weight;
    else if (n == "hrkx") v = &s->hrkx;
    else if (n == "hrky") v = &s->hrky;
    else if (n == "rk") v = &s->rk;
    else if (n == "spmf") v = &s->spmf;
    else if (n == "alk") v = &s->alk;
    else if (n == "kmag") v = &s->kmag;
    else if (n == "xtrig") v = &s->d2fft.xtrig;
    else if (n == "ytrig") v = &s->d2fft.ytrig;
    else if (n == "kmax") { if (out) out[0] = (double)s->kmax; return 1; }
    else if (n == "kmaxred") { if (out) out[0] = (double)s->kmaxred; return 1; }
    else if (n == "xfactors" || n == "yfactors") {
        const size_t* f = n == "xfactors" ? s->d2fft.xf : s->d2fft.yf;
        if (out) for (int i = 0; i < 5; i++) out[i] = (double)f[i];
        return 5;
    }
    if (!v) return -1;
    if (out) std::memcpy(out, v->data(), sizeof(double) * v->size());
    return (long)v->size();
}
int orc_main_invert(void* s, const double* qs, const double* ds, const double* gs, double* r, double* u,
                    double* v, double* zeta) {
    static_cast<Spectral*>(s)->main_invert(qs, ds, gs, r, u, v, zeta);
    return 0;
}
int orc_jacob(void* s, const double* aa, const double* bb, double* cs) {
    static_cast<Spectral*>(s)->jacob(aa, bb, cs);
    return 0;
}
int orc_divs(void* s, const double* aa, const double* bb, double* cs) {
    static_cast<Spectral*>(s)->divs(aa, bb, cs);
    return 0;
}
int orc_ptospc3d(void* s, const double* fp, double* fs, size_t izbeg, size_t izend) {
    static_cast<Spectral*>(s)->ptospc3d(fp, fs, izbeg, izend);
    return 0;
}
int orc_spctop3d(void* s, const double* fs, double* fp, size_t izbeg, size_t izend) {
    static_cast<Spectral*>(s)->spctop3d(fs, fp, izbeg, izend);
    return 0;
}
int orc_deal3d(void* s, double* fp) {
    static_cast<Spectral*>(s)->deal3d(fp);
    return 0;
}
int orc_deal2d(void* s, double* fp) {
    static_cast<Spectral*>(s)->deal2d(fp);
    return 0;
}
int orc_spec1d(void* s, const double* ss, double* spec) {
    static_cast<Spectral*>(s)->spec1d(ss, spec);
    return 0;
}
int orc_balinit(void* s, const double* qq_init, double* qq, double* dd, double* gg) {
    return balinit(*static_cast<Spectral*>(s), qq_init, qq, dd, gg);
}
int orc_vertical(void* s, double** fields) {
    State st = make_state(s, fields, 0.0, 0);
    vertical(st);
    return 0;
}
int orc_coeffs(void* s, double** fields, double* sigx, double* sigy, double* cpt1, double* cpt2) {
    coeffs(make_state(s, fields, 0.0, 0), sigx, sigy, cpt1, cpt2);
    return 0;
}
int orc_cpsource(void* s, double** fields, double* sp0) {
    cpsource(make_state(s, fields, 0.0, 0), sp0);
    return 0;
}
int orc_psolve(void* s, double** fields, int nitmax, int* iters, double* errp) {
    State st = make_state(s, fields, 0.0, 0);
    if (static_cast<Spectral*>(s)->nz < 3) return -1;
    return psolve(st, nitmax, iters, errp);
}
int orc_source(void* s, double** fields, double* sqs, double* sds, double* sgs) {
    source(make_state(s, fields, 0.0, 0), sqs, sds, sgs);
    return 0;
}
int orc_diagnose(void* s, double** fields, double t, double* out5) {
    diagnose(make_state(s, fields, t, 0), out5);
    return 0;
}
int orc_advance(void* s, double** fields, double* t, int ggen, int nitmax, double* diag5, int* iters3) {
    if (static_cast<Spectral*>(s)->nz < 3) return -1;
    State st = make_state(s, fields, *t, ggen);
    int rc = advance(st, nitmax, diag5, iters3);
    *t = st.t;
    return rc;
}
int orc_energy(void* s, double** fields, double t, double* out6) {
    energy(make_state(s, fields, t, 0), out6);
    return 0;
}
int orc_savegrid(void* s, double** fields, float* d3, float* d2, double* spec3) {
    savegrid(make_state(s, fields, 0.0, 0), d3, d2, spec3);
    return 0;
}
int orc_run(void* s, double** fields, double* t, double tsim, double tgsave, int nitmax, double* mon, int maxmon,
            int* nmon, double* eco, int maxeco, int* neco, int* iters_total) {
    if (static_cast<Spectral*>(s)->nz < 3 || !(tgsave > 0.0)) return -1;
    State st = make_state(s, fields, *t, 0);
    int rc = run(st, tsim, tgsave, nitmax, mon, maxmon, nmon, eco, maxeco, neco, iters_total);
    *t = st.t;
    return rc;
}
}
