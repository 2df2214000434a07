// CPU restatement of the reference's balanced initialisation (test infrastructure only: the parity
// checker; never linked into the GPU library).
//
// Follows src/balinit.rs:21-436: from a 2-D PV anomaly, the height anomaly, vorticity and velocity of
// the flow, then the fixed-point iteration for the balanced divergence (delta_t = 0) and acceleration
// divergence (gamma_t = 0), stopped when the energy-norm change falls below 1e-9 or starts to grow.
// Sums are sequential, as Rust's Iterator::sum over a Vec is; (a - b).powf(2.0) is pow(a - b, 2.0).
// D2FFT::ptospc/spctop destroy their input in the reference and in this restatement alike, so each
// call below is given the array the reference gives it.
#include "sw_oracle_internal.h"

#include <cmath>
#include <vector>

namespace orc {

typedef std::vector<double> Vec;

static double seq_sum(const Vec& a) {
    double s = 0.0;
    for (double x : a) s = s + x;
    return s;
}

// Returns the iteration count; qq, dd, gg: the three records of sw_init.r8 (balinit.rs:419-433)
int balinit(const Spectral& S, const double* qq_init, double* qq_out, double* dd_out, double* gg_out) {
    const size_t ng = S.ng, n2 = ng * ng;
    const D2FFT& f = S.d2fft;
    const double toler = 1.0e-9, small = 1.0e-12;
    const double dsumi = 1.0 / (double)n2;
    Vec zz(qq_init, qq_init + n2);
    Vec uu(n2, 0.0), vv(n2, 0.0), qq = zz, hh(n2, 0.0), dd(n2, 0.0), gg(n2, 0.0);
    Vec wkp(n2, 0.0), wkq(n2, 0.0), htot(n2, 0.0), ds(n2, 0.0), gs(n2, 0.0);
    Vec wka(n2, 0.0), wkb(n2, 0.0), wkc(n2, 0.0), wkd(n2, 0.0), wke(n2, 0.0);
    double qadd, qbar, fqbar, uio, vio, ddrmserr, ggrmserr, toterr, toterrpre;

    // height anomaly (balinit.rs:94-125)
    qadd = -dsumi * seq_sum(qq);
    for (auto& e : qq) e += qadd;
    qbar = dsumi * seq_sum(qq);
    for (size_t i = 0; i < n2; i++) wkp[i] = COF * qq[i];
    f.ptospc(wkp.data(), wkb.data());
    fqbar = COF * qbar;
    for (size_t p = 0; p < n2; p++) wka[p] = S.filt[p] * wkb[p] / (S.opak[p] - fqbar);
    f.spctop(wka.data(), hh.data());
    for (size_t i = 0; i < n2; i++) htot[i] = 1.0 + hh[i];

    // relative vorticity (:127-139)
    for (size_t i = 0; i < n2; i++) wkp[i] = htot[i] * qq[i];
    qadd = -dsumi * seq_sum(wkp);
    for (auto& e : qq) e += qadd;
    for (size_t i = 0; i < n2; i++) zz[i] = htot[i] * (COF + qq[i]) - COF;

    // velocity and its mean flow (:141-180)
    f.ptospc(zz.data(), wkb.data());
    for (size_t p = 0; p < n2; p++) {
        wka[p] = S.rlap[p] * wkb[p];
        wkb[p] *= S.filt[p];
    }
    f.spctop(wkb.data(), zz.data());
    f.xderiv(S.hrkx.data(), wka.data(), wkd.data());
    f.yderiv(S.hrky.data(), wka.data(), wkb.data());
    for (auto& e : wkb) e = -e;
    f.spctop(wkb.data(), uu.data());
    f.spctop(wkd.data(), vv.data());
    {
        double s = 0.0;
        for (size_t i = 0; i < n2; i++) s = s + hh[i] * uu[i];
        uio = -s * dsumi;
    }
    for (auto& e : uu) e += uio;

    // the iteration (:182-412)
    toterrpre = 1.0 / std::pow(small, 2.0);
    toterr = 1.0 / 2.0;
    Vec ddpre = dd, ggpre = gg;
    int iters = 0;
    auto rmserr = [&](const Vec& a, const Vec& pre) {
        double num = 0.0, den = 0.0;
        for (size_t i = 0; i < n2; i++) num = num + std::pow(a[i] - pre[i], 2.0);
        for (size_t i = 0; i < n2; i++) den = den + std::pow(pre[i], 2.0);
        return num / (den + small);
    };
    while (toterr > toler) {
        iters++;
        // balanced gamma (:184-223)
        S.jacob(uu.data(), vv.data(), wkb.data());
        for (size_t i = 0; i < n2; i++) wkp[i] = dd[i] * uu[i];
        for (size_t i = 0; i < n2; i++) wkq[i] = dd[i] * vv[i];
        S.divs(wkp.data(), wkq.data(), wka.data());
        for (size_t p = 0; p < n2; p++) gs[p] = S.filt[p] * (wka[p] - 2.0 * wkb[p]);
        wka = gs;
        f.spctop(wka.data(), gg.data());
        ggrmserr = rmserr(gg, ggpre);

        // balanced delta (:225-268)
        for (size_t i = 0; i < n2; i++) wkp[i] = hh[i] * uu[i];
        for (size_t i = 0; i < n2; i++) wkq[i] = hh[i] * vv[i];
        S.divs(wkp.data(), wkq.data(), wka.data());
        for (size_t i = 0; i < n2; i++) wkp[i] = zz[i] * uu[i];
        for (size_t i = 0; i < n2; i++) wkq[i] = zz[i] * vv[i];
        S.divs(wkp.data(), wkq.data(), wkb.data());
        for (size_t p = 0; p < n2; p++) ds[p] = S.helm[p] * (COF * wkb[p] - S.c2g2[p] * wka[p]);
        wka = ds;
        f.spctop(wka.data(), dd.data());
        ddrmserr = rmserr(dd, ddpre);

        // height anomaly (:270-302)
        for (size_t i = 0; i < n2; i++) htot[i] = 1.0 + hh[i];
        {
            double s = 0.0;
            for (size_t i = 0; i < n2; i++) s = s + qq[i] * htot[i];
            qadd = -dsumi * s;
        }
        for (auto& e : qq) e += qadd;
        qbar = dsumi * seq_sum(qq);
        for (size_t i = 0; i < n2; i++) wkp[i] = COF * (qq[i] + hh[i] * (qq[i] - qbar)) - gg[i];
        f.ptospc(wkp.data(), wkb.data());
        fqbar = COF * qbar;
        for (size_t p = 0; p < n2; p++) wka[p] = S.filt[p] * wkb[p] / (S.opak[p] - fqbar);
        f.spctop(wka.data(), hh.data());
        for (size_t i = 0; i < n2; i++) htot[i] = 1.0 + hh[i];

        // relative vorticity (:304-317)
        for (size_t i = 0; i < n2; i++) wkp[i] = qq[i] * htot[i];
        qadd = -dsumi * seq_sum(wkp);
        for (auto& e : qq) e += qadd;
        for (size_t i = 0; i < n2; i++) zz[i] = htot[i] * (COF + qq[i]) - COF;

        // velocity (:319-384)
        f.ptospc(zz.data(), wkb.data());
        for (size_t p = 0; p < n2; p++) wka[p] = S.rlap[p] * wkb[p];
        for (size_t p = 0; p < n2; p++) wkb[p] *= S.filt[p];
        f.spctop(wkb.data(), zz.data());
        f.xderiv(S.hrkx.data(), wka.data(), wkd.data());
        f.yderiv(S.hrky.data(), wka.data(), wkb.data());
        for (size_t p = 0; p < n2; p++) wke[p] = S.rlap[p] * ds[p];
        f.xderiv(S.hrkx.data(), wke.data(), wka.data());
        f.yderiv(S.hrky.data(), wke.data(), wkc.data());
        for (size_t i = 0; i < n2; i++) wkb[i] = wka[i] - wkb[i];
        for (size_t i = 0; i < n2; i++) wkd[i] += wkc[i];
        f.spctop(wkb.data(), uu.data());
        f.spctop(wkd.data(), vv.data());

        // mean flow (:386-395)
        {
            double su = 0.0, sv = 0.0;
            for (size_t i = 0; i < n2; i++) su = su + hh[i] * uu[i];
            for (size_t i = 0; i < n2; i++) sv = sv + hh[i] * vv[i];
            uio = -su * dsumi;
            vio = -sv * dsumi;
        }
        for (auto& e : uu) e += uio;
        for (auto& e : vv) e += vio;

        // overall error and the stopping rule (:397-411)
        toterr = (1.0 / 2.0) * (ddrmserr + ggrmserr);
        if (toterrpre <= toterr) break;
        ddpre = dd;
        ggpre = gg;
        toterrpre = toterr;
    }

    for (size_t i = 0; i < n2; i++) {
        qq_out[i] = zz[i] - COF * hh[i];
        dd_out[i] = dd[i];
        gg_out[i] = gg[i];
    }
    return iters;
}

}  // namespace orc
