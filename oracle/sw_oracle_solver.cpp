// CPU restatement of the reference's spectral operators and nhswps timestep kernels (test
// infrastructure only: the parity checker and the CPU baseline; never linked into the GPU library).
//
// Follows src/spectral/mod.rs (Spectral::new :65-299, main_invert :305-417, jacob :423-457,
// divs :461-475, ptospc3d :479-498, spctop3d :502-521, deal3d :524-541, deal2d :544-554,
// spec1d :558-591) and src/nhswps/{vertical.rs:8-143, coeffs.rs:8-101, cpsource.rs:13-261,
// psolve.rs:16-411, source.rs:19-138, advance.rs:13-279, mod.rs:666-711 (diagnose)}.
// Arrays use the reference's layout: element (i,j,iz) at i + ng*j + ng*ng*iz.  Levels are
// processed in parallel with OpenMP where the reference uses rayon; the reference's mutex
// accumulations (main_invert's mean flow) are summed here in level order.
#include "sw_oracle_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace orc {

typedef std::vector<double> Vec;

// ndarray's contiguous .sum(): eight interleaved partial sums (ndarray 0.15 numeric_util)
double unrolled_sum(const double* x, size_t n) {
    double acc = 0.0, p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0, p6 = 0, p7 = 0;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        p0 = p0 + x[i];
        p1 = p1 + x[i + 1];
        p2 = p2 + x[i + 2];
        p3 = p3 + x[i + 3];
        p4 = p4 + x[i + 4];
        p5 = p5 + x[i + 5];
        p6 = p6 + x[i + 6];
        p7 = p7 + x[i + 7];
    }
    acc = acc + (p0 + p4);
    acc = acc + (p1 + p5);
    acc = acc + (p2 + p6);
    acc = acc + (p3 + p7);
    for (; i < n; i++) acc = acc + x[i];
    return acc;
}

int Spectral::init(size_t ng_, size_t nz_) {
    ng = ng_;
    nz = nz_;
    if (ng < 2 || nz < 2) return -1;
    const size_t n2 = ng * ng;
    double dt = 1.0 / (double)ng;
    double dt2 = dt * (1.0 / 2.0);
    double dt2i = 1.0 / dt2;
    double dz = HBAR / (double)nz;
    double dzi = 1.0 / dz;
    double dzisq = std::pow(dzi, 2.0);

    for (Vec* a : {&hlap, &glap, &rlap, &helm, &c2g2, &simp, &rope, &fope, &filt, &diss, &opak, &rdis, &ap})
        a->assign(n2, 0.0);
    etdv.assign(n2 * nz, 0.0);
    htdv.assign(n2 * nz, 0.0);
    etd1.assign(nz, 0.0);
    htd1.assign(nz, 0.0);
    theta.assign(nz + 1, 0.0);
    weight.assign(nz + 1, 0.0);
    hrkx.assign(ng, 0.0);
    hrky.assign(ng, 0.0);
    rk.assign(ng, 0.0);
    spmf.assign(ng + 1, 0.0);
    alk.assign(ng, 0.0);
    kmag.assign(n2, 0.0);
    Vec a0(n2), a0b(n2), apb(n2);

    if (d2fft.init(ng, ng, 2.0 * PI, 2.0 * PI, hrkx.data(), hrky.data())) return -1;

    rk[0] = 0.0;
    for (size_t k = 1; k < ng / 2; k++) {
        rk[k] = hrkx[2 * k - 1];
        rk[ng - k] = hrkx[2 * k - 1];
    }
    rk[ng / 2] = hrkx[ng - 1];

    double rkmax = (double)(ng / 2);
    kmax = (size_t)std::ceil(rkmax * std::sqrt(2.0));
    for (size_t ky = 0; ky < ng; ky++)
        for (size_t kx = 0; kx < ng; kx++) {
            size_t k = (size_t)std::round(std::sqrt(std::pow(rk[kx], 2.0) + std::pow(rk[ky], 2.0)));
            kmag[kx + ng * ky] = (double)k;
            spmf[k] += 1.0;
        }
    double snorm = 4.0 * PI / (double)(ng * ng);
    spmf[0] = 0.0;
    for (size_t k = 1; k <= kmax; k++) {
        spmf[k] = snorm * (double)k / spmf[k];
        alk[k - 1] = std::log10((double)k);
    }
    kmaxred = ng / 2;

    double anu = CDAMP * COF / std::pow(rkmax, 2.0 * NNU);
    double rkfsq = std::pow((double)ng / 3.0, 2.0);
    for (size_t ky = 0; ky < ng; ky++)
        for (size_t kx = 0; kx < ng; kx++) {
            size_t p = kx + ng * ky;
            double rks = std::pow(rk[kx], 2.0) + std::pow(rk[ky], 2.0);
            hlap[p] = -rks;
            opak[p] = -(FSQ + CSQ * rks);
            diss[p] = anu * std::pow(rks, NNU);
            if (rks > rkfsq) {
                filt[p] = glap[p] = c2g2[p] = rlap[p] = helm[p] = rope[p] = rdis[p] = 0.0;
            } else {
                filt[p] = 1.0;
                glap[p] = GRAVITY * rks;
                c2g2[p] = -CSQ * rks;
                rlap[p] = -1.0 / (rks + 1.0E-20);
                helm[p] = 1.0 / opak[p];
                rope[p] = c2g2[p] * helm[p];
                rdis[p] = dt2i + diss[p];
            }
            double rrsq = std::pow(dt2i + diss[p], 2.0);
            fope[p] = -c2g2[p] / (rrsq - opak[p]);
            simp[p] = 1.0 / (rrsq + FSQ);
            diss[p] = 2.0 / (1.0 + dt2 * diss[p]);
        }
    rlap[0] = 0.0;

    for (size_t iz = 0; iz <= nz; iz++) {
        theta[iz] = dz * (double)iz;
        weight[iz] = 1.0;
    }
    weight[0] = 0.5;
    weight[nz] = 0.5;
    for (auto& e : weight) e /= (double)nz;

    for (size_t ky = 0; ky < ng; ky++)
        for (size_t kx = 0; kx < ng; kx++) {
            size_t p = kx + ng * ky;
            double rks = std::pow(rk[kx], 2.0) + std::pow(rk[ky], 2.0);
            a0[p] = -2.0 * dzisq - (5.0 / 6.0) * rks;
            a0b[p] = -dzisq - (1.0 / 3.0) * rks;
            ap[p] = dzisq - (1.0 / 12.0) * rks;
            apb[p] = dzisq - (1.0 / 6.0) * rks;
        }
    for (size_t p = 0; p < n2; p++) {
        htdv[p] = filt[p] / a0b[p];
        etdv[p] = -apb[p] * htdv[p];
    }
    for (size_t iz = 1; iz + 2 <= nz; iz++)
        for (size_t p = 0; p < n2; p++) {
            htdv[p + n2 * iz] = filt[p] / (a0[p] + ap[p] * etdv[p + n2 * (iz - 1)]);
            etdv[p + n2 * iz] = -ap[p] * htdv[p + n2 * iz];
        }
    for (size_t p = 0; p < n2; p++)
        htdv[p + n2 * (nz - 1)] = filt[p] / (a0[p] + ap[p] * etdv[p + n2 * (nz - 2)]);

    htd1[0] = 1.0 / (2.0 / 3.0);
    etd1[0] = -(1.0 / 6.0) * htd1[0];
    for (size_t iz = 2; iz < nz; iz++) {
        htd1[iz - 1] = 1.0 / ((2.0 / 3.0) + (1.0 / 6.0) * etd1[iz - 2]);
        etd1[iz - 1] = -(1.0 / 6.0) * htd1[iz - 1];
    }
    htd1[nz - 1] = 1.0 / ((2.0 / 3.0) + (1.0 / 3.0) * etd1[nz - 2]);
    return 0;
}

void Spectral::main_invert(const double* qs, const double* ds, const double* gs, double* r, double* u,
                           double* v, double* zeta) const {
    const size_t n2 = ng * ng, L = nz + 1;
    double dsumi = 1.0 / (double)(ng * ng);
    Vec es(n2 * L), wkh(n2, 0.0), ubar(L), vbar(L);
    for (size_t p = 0; p < n2 * L; p++) es[p] = COFI * (COFI * gs[p] - qs[p]);
    for (size_t iz = 0; iz < L; iz++)
        for (size_t p = 0; p < n2; p++) wkh[p] += weight[iz] * es[p + n2 * iz];
    for (size_t p = 0; p < n2; p++) wkh[p] *= rope[p];

#pragma omp parallel for schedule(dynamic)
    for (size_t iz = 0; iz < L; iz++) {
        Vec wka(n2), wkb(n2), wkc(n2), wkd(n2), wke(n2), wkf(n2), wkg(n2), prod(n2);
        const size_t o = n2 * iz;
        for (size_t p = 0; p < n2; p++) wka[p] = es[o + p] - wkh[p];
        for (size_t p = 0; p < n2; p++) wkb[p] = qs[o + p] + COF * wka[p];
        for (size_t p = 0; p < n2; p++) wkc[p] = rlap[p] * wkb[p];
        for (size_t p = 0; p < n2; p++) wkd[p] = rlap[p] * ds[o + p];
        d2fft.xderiv(hrkx.data(), wkd.data(), wke.data());
        d2fft.yderiv(hrky.data(), wkd.data(), wkf.data());
        d2fft.xderiv(hrkx.data(), wkc.data(), wkd.data());
        d2fft.yderiv(hrky.data(), wkc.data(), wkg.data());
        for (size_t p = 0; p < n2; p++) wke[p] -= wkg[p];
        for (size_t p = 0; p < n2; p++) wkf[p] += wkd[p];
        d2fft.spctop(wka.data(), r + o);
        d2fft.spctop(wkb.data(), zeta + o);
        d2fft.spctop(wke.data(), u + o);
        d2fft.spctop(wkf.data(), v + o);
        for (size_t p = 0; p < n2; p++) prod[p] = u[o + p] * r[o + p];
        ubar[iz] = weight[iz] * unrolled_sum(prod.data(), n2) * dsumi;
        for (size_t p = 0; p < n2; p++) prod[p] = v[o + p] * r[o + p];
        vbar[iz] = weight[iz] * unrolled_sum(prod.data(), n2) * dsumi;
    }
    double uio = 0.0, vio = 0.0;
    for (size_t iz = 0; iz < L; iz++) {
        uio -= ubar[iz];
        vio -= vbar[iz];
    }
    for (size_t p = 0; p < n2 * L; p++) {
        u[p] += uio;
        v[p] += vio;
    }
}

void Spectral::jacob(const double* aa, const double* bb, double* cs) const {
    const size_t n2 = ng * ng;
    Vec ax(n2), ay(n2), bx(n2), by(n2), wka(n2), wkb(aa, aa + n2);
    d2fft.ptospc(wkb.data(), wka.data());
    d2fft.xderiv(hrkx.data(), wka.data(), wkb.data());
    d2fft.spctop(wkb.data(), ax.data());
    d2fft.yderiv(hrky.data(), wka.data(), wkb.data());
    d2fft.spctop(wkb.data(), ay.data());
    wkb.assign(bb, bb + n2);
    d2fft.ptospc(wkb.data(), wka.data());
    d2fft.xderiv(hrkx.data(), wka.data(), wkb.data());
    d2fft.spctop(wkb.data(), bx.data());
    d2fft.yderiv(hrky.data(), wka.data(), wkb.data());
    d2fft.spctop(wkb.data(), by.data());
    for (size_t p = 0; p < n2; p++) wkb[p] = ax[p] * by[p] - ay[p] * bx[p];
    d2fft.ptospc(wkb.data(), cs);
}

void Spectral::divs(const double* aa, const double* bb, double* cs) const {
    const size_t n2 = ng * ng;
    Vec wkp(aa, aa + n2), wka(n2), wkb(n2);
    d2fft.ptospc(wkp.data(), wka.data());
    d2fft.xderiv(hrkx.data(), wka.data(), wkb.data());
    wkp.assign(bb, bb + n2);
    d2fft.ptospc(wkp.data(), wka.data());
    d2fft.yderiv(hrky.data(), wka.data(), cs);
    for (size_t p = 0; p < n2; p++) cs[p] += wkb[p];
}

void Spectral::ptospc3d(const double* fp, double* fs, size_t izbeg, size_t izend) const {
    const size_t n2 = ng * ng;
#pragma omp parallel for schedule(dynamic)
    for (size_t iz = izbeg; iz <= izend; iz++) {
        Vec wkp(fp + n2 * iz, fp + n2 * (iz + 1));
        d2fft.ptospc(wkp.data(), fs + n2 * iz);
    }
}

void Spectral::spctop3d(const double* fs, double* fp, size_t izbeg, size_t izend) const {
    const size_t n2 = ng * ng;
#pragma omp parallel for schedule(dynamic)
    for (size_t iz = izbeg; iz <= izend; iz++) {
        Vec wks(fs + n2 * iz, fs + n2 * (iz + 1));
        d2fft.spctop(wks.data(), fp + n2 * iz);
    }
}

void Spectral::deal3d(double* fp) const {
    const size_t n2 = ng * ng;
#pragma omp parallel for schedule(dynamic)
    for (size_t iz = 0; iz <= nz; iz++) {
        Vec wkp(fp + n2 * iz, fp + n2 * (iz + 1)), wks(n2);
        d2fft.ptospc(wkp.data(), wks.data());
        for (size_t p = 0; p < n2; p++) wks[p] *= filt[p];
        d2fft.spctop(wks.data(), fp + n2 * iz);
    }
}

void Spectral::deal2d(double* fp) const {
    const size_t n2 = ng * ng;
    Vec fs(n2);
    d2fft.ptospc(fp, fs.data());
    for (size_t p = 0; p < n2; p++) fs[p] *= filt[p];
    d2fft.spctop(fs.data(), fp);
}

void Spectral::spec1d(const double* ss, double* spec) const {
    for (size_t k = 0; k <= kmax; k++) spec[k] = 0.0;
    spec[(size_t)kmag[0]] += (1.0 / 4.0) * std::pow(ss[0], 2.0);
    for (size_t kx = 1; kx < ng; kx++) spec[(size_t)kmag[kx]] += (1.0 / 2.0) * std::pow(ss[kx], 2.0);
    for (size_t ky = 1; ky < ng; ky++)
        spec[(size_t)kmag[ng * ky]] += (1.0 / 2.0) * std::pow(ss[ng * ky], 2.0);
    for (size_t ky = 1; ky < ng; ky++)
        for (size_t kx = 1; kx < ng; kx++)
            spec[(size_t)kmag[kx + ng * ky]] += std::pow(ss[kx + ng * ky], 2.0);
}

// ---------------------------------------------------------------- nhswps kernels

void vertical(State& st) {
    const Spectral& S = *st.sp;
    const size_t ng = S.ng, nz = S.nz, n2 = ng * ng, L = nz + 1;
    double dz2 = (HBAR / (double)nz) / 2.0;
    double *z = st.f[Z], *zx = st.f[ZX], *zy = st.f[ZY], *r = st.f[R], *w = st.f[W];
    double *u = st.f[U], *v = st.f[V], *aa = st.f[AA], *ds = st.f[DS];
    Vec rsrc(n2 * L, 0.0);
    for (size_t p = 0; p < n2; p++) z[n2 + p] = dz2 * (r[p] + r[n2 + p]);
    for (size_t iz = 1; iz < nz; iz++)
        for (size_t p = 0; p < n2; p++)
            z[n2 * (iz + 1) + p] = z[n2 * iz + p] + dz2 * (r[n2 * iz + p] + r[n2 * (iz + 1) + p]);

#pragma omp parallel for schedule(dynamic)
    for (size_t iz = 1; iz <= nz; iz++) {
        Vec wkq(n2), wka(n2), wkb(n2);
        double th = S.theta[iz];
        for (size_t p = 0; p < n2; p++) z[n2 * iz + p] += th;
        for (size_t p = 0; p < n2; p++) wkq[p] = z[n2 * iz + p];
        S.d2fft.ptospc(wkq.data(), wka.data());
        S.d2fft.xderiv(S.hrkx.data(), wka.data(), wkb.data());
        S.d2fft.spctop(wkb.data(), zx + n2 * iz);
        S.d2fft.yderiv(S.hrky.data(), wka.data(), wkb.data());
        S.d2fft.spctop(wkb.data(), zy + n2 * iz);
    }

#pragma omp parallel for schedule(dynamic)
    for (size_t iz = 0; iz <= nz; iz++) {
        Vec wkq(n2), wka(n2), wkb(n2), wkc(n2);
        const size_t o = n2 * iz;
        for (size_t p = 0; p < n2; p++) wkq[p] = u[o + p] * r[o + p];
        S.d2fft.ptospc(wkq.data(), wka.data());
        S.d2fft.xderiv(S.hrkx.data(), wka.data(), wkb.data());
        for (size_t p = 0; p < n2; p++) wkq[p] = v[o + p] * r[o + p];
        S.d2fft.ptospc(wkq.data(), wka.data());
        S.d2fft.yderiv(S.hrky.data(), wka.data(), wkc.data());
        for (size_t p = 0; p < n2; p++) aa[o + p] = S.filt[p] * (wkb[p] + wkc[p]);
        for (size_t p = 0; p < n2; p++) wka[p] = aa[o + p] + ds[o + p];
        S.d2fft.spctop(wka.data(), wkq.data());
        for (size_t p = 0; p < n2; p++) rsrc[o + p] = -wkq[p];
    }

    for (size_t p = 0; p < n2; p++) w[n2 + p] = dz2 * (rsrc[p] + rsrc[n2 + p]);
    for (size_t iz = 1; iz < nz; iz++)
        for (size_t p = 0; p < n2; p++)
            w[n2 * (iz + 1) + p] = w[n2 * iz + p] + dz2 * (rsrc[n2 * iz + p] + rsrc[n2 * (iz + 1) + p]);

#pragma omp parallel for schedule(dynamic)
    for (size_t iz = 1; iz <= nz; iz++) {
        Vec wkq(n2);
        const size_t o = n2 * iz;
        for (size_t p = 0; p < n2; p++) wkq[p] = u[o + p] * zx[o + p] + v[o + p] * zy[o + p];
        S.deal2d(wkq.data());
        for (size_t p = 0; p < n2; p++) w[o + p] += wkq[p];
    }
}

void coeffs(const State& st, double* sigx, double* sigy, double* cpt1, double* cpt2) {
    const Spectral& S = *st.sp;
    const size_t ng = S.ng, nz = S.nz, n2 = ng * ng, L = nz + 1;
    double qdzi = (1.0 / 4.0) * (1.0 / (HBAR / (double)nz));
    const double *ri = st.f[RI], *zx = st.f[ZX], *zy = st.f[ZY];
    for (size_t p = 0; p < n2 * L; p++) sigx[p] = ri[p] * zx[p];
    for (size_t p = 0; p < n2 * L; p++) sigy[p] = ri[p] * zy[p];
    S.deal3d(sigx);
    S.deal3d(sigy);
    for (size_t p = 0; p < n2 * L; p++)
        cpt2[p] = 1.0 - std::pow(ri[p], 2.0) - std::pow(sigx[p], 2.0) - std::pow(sigy[p], 2.0);
    S.deal3d(cpt2);

#pragma omp parallel for schedule(dynamic)
    for (size_t iz = 0; iz <= nz; iz++) {
        const size_t o = n2 * iz;
        if (iz == 0) {
            for (size_t p = 0; p < n2; p++)
                cpt1[p] = qdzi * (4.0 * cpt2[n2 + p] - 3.0 * cpt2[p] - cpt2[2 * n2 + p]);
            continue;
        }
        Vec wka(n2), wkp(n2);
        S.divs(sigx + o, sigy + o, wka.data());
        S.d2fft.spctop(wka.data(), wkp.data());
        if (iz < nz) {
            for (size_t p = 0; p < n2; p++)
                cpt1[o + p] = qdzi * (cpt2[o + n2 + p] - cpt2[o - n2 + p]) + wkp[p];
        } else {
            for (size_t p = 0; p < n2; p++)
                cpt1[o + p] = qdzi * (3.0 * cpt2[o + p] + cpt2[o - 2 * n2 + p] - 4.0 * cpt2[o - n2 + p]) + wkp[p];
        }
    }
    for (size_t p = 0; p < n2 * L; p++) {
        sigx[p] *= 2.0;
        sigy[p] *= 2.0;
    }
}

void cpsource(const State& st, double* sp0) {
    const Spectral& S = *st.sp;
    const size_t ng = S.ng, nz = S.nz, n2 = ng * ng, L = nz + 1;
    double hdzi = (1.0 / 2.0) * (1.0 / (HBAR / (double)nz));
    const double *z = st.f[Z], *u = st.f[U], *v = st.f[V], *w = st.f[W], *ri = st.f[RI];
    const double *zeta = st.f[ZETA], *zx = st.f[ZX], *zy = st.f[ZY];
    Vec ut(n2 * L), vt(n2 * L), wt(n2 * L), hsrc(n2);
    {
        Vec wkp(z + n2 * nz, z + n2 * L), wka(n2);
        S.d2fft.ptospc(wkp.data(), wka.data());
        for (size_t p = 0; p < n2; p++) wka[p] *= S.glap[p];
        S.d2fft.spctop(wka.data(), hsrc.data());
    }
    for (size_t iz = 0; iz <= nz; iz++) {
        const size_t o = n2 * iz;
        const double* fs[3] = {u, v, w};
        double* ts[3] = {ut.data(), vt.data(), wt.data()};
        for (int q = 0; q < 3; q++) {
            const double* a = fs[q];
            double* t = ts[q];
            if (iz == 0) {
                for (size_t p = 0; p < n2; p++) t[p] = hdzi * (4.0 * a[n2 + p] - 3.0 * a[p] - a[2 * n2 + p]);
            } else if (iz == nz) {
                for (size_t p = 0; p < n2; p++)
                    t[o + p] = hdzi * (3.0 * a[o + p] + a[o - 2 * n2 + p] - 4.0 * a[o - n2 + p]);
            } else {
                for (size_t p = 0; p < n2; p++) t[o + p] = hdzi * (a[o + n2 + p] - a[o - n2 + p]);
            }
        }
    }

#pragma omp parallel for schedule(dynamic)
    for (size_t iz = 0; iz <= nz; iz++) {
        const size_t o = n2 * iz;
        Vec wka(n2), wkb(n2), wkp(n2), wkq(n2), wkr(n2), ux(n2), uy(n2), vx(n2), vy(n2), wx(n2), wy(n2);
        for (size_t p = 0; p < n2; p++) wkq[p] = u[o + p];
        S.d2fft.ptospc(wkq.data(), wka.data());
        S.d2fft.xderiv(S.hrkx.data(), wka.data(), wkb.data());
        S.d2fft.spctop(wkb.data(), ux.data());
        S.d2fft.yderiv(S.hrky.data(), wka.data(), wkb.data());
        S.d2fft.spctop(wkb.data(), uy.data());
        for (size_t p = 0; p < n2; p++) wkq[p] = v[o + p];
        S.d2fft.ptospc(wkq.data(), wka.data());
        S.d2fft.xderiv(S.hrkx.data(), wka.data(), wkb.data());
        S.d2fft.spctop(wkb.data(), vx.data());
        S.d2fft.yderiv(S.hrky.data(), wka.data(), wkb.data());
        S.d2fft.spctop(wkb.data(), vy.data());
        double* s = sp0 + o;
        if (iz == 0) {
            for (size_t p = 0; p < n2; p++) wkq[p] = ri[p] * wt[p];
            S.deal2d(wkq.data());
            for (size_t p = 0; p < n2; p++) s[p] = hsrc[p] + COF * zeta[p];
            for (size_t p = 0; p < n2; p++)
                s[p] += 2.0 * (ux[p] * vy[p] - uy[p] * vx[p] + wkq[p] * (ux[p] + vy[p]));
        } else {
            for (size_t p = 0; p < n2; p++) wkq[p] = w[o + p];
            S.d2fft.ptospc(wkq.data(), wka.data());
            S.d2fft.xderiv(S.hrkx.data(), wka.data(), wkb.data());
            S.d2fft.spctop(wkb.data(), wx.data());
            S.d2fft.yderiv(S.hrky.data(), wka.data(), wkb.data());
            S.d2fft.spctop(wkb.data(), wy.data());
            const double *utl = ut.data() + o, *vtl = vt.data() + o, *wtl = wt.data() + o;
            for (size_t p = 0; p < n2; p++) wkp[p] = vtl[p] * zx[o + p] - utl[p] * zy[o + p];
            S.deal2d(wkp.data());
            for (size_t p = 0; p < n2; p++) wkq[p] = uy[p] * vtl[p] - utl[p] * vy[p];
            S.deal2d(wkq.data());
            for (size_t p = 0; p < n2; p++) wkr[p] = utl[p] * vx[p] - ux[p] * vtl[p];
            S.deal2d(wkr.data());
            for (size_t p = 0; p < n2; p++) wkq[p] = wkq[p] * zx[o + p] + wkr[p] * zy[o + p] - wy[p] * vtl[p];
            for (size_t p = 0; p < n2; p++) wkq[p] += (ux[p] + vy[p]) * wtl[p] - wx[p] * utl[p];
            S.deal2d(wkq.data());
            for (size_t p = 0; p < n2; p++)
                s[p] = hsrc[p] + COF * (zeta[o + p] - ri[o + p] * wkp[p]) + 2.0 * (ri[o + p] * wkq[p]);
            for (size_t p = 0; p < n2; p++) s[p] += 2.0 * (ux[p] * vy[p] - uy[p] * vx[p]);
        }
    }
}

// Returns 0 on convergence, -2 when the iteration cap is reached, -3 when the error grows past 1.
int psolve(State& st, int nitmax, int* iters, double* errp_out) {
    const Spectral& S = *st.sp;
    const size_t ng = S.ng, nz = S.nz, n2 = ng * ng, L = nz + 1;
    const double toler = 1.0E-9;
    double dz = HBAR / (double)nz;
    double dzi = 1.0 / dz;
    double dz2 = dz / 2.0;
    double dz6 = dz / 6.0;
    double dzisq = std::pow(1.0 / dz, 2.0);
    double hdzi = (1.0 / 2.0) * (1.0 / (HBAR / (double)nz));
    double *ri = st.f[RI], *r = st.f[R], *pn = st.f[PN], *ps = st.f[PS], *dpn = st.f[DPN];

    Vec sp0(n2 * L), sigx(n2 * L), sigy(n2 * L), cpt1(n2 * L), cpt2(n2 * L);
    Vec dpdt(n2, 0.0), d2pdxt(n2), d2pdyt(n2), d2pdt2(n2, 0.0), wkp(n2), sp(n2 * L, 0.0), gg(n2 * L, 0.0);
    Vec wka(n2), wkb(n2), wkc(n2), wkd(n2), wkq(n2, 0.0);

    for (size_t p = 0; p < n2 * L; p++) ri[p] = 1.0 / (r[p] + 1.0);
    S.deal3d(ri);
    vertical(st);
    coeffs(st, sigx.data(), sigy.data(), cpt1.data(), cpt2.data());
    cpsource(st, sp0.data());

    Vec pna(pn, pn + n2 * L);
    double errp = 1.0;
    int iter = 0;
    while (errp > toler && iter < nitmax) {
        S.ptospc3d(pn, ps, 0, nz - 1);
        for (size_t p = 0; p < n2; p++) ps[n2 * nz + p] = 0.0;

        for (size_t p = 0; p < n2; p++)
            wkd[p] = (2.0 * ps[p] - 5.0 * ps[n2 + p] + 4.0 * ps[2 * n2 + p] - ps[3 * n2 + p]) * dzisq;
        S.d2fft.spctop(wkd.data(), d2pdt2.data());
        for (size_t p = 0; p < n2; p++) wkp[p] = sp0[p] + cpt2[p] * d2pdt2[p];
        S.d2fft.ptospc(wkp.data(), sp.data());

#pragma omp parallel for schedule(dynamic)
        for (size_t iz = 1; iz <= nz - 1; iz++) {
            const size_t o = n2 * iz;
            Vec la(n2), lb(n2), lc(n2), ld(n2), lp(n2), lx(n2), ly(n2), lt(n2), lt2(n2);
            for (size_t p = 0; p < n2; p++) la[p] = (ps[o + n2 + p] - ps[o - n2 + p]) * hdzi;
            for (size_t p = 0; p < n2; p++) ld[p] = (ps[o + n2 + p] - 2.0 * ps[o + p] + ps[o - n2 + p]) * dzisq;
            S.d2fft.xderiv(S.hrkx.data(), la.data(), lb.data());
            S.d2fft.yderiv(S.hrky.data(), la.data(), lc.data());
            S.d2fft.spctop(lb.data(), lx.data());
            S.d2fft.spctop(lc.data(), ly.data());
            for (size_t p = 0; p < n2; p++) lp[p] = sp0[o + p] + sigx[o + p] * lx[p] + sigy[o + p] * ly[p];
            S.d2fft.spctop(la.data(), lt.data());
            S.d2fft.spctop(ld.data(), lt2.data());
            if (iz == nz - 1) {
                dpdt = lt;
                d2pdt2 = lt2;
            } else if (iz == nz - 2) {
                wkq = lt2;
            }
            for (size_t p = 0; p < n2; p++) lp[p] += cpt2[o + p] * lt2[p] + cpt1[o + p] * lt[p];
            S.d2fft.ptospc(lp.data(), sp.data() + o);
        }

        for (size_t p = 0; p < n2; p++) dpdt[p] += dz2 * (3.0 * d2pdt2[p] - wkq[p]);
        for (size_t p = 0; p < n2; p++) d2pdt2[p] = 2.0 * d2pdt2[p] - wkq[p];
        wkp = dpdt;
        S.d2fft.ptospc(wkp.data(), wka.data());
        S.d2fft.xderiv(S.hrkx.data(), wka.data(), wkb.data());
        S.d2fft.yderiv(S.hrky.data(), wka.data(), wkc.data());
        S.d2fft.spctop(wkb.data(), d2pdxt.data());
        S.d2fft.spctop(wkc.data(), d2pdyt.data());
        {
            const size_t o = n2 * nz;
            for (size_t p = 0; p < n2; p++)
                wkp[p] = sp0[o + p] + sigx[o + p] * d2pdxt[p] + sigy[o + p] * d2pdyt[p];
            for (size_t p = 0; p < n2; p++) wkp[p] += cpt2[o + p] * d2pdt2[p] + cpt1[o + p] * dpdt[p];
            S.d2fft.ptospc(wkp.data(), sp.data() + o);
        }

        for (size_t p = 0; p < n2; p++) gg[p] = (1.0 / 3.0) * sp[p] + (1.0 / 6.0) * sp[n2 + p];
        for (size_t iz = 1; iz < nz; iz++) {
            const size_t o = n2 * iz;
            for (size_t p = 0; p < n2; p++)
                gg[o + p] = (1.0 / 12.0) * (sp[o - n2 + p] + sp[o + n2 + p]) + (5.0 / 6.0) * sp[o + p];
        }
        for (size_t p = 0; p < n2; p++) ps[p] = gg[p] * S.htdv[p];
        for (size_t iz = 1; iz < nz; iz++) {
            const size_t o = n2 * iz;
            for (size_t p = 0; p < n2; p++) ps[o + p] = (gg[o + p] - S.ap[p] * ps[o - n2 + p]) * S.htdv[o + p];
        }
        for (size_t iz = nz - 1; iz-- > 0;) {
            const size_t o = n2 * iz;
            for (size_t p = 0; p < n2; p++) ps[o + p] += S.etdv[o + p] * ps[o + n2 + p];
        }
        for (size_t p = 0; p < n2; p++) ps[n2 * nz + p] = 0.0;
        S.spctop3d(ps, pn, 0, nz - 1);
        for (size_t p = 0; p < n2; p++) pn[n2 * nz + p] = 0.0;

        // ndarray iter() visits the (ng,ng,nz+1) index space with the last index fastest
        double num = 0.0, den = 0.0;
        for (size_t i = 0; i < ng; i++)
            for (size_t j = 0; j < ng; j++)
                for (size_t iz = 0; iz <= nz; iz++) {
                    size_t p = i + ng * j + n2 * iz;
                    num += std::pow(pn[p] - pna[p], 2.0);
                }
        for (size_t i = 0; i < ng; i++)
            for (size_t j = 0; j < ng; j++)
                for (size_t iz = 0; iz <= nz; iz++) den += std::pow(pna[i + ng * j + n2 * iz], 2.0);
        errp = std::sqrt(num / (den + 1.0E-20));
        if (iter > 0 && errp > 1.0) {
            if (iters) *iters = iter + 1;
            if (errp_out) *errp_out = errp;
            return -3;
        }
        iter += 1;
        std::memcpy(pna.data(), pn, sizeof(double) * n2 * L);
    }
    if (iters) *iters = iter;
    if (errp_out) *errp_out = errp;
    if (iter >= nitmax) return -2;

    for (size_t iz = 1; iz < nz; iz++) {
        const size_t o = n2 * iz;
        for (size_t p = 0; p < n2; p++) gg[o + p] = (ps[o + n2 + p] - ps[o - n2 + p]) * hdzi;
    }
    {
        const size_t o = n2 * nz;
        for (size_t p = 0; p < n2; p++) gg[o + p] = dz6 * sp[o + p] - ps[o - n2 + p] * dzi;
    }
    for (size_t p = 0; p < n2; p++) gg[n2 + p] *= S.htd1[0];
    for (size_t iz = 2; iz < nz; iz++) {
        const size_t o = n2 * iz;
        for (size_t p = 0; p < n2; p++) gg[o + p] = (gg[o + p] - (1.0 / 6.0) * gg[o - n2 + p]) * S.htd1[iz - 1];
    }
    {
        const size_t o = n2 * nz;
        for (size_t p = 0; p < n2; p++) gg[o + p] = (gg[o + p] - (1.0 / 3.0) * gg[o - n2 + p]) * S.htd1[nz - 1];
    }
    for (size_t iz = nz - 1; iz >= 1; iz--) {
        const size_t o = n2 * iz;
        for (size_t p = 0; p < n2; p++) gg[o + p] += S.etd1[iz - 1] * gg[o + n2 + p];
    }
    S.spctop3d(gg.data(), dpn, 1, nz);
    return 0;
}

void source(const State& st, double* sqs, double* sds, double* sgs) {
    const Spectral& S = *st.sp;
    const size_t ng = S.ng, nz = S.nz, n2 = ng * ng;
    const double *aa = st.f[AA], *qs = st.f[QS], *ds = st.f[DS], *u = st.f[U], *v = st.f[V];
    const double *ri = st.f[RI], *dpn = st.f[DPN], *z = st.f[Z], *zx = st.f[ZX], *zy = st.f[ZY];
    const double* ps = st.f[PS];
    Vec wkd(n2, 0.0);
    for (size_t iz = 0; iz <= nz; iz++)
        for (size_t p = 0; p < n2; p++) wkd[p] += S.weight[iz] * aa[n2 * iz + p];
    for (size_t p = 0; p < n2; p++) wkd[p] *= S.c2g2[p];

#pragma omp parallel for schedule(dynamic)
    for (size_t iz = 0; iz <= nz; iz++) {
        const size_t o = n2 * iz;
        Vec dd(n2), ff(n2), wkp(n2), wkq(n2), wka(n2), wkb(n2), wkc(n2);
        for (size_t p = 0; p < n2; p++) wka[p] = qs[o + p];
        S.d2fft.spctop(wka.data(), wkq.data());
        for (size_t p = 0; p < n2; p++) wkp[p] = wkq[p] * u[o + p];
        for (size_t p = 0; p < n2; p++) wkq[p] *= v[o + p];
        S.divs(wkp.data(), wkq.data(), wka.data());
        for (size_t p = 0; p < n2; p++) ff[p] = ri[o + p] * dpn[o + p];
        S.deal2d(ff.data());
        for (size_t p = 0; p < n2; p++) wkq[p] = z[o + p];
        S.jacob(ff.data(), wkq.data(), wkb.data());
        for (size_t p = 0; p < n2; p++) sqs[o + p] = S.filt[p] * (wkb[p] - wka[p]);
        S.jacob(u + o, v + o, wkc.data());
        for (size_t p = 0; p < n2; p++) wka[p] = ds[o + p];
        S.d2fft.spctop(wka.data(), dd.data());
        for (size_t p = 0; p < n2; p++) wkp[p] = ff[p] * zx[o + p] - dd[p] * u[o + p];
        for (size_t p = 0; p < n2; p++) wkq[p] = ff[p] * zy[o + p] - dd[p] * v[o + p];
        S.divs(wkp.data(), wkq.data(), wkb.data());
        for (size_t p = 0; p < n2; p++)
            sds[o + p] = S.filt[p] * (2.0 * wkc[p] + wkb[p] - S.hlap[p] * ps[o + p]);
        for (size_t p = 0; p < n2; p++) sgs[o + p] = COF * sqs[o + p] + wkd[p] - FSQ * aa[o + p];
    }
}

// out = {t, 0.5*zrms^2, zrms, zmax, umax} (the monitor.asc line)
void diagnose(const State& st, double out[5]) {
    const Spectral& S = *st.sp;
    const size_t ng = S.ng, nz = S.nz, n2 = ng * ng;
    const double *u = st.f[U], *v = st.f[V], *zeta = st.f[ZETA];
    double umax = NAN, zmax = NAN;
    for (size_t p = 0; p < n2 * (nz + 1); p++) {
        double e = std::pow(u[p], 2.0) + std::pow(v[p], 2.0);
        umax = std::fmax(umax, e);
        zmax = std::fmax(zmax, std::fabs(zeta[p]));
    }
    umax = std::sqrt(umax);
    double vsumi = 1.0 / (double)(ng * ng * nz);
    double sum = 0.0;
    for (size_t i = 0; i < ng; i++)
        for (size_t j = 0; j < ng; j++) {
            size_t p = i + ng * j;
            sum += (1.0 / 2.0) * std::pow(zeta[p], 2.0);
            sum += (1.0 / 2.0) * std::pow(zeta[p + n2 * nz], 2.0);
            for (size_t k = 1; k < nz; k++) sum += std::pow(zeta[p + n2 * k], 2.0);
        }
    double zrms = std::sqrt(vsumi * sum);
    out[0] = st.t;
    out[1] = (1.0 / 2.0) * std::pow(zrms, 2.0);
    out[2] = zrms;
    out[3] = zmax;
    out[4] = umax;
}

int advance(State& st, int nitmax, double diag[5], int iters[3]) {
    const Spectral& S = *st.sp;
    const size_t ng = S.ng, nz = S.nz, n2 = ng * ng, L = nz + 1, N = n2 * L;
    double dt = 1.0 / (double)ng;
    double dt4 = dt / 4.0;
    double dt4i = 1.0 / dt4;
    const int niter = 2;
    double *qs = st.f[QS], *ds = st.f[DS], *gs = st.f[GS];
    Vec qsi(N), qsm(N), sqs(N), sds(N), nds(N), sgs(N), ngs(N), wka(n2), wkb(n2);
    int rc = 0, it = 0;
    double errp = 0.0;
    for (int i = 0; i < 3; i++) iters[i] = 0;

    if (st.ggen) {
        S.main_invert(qs, ds, gs, st.f[R], st.f[U], st.f[V], st.f[ZETA]);
        if ((rc = psolve(st, nitmax, &it, &errp))) return rc;
    }
    diagnose(st, diag);
    source(st, sqs.data(), sds.data(), sgs.data());

    qsi.assign(qs, qs + N);
    for (size_t p = 0; p < N; p++) qsm[p] = qs[p] + dt4 * sqs[p];
    for (size_t iz = 0; iz < L; iz++)
        for (size_t p = 0; p < n2; p++) {
            size_t q = p + n2 * iz;
            qs[q] = S.diss[p] * (qsm[q] + dt4 * sqs[q]) - qsi[q];
        }
    Vec dsi(ds, ds + N), gsi(gs, gs + N);
    for (size_t p = 0; p < N; p++) nds[p] = sds[p] + dt4i * dsi[p];
    for (size_t p = 0; p < N; p++) ngs[p] = sgs[p] + dt4i * gsi[p];

    for (int pass = 0; pass <= niter; pass++) {
        if (pass > 0) {
            S.main_invert(qs, ds, gs, st.f[R], st.f[U], st.f[V], st.f[ZETA]);
            if ((rc = psolve(st, nitmax, &it, &errp))) return rc;
            iters[pass] = it;
            source(st, sqs.data(), sds.data(), sgs.data());
            for (size_t iz = 0; iz < L; iz++)
                for (size_t p = 0; p < n2; p++) {
                    size_t q = p + n2 * iz;
                    qs[q] = S.diss[p] * (qsm[q] + dt4 * sqs[q]) - qsi[q];
                }
        } else {
            iters[0] = st.ggen ? it : 0;
        }
        for (size_t p = 0; p < N; p++) sds[p] += nds[p];
        for (size_t p = 0; p < N; p++) sgs[p] += ngs[p];
        std::fill(wka.begin(), wka.end(), 0.0);
        std::fill(wkb.begin(), wkb.end(), 0.0);
        for (size_t iz = 0; iz < L; iz++)
            for (size_t p = 0; p < n2; p++) {
                size_t q = p + n2 * iz;
                ds[q] = sgs[q] + S.rdis[p] * sds[q];
            }
        for (size_t iz = 0; iz < L; iz++) {
            for (size_t p = 0; p < n2; p++) wka[p] += S.weight[iz] * ds[p + n2 * iz];
            for (size_t p = 0; p < n2; p++) wkb[p] += S.weight[iz] * sds[p + n2 * iz];
        }
        for (size_t p = 0; p < n2; p++) wka[p] *= S.fope[p];
        for (size_t p = 0; p < n2; p++) wkb[p] *= S.c2g2[p];
        for (size_t iz = 0; iz < L; iz++)
            for (size_t p = 0; p < n2; p++) {
                size_t q = p + n2 * iz;
                ds[q] = S.simp[p] * (ds[q] - wka[p]) - dsi[q];
                gs[q] = wkb[p] - FSQ * sds[q] + S.rdis[p] * sgs[q];
            }
        std::fill(wka.begin(), wka.end(), 0.0);
        for (size_t iz = 0; iz < L; iz++)
            for (size_t p = 0; p < n2; p++) wka[p] += S.weight[iz] * gs[p + n2 * iz];
        for (size_t p = 0; p < n2; p++) wka[p] *= S.fope[p];
        for (size_t iz = 0; iz < L; iz++)
            for (size_t p = 0; p < n2; p++) {
                size_t q = p + n2 * iz;
                gs[q] = S.simp[p] * (gs[q] - wka[p]) - gsi[q];
            }
    }
    st.t += dt;
    return 0;
}

// The energies savegrid writes to ecomp.asc (src/nhswps/mod.rs:365-420): out = {t, 0, ekin, ekin, epot,
// etot}.  ekin = 1/2 sum (1+r)(u^2+v^2[+w^2]) with half weight on the bottom (no w) and top levels,
// times gvol/(2 HBAR); epot = 1/2 gl^2 csq sum ((1/HBAR) z_top - 1)^2.
void energy(const State& st, double out[6]) {
    const Spectral& S = *st.sp;
    const size_t ng = S.ng, nz = S.nz, n2 = ng * ng;
    const double *r = st.f[R], *u = st.f[U], *v = st.f[V], *w = st.f[W], *z = st.f[Z];
    double sb = 0.0, st_ = 0.0;
    for (size_t p = 0; p < n2; p++) sb += (1.0 + r[p]) * (std::pow(u[p], 2.0) + std::pow(v[p], 2.0));
    double ekin = (1.0 / 2.0) * sb;
    const size_t ot = n2 * nz;
    for (size_t p = 0; p < n2; p++)
        st_ += (1.0 + r[ot + p]) * (std::pow(u[ot + p], 2.0) + std::pow(v[ot + p], 2.0) + std::pow(w[ot + p], 2.0));
    ekin += (1.0 / 2.0) * st_;
    for (size_t iz = 1; iz < nz; iz++) {
        double sl = 0.0;
        const size_t o = n2 * iz;
        for (size_t p = 0; p < n2; p++)
            sl += (1.0 + r[o + p]) * (std::pow(u[o + p], 2.0) + std::pow(v[o + p], 2.0) + std::pow(w[o + p], 2.0));
        ekin += sl;
    }
    const double gl = (2.0 * PI) / (double)ng, dz = HBAR / (double)nz, gvol = gl * gl * dz;
    ekin *= (1.0 / 2.0) * gvol * (1.0 / HBAR);
    double sp = 0.0;
    for (size_t p = 0; p < n2; p++) sp += std::pow((1.0 / HBAR) * z[ot + p] - 1.0, 2.0);
    const double epot = (1.0 / 2.0) * (gl * gl) * CSQ * sp;
    out[0] = st.t;
    out[1] = 0.0;
    out[2] = ekin;
    out[3] = ekin;
    out[4] = epot;
    out[5] = ekin + epot;
}

// The file output of savegrid (src/nhswps/mod.rs:438-652) as numbers: spec3 = the vertically averaged
// 1-d spectra of zeta, delta and gamma (three arrays of ng + 1, each level's spec1d weighted by the
// trapezoid weights, then times spmf: :438-482; spectra.asc prints k = 1..kmaxred of them as log10 with
// alk), d3 = the 3-D fields ql, d, g (spctop of qs, ds, gs), r, w, pn as f32 (:491-573), and d2 = the
// vertically integrated 2-D fields d, zeta, q, g, h as f32 (:575-652), in the order savegrid writes them.
void savegrid(const State& st, float* d3, float* d2, double* spec3) {
    const Spectral& S = *st.sp;
    const size_t ng = S.ng, nz = S.nz, n2 = ng * ng, L = nz + 1, N = n2 * L;
    const double *zeta = st.f[ZETA], *r = st.f[R], *z = st.f[Z], *w = st.f[W];
    Vec wkp(n2), wks(n2), wkq(n2), tmp(ng + 1, 0.0);
    double *zsp = spec3, *dsp = spec3 + (ng + 1), *gsp = spec3 + 2 * (ng + 1);
    for (size_t k = 0; k < 3 * (ng + 1); k++) spec3[k] = 0.0;
    for (size_t iz = 0; iz < L; iz++) {
        std::memcpy(wkp.data(), zeta + n2 * iz, n2 * sizeof(double));
        S.d2fft.ptospc(wkp.data(), wks.data());
        S.spec1d(wks.data(), tmp.data());
        for (size_t k = 0; k <= ng; k++) zsp[k] = zsp[k] + S.weight[iz] * tmp[k];
        S.spec1d(st.f[DS] + n2 * iz, tmp.data());
        for (size_t k = 0; k <= ng; k++) dsp[k] = dsp[k] + S.weight[iz] * tmp[k];
        S.spec1d(st.f[GS] + n2 * iz, tmp.data());
        for (size_t k = 0; k <= ng; k++) gsp[k] = gsp[k] + S.weight[iz] * tmp[k];
    }
    for (size_t k = 0; k <= ng; k++) {
        zsp[k] *= S.spmf[k];
        dsp[k] *= S.spmf[k];
        gsp[k] *= S.spmf[k];
    }
    // 3-D fields: ql, d, g in physical space, then r, w, pn
    const int spec_fields[3] = {QS, DS, GS};
    for (int q = 0; q < 3; q++)
        for (size_t iz = 0; iz < L; iz++) {
            std::memcpy(wks.data(), st.f[spec_fields[q]] + n2 * iz, n2 * sizeof(double));
            S.d2fft.spctop(wks.data(), wkp.data());
            for (size_t p = 0; p < n2; p++) d3[q * N + n2 * iz + p] = (float)wkp[p];
        }
    const int phys_fields[3] = {R, W, PN};
    for (int q = 0; q < 3; q++)
        for (size_t e = 0; e < N; e++) d3[(3 + q) * N + e] = (float)st.f[phys_fields[q]][e];
    // 2-D fields
    const double* zt = z + n2 * nz;
    const double* wt = w + n2 * nz;
    for (size_t p = 0; p < n2; p++) d2[p] = (float)(-wt[p] / zt[p]);
    for (size_t p = 0; p < n2; p++) wkp[p] = 0.0;
    for (size_t iz = 0; iz < L; iz++)
        for (size_t p = 0; p < n2; p++) wkp[p] += S.weight[iz] * zeta[n2 * iz + p] * (1.0 + r[n2 * iz + p]);
    for (size_t p = 0; p < n2; p++) wkp[p] *= HBAR / zt[p];
    for (size_t p = 0; p < n2; p++) d2[n2 + p] = (float)wkp[p];
    for (size_t p = 0; p < n2; p++) wkp[p] = HBAR * (wkp[p] + COF) / zt[p] - COF;
    for (size_t p = 0; p < n2; p++) d2[2 * n2 + p] = (float)wkp[p];
    for (size_t p = 0; p < n2; p++) wkp[p] = 0.0;
    for (size_t iz = 0; iz < L; iz++) {
        std::memcpy(wks.data(), st.f[GS] + n2 * iz, n2 * sizeof(double));
        S.d2fft.spctop(wks.data(), wkq.data());
        for (size_t p = 0; p < n2; p++) wkp[p] += S.weight[iz] * wkq[p] * (1.0 + r[n2 * iz + p]);
    }
    for (size_t p = 0; p < n2; p++) wkp[p] *= HBAR / zt[p];
    for (size_t p = 0; p < n2; p++) d2[3 * n2 + p] = (float)wkp[p];
    for (size_t p = 0; p < n2; p++) d2[4 * n2 + p] = (float)((1.0 / HBAR) * zt[p] - 1.0);
}

// The time loop of nhswps (src/nhswps/mod.rs:285-345) from a state whose prognostic fields qs, ds, gs
// are set: an inversion, then while t <= tsim: at save steps (itime = round(t/dt) a multiple of
// ngsave = round(tgsave/dt)) main_invert, psolve and the savegrid energies, ggen = false, otherwise
// ggen = true; then advance (one monitor row).  After the loop the save test is repeated with
// itime = floor(t/dt).  Rows beyond the given capacities are counted but not stored.
int run(State& st, double tsim, double tgsave, int nitmax, double* mon, int maxmon, int* nmon, double* eco, int maxeco,
        int* neco, int* iters_total) {
    const Spectral& S = *st.sp;
    const double dt = 1.0 / (double)S.ng;
    double *qs = st.f[QS], *ds = st.f[DS], *gs = st.f[GS];
    int rc, it = 0;
    double errp = 0.0;
    *nmon = *neco = *iters_total = 0;
    S.main_invert(qs, ds, gs, st.f[R], st.f[U], st.f[V], st.f[ZETA]);
    const size_t ngsave = (size_t)std::round(tgsave / dt);
    auto save = [&]() -> int {
        S.main_invert(qs, ds, gs, st.f[R], st.f[U], st.f[V], st.f[ZETA]);
        if ((rc = psolve(st, nitmax, &it, &errp))) return rc;
        *iters_total += it;
        double e[6];
        energy(st, e);
        if (*neco < maxeco) std::memcpy(eco + 6 * (size_t)*neco, e, sizeof(e));
        (*neco)++;
        return 0;
    };
    while (st.t <= tsim) {
        const size_t itime = (size_t)std::round(st.t / dt), jtime = itime / ngsave;
        if (ngsave * jtime == itime) {
            if ((rc = save())) return rc;
            st.ggen = 0;
        } else {
            st.ggen = 1;
        }
        double d[5];
        int its[3];
        if ((rc = advance(st, nitmax, d, its))) return rc;
        *iters_total += its[0] + its[1] + its[2];
        if (*nmon < maxmon) std::memcpy(mon + 5 * (size_t)*nmon, d, sizeof(d));
        (*nmon)++;
    }
    const size_t itime = (size_t)(st.t / dt), jtime = itime / ngsave;
    if (ngsave * jtime == itime)
        if ((rc = save())) return rc;
    return 0;
}

}  // namespace orc
