// Host-side setup shared by the library: the reference's factorisation rule, its trig table, the
// spectral operators of Spectral::new, and the GPU FFT plan (radix passes, rows per block).
#pragma once
#include <cmath>
#include <cstddef>
#include <vector>

namespace sw {

const double PI = 3.14159265358979323846;
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

// Greedy factorisation into 6, 4, 2, 3, 5 (src/stafft/mod.rs:97-150).  Returns false when a prime
// factor >= 7 remains (the reference exits).
inline bool factorise_ref(size_t n, size_t f[5]) {
    const size_t fac[5] = {6, 4, 2, 3, 5};
    for (int i = 0; i < 5; i++) f[i] = 0;
    if (n < 2) return false;
    size_t rem = n;
    for (int i = 0; i < 5; i++)
        while (rem % fac[i] == 0) {
            f[i]++;
            rem /= fac[i];
            if (rem == 1) return true;
        }
    return false;
}

// Lengths whose reference transforms are not a DFT: a radix-2 stage after a radix-5 stage panics
// (src/stafft/forward.rs:729-743), and two or more radix-5 stages read a wrong twiddle in revrdx5
// (src/stafft/reverse.rs:361-392).  The library rejects them rather than return different numbers.
inline bool reference_transform_is_dft(const size_t f[5]) { return !(f[2] > 0 && f[4] > 0) && f[4] <= 1; }

// The reference's trig table (src/stafft/mod.rs:67-95): used only for the sw_initfft export.
inline void initfft_ref(size_t n, const size_t f[5], double* trig) {
    const size_t fac[5] = {6, 4, 2, 3, 5};
    double ftwopin = 2.0 * PI / (double)n;
    size_t rem = n, m = 1;
    for (int i = 0; i < 5; i++)
        for (size_t e = 1; e <= f[i]; e++) {
            rem /= fac[i];
            for (size_t k = 1; k < fac[i]; k++)
                for (size_t l = 0; l < rem; l++) trig[(m++) - 1] = ftwopin * (double)(k * l);
            ftwopin *= (double)fac[i];
        }
    for (size_t i = 1; i < n; i++) {
        trig[i + n - 1] = -std::sin(trig[i - 1]);
        trig[i - 1] = std::cos(trig[i - 1]);
    }
}

// Radix passes of the GPU transform: 8s first, then 4, 2, 3, 5.
inline std::vector<int> gpu_radices(size_t n) {
    std::vector<int> r;
    while (n % 8 == 0) { r.push_back(8); n /= 8; }
    while (n % 4 == 0) { r.push_back(4); n /= 4; }
    while (n % 2 == 0) { r.push_back(2); n /= 2; }
    while (n % 3 == 0) { r.push_back(3); n /= 3; }
    while (n % 5 == 0) { r.push_back(5); n /= 5; }
    if (n != 1) r.clear();
    return r;
}

// Spectral operators of Spectral::new (src/spectral/mod.rs:65-299), computed on the host in FP64
// with the reference's formulas and operation order.
struct HostOps {
    size_t ng = 0, nz = 0;
    std::vector<double> hlap, glap, rlap, helm, c2g2, simp, rope, fope, filt, diss, opak, rdis, ap;
    std::vector<double> etdv, htdv;  // (kx,ky,iz) at kx + ng*ky + ng*ng*iz
    std::vector<double> etd1, htd1, theta, weight, hrkx, hrky, rk, spmf, alk, kmag;
    size_t kmax = 0, kmaxred = 0;
    size_t xf[5] = {0, 0, 0, 0, 0};
    std::vector<double> xtrig;

    bool build(size_t ng_, size_t nz_) {
        ng = ng_;
        nz = nz_;
        const size_t n2 = ng * ng;
        if (!factorise_ref(ng, xf)) return false;
        xtrig.assign(2 * ng, 0.0);
        initfft_ref(ng, xf, xtrig.data());
        double dt = 1.0 / (double)ng, dt2 = dt * (1.0 / 2.0), dt2i = 1.0 / dt2;
        double dz = HBAR / (double)nz, dzi = 1.0 / dz, dzisq = std::pow(dzi, 2.0);
        for (auto* a : {&hlap, &glap, &rlap, &helm, &c2g2, &simp, &rope, &fope, &filt, &diss, &opak, &rdis, &ap})
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
        double sc = PI / (2.0 * PI);  // D2FFT::new with lx = ly = 2 pi (sta2dfft.rs:66-83)
        for (size_t k = 1; k <= ng; k++) hrkx[k - 1] = hrky[k - 1] = sc * (double)k;
        rk[0] = 0.0;
        for (size_t k = 1; k < ng / 2; k++) rk[k] = rk[ng - k] = hrkx[2 * k - 1];
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
        weight[0] = weight[nz] = 0.5;
        for (auto& e : weight) e /= (double)nz;
        std::vector<double> a0(n2), a0b(n2), apb(n2);
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
        return true;
    }
};

}  // namespace sw
