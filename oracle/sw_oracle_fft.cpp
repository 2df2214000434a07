// CPU restatement of the reference's 1-D and 2-D real FFTs (test infrastructure only: the parity
// checker and the CPU baseline; never linked into the GPU library).
//
// Follows src/stafft/mod.rs (initfft :67-95, factorisen :97-150, forfft :158-293, revfft :301-452),
// src/stafft/forward.rs (forrdx6 :9, forrdx5 :245, forrdx4 :438, forrdx3 :582, forrdx2 :689),
// src/stafft/reverse.rs (revrdx6 :9, revrdx5 :237, revrdx4 :424, revrdx3 :562, revrdx2 :666)
// and src/sta2dfft.rs (D2FFT::new :37-85, ptospc :91-107, spctop :113-123, xderiv :128-152,
// yderiv :157-184).  Stage order, butterfly arithmetic and operand order are kept identical so
// that the reference's exact (assert_eq!) fixtures are reproduced bit for bit; build with
// -ffp-contract=off.
//
// Two places where the reference's arithmetic is not a plain DFT are reproduced on purpose and
// marked "reference slip": both only occur for shapes the solver never uses.
#include "sw_oracle_internal.h"

#include <cmath>
#include <vector>

namespace orc {

const double SINFPI3 = 0.8660254037844386;
const double RTF516 = 0.5590169943749475;
const double SINF2PI5 = 0.9510565162951535;
const double SINRAT = 0.6180339887498949;
const double FRAC_1_SQRT_2 = 0.70710678118654752440;

int factorisen(size_t n, size_t f[5]) {
    size_t rem = n;
    for (int i = 0; i < 5; i++) f[i] = 0;
    const size_t fac[5] = {6, 4, 2, 3, 5};
    if (n < 2) return -1;
    for (int i = 0; i < 5; i++) {
        while (rem % fac[i] == 0) {
            f[i]++;
            rem /= fac[i];
            if (rem == 1) return 0;
        }
    }
    return -1;  // reference: "Factorization failed", exit(1)
}

int initfft(size_t n, size_t f[5], double* trig) {
    const size_t fac[5] = {6, 4, 2, 3, 5};
    if (factorisen(n, f) != 0) return -1;
    double ftwopin = 2.0 * PI / (double)n;
    size_t rem = n, m = 1;
    for (int i = 0; i < 5; i++) {
        for (size_t e = 1; e <= f[i]; e++) {
            rem /= fac[i];
            for (size_t k = 1; k < fac[i]; k++)
                for (size_t l = 0; l < rem; l++) {
                    trig[m - 1] = ftwopin * (double)(k * l);
                    m++;
                }
            ftwopin *= (double)fac[i];
        }
    }
    for (size_t i = 1; i < n; i++) {
        trig[i + n - 1] = -std::sin(trig[i - 1]);
        trig[i - 1] = std::cos(trig[i - 1]);
    }
    return 0;
}

// ---------------------------------------------------------------- forward radix kernels
// a(i,l,k) = a[i + nv*(l + R*k)], b(i,k,l) = b[i + nv*(k + lv*l)], trig views (k,l) = t[k + lv*l]
#define FA(i, l, k) a[(i) + nv * ((l) + R * (k))]
#define FB(i, k, l) b[(i) + nv * ((k) + lv * (l))]
#define TC(k, l) cs[(k) + lv * (l)]
#define TS(k, l) sn[(k) + lv * (l)]

static void forrdx6(const double* a, double* b, size_t nv, size_t lv, const double* cs, const double* sn) {
    const size_t R = 6;
    for (size_t i = 0; i < nv; i++) {
        double t1r = FA(i, 2, 0) + FA(i, 4, 0);
        double t2r = FA(i, 0, 0) - 0.5 * t1r;
        double t3r = SINFPI3 * (FA(i, 4, 0) - FA(i, 2, 0));
        double u0r = FA(i, 0, 0) + t1r;
        double t1i = FA(i, 5, 0) + FA(i, 1, 0);
        double t2i = FA(i, 3, 0) - 0.5 * t1i;
        double t3i = SINFPI3 * (FA(i, 5, 0) - FA(i, 1, 0));
        double v0r = FA(i, 3, 0) + t1i;
        FB(i, 0, 0) = u0r + v0r;
        FB(i, 0, 1) = t2r - t2i;
        FB(i, 0, 2) = t2r + t2i;
        FB(i, 0, 3) = u0r - v0r;
        FB(i, 0, 4) = t3i - t3r;
        FB(i, 0, 5) = t3r + t3i;
    }
    const bool small = nv <= (lv - 1) / 2;
    for (size_t k = 1; k <= (lv - 1) / 2; k++) {
        size_t kc = lv - k;
        double c1k = TC(k, 0), s1k = TS(k, 0), c2k = TC(k, 1), s2k = TS(k, 1);
        double c3k = TC(k, 2), s3k = TS(k, 2), c4k = TC(k, 3), s4k = TS(k, 3);
        // reference slip (forward.rs:103-106): the small-batch branch reads cosine column 3
        double c5k = small ? TC(k, 3) : TC(k, 4), s5k = TS(k, 4);
        for (size_t i = 0; i < nv; i++) {
            double x1p = c1k * FA(i, 1, k) - s1k * FA(i, 1, kc);
            double y1p = c1k * FA(i, 1, kc) + s1k * FA(i, 1, k);
            double x2p = c2k * FA(i, 2, k) - s2k * FA(i, 2, kc);
            double y2p = c2k * FA(i, 2, kc) + s2k * FA(i, 2, k);
            double x3p = c3k * FA(i, 3, k) - s3k * FA(i, 3, kc);
            double y3p = c3k * FA(i, 3, kc) + s3k * FA(i, 3, k);
            double x4p = c4k * FA(i, 4, k) - s4k * FA(i, 4, kc);
            double y4p = c4k * FA(i, 4, kc) + s4k * FA(i, 4, k);
            double x5p = c5k * FA(i, 5, k) - s5k * FA(i, 5, kc);
            double y5p = c5k * FA(i, 5, kc) + s5k * FA(i, 5, k);
            double t1r = x2p + x4p;
            double t1i = y2p + y4p;
            double t2r = FA(i, 0, k) - 0.5 * t1r;
            double t2i = FA(i, 0, kc) - 0.5 * t1i;
            double t3r = SINFPI3 * (x2p - x4p);
            double t3i = SINFPI3 * (y2p - y4p);
            double u0r = FA(i, 0, k) + t1r;
            double u0i = FA(i, 0, kc) + t1i;
            double u1r = t2r + t3i;
            double u1i = t2i - t3r;
            double u2r = t2r - t3i;
            double u2i = t2i + t3r;
            t1r = x5p + x1p;
            t1i = y5p + y1p;
            t2r = x3p - 0.5 * t1r;
            t2i = y3p - 0.5 * t1i;
            t3r = SINFPI3 * (x5p - x1p);
            t3i = SINFPI3 * (y5p - y1p);
            double v0r = x3p + t1r;
            double v0i = y3p + t1i;
            double v1r = t2r + t3i;
            double v1i = t3r - t2i;
            double v2r = t2r - t3i;
            double v2i = t2i + t3r;
            FB(i, k, 0) = u0r + v0r;
            FB(i, kc, 0) = u2r - v2r;
            FB(i, k, 1) = u1r - v1r;
            FB(i, kc, 1) = u1r + v1r;
            FB(i, k, 2) = u2r + v2r;
            FB(i, kc, 2) = u0r - v0r;
            FB(i, k, 3) = v0i - u0i;
            FB(i, kc, 3) = u2i + v2i;
            FB(i, k, 4) = v1i - u1i;
            FB(i, kc, 4) = u1i + v1i;
            FB(i, k, 5) = v2i - u2i;
            FB(i, kc, 5) = u0i + v0i;
        }
    }
    if (lv % 2 == 0) {
        size_t h = lv / 2;
        for (size_t i = 0; i < nv; i++) {
            double q1 = FA(i, 2, h) - FA(i, 4, h);
            double q2 = FA(i, 0, h) + 0.5 * q1;
            double q3 = SINFPI3 * (FA(i, 2, h) + FA(i, 4, h));
            double q4 = FA(i, 1, h) + FA(i, 5, h);
            double q5 = -FA(i, 3, h) - 0.5 * q4;
            double q6 = SINFPI3 * (FA(i, 1, h) - FA(i, 5, h));
            FB(i, h, 0) = q2 + q6;
            FB(i, h, 1) = FA(i, 0, h) - q1;
            FB(i, h, 2) = q2 - q6;
            FB(i, h, 3) = q5 + q3;
            FB(i, h, 4) = FA(i, 3, h) - q4;
            FB(i, h, 5) = q5 - q3;
        }
    }
}

static void forrdx5(const double* a, double* b, size_t nv, size_t lv, const double* cs, const double* sn) {
    const size_t R = 5;
    for (size_t i = 0; i < nv; i++) {
        double t1r = FA(i, 1, 0) + FA(i, 4, 0);
        double t2r = FA(i, 2, 0) + FA(i, 3, 0);
        double t3r = SINF2PI5 * (FA(i, 4, 0) - FA(i, 1, 0));
        double t4r = SINF2PI5 * (FA(i, 2, 0) - FA(i, 3, 0));
        double t5r = t1r + t2r;
        double t6r = RTF516 * (t1r - t2r);
        double t7r = FA(i, 0, 0) - 0.25 * t5r;
        FB(i, 0, 0) = FA(i, 0, 0) + t5r;
        FB(i, 0, 1) = t7r + t6r;
        FB(i, 0, 2) = t7r - t6r;
        FB(i, 0, 3) = t4r + SINRAT * t3r;
        FB(i, 0, 4) = t3r - SINRAT * t4r;
    }
    for (size_t k = 1; k <= (lv - 1) / 2; k++) {
        size_t kc = lv - k;
        double c1k = TC(k, 0), s1k = TS(k, 0), c2k = TC(k, 1), s2k = TS(k, 1);
        double c3k = TC(k, 2), s3k = TS(k, 2), c4k = TC(k, 3), s4k = TS(k, 3);
        for (size_t i = 0; i < nv; i++) {
            double x1p = c1k * FA(i, 1, k) - s1k * FA(i, 1, kc);
            double y1p = c1k * FA(i, 1, kc) + s1k * FA(i, 1, k);
            double x2p = c2k * FA(i, 2, k) - s2k * FA(i, 2, kc);
            double y2p = c2k * FA(i, 2, kc) + s2k * FA(i, 2, k);
            double x3p = c3k * FA(i, 3, k) - s3k * FA(i, 3, kc);
            double y3p = c3k * FA(i, 3, kc) + s3k * FA(i, 3, k);
            double x4p = c4k * FA(i, 4, k) - s4k * FA(i, 4, kc);
            double y4p = c4k * FA(i, 4, kc) + s4k * FA(i, 4, k);
            double t1r = x1p + x4p;
            double t1i = y1p + y4p;
            double t2r = x2p + x3p;
            double t2i = y2p + y3p;
            double t3r = SINF2PI5 * (x1p - x4p);
            double t3i = SINF2PI5 * (y1p - y4p);
            double t4r = SINF2PI5 * (x2p - x3p);
            double t4i = SINF2PI5 * (y2p - y3p);
            double t5r = t1r + t2r;
            double t5i = t1i + t2i;
            double t6r = RTF516 * (t1r - t2r);
            double t6i = RTF516 * (t1i - t2i);
            double t7r = FA(i, 0, k) - 0.25 * t5r;
            double t7i = FA(i, 0, kc) - 0.25 * t5i;
            double t8r = t7r + t6r;
            double t8i = t7i + t6i;
            double t9r = t7r - t6r;
            double t9i = t7i - t6i;
            double t10r = t3r + SINRAT * t4r;
            double t10i = t3i + SINRAT * t4i;
            double t11r = t4r - SINRAT * t3r;
            double t11i = SINRAT * t3i - t4i;
            FB(i, k, 0) = FA(i, 0, k) + t5r;
            FB(i, kc, 0) = t8r - t10i;
            FB(i, k, 1) = t8r + t10i;
            FB(i, kc, 1) = t9r - t11i;
            FB(i, k, 2) = t9r + t11i;
            FB(i, kc, 2) = t9i + t11r;
            FB(i, k, 3) = t11r - t9i;
            FB(i, kc, 3) = t8i - t10r;
            FB(i, k, 4) = -t8i - t10r;
            FB(i, kc, 4) = FA(i, 0, kc) + t5i;
        }
    }
}

static void forrdx4(const double* a, double* b, size_t nv, size_t lv, const double* cs, const double* sn) {
    const size_t R = 4;
    for (size_t i = 0; i < nv; i++) {
        double t1r = FA(i, 0, 0) + FA(i, 2, 0);
        double t2r = FA(i, 1, 0) + FA(i, 3, 0);
        FB(i, 0, 0) = t1r + t2r;
        FB(i, 0, 1) = FA(i, 0, 0) - FA(i, 2, 0);
        FB(i, 0, 2) = t1r - t2r;
        FB(i, 0, 3) = FA(i, 3, 0) - FA(i, 1, 0);
    }
    for (size_t k = 1; k <= (lv - 1) / 2; k++) {
        size_t kc = lv - k;
        double c1k = TC(k, 0), s1k = TS(k, 0), c2k = TC(k, 1), s2k = TS(k, 1);
        double c3k = TC(k, 2), s3k = TS(k, 2);
        for (size_t i = 0; i < nv; i++) {
            double x1p = c1k * FA(i, 1, k) - s1k * FA(i, 1, kc);
            double y1p = c1k * FA(i, 1, kc) + s1k * FA(i, 1, k);
            double x2p = c2k * FA(i, 2, k) - s2k * FA(i, 2, kc);
            double y2p = c2k * FA(i, 2, kc) + s2k * FA(i, 2, k);
            double x3p = c3k * FA(i, 3, k) - s3k * FA(i, 3, kc);
            double y3p = c3k * FA(i, 3, kc) + s3k * FA(i, 3, k);
            double t1r = FA(i, 0, k) + x2p;
            double t1i = FA(i, 0, kc) + y2p;
            double t2r = x1p + x3p;
            double t2i = y1p + y3p;
            double t3r = FA(i, 0, k) - x2p;
            double t3i = FA(i, 0, kc) - y2p;
            double t4r = x3p - x1p;
            double t4i = y1p - y3p;
            FB(i, k, 0) = t1r + t2r;
            FB(i, kc, 0) = t3r - t4i;
            FB(i, k, 1) = t3r + t4i;
            FB(i, kc, 1) = t1r - t2r;
            FB(i, k, 2) = t2i - t1i;
            FB(i, kc, 2) = t3i + t4r;
            FB(i, k, 3) = t4r - t3i;
            FB(i, kc, 3) = t1i + t2i;
        }
    }
    if (lv % 2 == 0) {
        size_t h = lv / 2;
        for (size_t i = 0; i < nv; i++) {
            double q1 = FRAC_1_SQRT_2 * (FA(i, 1, h) - FA(i, 3, h));
            double q2 = FRAC_1_SQRT_2 * (FA(i, 1, h) + FA(i, 3, h));
            FB(i, h, 0) = FA(i, 0, h) + q1;
            FB(i, h, 1) = FA(i, 0, h) - q1;
            FB(i, h, 2) = FA(i, 2, h) - q2;
            FB(i, h, 3) = -FA(i, 2, h) - q2;
        }
    }
}

static void forrdx3(const double* a, double* b, size_t nv, size_t lv, const double* cs, const double* sn) {
    const size_t R = 3;
    for (size_t i = 0; i < nv; i++) {
        double t1r = FA(i, 1, 0) + FA(i, 2, 0);
        FB(i, 0, 0) = FA(i, 0, 0) + t1r;
        FB(i, 0, 1) = FA(i, 0, 0) - 0.5 * t1r;
        FB(i, 0, 2) = SINFPI3 * (FA(i, 2, 0) - FA(i, 1, 0));
    }
    for (size_t k = 1; k <= (lv - 1) / 2; k++) {
        size_t kc = lv - k;
        double c1k = TC(k, 0), s1k = TS(k, 0), c2k = TC(k, 1), s2k = TS(k, 1);
        for (size_t i = 0; i < nv; i++) {
            double x1p = c1k * FA(i, 1, k) - s1k * FA(i, 1, kc);
            double y1p = c1k * FA(i, 1, kc) + s1k * FA(i, 1, k);
            double x2p = c2k * FA(i, 2, k) - s2k * FA(i, 2, kc);
            double y2p = c2k * FA(i, 2, kc) + s2k * FA(i, 2, k);
            double t1r = x1p + x2p;
            double t1i = y1p + y2p;
            double t2r = FA(i, 0, k) - 0.5 * t1r;
            double t2i = 0.5 * t1i - FA(i, 0, kc);
            double t3r = SINFPI3 * (x2p - x1p);
            double t3i = SINFPI3 * (y1p - y2p);
            FB(i, k, 0) = FA(i, 0, k) + t1r;
            FB(i, kc, 0) = t2r - t3i;
            FB(i, k, 1) = t2r + t3i;
            FB(i, kc, 1) = t3r - t2i;
            FB(i, k, 2) = t2i + t3r;
            FB(i, kc, 2) = FA(i, 0, kc) + t1i;
        }
    }
}

static int forrdx2(const double* a, double* b, size_t nv, size_t lv, const double* cs, const double* sn) {
    const size_t R = 2;
    for (size_t i = 0; i < nv; i++) {
        FB(i, 0, 0) = FA(i, 0, 0) + FA(i, 1, 0);
        FB(i, 0, 1) = FA(i, 0, 0) - FA(i, 1, 0);
    }
    // reference slip (forward.rs:729-732,742-743): the twiddle is read at cosine[[0,k-1]], a
    // bounds-checked index that is valid only for k = 1 and panics for k >= 2.
    if ((lv - 1) / 2 >= 2) return -1;
    for (size_t k = 1; k <= (lv - 1) / 2; k++) {
        size_t kc = lv - k;
        double c1k = cs[0], s1k = sn[0];
        for (size_t i = 0; i < nv; i++) {
            double x1 = c1k * FA(i, 1, k) - s1k * FA(i, 1, kc);
            double y1 = c1k * FA(i, 1, kc) + s1k * FA(i, 1, k);
            FB(i, k, 0) = FA(i, 0, k) + x1;
            FB(i, kc, 0) = FA(i, 0, k) - x1;
            FB(i, k, 1) = y1 - FA(i, 0, kc);
            FB(i, kc, 1) = FA(i, 0, kc) + y1;
        }
    }
    return 0;
}
#undef FA
#undef FB

// ---------------------------------------------------------------- reverse radix kernels
// a(i,k,l) = a[i + nv*(k + lv*l)], b(i,l,k) = b[i + nv*(l + R*k)]
#define RA(i, k, l) a[(i) + nv * ((k) + lv * (l))]
#define RB(i, l, k) b[(i) + nv * ((l) + R * (k))]

static void revrdx6(const double* a, double* b, size_t nv, size_t lv, const double* cs, const double* sn) {
    const size_t R = 6;
    for (size_t i = 0; i < nv; i++) {
        double t2r = RA(i, 0, 0) - 0.5 * RA(i, 0, 2);
        double t3r = SINFPI3 * RA(i, 0, 4);
        double u0r = RA(i, 0, 0) + RA(i, 0, 2);
        double u1r = t2r + t3r;
        double u2r = t2r - t3r;
        double t2i = RA(i, 0, 3) - 0.5 * RA(i, 0, 1);
        double t3i = -SINFPI3 * RA(i, 0, 5);
        double v0r = RA(i, 0, 3) + RA(i, 0, 1);
        double v1r = t2i + t3i;
        double v2r = t2i - t3i;
        RB(i, 0, 0) = u0r + v0r;
        RB(i, 1, 0) = u1r - v1r;
        RB(i, 2, 0) = u2r + v2r;
        RB(i, 3, 0) = u0r - v0r;
        RB(i, 4, 0) = u1r + v1r;
        RB(i, 5, 0) = u2r - v2r;
    }
    const bool small = nv <= (lv - 1) / 2;
    for (size_t k = 1; k <= (lv - 1) / 2; k++) {
        size_t kc = lv - k;
        double c1k = TC(k, 0), s1k = TS(k, 0), c2k = TC(k, 1), s2k = TS(k, 1);
        double c3k = TC(k, 2), s3k = TS(k, 2);
        // reference slip (reverse.rs:143-148): the small-batch branch reads columns 4 and 5
        double c4k = small ? TC(k, 4) : TC(k, 3), s4k = small ? TS(k, 4) : TS(k, 3);
        double c5k = small ? TC(k, 5) : TC(k, 4), s5k = small ? TS(k, 5) : TS(k, 4);
        for (size_t i = 0; i < nv; i++) {
            double t1r = RA(i, k, 2) + RA(i, kc, 1);
            double t1i = RA(i, kc, 3) - RA(i, k, 4);
            double t2r = RA(i, k, 0) - 0.5 * t1r;
            double t2i = RA(i, kc, 5) - 0.5 * t1i;
            double t3r = SINFPI3 * (RA(i, k, 2) - RA(i, kc, 1));
            double t3i = SINFPI3 * (RA(i, kc, 3) + RA(i, k, 4));
            double u0r = RA(i, k, 0) + t1r;
            double u0i = RA(i, kc, 5) + t1i;
            double u1r = t2r + t3i;
            double u1i = t2i - t3r;
            double u2r = t2r - t3i;
            double u2i = t2i + t3r;
            t1r = RA(i, kc, 0) + RA(i, k, 1);
            t1i = RA(i, kc, 4) - RA(i, k, 5);
            t2r = RA(i, kc, 2) - 0.5 * t1r;
            t2i = -RA(i, k, 3) - 0.5 * t1i;
            t3r = SINFPI3 * (RA(i, kc, 0) - RA(i, k, 1));
            t3i = SINFPI3 * (-RA(i, k, 5) - RA(i, kc, 4));
            double v0r = RA(i, kc, 2) + t1r;
            double v0i = t1i - RA(i, k, 3);
            double v1r = t2r + t3i;
            double v1i = t2i - t3r;
            double v2r = t2r - t3i;
            double v2i = t2i + t3r;
            double x1p = u1r - v1r;
            double y1p = u1i - v1i;
            double x2p = u2r + v2r;
            double y2p = u2i + v2i;
            double x3p = u0r - v0r;
            double y3p = u0i - v0i;
            double x4p = u1r + v1r;
            double y4p = u1i + v1i;
            double x5p = u2r - v2r;
            double y5p = u2i - v2i;
            RB(i, 0, k) = u0r + v0r;
            RB(i, 0, kc) = u0i + v0i;
            RB(i, 1, k) = c1k * x1p - s1k * y1p;
            RB(i, 1, kc) = c1k * y1p + s1k * x1p;
            RB(i, 2, k) = c2k * x2p - s2k * y2p;
            RB(i, 2, kc) = c2k * y2p + s2k * x2p;
            RB(i, 3, k) = c3k * x3p - s3k * y3p;
            RB(i, 3, kc) = c3k * y3p + s3k * x3p;
            RB(i, 4, k) = c4k * x4p - s4k * y4p;
            RB(i, 4, kc) = c4k * y4p + s4k * x4p;
            RB(i, 5, k) = c5k * x5p - s5k * y5p;
            RB(i, 5, kc) = c5k * y5p + s5k * x5p;
        }
    }
    if (lv % 2 == 0) {
        size_t h = lv / 2;
        for (size_t i = 0; i < nv; i++) {
            double q1 = RA(i, h, 0) + RA(i, h, 2);
            double q2 = RA(i, h, 5) + RA(i, h, 3);
            double q3 = RA(i, h, 1) - 0.5 * q1;
            double q4 = RA(i, h, 4) + 0.5 * q2;
            double q5 = SINFPI3 * (RA(i, h, 0) - RA(i, h, 2));
            double q6 = SINFPI3 * (RA(i, h, 5) - RA(i, h, 3));
            RB(i, 0, h) = RA(i, h, 1) + q1;
            RB(i, 1, h) = q4 + q5;
            RB(i, 2, h) = q6 - q3;
            RB(i, 3, h) = q2 - RA(i, h, 4);
            RB(i, 4, h) = q3 + q6;
            RB(i, 5, h) = q4 - q5;
        }
    }
}

static void revrdx5(const double* a, double* b, size_t nv, size_t lv, const double* cs, const double* sn) {
    const size_t R = 5;
    for (size_t i = 0; i < nv; i++) {
        double t3r = SINF2PI5 * RA(i, 0, 4);
        double t4r = SINF2PI5 * RA(i, 0, 3);
        double t5r = RA(i, 0, 1) + RA(i, 0, 2);
        double t6r = RTF516 * (RA(i, 0, 1) - RA(i, 0, 2));
        double t7r = RA(i, 0, 0) - 0.25 * t5r;
        double t8r = t7r + t6r;
        double t9r = t7r - t6r;
        double t10r = t3r + SINRAT * t4r;
        double t11r = SINRAT * t3r - t4r;
        RB(i, 0, 0) = RA(i, 0, 0) + t5r;
        RB(i, 1, 0) = t8r + t10r;
        RB(i, 2, 0) = t9r + t11r;
        RB(i, 3, 0) = t9r - t11r;
        RB(i, 4, 0) = t8r - t10r;
    }
    for (size_t k = 1; k <= (lv - 1) / 2; k++) {
        size_t kc = lv - k;
        double c1k = TC(k, 0), s1k = TS(k, 0), c2k = TC(k, 1), s2k = TS(k, 1);
        double c3k = TC(k, 2), s3k = TS(k, 2);
        // reference slip (reverse.rs:361-362,374-375,391-392): the fourth twiddle is read at
        // column 4 of a 4-column view in both branches (only reached when 25 divides n).
        double c4k = TC(k, 4), s4k = TS(k, 4);
        for (size_t i = 0; i < nv; i++) {
            double t1r = RA(i, k, 1) + RA(i, kc, 0);
            double t1i = RA(i, kc, 3) - RA(i, k, 4);
            double t2r = RA(i, k, 2) + RA(i, kc, 1);
            double t2i = RA(i, kc, 2) - RA(i, k, 3);
            double t3r = SINF2PI5 * (RA(i, k, 1) - RA(i, kc, 0));
            double t3i = SINF2PI5 * (RA(i, kc, 3) + RA(i, k, 4));
            double t4r = SINF2PI5 * (RA(i, k, 2) - RA(i, kc, 1));
            double t4i = SINF2PI5 * (RA(i, kc, 2) + RA(i, k, 3));
            double t5r = t1r + t2r;
            double t5i = t1i + t2i;
            double t6r = RTF516 * (t1r - t2r);
            double t6i = RTF516 * (t1i - t2i);
            double t7r = RA(i, k, 0) - 0.25 * t5r;
            double t7i = RA(i, kc, 4) - 0.25 * t5i;
            double t8r = t7r + t6r;
            double t8i = t7i + t6i;
            double t9r = t7r - t6r;
            double t9i = t7i - t6i;
            double t10r = t3r + SINRAT * t4r;
            double t10i = t3i + SINRAT * t4i;
            double t11r = SINRAT * t3r - t4r;
            double t11i = SINRAT * t3i - t4i;
            double x1p = t8r + t10i;
            double y1p = t8i - t10r;
            double x2p = t9r + t11i;
            double y2p = t9i - t11r;
            double x3p = t9r - t11i;
            double y3p = t9i + t11r;
            double x4p = t8r - t10i;
            double y4p = t8i + t10r;
            RB(i, 0, k) = RA(i, k, 0) + t5r;
            RB(i, 0, kc) = RA(i, kc, 4) + t5i;
            RB(i, 1, k) = c1k * x1p - s1k * y1p;
            RB(i, 1, kc) = c1k * y1p + s1k * x1p;
            RB(i, 2, k) = c2k * x2p - s2k * y2p;
            RB(i, 2, kc) = c2k * y2p + s2k * x2p;
            RB(i, 3, k) = c3k * x3p - s3k * y3p;
            RB(i, 3, kc) = c3k * y3p + s3k * x3p;
            RB(i, 4, k) = c4k * x4p - s4k * y4p;
            RB(i, 4, kc) = c4k * y4p + s4k * x4p;
        }
    }
}

static void revrdx4(const double* a, double* b, size_t nv, size_t lv, const double* cs, const double* sn) {
    const size_t R = 4;
    for (size_t i = 0; i < nv; i++) {
        double t1r = RA(i, 0, 0) + RA(i, 0, 2);
        double t2r = RA(i, 0, 1);
        double t3r = RA(i, 0, 0) - RA(i, 0, 2);
        double t4r = RA(i, 0, 3);
        RB(i, 0, 0) = t1r + t2r;
        RB(i, 1, 0) = t3r + t4r;
        RB(i, 2, 0) = t1r - t2r;
        RB(i, 3, 0) = t3r - t4r;
    }
    for (size_t k = 1; k <= (lv - 1) / 2; k++) {
        size_t kc = lv - k;
        double c1k = TC(k, 0), s1k = TS(k, 0), c2k = TC(k, 1), s2k = TS(k, 1);
        double c3k = TC(k, 2), s3k = TS(k, 2);
        for (size_t i = 0; i < nv; i++) {
            double t1r = RA(i, k, 0) + RA(i, kc, 1);
            double t1i = RA(i, kc, 3) - RA(i, k, 2);
            double t2r = RA(i, k, 1) + RA(i, kc, 0);
            double t2i = RA(i, kc, 2) - RA(i, k, 3);
            double t3r = RA(i, k, 0) - RA(i, kc, 1);
            double t3i = RA(i, kc, 3) + RA(i, k, 2);
            double t4r = RA(i, k, 1) - RA(i, kc, 0);
            double t4i = RA(i, kc, 2) + RA(i, k, 3);
            double x1p = t3r + t4i;
            double y1p = t3i - t4r;
            double x2p = t1r - t2r;
            double y2p = t1i - t2i;
            double x3p = t3r - t4i;
            double y3p = t3i + t4r;
            RB(i, 0, k) = t1r + t2r;
            RB(i, 0, kc) = t1i + t2i;
            RB(i, 1, k) = c1k * x1p - s1k * y1p;
            RB(i, 1, kc) = c1k * y1p + s1k * x1p;
            RB(i, 2, k) = c2k * x2p - s2k * y2p;
            RB(i, 2, kc) = c2k * y2p + s2k * x2p;
            RB(i, 3, k) = c3k * x3p - s3k * y3p;
            RB(i, 3, kc) = c3k * y3p + s3k * x3p;
        }
    }
    if (lv % 2 == 0) {
        size_t h = lv / 2;
        for (size_t i = 0; i < nv; i++) {
            RB(i, 0, h) = RA(i, h, 0) + RA(i, h, 1);
            RB(i, 2, h) = RA(i, h, 3) - RA(i, h, 2);
            double t3r = RA(i, h, 0) - RA(i, h, 1);
            double t4r = RA(i, h, 3) + RA(i, h, 2);
            RB(i, 1, h) = FRAC_1_SQRT_2 * (t3r + t4r);
            RB(i, 3, h) = FRAC_1_SQRT_2 * (t4r - t3r);
        }
    }
}

static void revrdx3(const double* a, double* b, size_t nv, size_t lv, const double* cs, const double* sn) {
    const size_t R = 3;
    for (size_t i = 0; i < nv; i++) {
        double t1r = RA(i, 0, 1);
        double t2r = RA(i, 0, 0) - 0.5 * t1r;
        double t3r = SINFPI3 * RA(i, 0, 2);
        RB(i, 0, 0) = RA(i, 0, 0) + t1r;
        RB(i, 1, 0) = t2r + t3r;
        RB(i, 2, 0) = t2r - t3r;
    }
    for (size_t k = 1; k <= (lv - 1) / 2; k++) {
        size_t kc = lv - k;
        double c1k = TC(k, 0), s1k = TS(k, 0), c2k = TC(k, 1), s2k = TS(k, 1);
        for (size_t i = 0; i < nv; i++) {
            double t1r = RA(i, k, 1) + RA(i, kc, 0);
            double t1i = RA(i, kc, 1) - RA(i, k, 2);
            double t2r = RA(i, k, 0) - 0.5 * t1r;
            double t2i = RA(i, kc, 2) - 0.5 * t1i;
            double t3r = SINFPI3 * (RA(i, k, 1) - RA(i, kc, 0));
            double t3i = SINFPI3 * (RA(i, kc, 1) + RA(i, k, 2));
            double x1p = t2r + t3i;
            double y1p = t2i - t3r;
            double x2p = t2r - t3i;
            double y2p = t2i + t3r;
            RB(i, 0, k) = RA(i, k, 0) + t1r;
            RB(i, 0, kc) = RA(i, kc, 2) + t1i;
            RB(i, 1, k) = c1k * x1p - s1k * y1p;
            RB(i, 1, kc) = s1k * x1p + c1k * y1p;
            RB(i, 2, k) = c2k * x2p - s2k * y2p;
            RB(i, 2, kc) = s2k * x2p + c2k * y2p;
        }
    }
}

static void revrdx2(const double* a, double* b, size_t nv, size_t lv, const double* cs, const double* sn) {
    const size_t R = 2;
    for (size_t i = 0; i < nv; i++) {
        RB(i, 0, 0) = RA(i, 0, 0) + RA(i, 0, 1);
        RB(i, 1, 0) = RA(i, 0, 0) - RA(i, 0, 1);
    }
    for (size_t k = 1; k <= (lv - 1) / 2; k++) {
        size_t kc = lv - k;
        double c1k = TC(k, 0), s1k = TS(k, 0);
        for (size_t i = 0; i < nv; i++) {
            double x1p = RA(i, k, 0) - RA(i, kc, 0);
            double y1p = RA(i, kc, 1) + RA(i, k, 1);
            RB(i, 0, k) = RA(i, k, 0) + RA(i, kc, 0);
            RB(i, 0, kc) = RA(i, kc, 1) - RA(i, k, 1);
            RB(i, 1, k) = c1k * x1p - s1k * y1p;
            RB(i, 1, kc) = c1k * y1p + s1k * x1p;
        }
    }
}
#undef RA
#undef RB
#undef TC
#undef TS

// ---------------------------------------------------------------- batched transforms
int forfft(size_t m, size_t n, double* xs, const double* trig, const size_t f[5]) {
    std::vector<double> wk(m * n, 0.0);
    size_t rem = n, cum = 1;
    bool orig = true;
    // stage order 5, 3, 2, 4, 6 (mod.rs:173-281)
    const int order[5] = {4, 3, 2, 1, 0};
    const size_t fac[5] = {6, 4, 2, 3, 5};
    for (int oi = 0; oi < 5; oi++) {
        int fi = order[oi];
        size_t R = fac[fi];
        for (size_t e = 0; e < f[fi]; e++) {
            rem /= R;
            size_t iloc = (rem - 1) * R * cum;
            const double* cs = trig + iloc;
            const double* sn = trig + n + iloc;
            const double* a = orig ? xs : wk.data();
            double* b = orig ? wk.data() : xs;
            size_t nv = m * rem;
            switch (R) {
                case 5: forrdx5(a, b, nv, cum, cs, sn); break;
                case 3: forrdx3(a, b, nv, cum, cs, sn); break;
                case 2: if (forrdx2(a, b, nv, cum, cs, sn)) return -1; break;
                case 4: forrdx4(a, b, nv, cum, cs, sn); break;
                case 6: forrdx6(a, b, nv, cum, cs, sn); break;
            }
            orig = !orig;
            cum *= R;
        }
    }
    double normfac = 1.0 / std::sqrt((double)n);
    for (size_t i = 0; i < m * n; i++) xs[i] = orig ? xs[i] * normfac : wk[i] * normfac;
    return 0;
}

int revfft(size_t m, size_t n, double* xs, const double* trig, const size_t f[5]) {
    std::vector<double> wk(m * n, 0.0);
    size_t rem = n, cum = 1;
    bool orig = true;
    for (size_t i = (n / 2 + 1) * m; i < m * n; i++) xs[i] = -xs[i];
    for (size_t i = 0; i < m; i++) xs[i] *= 0.5;
    if (n % 2 == 0) {
        size_t k = m * n / 2;
        for (size_t i = 0; i < m; i++) xs[k + i] *= 0.5;
    }
    // stage order 6, 4, 2, 3, 5 (mod.rs:332-440)
    const size_t fac[5] = {6, 4, 2, 3, 5};
    for (int fi = 0; fi < 5; fi++) {
        size_t R = fac[fi];
        for (size_t e = 0; e < f[fi]; e++) {
            rem /= R;
            size_t iloc = (cum - 1) * R * rem;
            const double* cs = trig + iloc;
            const double* sn = trig + n + iloc;
            const double* a = orig ? xs : wk.data();
            double* b = orig ? wk.data() : xs;
            size_t nv = m * cum;
            switch (R) {
                case 6: revrdx6(a, b, nv, rem, cs, sn); break;
                case 4: revrdx4(a, b, nv, rem, cs, sn); break;
                case 2: revrdx2(a, b, nv, rem, cs, sn); break;
                case 3: revrdx3(a, b, nv, rem, cs, sn); break;
                case 5: revrdx5(a, b, nv, rem, cs, sn); break;
            }
            orig = !orig;
            cum *= R;
        }
    }
    double normfac = 2.0 / std::sqrt((double)n);
    for (size_t i = 0; i < m * n; i++) xs[i] = orig ? xs[i] * normfac : wk[i] * normfac;
    return 0;
}

// ---------------------------------------------------------------- 2-D transforms
int D2FFT::init(size_t nx_, size_t ny_, double lx, double ly, double* kx, double* ky) {
    nx = nx_;
    ny = ny_;
    xtrig.assign(2 * nx, 0.0);
    ytrig.assign(2 * ny, 0.0);
    if (initfft(nx, xf, xtrig.data()) || initfft(ny, yf, ytrig.data())) return -1;
    if (lx == 0.0) return -1;
    double sc = PI / lx;
    for (size_t k = 1; k <= nx; k++) kx[k - 1] = sc * (double)k;
    if (ly != 0.0) {
        sc = PI / ly;
        for (size_t k = 1; k <= ny; k++) ky[k - 1] = sc * (double)k;
    }
    return 0;
}

// rvar(ny,nx) -> svar(nx,ny); rvar is destroyed (sta2dfft.rs:91-107)
void D2FFT::ptospc(double* rvar, double* svar) const {
    forfft(ny, nx, rvar, xtrig.data(), xf);
    for (size_t j = 0; j < nx; j++)
        for (size_t i = 0; i < ny; i++) svar[j + nx * i] = rvar[i + ny * j];
    forfft(nx, ny, svar, ytrig.data(), yf);
}

// svar(nx,ny) -> rvar(ny,nx); svar is destroyed (sta2dfft.rs:113-123)
void D2FFT::spctop(double* svar, double* rvar) const {
    revfft(nx, ny, svar, ytrig.data(), yf);
    for (size_t i = 0; i < ny; i++)
        for (size_t j = 0; j < nx; j++) rvar[i + ny * j] = svar[j + nx * i];
    revfft(ny, nx, rvar, xtrig.data(), xf);
}

void D2FFT::xderiv(const double* rkx, const double* var, double* der) const {
    size_t nwx = nx / 2, nxp2 = nx + 2;
    for (size_t ky = 0; ky < ny; ky++) {
        der[0 + nx * ky] = 0.0;
        for (size_t kx = 2; kx <= nx - nwx; kx++) {
            size_t dkx = 2 * (kx - 1), kxc = nxp2 - kx;
            der[(kx - 1) + nx * ky] = -rkx[dkx - 1] * var[(kxc - 1) + nx * ky];
            der[(kxc - 1) + nx * ky] = rkx[dkx - 1] * var[(kx - 1) + nx * ky];
        }
    }
    if (nx % 2 == 0) {
        size_t kxc = nwx + 1;
        for (size_t ky = 0; ky < ny; ky++) der[(kxc - 1) + nx * ky] = 0.0;
    }
}

void D2FFT::yderiv(const double* rky, const double* var, double* der) const {
    size_t nwy = ny / 2, nyp2 = ny + 2;
    for (size_t kx = 0; kx < nx; kx++) der[kx] = 0.0;
    for (size_t ky = 2; ky <= ny - nwy; ky++) {
        size_t kyc = nyp2 - ky;
        double fac = rky[2 * (ky - 1) - 1];
        for (size_t kx = 0; kx < nx; kx++) {
            der[kx + nx * (ky - 1)] = -fac * var[kx + nx * (kyc - 1)];
            der[kx + nx * (kyc - 1)] = fac * var[kx + nx * (ky - 1)];
        }
    }
    if (ny % 2 == 0) {
        size_t kyc = nwy + 1;
        for (size_t kx = 0; kx < nx; kx++) der[kx + nx * (kyc - 1)] = 0.0;
    }
}

}  // namespace orc
