// sd_host.h - host-side material table and scalar physics of the engine.
// Parameters and formulas restate the reference (semiconductor_3d.c):
// mat_init :423-629, mat_prepare :631-642, mob_low :646-663,
// bgn_half :839-843, B :670-675, recomb :686-697.
#pragma once
#include "sd_common.h"
#include <cmath>
#include <cstring>
#include <cstdio>

struct HostMat {
    int id, insulator;
    double T, VT, Eg, chi, eps_r, Nc, Nv, ni;
    double mu_max[2], mu_min[2], Nref[2], alpha[2];
    double mt_min, mt_max[2], mt_Nref, mt_alpha;
    int mas_on;
    double mas_min1[2], mas_min2[2], mas_mu1[2], mas_Pc[2], mas_Cr[2], mas_Cs[2], mas_a[2], mas_b[2];
    double vsat_n, vsat_p, beta_n, beta_p;
    double tau_n0, tau_p0, Cn, Cp, Brad, tau_Nref;
    double bgn_E, bgn_N, bgn_C;
    double ani_n, ani_p;
    int ii_on;
    // Lombardi surface mobility (Sentaurus form, cm units as in the reference's
    // mat_init :462-465): lb_on bit 0 electrons, bit 1 holes; [0] n, [1] p
    int lb_on;
    double lbB[2], lbC[2], lbL[2], lbD[2], lbE[2];
    // MLDA confinement masses (units of m0) and valley weights (:146, :457, :619-623)
    double mq_e1, mq_e2, wq_e1, mq_h1, mq_h2, wq_h1;
    // gate tunnelling (:147-151): supply masses of a semiconductor, tunnelling
    // masses of an insulator (units of m0; 0 = none)
    double ms_e, ms_h, mt_e, mt_h;
    double ii_a[2][2], ii_b[2][2], ii_E0[2], ii_m[2], ii_hw, ii_g;
    double lnc, lnv;
};

// Material parameters at temperature T (SI units).  Returns false for an id
// outside the table (the caller falls back to silicon, as the reference).
inline bool host_mat_init(HostMat *m, int id, double T) {
    memset(m, 0, sizeof(*m));
    m->id = id; m->T = T; m->VT = SD_KB * T / SD_Q;
    const double tr = T / 300.;
    m->mt_min = -0.57; m->mt_max[0] = -2.33; m->mt_max[1] = -2.23; m->mt_Nref = 2.4; m->mt_alpha = -0.146;
    auto mob = [&](double mxn, double mxp, double mnn, double mnp, double rn, double rp, double an, double ap) {
        m->mu_max[0] = mxn; m->mu_max[1] = mxp; m->mu_min[0] = mnn; m->mu_min[1] = mnp;
        m->Nref[0] = rn; m->Nref[1] = rp; m->alpha[0] = an; m->alpha[1] = ap;
    };
    auto texp = [&](double mn, double mxn, double mxp, double nr, double al) {
        m->mt_min = mn; m->mt_max[0] = mxn; m->mt_max[1] = mxp; m->mt_Nref = nr; m->mt_alpha = al;
    };
    auto rec = [&](double tn, double tp, double cn, double cp, double b) {
        m->tau_n0 = tn; m->tau_p0 = tp; m->Cn = cn; m->Cp = cp; m->Brad = b;
    };
    auto iiset = [&](double an, double bn, double ap, double bp, double mn, double mp) {
        m->ii_on = 1; m->ii_E0[0] = m->ii_E0[1] = 1e30; m->ii_m[0] = mn; m->ii_m[1] = mp;
        m->ii_a[0][0] = m->ii_a[0][1] = an; m->ii_b[0][0] = m->ii_b[0][1] = bn;
        m->ii_a[1][0] = m->ii_a[1][1] = ap; m->ii_b[1][0] = m->ii_b[1][1] = bp;
    };
    switch (id) {
    case 0:  // Si: Masetti mobility, Canali saturation, Klaassen BGN, van Overstraeten impact ionisation
        m->Eg = 1.170 - 4.73e-4 * T * T / (T + 636.); m->chi = 4.05; m->eps_r = 11.7;
        m->Nc = 2.86e25 * pow(tr, 1.58); m->Nv = 3.10e25 * pow(tr, 1.85);
        m->mas_on = 1; m->mt_max[0] = -2.5; m->mt_max[1] = -2.2;
        mob(0.1417, 0.04705, 0.00522, 0.00449, 9.68e22, 2.23e23, 0.680, 0.719);
        m->mas_min1[0] = 0.00522; m->mas_min2[0] = 0.00522; m->mas_mu1[0] = 0.00434; m->mas_Pc[0] = 0.;
        m->mas_Cr[0] = 9.68e22; m->mas_Cs[0] = 3.43e26; m->mas_a[0] = 0.680; m->mas_b[0] = 2.0;
        m->mas_min1[1] = 0.00449; m->mas_min2[1] = 0.; m->mas_mu1[1] = 0.00290; m->mas_Pc[1] = 9.23e22;
        m->mas_Cr[1] = 2.23e23; m->mas_Cs[1] = 6.10e26; m->mas_a[1] = 0.719; m->mas_b[1] = 2.0;
        m->vsat_n = 1.07e5 * pow(tr, -0.87); m->vsat_p = 8.37e4 * pow(tr, -0.52);
        m->beta_n = 1.109 * pow(tr, 0.66); m->beta_p = 1.213 * pow(tr, 0.17);
        rec(1e-5, 1e-5, 2.8e-43, 9.9e-44, 4.73e-21);
        m->tau_Nref = 1e22;
        m->bgn_E = 6.92e-3; m->bgn_N = 1.3e23; m->bgn_C = 0.5;
        m->ii_on = 1; m->ii_hw = 0.063; m->ii_m[0] = m->ii_m[1] = 1.;
        m->ii_a[0][0] = m->ii_a[0][1] = 7.03e7; m->ii_b[0][0] = m->ii_b[0][1] = 1.231e8; m->ii_E0[0] = 4e7;
        m->ii_a[1][0] = 1.582e8; m->ii_b[1][0] = 2.036e8; m->ii_a[1][1] = 6.71e7; m->ii_b[1][1] = 1.693e8; m->ii_E0[1] = 4e7;
        m->lb_on = 3;                                    // Lombardi (1988): B/F + C N^l / F^(1/3); delta/F^2
        m->lbB[0] = 4.75e7; m->lbC[0] = 580.; m->lbL[0] = 0.125; m->lbD[0] = 5.82e14;
        m->lbB[1] = 9.925e6; m->lbC[1] = 2947.; m->lbL[1] = 0.0317; m->lbD[1] = 2.0546e14;
        // MLDA: electrons in the 2 Delta-2 (0.916) and 4 Delta-4 (0.19) valleys, holes heavy / light
        m->mq_e1 = 0.916; m->mq_e2 = 0.19; m->wq_e1 = 1. / 3.; m->mq_h1 = 0.29; m->mq_h2 = 0.20; m->wq_h1 = 0.84;
        m->ms_e = 0.50; m->ms_h = 0.50; m->mt_e = 0.26; m->mt_h = 0.39;
        break;
    case 1:  // Ge
        m->Eg = 0.742 - 4.77e-4 * T * T / (T + 235.); m->chi = 4.00; m->eps_r = 16.2;
        m->Nc = 1.04e25 * pow(tr, 1.5); m->Nv = 6.0e24 * pow(tr, 1.5);
        mob(0.39, 0.19, 0.040, 0.020, 1.30e23, 1.00e23, 0.56, 0.43); texp(0., -1.66, -2.33, 0., 0.);
        m->vsat_n = 6e4; m->vsat_p = 5.4e4; m->beta_n = 1.; m->beta_p = 1.;
        rec(1e-3, 1e-3, 1e-42, 1e-42, 6.4e-20);
        break;
    case 2:  // GaAs
        m->Eg = 1.519 - 5.405e-4 * T * T / (T + 204.); m->chi = 4.07; m->eps_r = 12.9;
        m->Nc = 4.37e23 * pow(tr, 1.5); m->Nv = 8.68e24 * pow(tr, 1.5);
        mob(0.94, 0.040, 0.050, 0.0020, 6.0e22, 2.75e23, 0.394, 0.394); texp(0., -2.1, -2.1, 3.0, 0.);
        m->vsat_n = 7.7e4; m->vsat_p = 7.7e4; m->beta_n = 1.; m->beta_p = 1.;
        rec(1e-9, 1e-9, 1e-43, 1e-43, 1.5e-16);
        iiset(2.99e7, 6.848e7, 2.215e7, 6.57e7, 1.6, 1.75);
        break;
    case 3:  // InP
        m->Eg = 1.4236 - 3.63e-4 * T * T / (T + 162.); m->chi = 4.38; m->eps_r = 12.5;
        m->Nc = 5.7e23 * pow(tr, 1.5); m->Nv = 1.1e25 * pow(tr, 1.5);
        mob(0.52, 0.015, 0.040, 0.0010, 3.00e23, 4.90e23, 0.47, 0.40); texp(0., -2.0, -2.0, 3.25, 0.);
        m->vsat_n = 1e5; m->vsat_p = 5e4; m->beta_n = 1.; m->beta_p = 1.;
        rec(1e-8, 1e-8, 1e-43, 1e-43, 8.5e-17);
        break;
    case 4:  // 4H-SiC (surface-mobility and incomplete-ionisation data are not carried: models 4/9 are rejected)
        m->Eg = 3.265 - 6.5e-4 * T * T / (T + 1300.); m->chi = 3.7; m->eps_r = 9.7;
        m->Nc = 1.73e25 * pow(tr, 1.5); m->Nv = 2.64e25 * pow(tr, 1.5);
        mob(0.095, 0.0124, 0.0040, 0.00159, 1.94e23, 1.76e25, 0.61, 0.34); texp(-0.5, -2.40, -2.15, 0., 0.);
        m->vsat_n = 2.2e5; m->vsat_p = 1.6e5; m->beta_n = 1.2; m->beta_p = 1.;
        rec(1e-6, 1e-7, 5e-44, 2e-44, 1.5e-18);
        iiset(1.76e10, 3.30e9, 3.41e10, 2.50e9, 1., 1.);
        m->ani_n = 1. / 0.83; m->ani_p = 1. / 1.15;
        break;
    case 5:  // GaN
        m->Eg = 3.507 - 9.09e-4 * T * T / (T + 830.); m->chi = 4.1; m->eps_r = 8.9;
        m->Nc = 2.3e24 * pow(tr, 1.5); m->Nv = 4.6e25 * pow(tr, 1.5);
        mob(0.120, 0.0030, 0.0055, 0.00030, 1.00e23, 3.00e23, 1.0, 2.0); texp(-1.02, -3.84, -3.84, 0., 0.);
        m->vsat_n = 2.5e5; m->vsat_p = 2e5; m->beta_n = 1.; m->beta_p = 1.;
        rec(1e-8, 1e-8, 1e-43, 1e-43, 1.1e-16);
        iiset(2.11e11, 3.689e9, 4.39e8, 1.8e9, 1., 1.);
        break;
    case 6:  // Al0.3Ga0.7As
        m->Eg = 1.882 - 5.58e-4 * T * T / (T + 295.); m->chi = 3.827; m->eps_r = 12.24;
        m->Nc = 7.2e23 * pow(tr, 1.5); m->Nv = 1.05e25 * pow(tr, 1.5);
        mob(0.35, 0.025, 0.020, 0.0015, 6.00e22, 1.00e23, 0.394, 0.394); texp(0., -2.1, -2.1, 3.0, 0.);
        m->vsat_n = 6e4; m->vsat_p = 5e4; m->beta_n = 1.; m->beta_p = 1.;
        rec(1e-9, 1e-9, 1e-43, 1e-43, 5e-17);
        break;
    case 7:  // Al0.25Ga0.75N
        m->Eg = 3.990 - 9.09e-4 * T * T / (T + 830.); m->chi = 3.76; m->eps_r = 8.8;
        m->Nc = 3.2e24 * pow(tr, 1.5); m->Nv = 4.6e25 * pow(tr, 1.5);
        mob(0.030, 0.0010, 0.0030, 0.00020, 1.00e23, 3.00e23, 1.0, 2.0); texp(-1.02, -3.84, -3.84, 0., 0.);
        m->vsat_n = 1.5e5; m->vsat_p = 1e5; m->beta_n = 1.; m->beta_p = 1.;
        rec(1e-9, 1e-9, 1e-43, 1e-43, 1e-16);
        break;
    case 8: m->insulator = 1; m->Eg = 9.0; m->chi = 0.95; m->eps_r = 3.9; m->mt_e = 0.42; m->mt_h = 0.33; break;     // SiO2
    case 9: m->insulator = 1; m->Eg = 6.8; m->chi = 1.35; m->eps_r = 9.0; m->mt_e = 0.35; m->mt_h = 0.35; break;     // Al2O3
    case 10: m->insulator = 1; m->Eg = 5.8; m->chi = 2.55; m->eps_r = 22.0; m->mt_e = 0.18; m->mt_h = 0.2; break;   // HfO2
    case 11: m->insulator = 1; m->Eg = 5.1; m->chi = 2.05; m->eps_r = 7.5; m->mt_e = 0.5; m->mt_h = 0.5; break;    // Si3N4
    case 12: m->insulator = 1; m->Eg = 10.; m->chi = 0.; m->eps_r = 1.0; m->mt_e = 1.0; m->mt_h = 1.0; break;      // void
    case 13: m->insulator = 1; m->Eg = 10.; m->chi = 0.; m->eps_r = 1.0e3; m->mt_e = 1.0; m->mt_h = 1.0; break;    // floating metal inclusion
    default: return false;
    }
    if (!m->insulator && !(m->mq_e1 > 0.) && id >= 0 && id < 8) {   // single DOS-like masses
        static const double me[8] = {0.916, 0.082, 0.067, 0.08, 0.37, 0.20, 0.09, 0.23};
        static const double mh[8] = {0.29, 0.28, 0.50, 0.60, 1.00, 1.00, 0.55, 1.00};
        m->mq_e1 = m->mq_e2 = me[id]; m->wq_e1 = 1.; m->mq_h1 = m->mq_h2 = mh[id]; m->wq_h1 = 1.;
        m->ms_e = me[id] < 0.5 ? me[id] : 0.5; m->ms_h = 0.5;
    }
    m->ni = m->insulator ? 0. : sqrt(m->Nc * m->Nv) * exp(-m->Eg / (2. * m->VT));
    m->ii_g = m->ii_hw > 0. ? tanh(m->ii_hw / (2. * SD_KB * 300. / SD_Q)) / tanh(m->ii_hw / (2. * m->VT)) : 1.;
    return true;
}

// Masetti (mas_on) or Arora low-field mobility (m^2/Vs)
inline double host_mob_low(const HostMat &m, double Ntot, int hole) {
    if (Ntot < 1.) Ntot = 1.;
    double tr = m.T / 300.;
    int c = hole ? 1 : 0;
    if (m.mas_on) {
        double muL = m.mu_max[c] * pow(tr, m.mt_max[c]);
        double mu = m.mas_min1[c] * exp(-m.mas_Pc[c] / Ntot) + (muL - m.mas_min2[c]) / (1. + pow(Ntot / m.mas_Cr[c], m.mas_a[c]))
                    - m.mas_mu1[c] / (1. + pow(m.mas_Cs[c] / Ntot, m.mas_b[c]));
        return mu > 1e-6 ? mu : 1e-6;
    }
    double mumin = m.mu_min[c] * pow(tr, m.mt_min);
    double mumax = m.mu_max[c] * pow(tr, m.mt_max[c]);
    double Nref = m.Nref[c] * pow(tr, m.mt_Nref);
    double alpha = m.alpha[c] * pow(tr, m.mt_alpha);
    if (mumax < mumin) mumax = mumin;
    return mumin + (mumax - mumin) / (1. + pow(Ntot / Nref, alpha));
}

// half the band-gap narrowing in units of VT (Klaassen form)
inline double host_bgn_half(const HostMat &m, double N) {
    if (!(m.bgn_E > 0.) || N < 1.) return 0.;
    double L = log(N / m.bgn_N);
    return 0.5 * m.bgn_E * (L + sqrt(L * L + m.bgn_C)) / m.VT;
}

inline double host_B(double x) {
    if (fabs(x) < 1e-6) return 1. - x / 2. + x * x / 12.;
    if (x > 700.) return 0.;
    if (x < -700.) return -x;
    return x / expm1(x);
}

inline void host_recomb(const HostMat &m, double ni, double tf, double n, double p, double *R, double *dRn, double *dRp) {
    double ni2 = ni * ni, np = n * p - ni2;
    double den = tf * (m.tau_p0 * (n + ni) + m.tau_n0 * (p + ni)) + 1e-300;
    double Rs = np / den;
    double dsn = (p * den - np * tf * m.tau_p0) / (den * den);
    double dsp = (n * den - np * tf * m.tau_n0) / (den * den);
    double Au = m.Cn * n + m.Cp * p;
    *R = Rs + Au * np + m.Brad * np;
    *dRn = dsn + m.Cn * np + Au * p + m.Brad * p;
    *dRp = dsp + m.Cp * np + Au * n + m.Brad * n;
}

inline MatDev host_mat_dev(const HostMat &m) {
    MatDev d;
    memset(&d, 0, sizeof d);
    d.VT = m.VT; d.lnc = m.lnc; d.lnv = m.lnv; d.ni = m.ni; d.Eg = m.Eg; d.chi = m.chi;
    d.eps_r = m.eps_r; d.Nc = m.Nc; d.Nv = m.Nv; d.ms_e = m.ms_e; d.ms_h = m.ms_h; d.mt_e = m.mt_e; d.mt_h = m.mt_h;
    d.tau_n0 = m.tau_n0; d.tau_p0 = m.tau_p0; d.Cn = m.Cn; d.Cp = m.Cp; d.Brad = m.Brad; d.tau_Nref = m.tau_Nref;
    d.vsat_n = m.vsat_n; d.vsat_p = m.vsat_p; d.beta_n = m.beta_n; d.beta_p = m.beta_p;
    d.ani_n = m.ani_n; d.ani_p = m.ani_p;
    memcpy(d.ii_a, m.ii_a, sizeof d.ii_a); memcpy(d.ii_b, m.ii_b, sizeof d.ii_b);
    d.ii_E0[0] = m.ii_E0[0]; d.ii_E0[1] = m.ii_E0[1]; d.ii_m[0] = m.ii_m[0]; d.ii_m[1] = m.ii_m[1]; d.ii_g = m.ii_g;
    d.insulator = m.insulator; d.ii_on = m.ii_on;
    d.T = m.T; d.lb_on = m.lb_on;
    for (int c = 0; c < 2; c++) {
        d.lbB[c] = m.lbB[c]; d.lbC[c] = m.lbC[c]; d.lbL[c] = m.lbL[c]; d.lbD[c] = m.lbD[c]; d.lbE[c] = m.lbE[c];
    }
    return d;
}
