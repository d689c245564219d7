// psr_chem.hip -- per-particle ionization / temperature update on gfx950
// (update_pars, src/loop.c:647-776; step_evolve, src/solve.c:76-224).
//
// One thread per particle in FP64.  The right-hand side is the reference's
// 7-variable system (src/solve.c:42-141); variables the run flags freeze
// (zero derivative) are left out of the implicit solve, so the linear algebra
// is NV x NV with NV = 1 for an isothermal hydrogen-only on-the-spot run and
// NV = 4 for H+He with temperature.  The integrator is the three-stage
// L-stable Rosenbrock method ROS3 (Sandu et al. 1997, order 3) with an
// embedded second-order error estimate and the reference tolerances (rtol
// 1e-4, atol 1e-7 / 1 for ie);
// the Jacobian is exact (dual numbers through the right-hand side and the
// rate table's interpolation), one table pass per step where finite
// differences took N + 1.  The x_HI variable and x_HII = 1 - x_HI are both
// kept in FP64.
#include <mutex>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "psr_internal.h"
#include "psr_mlp.h"

namespace {

enum {
    AR_HI_CI = 0, AR_HII_RC_A, AR_HII_RCC_A, AR_HII_RC_B, AR_HII_RCC_B, AR_HEI_CI, AR_HEII_CI,
    AR_HEII_RC_A, AR_HEII_RC_B, AR_HEIII_RC_A, AR_HEIII_RC_B, AR_HI_CIC, AR_HI_CEC, AR_E_BREMC,
    AR_E_COMPC
};

struct Rates {
    double gamma_HI, alpha_HII_A, alpha_HII_B, alpha_HII_AB, gamma_HeI, gamma_HeII, alpha_HeII_A,
        alpha_HeII_B, alpha_HeII_AB, alpha_HeIII_A, alpha_HeIII_B, alpha_HeIII_AB, eta_HII, psi_HI,
        zeta_HI, beta, chi;
};

// ar_get for every column at one logT (src/tabfun.c:38-58), times nH as
// FETCH_VARS does (src/solve.c:56-72)
template <bool OTS>
__device__ __forceinline__ void rates_at(const DevTables &T, double logT, double nH, Rates &R) {
    const int n = T.ar_nrows;
    const long long idx = (long long)((logT - T.ar_min) * T.ar_step_inv);
    double lw = 0, rw = 0;
    int i0 = 0;
    int mode;  // 0 clamp low, 1 interior, 2 clamp high
    if (idx < 0) {
        mode = 0;
    } else if (idx >= n - 1) {
        mode = 2;
        i0 = n - 1;
    } else {
        mode = 1;
        i0 = (int)idx;
        lw = T.ar[(size_t)(i0 + 1) * 16] - logT;
        rw = logT - T.ar[(size_t)i0 * 16];
    }
    const double *r0 = T.ar + (size_t)i0 * 16;
    auto col = [&](int k) -> double {
        if (mode != 1) return r0[k + 1];
        return (lw * r0[k + 1] + rw * r0[16 + k + 1]) * T.ar_step_inv;
    };
    R.gamma_HI = col(AR_HI_CI) * nH;
    R.alpha_HII_A = col(AR_HII_RC_A) * nH;
    R.alpha_HII_B = col(AR_HII_RC_B) * nH;
    R.alpha_HII_AB = fdim(R.alpha_HII_A, R.alpha_HII_B);
    R.gamma_HeI = col(AR_HEI_CI) * nH;
    R.gamma_HeII = col(AR_HEII_CI) * nH;
    R.alpha_HeII_A = col(AR_HEII_RC_A) * nH;
    R.alpha_HeII_B = col(AR_HEII_RC_B) * nH;
    R.alpha_HeII_AB = fdim(R.alpha_HeII_A, R.alpha_HeII_B);
    R.alpha_HeIII_A = col(AR_HEIII_RC_A) * nH;
    R.alpha_HeIII_B = col(AR_HEIII_RC_B) * nH;
    R.alpha_HeIII_AB = fdim(R.alpha_HeIII_A, R.alpha_HeIII_B);
    R.eta_HII = (OTS ? col(AR_HII_RCC_B) : col(AR_HII_RCC_A)) * nH;
    R.psi_HI = col(AR_HI_CEC) * nH;
    R.zeta_HI = col(AR_HI_CIC) * nH;
    R.beta = col(AR_E_BREMC) * nH;
    R.chi = col(AR_E_COMPC);
}

template <bool ISO, bool ADI, bool OTS, bool HONLY>
struct Sys {
    static constexpr int NV = 1 + (HONLY ? 0 : 2) + (OTS ? 0 : (HONLY ? 1 : 3)) + ((ADI || ISO) ? 0 : 1);
    __device__ static constexpr int act(int k) {
        // evolving variables in increasing index order
        int idx[7] = {0, 0, 0, 0, 0, 0, 0};
        int n = 0;
        idx[n++] = 0;
        if (!HONLY) { idx[n++] = 1; idx[n++] = 2; }
        if (!OTS) {
            idx[n++] = 3;
            if (!HONLY) { idx[n++] = 4; idx[n++] = 5; }
        }
        if (!(ADI || ISO)) idx[n++] = 6;
        return idx[k];
    }
};

struct StepIn {
    double yGdepHI, yGdepHeI, yGdepHeII, heat, time, nH, T, yeMET;
    // the deposits per unit time, the constant terms of the right-hand side
    // (the same quotients it forms, computed once per step instead of per
    // evaluation): set by rates_per_time()
    double gHI, gHeI, gHeII, qheat;
};

__device__ __forceinline__ void rates_per_time(StepIn &s) {
    s.gHI = s.yGdepHI / s.time;
    s.gHeI = s.yGdepHeI / s.time;
    s.gHeII = s.yGdepHeII / s.time;
    s.qheat = s.heat / s.time;
}

// f_i / time of src/solve.c:76-141 from the fractions, ye and the rates, in
// the reference's expression order (rhs and rhs_jac share it, so their f
// values are the same bits)
template <bool ISO, bool ADI, bool OTS, bool HONLY>
__device__ __forceinline__ void rhs_rows(const double x[7], double ye, const Rates &R, const StepIn &s,
                                         const DevPhys &P, double f[7]) {
    const double xHI = x[0];
    const double xHII = 1.0 - x[0];
    const double xHeI = x[1];
    const double xHeII = x[2];
    const double xHeIII = 1.0 - x[2] - x[1];
    if (!OTS) {
        f[0] = -s.gHI + (-R.gamma_HI * ye * xHI + R.alpha_HII_A * ye * xHII);
        f[3] = R.alpha_HII_AB * ye * xHII;
        if (HONLY) {
            f[1] = f[2] = f[4] = f[5] = 0.0;
        } else {
            f[1] = -s.gHeI + (-R.gamma_HeI * ye * xHeI + R.alpha_HeII_A * ye * xHeII);
            f[2] = -s.gHeII +
                   (-R.gamma_HeII * ye * xHeII - R.alpha_HeII_A * ye * xHeII + R.alpha_HeIII_A * ye * xHeIII);
            f[4] = R.alpha_HeII_AB * ye * xHeII;
            f[5] = R.alpha_HeIII_AB * ye * xHeIII;
        }
    } else {
        f[0] = -s.gHI + (-R.gamma_HI * ye * xHI + R.alpha_HII_B * ye * xHII);
        if (HONLY) {
            f[1] = f[2] = 0.0;
        } else {
            // x_HI in the He lines, as src/solve.c:103-104
            f[1] = -s.gHeI + (-R.gamma_HeI * ye * xHI + R.alpha_HeII_B * ye * xHeII);
            f[2] = -s.gHeII +
                   (-R.gamma_HeII * ye * xHI - R.alpha_HeII_B * ye * xHeII + R.alpha_HeIII_B * ye * xHeIII);
        }
        f[3] = f[4] = f[5] = 0;
    }
    if (ADI || ISO) {
        f[6] = 0.0;
    } else {
        const double L = P.C_H_PER_MASS * ((R.zeta_HI + R.psi_HI) * ye * xHI + (R.eta_HII * ye * xHII) +
                                           R.beta * ye * xHII + R.chi * ye);
        f[6] = s.qheat - L;
    }
}

__device__ __forceinline__ double ye_of(const double x[7], const StepIn &s, const DevPhys &P) {
    return (1.0 - x[0]) + P.C_HEMF / P.C_HMF * 0.25 * (x[2] + 2 * (1.0 - x[2] - x[1])) + s.yeMET;
}

// RHS of src/solve.c:76-141, scaled by the step time
template <bool ISO, bool ADI, bool OTS, bool HONLY>
__device__ void rhs(const double x[7], double f[7], const StepIn &s, const DevTables &T,
                    const DevPhys &P, const Rates &Riso) {
    const double ye = ye_of(x, s, P);
    Rates R;
    if (ISO) {
        R = Riso;
    } else {
        const double Tk = P.U_MPROTON_OVER_C_BOLTZMANN * x[6] / (P.C_HEMF * 0.25 + P.C_HMF + P.C_HMF * ye) * 2.0 / 3.0;
        rates_at<OTS>(T, log10(Tk), s.nH, R);
    }
    rhs_rows<ISO, ADI, OTS, HONLY>(x, ye, R, s, P, f);
#pragma unroll
    for (int i = 0; i < 7; i++) f[i] *= s.time;
}

// the rates at log10 T (as rates_at) and d rate / d log10 T
template <bool OTS>
__device__ __forceinline__ void rates_slopes_at(const DevTables &T, double logT, double nH, Rates &R, Rates &S) {
    const int n = T.ar_nrows;
    const long long idx = (long long)((logT - T.ar_min) * T.ar_step_inv);
    double lw = 0, rw = 0;
    int i0 = 0;
    int mode;  // 0 clamp low, 1 interior, 2 clamp high
    if (idx < 0) {
        mode = 0;
    } else if (idx >= n - 1) {
        mode = 2;
        i0 = n - 1;
    } else {
        mode = 1;
        i0 = (int)idx;
        lw = T.ar[(size_t)(i0 + 1) * 16] - logT;
        rw = logT - T.ar[(size_t)i0 * 16];
    }
    const double *r0 = T.ar + (size_t)i0 * 16;
    auto col = [&](int k, double &sl) -> double {
        if (mode != 1) {
            sl = 0.0;
            return r0[k + 1];
        }
        sl = (r0[16 + k + 1] - r0[k + 1]) * T.ar_step_inv;
        return (lw * r0[k + 1] + rw * r0[16 + k + 1]) * T.ar_step_inv;
    };
    double sl;
    R.gamma_HI = col(AR_HI_CI, sl) * nH;
    S.gamma_HI = sl * nH;
    R.alpha_HII_A = col(AR_HII_RC_A, sl) * nH;
    S.alpha_HII_A = sl * nH;
    R.alpha_HII_B = col(AR_HII_RC_B, sl) * nH;
    S.alpha_HII_B = sl * nH;
    R.gamma_HeI = col(AR_HEI_CI, sl) * nH;
    S.gamma_HeI = sl * nH;
    R.gamma_HeII = col(AR_HEII_CI, sl) * nH;
    S.gamma_HeII = sl * nH;
    R.alpha_HeII_A = col(AR_HEII_RC_A, sl) * nH;
    S.alpha_HeII_A = sl * nH;
    R.alpha_HeII_B = col(AR_HEII_RC_B, sl) * nH;
    S.alpha_HeII_B = sl * nH;
    R.alpha_HeIII_A = col(AR_HEIII_RC_A, sl) * nH;
    S.alpha_HeIII_A = sl * nH;
    R.alpha_HeIII_B = col(AR_HEIII_RC_B, sl) * nH;
    S.alpha_HeIII_B = sl * nH;
    R.eta_HII = (OTS ? col(AR_HII_RCC_B, sl) : col(AR_HII_RCC_A, sl)) * nH;
    S.eta_HII = sl * nH;
    R.psi_HI = col(AR_HI_CEC, sl) * nH;
    S.psi_HI = sl * nH;
    R.zeta_HI = col(AR_HI_CIC, sl) * nH;
    S.zeta_HI = sl * nH;
    R.beta = col(AR_E_BREMC, sl) * nH;
    S.beta = sl * nH;
    R.chi = col(AR_E_COMPC, sl);
    S.chi = sl;
    // fdim(a, b) = a - b where a > b, else 0, and so its slope
    R.alpha_HII_AB = fdim(R.alpha_HII_A, R.alpha_HII_B);
    S.alpha_HII_AB = R.alpha_HII_A > R.alpha_HII_B ? S.alpha_HII_A - S.alpha_HII_B : 0.0;
    R.alpha_HeII_AB = fdim(R.alpha_HeII_A, R.alpha_HeII_B);
    S.alpha_HeII_AB = R.alpha_HeII_A > R.alpha_HeII_B ? S.alpha_HeII_A - S.alpha_HeII_B : 0.0;
    R.alpha_HeIII_AB = fdim(R.alpha_HeIII_A, R.alpha_HeIII_B);
    S.alpha_HeIII_AB = R.alpha_HeIII_A > R.alpha_HeIII_B ? S.alpha_HeIII_A - S.alpha_HeIII_B : 0.0;
}

// ---- the Jacobian of the right-hand side ----
// Every term of row i carries one factor ye: f_i = c_i + ye p_i(x, r(log10 T)),
// with the rates r piecewise linear in log10 T (slope s_r on the segment).
// Only x_HI, x_HeI, x_HeII appear in ye and in the p_i, and log10 T depends
// on ye and ie, so
//     d f_i / d x_q = (d ye / d x_q) p_i
//                     + ye (dp_i/d log10 T · d log10 T / d x_q + dp_i/dx_q |_T).
// Per row that is the value p_i, its slope in log10 T and three partials at
// fixed T: a few products instead of carrying every intermediate value's
// derivatives through the right-hand side.
struct RowD {
    double p, ps, d0, d1, d2;  // p_i, d p_i / d log10 T, d p_i / d x_{0,1,2}
};

template <bool ISO, bool OTS>
__device__ __forceinline__ RowD row_d(int i, const double x[7], const Rates &R, const Rates &S, double C) {
    const double xHI = x[0], xHII = 1.0 - x[0], xHeI = x[1], xHeII = x[2], xHeIII = 1.0 - x[2] - x[1];
    RowD r{0.0, 0.0, 0.0, 0.0, 0.0};
    switch (i) {
    case 0:
        if (OTS) {
            r.p = -R.gamma_HI * xHI + R.alpha_HII_B * xHII;
            if (!ISO) r.ps = -S.gamma_HI * xHI + S.alpha_HII_B * xHII;
            r.d0 = -R.gamma_HI - R.alpha_HII_B;
        } else {
            r.p = -R.gamma_HI * xHI + R.alpha_HII_A * xHII;
            if (!ISO) r.ps = -S.gamma_HI * xHI + S.alpha_HII_A * xHII;
            r.d0 = -R.gamma_HI - R.alpha_HII_A;
        }
        break;
    case 1:
        if (OTS) {
            r.p = -R.gamma_HeI * xHI + R.alpha_HeII_B * xHeII;
            if (!ISO) r.ps = -S.gamma_HeI * xHI + S.alpha_HeII_B * xHeII;
            r.d0 = -R.gamma_HeI;
            r.d2 = R.alpha_HeII_B;
        } else {
            r.p = -R.gamma_HeI * xHeI + R.alpha_HeII_A * xHeII;
            if (!ISO) r.ps = -S.gamma_HeI * xHeI + S.alpha_HeII_A * xHeII;
            r.d1 = -R.gamma_HeI;
            r.d2 = R.alpha_HeII_A;
        }
        break;
    case 2:
        if (OTS) {
            r.p = -R.gamma_HeII * xHI - R.alpha_HeII_B * xHeII + R.alpha_HeIII_B * xHeIII;
            if (!ISO) r.ps = -S.gamma_HeII * xHI - S.alpha_HeII_B * xHeII + S.alpha_HeIII_B * xHeIII;
            r.d0 = -R.gamma_HeII;
            r.d1 = -R.alpha_HeIII_B;
            r.d2 = -R.alpha_HeII_B - R.alpha_HeIII_B;
        } else {
            r.p = -R.gamma_HeII * xHeII - R.alpha_HeII_A * xHeII + R.alpha_HeIII_A * xHeIII;
            if (!ISO) r.ps = -S.gamma_HeII * xHeII - S.alpha_HeII_A * xHeII + S.alpha_HeIII_A * xHeIII;
            r.d1 = -R.alpha_HeIII_A;
            r.d2 = -R.gamma_HeII - R.alpha_HeII_A - R.alpha_HeIII_A;
        }
        break;
    case 3:
        r.p = R.alpha_HII_AB * xHII;
        if (!ISO) r.ps = S.alpha_HII_AB * xHII;
        r.d0 = -R.alpha_HII_AB;
        break;
    case 4:
        r.p = R.alpha_HeII_AB * xHeII;
        if (!ISO) r.ps = S.alpha_HeII_AB * xHeII;
        r.d2 = R.alpha_HeII_AB;
        break;
    case 5:
        r.p = R.alpha_HeIII_AB * xHeIII;
        if (!ISO) r.ps = S.alpha_HeIII_AB * xHeIII;
        r.d1 = -R.alpha_HeIII_AB;
        r.d2 = -R.alpha_HeIII_AB;
        break;
    default:  // 6: the cooling, f_6 = heat / time - C ye (...)
        r.p = -C * ((R.zeta_HI + R.psi_HI) * xHI + (R.eta_HII + R.beta) * xHII + R.chi);
        if (!ISO) r.ps = -C * ((S.zeta_HI + S.psi_HI) * xHI + (S.eta_HII + S.beta) * xHII + S.chi);
        r.d0 = -C * ((R.zeta_HI + R.psi_HI) - (R.eta_HII + R.beta));
        break;
    }
    return r;
}

// f = rhs(x) and J[i][j] = d f_act(i) / d x_act(j) over the evolving variables
template <bool ISO, bool ADI, bool OTS, bool HONLY>
__device__ void rhs_jac(const double x[7], double f[7], double *J, const StepIn &s, const DevTables &T,
                        const DevPhys &P, const Rates &Riso) {
    using SY = Sys<ISO, ADI, OTS, HONLY>;
    constexpr int N = SY::NV;
    const double cHe = P.C_HEMF / P.C_HMF * 0.25;
    const double ye = ye_of(x, s, P);
    // d ye / d x_q for q = 0, 1, 2
    const double dye[3] = {-1.0, -2.0 * cHe, -cHe};
    Rates R, Sl;
    double dL[N];  // d log10 T / d x_act(j)
#pragma unroll
    for (int jj = 0; jj < N; jj++) dL[jj] = 0.0;
    if (ISO) {
        R = Riso;
    } else {
        const double den = P.C_HEMF * 0.25 + P.C_HMF + P.C_HMF * ye;
        const double Tk = P.U_MPROTON_OVER_C_BOLTZMANN * x[6] / den * 2.0 / 3.0;
        rates_slopes_at<OTS>(T, log10(Tk), s.nH, R, Sl);
        // log10 T = log10 ie - log10 den + const
        const double dyefac = -P.C_HMF * 0.43429448190325182765 / den;
#pragma unroll
        for (int jj = 0; jj < N; jj++) {
            const int q = SY::act(jj);
            if (q < 3) dL[jj] = dyefac * dye[q];
            if (q == 6) dL[jj] = 0.43429448190325182765 / x[6];
        }
    }
    rhs_rows<ISO, ADI, OTS, HONLY>(x, ye, R, s, P, f);
#pragma unroll
    for (int i = 0; i < 7; i++) f[i] *= s.time;
#pragma unroll
    for (int ii = 0; ii < N; ii++) {
        const RowD r = row_d<ISO, OTS>(SY::act(ii), x, R, Sl, P.C_H_PER_MASS);
        const double yt = ye * s.time, pt = r.p * s.time, pst = r.ps * yt;
#pragma unroll
        for (int jj = 0; jj < N; jj++) {
            const int q = SY::act(jj);
            double v = ISO ? 0.0 : pst * dL[jj];
            if (q < 3) v += dye[q] * pt + yt * (q == 0 ? r.d0 : q == 1 ? r.d1 : r.d2);
            J[ii * N + jj] = v;
        }
    }
}

// ROS3 of Sandu, Verwer, Blom, Spee, Carmichael & Potra, Atmos. Environ. 31
// (1997) 3459-3472: three stages, two evaluations of F, order 3 with an
// embedded order-2 solution, L-stable; coefficients of the form
// G = I/(gamma h) - J, G K_i = F(y + sum a_ij K_j) + sum (c_ij/h) K_j
#define ROS3_GAMMA 0.43586652150845899941601945119356
#define ROS3_C21 (-1.0156171083877702091975600115545)
#define ROS3_C31 4.0759956452537699824805835358067
#define ROS3_C32 9.2076794298330791242156818474003
#define ROS3_M1 1.0
#define ROS3_M2 6.1697947043828245592553615689730
#define ROS3_M3 (-0.42772256543218573326238373806514)
#define ROS3_E1 0.5
#define ROS3_E2 (-2.9079558716805469821718236208017)
#define ROS3_E3 0.22354069897811569627360909276199
#define ROS3_MAXSTEPS 100000

// v^(-1/3) for the step-size controller, v in [1e-10, 1e10]: with v = m 2^e,
// m in [0.5, 1), e = 3q + r, v^(-1/3) = m^(-1/3) 2^(-r/3) 2^(-q), and
// m^(-1/3) by Newton's iteration for the reciprocal cube root, y += y (1 -
// m y^3) / 3 (no division; 4 steps from the linear guess reach 1 ulp).
// Correctly rounded operations only, so the host oracle computes the same bits
__device__ __forceinline__ double ctl_rcbrt(double v) {
    int e;
    const double m = frexp(v, &e);
    const int r = ((e % 3) + 3) % 3;
    const int q = (e - r) / 3;
    double y = 1.4 - 0.4 * m;
#pragma unroll
    for (int k = 0; k < 4; k++) y = y + y * (1.0 - m * (y * y * y)) * (1.0 / 3.0);
    const double c = r == 0 ? 1.0 : (r == 1 ? 0.79370052598409973738 : 0.62996052494743658238);
    return ldexp(y * c, -q);
}

// Isothermal hydrogen-only on-the-spot: with the rates fixed over the step
// the rate equation (x = x_HI, tau = t / time) is the Riccati equation
//     dx/dtau = F(x) = -g + (1 + m - x)(A - c x),  c = A + G,
// G = gamma_HI time, A = alpha_B time, g = yGdepHI, m = yeMET; F = a + b x +
// c x^2 with real roots r1 = a/q <= r2 = q/c, q = (lambda - b)/2, lambda^2 =
// ((1+m)G + mA)^2 + 4cg.  Exact solution over the step:
//     x(1) = x0 + F(x0) phi / D,  phi = (1 - e^-lambda)/lambda,
//     D = e^-lambda + w (1 - e^-lambda),  w = c (r2 - x0) / lambda,
// w from F(x0) = -c u0 (r2 - x0) (u0 = x0 - r1) when x0 is nearer r2.  The
// oracle's or_riccati_tick evaluates the same expressions in the same order.
template <bool ISO, bool OTS, bool HONLY>
constexpr bool closed_form() {
    return ISO && OTS && HONLY;
}

__device__ double riccati_tick(double x0, double g, double m, double G, double A) {
    const double c = A + G;
    const double F0 = -g + (1.0 + m - x0) * (A - c * x0);
    if (F0 == 0.0) return x0;
    const double b = -((1.0 + m) * c + A);
    const double a = (1.0 + m) * A - g;
    const double dd = (1.0 + m) * G + m * A;
    const double lam = sqrt(dd * dd + 4.0 * c * g);
    const double q = 0.5 * (lam - b);
    if (lam < 1e-8) {
        const double cu0 = q > 0 ? c * x0 - a * c / q : 0.0;
        const double phi = 1.0 - 0.5 * lam;
        return x0 + F0 * phi / (1.0 - cu0 * phi);
    }
    const double u0 = x0 - a / q;
    const double s = -expm1(-lam);
    const double w = (c * u0 > 0.5 * lam) ? -F0 / (u0 * lam) : (q - c * x0) / lam;
    const double D = (1.0 - s) + fmax(w, 0.0) * s;
    return x0 + F0 * (s / lam) / D;
}

// one step in closed form, then the clamps of src/solve.c:208-213
__device__ __forceinline__ void riccati_step(const double x0[7], const StepIn &s, const DevTables &T,
                                             double x[7]) {
    Rates R;
    rates_at<true>(T, log10(s.T), s.nH, R);
    const double G = R.gamma_HI * s.time;
    const double A = R.alpha_HII_B * s.time;
#pragma unroll
    for (int q = 0; q < 7; q++) x[q] = x0[q];
    x[0] = riccati_tick(x0[0], s.yGdepHI, s.yeMET, G, A);
    x[0] = fmin(1., fmax(0., x[0]));
    x[1] = fmin(1., fmax(0., x[1]));
    x[2] = fmin(1. - x[1], fmax(0., x[2]));
    x[3] = fmax(0., x[3]);
    x[4] = fmax(0., x[4]);
    x[5] = fmax(0., x[5]);
}

// LU with partial pivoting (first strictly larger pivot, as the oracle);
// every index is static after unrolling so the matrix stays in registers
template <int N>
__device__ __forceinline__ bool lu_factor(double a[N * N], int piv[N]) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        int p = k;
        double m = fabs(a[k * N + k]);
#pragma unroll
        for (int i = k + 1; i < N; i++)
            if (fabs(a[i * N + k]) > m) {
                m = fabs(a[i * N + k]);
                p = i;
            }
        piv[k] = p;
        if (m == 0.0) return false;
#pragma unroll
        for (int q = k + 1; q < N; q++)
            if (q == p) {
#pragma unroll
                for (int j = 0; j < N; j++) {
                    const double t = a[k * N + j];
                    a[k * N + j] = a[q * N + j];
                    a[q * N + j] = t;
                }
            }
        const double inv = 1.0 / a[k * N + k];
        a[k * N + k] = inv;  // the solves multiply by the pivot's reciprocal
#pragma unroll
        for (int i = k + 1; i < N; i++) {
            const double l = a[i * N + k] * inv;
            a[i * N + k] = l;
#pragma unroll
            for (int j = k + 1; j < N; j++) a[i * N + j] -= l * a[k * N + j];
        }
    }
    return true;
}

template <int N>
__device__ __forceinline__ void lu_solve(const double a[N * N], const int piv[N], double b[N]) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        // the row exchange as selects over the candidate rows, so b stays in
        // registers (an exchange written with the pivot as an index puts b
        // in scratch memory)
        const int p = piv[k];
        const double bk = b[k];
        double nk = bk;
#pragma unroll
        for (int q = k + 1; q < N; q++) {
            const bool sw = q == p;
            nk = sw ? b[q] : nk;
            b[q] = sw ? bk : b[q];
        }
        b[k] = nk;
#pragma unroll
        for (int i = k + 1; i < N; i++) b[i] -= a[i * N + k] * b[k];
    }
#pragma unroll
    for (int k = N - 1; k >= 0; k--) {
        double s = b[k];
#pragma unroll
        for (int j = k + 1; j < N; j++) s -= a[k * N + j] * b[j];
        b[k] = s * a[k * N + k];
    }
}

// Integrator state of one particle: normalized time 0..1 of
// src/solve.c:170-224.  attempt() performs one Rosenbrock step attempt (the
// right-hand side and Jacobian are re-evaluated after every accepted step),
// so a caller can interleave particles lane by lane.
template <bool ISO, bool ADI, bool OTS, bool HONLY>
struct Integ {
    using S = Sys<ISO, ADI, OTS, HONLY>;
    static constexpr int N = S::NV;
    double x[7];
    StepIn s;
    Rates Riso;
    double t, h;
    bool rejected, need_fj, dead;
    int nsteps;
    double f0[7];
    double J[N * N];

    __device__ void init(const double x0[7], const StepIn &s0, const DevTables &T) {
#pragma unroll
        for (int q = 0; q < 7; q++) x[q] = x0[q];
        s = s0;
        if (ISO) rates_at<OTS>(T, log10(s.T), s.nH, Riso);
        t = 0.0;
        h = 1.0;
        rejected = false;
        need_fj = true;
        dead = false;
        nsteps = 0;
    }

    // returns true when the integration has ended (t = 1 or the step limit)
    __device__ bool attempt(const DevTables &T, const DevPhys &P, double rtol) {
        const double atol[7] = {1e-7, 1e-7, 1e-7, 1e-7, 1e-7, 1e-7, 1.};
        double f1[7], xt[7], M[N * N], k1[N], k2[N], k3[N];
        int piv[N];
        if (need_fj) {
            if (dead || !(t < 1.0 && nsteps < ROS3_MAXSTEPS)) return true;
            if (h > 1.0 - t) h = 1.0 - t;
            rhs_jac<ISO, ADI, OTS, HONLY>(x, f0, J, s, T, P, Riso);
            need_fj = false;
        }
        nsteps++;
        const double ghinv = 1.0 / (ROS3_GAMMA * h);
#pragma unroll
        for (int ii = 0; ii < N; ii++)
#pragma unroll
            for (int jj = 0; jj < N; jj++) M[ii * N + jj] = (ii == jj ? ghinv : 0.0) - J[ii * N + jj];
        const bool ok = lu_factor<N>(M, piv);
        double E = INFINITY;
        double xn[7];
        if (ok) {
            const double c21 = ROS3_C21 / h, c31 = ROS3_C31 / h, c32 = ROS3_C32 / h;
#pragma unroll
            for (int ii = 0; ii < N; ii++) k1[ii] = f0[S::act(ii)];
            lu_solve<N>(M, piv, k1);
#pragma unroll
            for (int q = 0; q < 7; q++) xt[q] = x[q];
#pragma unroll
            for (int ii = 0; ii < N; ii++) xt[S::act(ii)] = x[S::act(ii)] + k1[ii];
            rhs<ISO, ADI, OTS, HONLY>(xt, f1, s, T, P, Riso);
#pragma unroll
            for (int ii = 0; ii < N; ii++) k2[ii] = f1[S::act(ii)] + c21 * k1[ii];
            lu_solve<N>(M, piv, k2);
#pragma unroll
            for (int ii = 0; ii < N; ii++) k3[ii] = f1[S::act(ii)] + c31 * k1[ii] + c32 * k2[ii];
            lu_solve<N>(M, piv, k3);
#pragma unroll
            for (int q = 0; q < 7; q++) xn[q] = x[q];
            double acc = 0.0;
#pragma unroll
            for (int ii = 0; ii < N; ii++) {
                const int i = S::act(ii);
                xn[i] = x[i] + ROS3_M1 * k1[ii] + ROS3_M2 * k2[ii] + ROS3_M3 * k3[ii];
                const double e = ROS3_E1 * k1[ii] + ROS3_E2 * k2[ii] + ROS3_E3 * k3[ii];
                const double sc = atol[i] + rtol * fmax(fabs(x[i]), fabs(xn[i]));
                acc += (e / sc) * (e / sc);
            }
            E = sqrt(acc / N);
            if (!(E == E)) E = INFINITY;
        }
        // a singular matrix or a non-finite error estimate never advances
        // the state; at the smallest step it ends the integration with x
        // unchanged (t < 1 counts it as a solver failure)
        const bool good = ok && E < INFINITY;
        if (!good && h < 1e-14) {
            dead = true;
            return true;
        }
        if (good && (E <= 1.0 || h < 1e-14)) {
            t += h;
#pragma unroll
            for (int q = 0; q < 7; q++) x[q] = xn[q];
            double fac = 0.9 * ctl_rcbrt(fmin(fmax(E, 1e-10), 1e10));
            fac = fmin(5.0, fmax(0.2, fac));
            if (rejected) fac = fmin(fac, 1.0);
            rejected = false;
            h *= fac;
            if (1.0 - t < 1e-13) t = 1.0;
            need_fj = true;
            return !(t < 1.0 && nsteps < ROS3_MAXSTEPS);
        }
        rejected = true;
        const double fac = (E == INFINITY) ? 0.2 : fmax(0.2, 0.9 * ctl_rcbrt(fmin(E, 1e10)));
        h *= fac;
        return false;
    }

    // state between attempts: the evolving variables, t, h, the attempt
    // count (negated after a rejected attempt: the next accepted step may
    // then not grow h).  f and J are recomputed at x on restore, the same
    // bits as the values the uninterrupted integration keeps.
    static constexpr int SAVE_WORDS = N + 3;
    __device__ void save(double *w) const {
#pragma unroll
        for (int jj = 0; jj < N; jj++) w[jj] = x[S::act(jj)];
        w[N] = t;
        w[N + 1] = h;
        w[N + 2] = rejected ? -(double)nsteps : (double)nsteps;
    }
    __device__ void restore(const double *w) {
#pragma unroll
        for (int jj = 0; jj < N; jj++) x[S::act(jj)] = w[jj];
        t = w[N];
        h = w[N + 1];
        nsteps = (int)fabs(w[N + 2]);
        rejected = w[N + 2] < 0.0;
        need_fj = true;
        dead = false;
    }

    // clamps of src/solve.c:208-213
    __device__ void clamp() {
        x[0] = fmin(1., fmax(0., x[0]));
        x[1] = fmin(1., fmax(0., x[1]));
        x[2] = fmin(1. - x[1], fmax(0., x[2]));
        x[3] = fmax(0., x[3]);
        x[4] = fmax(0., x[4]);
        x[5] = fmax(0., x[5]);
    }
};


// Step set-up of one particle, src/loop.c:662-711
__device__ __forceinline__ void load_step(const DevState &S, uint32_t i, int64_t tick, double tick_time,
                                          const DevPhys &P, double x0[7], StepIn &s) {
    const double lH = S.lH[i], lHeI = S.lHeI[i], lHeII = S.lHeII[i];
    const double rho = S.rho[i] * P.scaling_fac3_inv;
    s.time = (tick - S.lasthit[i]) * tick_time;
    const double ie0 = S.ie[i];
    s.T = P.fake_temperature > 0 ? P.fake_temperature : ieye2T(P, ie0, lambda_ye(P, lH, lHeI, lHeII));
    s.yGdepHI = S.yGdepHI[i];
    s.yGdepHeI = S.yGdepHeI[i];
    s.yGdepHeII = S.yGdepHeII[i];
    s.heat = S.heat[i];
    s.yeMET = S.yeMET[i];
    s.nH = P.C_H_PER_MASS * rho;
    rates_per_time(s);
    x0[0] = lH;
    x0[1] = lHeI;
    x0[2] = lHeII;
    x0[3] = (double)S.yGrecHII[i];
    x0[4] = (double)S.yGrecHeII[i];
    x0[5] = (double)S.yGrecHeIII[i];
    x0[6] = ie0;
}

// write-back, src/loop.c:745-766 (step_remain = 0 on this path)
template <bool ISO, bool ADI>
__device__ __forceinline__ void store_step(const DevState &S, uint32_t i, int64_t tick, const DevPhys &P,
                                           const double x[7], double T) {
    S.yGdepHI[i] = 0.0f;
    S.yGdepHeI[i] = 0.0f;
    S.yGdepHeII[i] = 0.0f;
    S.yGrecHII[i] = (float)x[3];
    S.yGrecHeII[i] = (float)x[4];
    S.yGrecHeIII[i] = (float)x[5];
    const double nlH = fmax(0.0, fmin(1.0, x[0]));
    const double nlHeI = fmax(0.0, fmin(1.0, x[1]));
    const double nlHeII = fmax(0.0, fmin(1.0 - nlHeI, x[2]));
    S.lH[i] = nlH;
    S.lHeI[i] = nlHeI;
    S.lHeII[i] = nlHeII;
    if (S.dv_he == 1) {
        DepViewHe *v = (DepViewHe *)S.dv + i;
        v->xHI = nlH;
        v->xHeI = nlHeI;
        v->xHeII = nlHeII;
    } else if (S.dv_he == 0) {
        ((DepViewH *)S.dv)[i].xHI = nlH;
    }
    float ie_new = ADI ? S.ie[i] : (float)x[6];
    S.lasthit[i] = tick;
    if (ISO) ie_new = (float)Tye2ie(P, T, lambda_ye(P, nlH, nlHeI, nlHeII));
    S.ie[i] = ie_new;
    S.heat[i] = 0.0f;
    S.active[i] = 0;
}

// block totals of the solver counters, then one atomic per counter and block
// (every thread of the 256-thread block must call this)
__device__ __forceinline__ void count_steps(DevStats *st, int steps, int fail, int done) {
    __shared__ int part[3][4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        steps += __shfl_xor(steps, o, 64);
        fail += __shfl_xor(fail, o, 64);
        done += __shfl_xor(done, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][w] = steps;
        part[1][w] = fail;
        part[2][w] = done;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        steps = part[0][0] + part[0][1] + part[0][2] + part[0][3];
        fail = part[1][0] + part[1][1] + part[1][2] + part[1][3];
        done = part[2][0] + part[2][1] + part[2][2] + part[2][3];
        if (done) {
            DevStats *s = stat_slot(st);
            atomicAdd(&s->active_count, (unsigned long long)done);
            atomicAdd(&s->solver_steps, (unsigned long long)steps);
            if (fail) atomicAdd(&s->solver_failures, (unsigned long long)fail);
        }
    }
}

// Capture of exact steps for training the surrogate.  Before the update,
// k_cap_pre writes the features of every active particle into a claimed row
// (one slot claim per wavefront) and remembers the row; after it,
// k_cap_post adds the targets from what store_step left: the fractions and
// the temperature after the step.
__global__ __launch_bounds__(256) void k_cap_pre(DevState S, int64_t npar, const DevTick *__restrict__ tk,
                                                 DevPhys P, MlpDev M, uint32_t *__restrict__ row_of,
                                                 const int32_t *halt) {
    if (halted(halt)) return;
    const int64_t tick = tk->tick;
    const double tick_time = tk->tick_time;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool on = i < npar && S.active[i];
    if (i < npar) row_of[i] = 0xFFFFFFFFu;
    const unsigned long long m = __ballot(on);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(M.cap_n, (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    if (!on) return;
    const unsigned long long slot = base + __popcll(m & ((1ull << lane) - 1));
    if (slot >= M.cap_max) return;
    double x0[7];
    StepIn s;
    load_step(S, (uint32_t)i, tick, tick_time, P, x0, s);
    float f[MLP_IN];
    mlp_features(x0[0], x0[1], x0[2], s.T, s.nH, s.yGdepHI, s.yGdepHeI, s.yGdepHeII, s.heat, s.yeMET, s.time, f);
    float *r = M.cap + slot * MLP_ROW;
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) r[k] = f[k];
    row_of[i] = (uint32_t)slot;
}

__global__ __launch_bounds__(256) void k_cap_post(DevState S, int64_t npar, DevPhys P, MlpDev M,
                                                  const uint32_t *__restrict__ row_of, int iso, const int32_t *halt) {
    if (halted(halt)) return;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= npar) return;
    const uint32_t slot = row_of[i];
    if (slot == 0xFFFFFFFFu) return;
    float *r = M.cap + (size_t)slot * MLP_ROW;
    const double lH = S.lH[i], lHeI = S.lHeI[i], lHeII = S.lHeII[i];
    // isothermal: T is the step's T (feature 4); otherwise from ie and ye
    const double T1 = iso ? exp10((double)r[4]) : ieye2T(P, S.ie[i], lambda_ye(P, lH, lHeI, lHeII));
    r[MLP_IN + 0] = (float)lH;
    r[MLP_IN + 1] = (float)lHeI;
    r[MLP_IN + 2] = (float)lHeII;
    r[MLP_IN + 3] = iso ? r[4] : mlp_lg(T1, 1e-30);
}

// Chemistry of the active set, phase 1: one thread per particle in tree
// order, exactly one step attempt (the whole step, h = 1).  Most particles
// finish with it; the others (attempt rejected, or accepted short of t = 1)
// save their integrator state and are marked active = 2 for phase 2, whose
// lane-persistent waves do not wait for their slowest lane.  With a single
// attempt no lane of a phase-1 wavefront waits for another: attempt counts
// range from 1 to over 100, and continuing to the first accepted step here
// made each wavefront run as many attempts as its slowest lane (256^3 H + He:
// phase 1 4.28 ms that way).
template <bool ISO, bool ADI, bool OTS, bool HONLY>
__global__ __launch_bounds__(256) void k_update(DevState S, int64_t npar, const DevTick *__restrict__ tk,
                                                DevTables T, DevPhys P, DevStats *__restrict__ st,
                                                double *__restrict__ saved, uint32_t *__restrict__ list,
                                                uint32_t *__restrict__ nlist) {
    using G = Integ<ISO, ADI, OTS, HONLY>;
    if (halted(&tk->flags)) return;  // this tick's trace overflowed: replayed later
    const int64_t tick = tk->tick;
    const double tick_time = tk->tick_time;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int steps = 0, fail = 0, done = 0;
    if (closed_form<ISO, OTS, HONLY>()) {
        if (i < npar && S.active[i]) {
            double x0[7], x[7];
            StepIn s;
            load_step(S, (uint32_t)i, tick, tick_time, P, x0, s);
            riccati_step(x0, s, T, x);
            steps = 1;
            done = 1;
            store_step<ISO, ADI>(S, (uint32_t)i, tick, P, x, s.T);
        }
    } else if (i < npar && S.active[i]) {
        double x0[7];
        StepIn s;
        load_step(S, (uint32_t)i, tick, tick_time, P, x0, s);
        G g;
        g.init(x0, s, T);
        if (g.attempt(T, P, 1e-4)) {
            g.clamp();
            steps = g.nsteps;
            fail = g.t < 1.0;
            done = 1;
            store_step<ISO, ADI>(S, (uint32_t)i, tick, P, g.x, s.T);
        } else {
            g.save(saved + (size_t)i * G::SAVE_WORDS);
            S.active[i] = 2;
        }
    }
    // the unfinished particles into the phase-2 list, one claim per wavefront
    // (the list's order does not matter: each particle is independent)
    if (!closed_form<ISO, OTS, HONLY>()) {
        const bool left = i < npar && S.active[i] == 2;
        const unsigned long long m = __ballot(left);
        if (m) {
            const int lane = threadIdx.x & 63;
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(nlist, (uint32_t)__popcll(m));
            base = __shfl(base, leader, 64);
            if (left) list[base + __popcll(m & ((1ull << lane) - 1))] = (uint32_t)i;
        }
    }
    count_steps(st, steps, fail, done);
}

// phase 2: the listed particles, lane by lane.  Integrations need from one
// to over a hundred steps, so a wavefront that ran one particle per lane to
// the end would wait for its slowest lane; here a lane whose particle has
// finished takes the next one from the list (one claim per wavefront for all
// its idle lanes), and every iteration is one step attempt on every busy
// lane.  nlist[0] is the list's length, nlist[1] the next unclaimed entry.
template <bool ISO, bool ADI, bool OTS, bool HONLY>
__global__ __launch_bounds__(256) void k_update_rest(DevState S, const uint32_t *__restrict__ list,
                                                     uint32_t *__restrict__ nlist,
                                                     const DevTick *__restrict__ tk, DevTables T, DevPhys P,
                                                     DevStats *__restrict__ st, const double *__restrict__ saved) {
    using G = Integ<ISO, ADI, OTS, HONLY>;
    if (halted(&tk->flags)) return;
    const int64_t tick = tk->tick;
    const double tick_time = tk->tick_time;
    const uint32_t n = nlist[0];
    const int lane = threadIdx.x & 63;
    int steps = 0, fail = 0, done = 0;
    G g;
    uint32_t i = 0;
    bool have = false, drained = false;
    for (;;) {
        const unsigned long long idle = __ballot(!have);
        if (idle && !drained) {
            const int leader = __ffsll((long long)idle) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&nlist[1], (uint32_t)__popcll(idle));
            base = __shfl(base, leader, 64);
            drained = base + (uint32_t)__popcll(idle) >= n;
            const uint32_t k = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1));
            // an entry the surrogate has stored since (active 0) is skipped
            if (!have && k < n && S.active[list[k]] == 2) {
                i = list[k];
                double x0[7];
                StepIn s;
                load_step(S, i, tick, tick_time, P, x0, s);
                g.init(x0, s, T);
                g.restore(saved + (size_t)i * G::SAVE_WORDS);
                have = true;
            }
        }
        if (!__ballot(have)) break;
        if (have && g.attempt(T, P, 1e-4)) {
            g.clamp();
            steps += g.nsteps;
            fail += g.t < 1.0;
            done++;
            store_step<ISO, ADI>(S, i, tick, P, g.x, g.s.T);
            have = false;
        }
    }
    count_steps(st, steps, fail, done);
}

template <bool ISO, bool ADI, bool OTS, bool HONLY>
__global__ __launch_bounds__(256) void k_solve(psr_step *__restrict__ steps, int64_t n, DevTables T, DevPhys P,
                                               double rtol) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    psr_step st = steps[i];
    StepIn s;
    s.yGdepHI = st.yGdepHI;
    s.yGdepHeI = st.yGdepHeI;
    s.yGdepHeII = st.yGdepHeII;
    s.heat = st.heat;
    s.time = st.time;
    s.nH = st.nH;
    s.T = st.T;
    s.yeMET = st.yeMET;
    rates_per_time(s);
    const double x0[7] = {st.lambdaH, st.lambdaHeI, st.lambdaHeII, st.yGrecHII, st.yGrecHeII, st.yGrecHeIII,
                          st.ie};
    double xc[7];
    const double *x = xc;
    int ns = 1;
    Integ<ISO, ADI, OTS, HONLY> g;
    if (closed_form<ISO, OTS, HONLY>()) {
        riccati_step(x0, s, T, xc);
    } else {
        g.init(x0, s, T);
        while (!g.attempt(T, P, rtol)) {
        }
        g.clamp();
        x = g.x;
        ns = g.t >= 1.0 ? g.nsteps : -g.nsteps;
    }
    st.lambdaH = x[0];
    st.lambdaHeI = x[1];
    st.lambdaHeII = x[2];
    st.yGrecHII = x[3];
    st.yGrecHeII = x[4];
    st.yGrecHeIII = x[5];
    st.ie = x[6];
    st.step_remain = 0.0;
    st.nsteps = ns;
    steps[i] = st;
}

// Surrogate chemistry (on-the-spot runs): 64 active particles per
// wavefront per round; each lane builds its particle's step and features,
// the wavefront runs the network over the 64 rows on MFMA (weights held in
// registers), and each lane then checks its prediction x1:
//   - the features lie inside the training domain;
//   - the charge / particle-conservation residual, x_HeI + x_HeII - 1
//     (x_HeIII >= 0; x_HII = 1 - x_HI and the electron count follow from the
//     fractions), is at most tol;
//   - with BALANCE, also the ionisation-balance residual of every evolved
//     variable, d = x1 - x0 - (f(x0) + f(x1)) / 2, the trapezoidal defect of
//     dx/dtau = f (src/solve.c:76-141, photo-ionisations as the constant
//     deposit rates): at most tol for the fractions (one-sided where x1 sits
//     at a bound the exact step clamps to, src/solve.c:208-213) and 10 tol
//     relative for ie.
// Accepted steps are stored as step_evolve's would be (src/loop.c:745-766);
// the others keep their active mark for the exact solver that runs next.
// LIST: only the particles the exact solver's first attempt left unfinished
// (the phase-2 list, active = 2; those it rejects continue from their saved
// integrator state); else every active particle (active = 1).
template <bool ISO, bool ADI, bool HONLY, bool BALANCE, bool LIST>
__global__ __launch_bounds__(256) void k_mlp_update(DevState S, int64_t npar, const DevTick *__restrict__ tk,
                                                    DevTables T, DevPhys P, DevStats *__restrict__ st,
                                                    const float *__restrict__ Pm, const float *__restrict__ norm,
                                                    double tol, const uint32_t *__restrict__ list,
                                                    const uint32_t *__restrict__ nlist) {
    __shared__ float s_bias[3 * MLP_HID + MLP_OUT];
    __shared__ float s_feat[4][64][MLP_IN + 1];
    __shared__ float s_out[4][64][MLP_OUT];
    __shared__ int part[2][4];
    if (halted(&tk->flags)) return;
    const int64_t tick = tk->tick;
    const double tick_time = tk->tick_time;
    for (int k = threadIdx.x; k < 3 * MLP_HID + MLP_OUT; k += blockDim.x) {
        const int l = k / MLP_HID, o = k % MLP_HID;
        s_bias[k] = Pm[(l == 0 ? MLP_B1 : l == 1 ? MLP_B2 : l == 2 ? MLP_B3 : MLP_B4) + o];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    MlpW w;
    mlp_load_w(Pm, w, lane);
    int nacc = 0, nfb = 0;
    const int64_t nwork = LIST ? min((int64_t)nlist[0], npar) : npar;
    for (int64_t base = ((int64_t)blockIdx.x * 4 + wv) * 64; base < nwork; base += (int64_t)gridDim.x * 256) {
        const int64_t q = base + lane;
        const int64_t i = LIST ? (q < nwork ? (int64_t)list[q] : npar) : q;
        const bool act = i < npar && S.active[i] == (LIST ? 2 : 1);
        if (!__ballot(act)) continue;
        double x0[7];
        StepIn s;
        float f[MLP_IN];
        bool inb = true;
        if (act) {
            load_step(S, (uint32_t)i, tick, tick_time, P, x0, s);
            mlp_features(x0[0], x0[1], x0[2], s.T, s.nH, s.yGdepHI, s.yGdepHeI, s.yGdepHeII, s.heat, s.yeMET,
                         s.time, f);
        }
#pragma unroll
        for (int k = 0; k < MLP_IN; k++) {
            const float v = act ? f[k] : 0.0f;
            inb = inb && v >= norm[2 * MLP_IN + k] && v <= norm[3 * MLP_IN + k];
            s_feat[wv][lane][k] = act ? (v - norm[k]) * norm[MLP_IN + k] : 0.0f;
        }
#pragma unroll
        for (int sub = 0; sub < 4; sub++) {
            mlp_bf16x4 xin[2];
            mlp_inputs(&s_feat[wv][16 * sub][0], MLP_IN + 1, lane, xin);
            const mlp_f32x4 o = mlp_forward16(w, s_bias, xin, lane);
            if (lane < 16)
#pragma unroll
                for (int j = 0; j < MLP_OUT; j++) s_out[wv][16 * sub + lane][j] = o[j];
        }
        if (act) {
            double pr[MLP_OUT];
            mlp_predict(f, s_out[wv][lane], pr);
            double x1[7] = {pr[0], HONLY ? x0[1] : pr[1], HONLY ? x0[2] : pr[2], x0[3], x0[4], x0[5], x0[6]};
            if (!ISO && !ADI) x1[6] = Tye2ie(P, exp10(pr[3]), lambda_ye(P, x1[0], x1[1], x1[2]));
            bool ok = inb && isfinite(x1[0]) && isfinite(x1[6]);
            if (!HONLY) ok = ok && x1[1] + x1[2] - 1.0 <= tol;
            if (BALANCE) {
                Rates Riso;
                if (ISO) rates_at<true>(T, log10(s.T), s.nH, Riso);
                double f0[7], f1[7];
                rhs<ISO, ADI, true, HONLY>(x0, f0, s, T, P, Riso);
                rhs<ISO, ADI, true, HONLY>(x1, f1, s, T, P, Riso);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    if (HONLY && k > 0) break;
                    const double d = x1[k] - x0[k] - 0.5 * (f0[k] + f1[k]);
                    const bool lo = x1[k] <= 1e-6, hi = x1[k] >= 1.0 - 1e-6;
                    ok = ok && isfinite(d) && (lo ? d >= -tol : hi ? d <= tol : fabs(d) <= tol);
                }
                if (!ISO && !ADI) {
                    const double d = x1[6] - x0[6] - 0.5 * (f0[6] + f1[6]);
                    ok = ok && isfinite(d) && fabs(d) <= 10.0 * tol * fmax(x0[6], x1[6]);
                }
            }
    if (ok) {
                store_step<ISO, ADI>(S, (uint32_t)i, tick, P, x1, s.T);
                nacc++;
            } else {
                nfb++;
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        nacc += __shfl_xor(nacc, o, 64);
        nfb += __shfl_xor(nfb, o, 64);
    }
    if (lane == 0) {
        part[0][wv] = nacc;
        part[1][wv] = nfb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        nacc = part[0][0] + part[0][1] + part[0][2] + part[0][3];
        nfb = part[1][0] + part[1][1] + part[1][2] + part[1][3];
        DevStats *sp = stat_slot(st);
        if (nacc) {
            atomicAdd(&sp->surrogate_count, (unsigned long long)nacc);
            atomicAdd(&sp->active_count, (unsigned long long)nacc);
        }
        if (nfb) atomicAdd(&sp->surrogate_fallbacks, (unsigned long long)nfb);
    }
}

}  // namespace

__global__ void k_zero_u32x2(uint32_t *p) {
    if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

#define CHEM_DISPATCH(KERNEL, ...)                                                                       \
    do {                                                                                                 \
        const int key_ = (c->F.isothermal ? 8 : 0) | (c->F.adiabatic ? 4 : 0) |                          \
                         (c->F.on_the_spot ? 2 : 0) | (c->F.h_only ? 1 : 0);                             \
        switch (key_) {                                                                                  \
        case 0: KERNEL<false, false, false, false><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break;     \
        case 1: KERNEL<false, false, false, true><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break;      \
        case 2: KERNEL<false, false, true, false><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break;      \
        case 3: KERNEL<false, false, true, true><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break;       \
        case 4: KERNEL<false, true, false, false><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break;      \
        case 5: KERNEL<false, true, false, true><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break;       \
        case 6: KERNEL<false, true, true, false><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break;       \
        case 7: KERNEL<false, true, true, true><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break;        \
        case 8: case 12: KERNEL<true, false, false, false><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break; \
        case 9: case 13: KERNEL<true, false, false, true><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break;  \
        case 10: case 14: KERNEL<true, false, true, false><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break; \
        default: KERNEL<true, false, true, true><<<grid, 256, 0, c->stream>>>(__VA_ARGS__); break;           \
        }                                                                                                \
    } while (0)

int psr_prepare_update(psr_ctx *c) {
    const int64_t n = c->npar;
    if (c->qlist_cap < n) {
        if (int rc = psr_join(c)) return rc;
        PSR_HIP(c, hipStreamSynchronize(c->stream));  // earlier ticks may still use them
        if (c->d_qlist) PSR_HIP(c, hipFree(c->d_qlist));
        if (c->d_saved) PSR_HIP(c, hipFree(c->d_saved));
        c->d_qlist = nullptr;
        c->d_saved = nullptr;
        c->qlist_cap = 0;
        PSR_HIP(c, hipMalloc(&c->d_qlist, sizeof(uint32_t) * n));
        PSR_HIP(c, hipMalloc(&c->d_saved, sizeof(double) * 10 * n));  // SAVE_WORDS <= 10
        c->qlist_cap = n;
    }
    if (!c->d_q) PSR_HIP(c, hipMalloc(&c->d_q, sizeof(uint32_t) * 2));
    if (psr_mlp_dev(c).cap && !psr_mlp_rowof(c, n)) return PSR_ERR_HIP;
    return PSR_OK;
}

// phase 1 (k_update: the closed form, or one exact attempt per active
// particle, listing the unfinished) or phase 2 (k_update_rest over the list);
// no training capture
int psr_launch_update_phase(psr_ctx *c, int phase) {
    if (c->npar == 0) return PSR_OK;
    const DevPhys P = psr_dev_phys(c);
    const DevState S = c->state();
    const int64_t n = c->npar;
    const DevTick *tk = c->d_tick;
    int rc = psr_prepare_update(c);
    if (rc) return rc;
    const bool closed = c->F.isothermal && c->F.on_the_spot && c->F.h_only;
    unsigned grid;
    if (phase == 1) {
        // the list starts empty (the closed form lists nothing)
        k_zero_u32x2<<<1, 64, 0, c->stream>>>(c->d_q);
        PSR_HIP(c, hipGetLastError());
        grid = (unsigned)((n + 255) / 256);
        CHEM_DISPATCH(k_update, S, n, tk, c->T, P, c->d_stats, c->d_saved, c->d_qlist, c->d_q);
    } else {
        if (closed) return PSR_OK;
        grid = (unsigned)std::max(1, c->cu_count * 8);
        CHEM_DISPATCH(k_update_rest, S, c->d_qlist, c->d_q, tk, c->T, P, c->d_stats, c->d_saved);
    }
    PSR_HIP(c, hipGetLastError());
    return PSR_OK;
}

int psr_launch_update(psr_ctx *c, bool timed) {
    if (c->npar == 0) return PSR_OK;
    const DevPhys P = psr_dev_phys(c);
    const DevState S = c->state();
    const MlpDev M = psr_mlp_dev(c);
    const int64_t n = c->npar;
    const DevTick *tk = c->d_tick;
    int rc = psr_prepare_update(c);
    if (rc) return rc;
    if (timed) PSR_EV(c, 3, 0);
    unsigned grid = (unsigned)((n + 255) / 256);
    const int32_t *halt = &c->d_tick->flags;  // as this tick's deposit left it
    uint32_t *row_of = nullptr;
    if (M.cap) {
        row_of = psr_mlp_rowof(c, n);
        k_cap_pre<<<grid, 256, 0, c->stream>>>(S, n, tk, P, M, row_of, halt);
        PSR_HIP(c, hipGetLastError());
    }
    const bool closed = c->F.isothermal && c->F.on_the_spot && c->F.h_only;
    if (!closed) {
        k_zero_u32x2<<<1, 64, 0, c->stream>>>(c->d_q);
        PSR_HIP(c, hipGetLastError());
    }
    CHEM_DISPATCH(k_update, S, n, tk, c->T, P, c->d_stats, c->d_saved, c->d_qlist, c->d_q);
    PSR_HIP(c, hipGetLastError());
    if (!closed) {
        // phase 2: the particles k_update listed as unfinished
        grid = (unsigned)std::max(1, c->cu_count * 8);
        CHEM_DISPATCH(k_update_rest, S, c->d_qlist, c->d_q, tk, c->T, P, c->d_stats, c->d_saved);
        PSR_HIP(c, hipGetLastError());
    }
    if (M.cap) {
        grid = (unsigned)((n + 255) / 256);
        k_cap_post<<<grid, 256, 0, c->stream>>>(S, n, P, M, row_of, c->F.isothermal ? 1 : 0, halt);
        PSR_HIP(c, hipGetLastError());
    }
    if (timed) PSR_EV(c, 3, 1);
    return PSR_OK;
}

// surrogate chemistry for the active set, then the exact solver for the
// particles whose prediction was rejected (and for every particle when the
// run is not on-the-spot, whose recombination counters the network does not
// predict)
static int mlp_launch(psr_ctx *c, bool list) {
    const DevPhys P = psr_dev_phys(c);
    const DevState S = c->state();
    const MlpDev M = psr_mlp_dev(c);
    const int64_t n = c->npar;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, c->cu_count * 4));
    const int key = (c->F.isothermal ? 4 : 0) | (c->F.adiabatic ? 2 : 0) | (c->F.h_only ? 1 : 0);
#define MLP_K(I, A, H, L)                                                                                   \
    do {                                                                                                    \
        if (c->mlp_balance)                                                                                 \
            k_mlp_update<I, A, H, true, L><<<grid, 256, 0, c->stream>>>(S, n, c->d_tick, c->T, P,          \
                                                                        c->d_stats, M.params, M.norm,      \
                                                                        c->mlp_residual, c->d_qlist, c->d_q); \
        else                                                                                                \
            k_mlp_update<I, A, H, false, L><<<grid, 256, 0, c->stream>>>(S, n, c->d_tick, c->T, P,         \
                                                                         c->d_stats, M.params, M.norm,     \
                                                                         c->mlp_residual, c->d_qlist, c->d_q); \
    } while (0)
#define MLP_KL(I, A, H)              \
    do {                             \
        if (list) MLP_K(I, A, H, true);  \
        else MLP_K(I, A, H, false);      \
    } while (0)
    switch (key) {
    case 0: MLP_KL(false, false, false); break;
    case 1: MLP_KL(false, false, true); break;
    case 2: MLP_KL(false, true, false); break;
    case 3: MLP_KL(false, true, true); break;
    case 4: case 6: MLP_KL(true, false, false); break;
    default: MLP_KL(true, false, true); break;
    }
#undef MLP_KL
#undef MLP_K
    PSR_HIP(c, hipGetLastError());
    return PSR_OK;
}

// surrogate chemistry (on-the-spot runs; otherwise the exact solver, since
// the network does not predict the recombination counters).  stiff: the
// exact solver's first attempt runs for every active particle, the network
// for the ones it leaves unfinished, the exact solver again for those the
// network's checks reject; else the network for every active particle and
// the exact solver for the rejected
int psr_mlp_update_dev(psr_ctx *c, bool stiff) {
    if (!psr_mlp_ready(c)) PSR_FAIL(c, PSR_ERR_STATE, "surrogate chemistry: no trained network (psr_mlp_train)");
    if (c->npar == 0) return PSR_OK;
    if (!c->F.on_the_spot) return psr_launch_update(c);
    PSR_EV(c, 3, 0);
    int rc;
    if (stiff) {
        // the closed form (isothermal hydrogen on the spot) finishes every
        // particle in phase 1: nothing is left for the network
        const bool closed = c->F.isothermal && c->F.on_the_spot && c->F.h_only;
        if ((rc = psr_launch_update_phase(c, 1))) return rc;
        if (!closed) {
            if ((rc = mlp_launch(c, true))) return rc;
            rc = psr_launch_update_phase(c, 2);
        }
    } else {
        if ((rc = mlp_launch(c, false))) return rc;
        rc = psr_launch_update(c, false);
    }
    PSR_EV(c, 3, 1);
    return rc;
}

int psr_solve_steps_dev(psr_ctx *c, psr_step *d_steps, int64_t n, double rtol) {
    if (n == 0) return PSR_OK;
    const DevPhys P = psr_dev_phys(c);
    const unsigned grid = (unsigned)((n + 255) / 256);
    CHEM_DISPATCH(k_solve, d_steps, n, c->T, P, rtol);
    PSR_HIP(c, hipGetLastError());
    return PSR_OK;
}

// ---- allocation-free step_evolve batches (psr_solve_steps_noalloc)
// The steps travel through this fixed device array instead of a hipMalloc'd
// staging buffer.  Each GPU holds its own copy for the life of the process;
// host code never uses the symbol itself, only the per-device address that
// psr_noalloc_init resolves (hipGetSymbolAddress) when a context is created.
static __device__ __align__(256) psr_step scratch[PSR_NOALLOC_MAX_STEPS];
static constexpr int kMaxDevices = 64;
static void *g_scratch_addr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// the calling thread's current device must be `device`
int psr_noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return PSR_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_addr[device]) return PSR_OK;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return PSR_ERR_HIP;
    g_scratch_addr[device] = ptr;
    return PSR_OK;
}

extern "C" int psr_solve_steps_noalloc(psr_ctx *c, psr_step *steps, int64_t n, double rtol) {
    if (!c || (n > 0 && !steps) || n < 0) return PSR_ERR_ARG;
    if (n > PSR_NOALLOC_MAX_STEPS)
        PSR_FAIL(c, PSR_ERR_CAPACITY, "psr_solve_steps_noalloc: at most PSR_NOALLOC_MAX_STEPS (32768) "
                                      "steps per call");
    if (!c->have_physics) PSR_FAIL(c, PSR_ERR_STATE, "no physics: call psr_set_physics");
    if (rtol <= 0) rtol = 1e-4;
    if (n == 0) return PSR_OK;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    psr_step *d = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (c->device < kMaxDevices) d = static_cast<psr_step *>(g_scratch_addr[c->device]);
    }
    if (!d) PSR_FAIL(c, PSR_ERR_STATE, "psr_solve_steps_noalloc: device scratch address not resolved");
    // every entry the kernel reads is overwritten by this call's input first
    hipError_t e = hipMemcpyAsync(d, steps, sizeof(psr_step) * n, hipMemcpyHostToDevice, c->stream);
    int rc = PSR_OK;
    if (e == hipSuccess) rc = psr_solve_steps_dev(c, d, n, rtol);
    if (e == hipSuccess && rc == PSR_OK)
        e = hipMemcpyAsync(steps, d, sizeof(psr_step) * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
        PSR_FAIL(c, PSR_ERR_HIP, std::string("psr_solve_steps_noalloc: ") + hipGetErrorString(e));
    if (rc) return rc;
    for (int64_t i = 0; i < n; i++)
        if (steps[i].nsteps < 0) {
            c->err = "psr_solve_steps_noalloc: a step did not reach t = 1";
            return PSR_ERR_SOLVER;
        }
    return PSR_OK;
}

extern "C" int psr_solve_steps(psr_ctx *c, psr_step *steps, int64_t n, double rtol) {
    if (!c || (n > 0 && !steps) || n < 0) return PSR_ERR_ARG;
    if (!c->have_physics) PSR_FAIL(c, PSR_ERR_STATE, "no physics: call psr_set_physics");
    if (rtol <= 0) rtol = 1e-4;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    psr_step *d = nullptr;
    PSR_HIP(c, hipMalloc(&d, sizeof(psr_step) * std::max<int64_t>(n, 1)));
    hipError_t e = hipMemcpyAsync(d, steps, sizeof(psr_step) * n, hipMemcpyHostToDevice, c->stream);
    int rc = PSR_OK;
    if (e == hipSuccess) rc = psr_solve_steps_dev(c, d, n, rtol);
    if (e == hipSuccess && rc == PSR_OK)
        e = hipMemcpyAsync(steps, d, sizeof(psr_step) * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) PSR_FAIL(c, PSR_ERR_HIP, std::string("psr_solve_steps: ") + hipGetErrorString(e));
    if (rc) return rc;
    for (int64_t i = 0; i < n; i++)
        if (steps[i].nsteps < 0) {
            c->err = "psr_solve_steps: a step did not reach t = 1";
            return PSR_ERR_SOLVER;
        }
    return PSR_OK;
}
