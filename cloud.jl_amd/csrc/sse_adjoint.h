// sse_adjoint.h -- reverse-mode (adjoint) derivatives of the Euler physics of
// sse_physics.h, written by hand from the forward formulas:
//   logmean / inv_logmean           src/ConservationLaws/ConservationLaws.jl:132-156
//   EC two-point flux (contracted)  src/ConservationLaws/euler_navierstokes.jl:358-401
//   Lax-Friedrichs interface flux   src/ConservationLaws/ConservationLaws.jl:75-101,
//                                   wave speed euler_navierstokes.jl:339-356
//   conservative_to_entropy!        euler_navierstokes.jl:306-319
//   entropy_to_conservative!        euler_navierstokes.jl:321-337
// Each *_adj function adds the pull-back of an output cotangent to the input
// cotangents. Derivatives follow the branch the forward evaluation takes
// (series or logarithm in logmean; the larger side in the wave speed, the
// second one on ties), so they are exact derivatives of the evaluated function.
#pragma once

#include "sse_physics.h"

namespace sse {

// d logmean(x, y) / dx and / dy
__device__ __forceinline__ void logmean_grad(double x, double y, double& dx, double& dy) {
    const double sum = x + y;
    const double q = (x - y) / sum;
    const double f2 = (x * (x - 2.0 * y) + y * y) / (x * (x + 2.0 * y) + y * y);
    if (f2 < 1.0e-4) {
        const double den = 210.0 + f2 * (70.0 + f2 * (42.0 + f2 * 30.0));
        const double dden = 70.0 + f2 * (84.0 + f2 * 90.0);
        const double dqx = 2.0 * y / (sum * sum), dqy = -2.0 * x / (sum * sum);
        const double base = 105.0 / den, slope = -sum * 105.0 / (den * den) * dden * 2.0 * q;
        dx = base + slope * dqx;
        dy = base + slope * dqy;
    } else {
        const double L = log(y / x);
        dx = (-L + (y - x) / x) / (L * L);
        dy = (L - (y - x) / y) / (L * L);
    }
}

// d inv_logmean(x, y) / dx and / dy
__device__ __forceinline__ void inv_logmean_grad(double x, double y, double& dx, double& dy) {
    const double sum = x + y;
    const double q = (x - y) / sum;
    const double f2 = (x * (x - 2.0 * y) + y * y) / (x * (x + 2.0 * y) + y * y);
    if (f2 < 1.0e-4) {
        const double den = 210.0 + f2 * (70.0 + f2 * (42.0 + f2 * 30.0));
        const double dden = 70.0 + f2 * (84.0 + f2 * 90.0);
        const double dqx = 2.0 * y / (sum * sum), dqy = -2.0 * x / (sum * sum);
        const double a = dden * 2.0 * q / (sum * 105.0), b = -den / (sum * sum * 105.0);
        dx = a * dqx + b;
        dy = a * dqy + b;
    } else {
        const double L = log(y / x), r = 1.0 / (y - x);
        dx = -r / x + L * r * r;
        dy = r / y - L * r * r;
    }
}

template <int D>
struct EulerAdj {
    static constexpr int NC = D + 2;

    // (rho, v, p) cotangents -> conservative cotangents, added into ub
    __device__ static void prim_to_cons(const LawParams& P, const double* u, double rb, const double* vb, double pb,
                                        double* ub) {
        const double rho = u[0];
        double v[D], s = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            v[m] = u[m + 1] / rho;
            s += v[m] * v[m];
        }
        // p = gm1 (E - 1/2 rho |v|^2)
        double vbt[D];
        ub[D + 1] += P.gm1 * pb;
        rb += -0.5 * P.gm1 * s * pb;
#pragma unroll
        for (int m = 0; m < D; ++m) vbt[m] = vb[m] - P.gm1 * rho * v[m] * pb;
        // v = m / rho
        double vdot = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            ub[m + 1] += vbt[m] / rho;
            vdot += vbt[m] * v[m];
        }
        ub[0] += rb - vdot / rho;
    }

    // G = sum_n a_n F_n(uL, uR) (entropy-conservative flux of Ranocha); Gb -> uLb, uRb
    __device__ static void ec_dot_adj(const LawParams& P, const double* uL, const double* uR, const double* a,
                                      const double* Gb, double* uLb, double* uRb) {
        double vL[D], vR[D], Vb[D];
        double sL = 0.0, sR = 0.0, vv = 0.0, av = 0.0, avL = 0.0, avR = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            vL[m] = uL[m + 1] / uL[0];
            vR[m] = uR[m + 1] / uR[0];
        }
#pragma unroll
        for (int m = 0; m < D; ++m) {
            sL += vL[m] * vL[m];
            sR += vR[m] * vR[m];
            vv += vL[m] * vR[m];
            Vb[m] = 0.5 * (vL[m] + vR[m]);
            av += a[m] * Vb[m];
            avL += a[m] * vL[m];
            avR += a[m] * vR[m];
        }
        const double pL = P.gm1 * (uL[D + 1] - 0.5 * uL[0] * sL);
        const double pR = P.gm1 * (uR[D + 1] - 0.5 * uR[0] * sR);
        const double bL = uL[0] / pL, bR = uR[0] / pR;
        const double rho_avg = logmean(uL[0], uR[0]);
        const double Cc = 0.5 * vv + P.inv_gm1 * inv_logmean(bL, bR);
        const double f = rho_avg * av;
        // reverse
        double fb = Gb[0] + Gb[D + 1] * Cc;
        double pavgb = 0.0;
        double Vbb[D];
#pragma unroll
        for (int m = 0; m < D; ++m) {
            fb += Gb[m + 1] * Vb[m];
            Vbb[m] = Gb[m + 1] * f;
            pavgb += Gb[m + 1] * a[m];
        }
        const double Ccb = Gb[D + 1] * f;
        double pLb = 0.5 * Gb[D + 1] * avR + 0.5 * pavgb;
        double pRb = 0.5 * Gb[D + 1] * avL + 0.5 * pavgb;
        double vLb[D], vRb[D];
        const double rhoavgb = fb * av, avb = fb * rho_avg;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            Vbb[m] += avb * a[m];
            vLb[m] = 0.5 * Gb[D + 1] * pR * a[m] + 0.5 * Vbb[m] + 0.5 * Ccb * vR[m];
            vRb[m] = 0.5 * Gb[D + 1] * pL * a[m] + 0.5 * Vbb[m] + 0.5 * Ccb * vL[m];
        }
        double gx, gy;
        inv_logmean_grad(bL, bR, gx, gy);
        const double bLb = P.inv_gm1 * Ccb * gx, bRb = P.inv_gm1 * Ccb * gy;
        logmean_grad(uL[0], uR[0], gx, gy);
        double rLb = rhoavgb * gx + bLb / pL;
        double rRb = rhoavgb * gy + bRb / pR;
        pLb += -bLb * uL[0] / (pL * pL);
        pRb += -bRb * uR[0] / (pR * pR);
        prim_to_cons(P, uL, rLb, vLb, pLb, uLb);
        prim_to_cons(P, uR, rRb, vRb, pRb, uRb);
    }

    // f* = F(uin, uout) . n + halfλ λmax (uin - uout) (Lax-Friedrichs, or only the
    // first term); fb -> uinb, uoutb, and returns d f* / d halfλ . fb
    __device__ static double lf_adj(const LawParams& P, bool lf, const double* ui, const double* uo, const double* n,
                                    const double* fb, double* uib, double* uob) {
        ec_dot_adj(P, ui, uo, n, fb, uib, uob);
        if (!lf) return 0.0;
        // wave speed (euler_navierstokes.jl:339-356)
        double si = 0.0, so = 0.0, vni = 0.0, vno = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            si += ui[m + 1] * ui[m + 1];
            so += uo[m + 1] * uo[m + 1];
            vni += (ui[m + 1] / ui[0]) * n[m];
            vno += (uo[m + 1] / uo[0]) * n[m];
        }
        const double p_in = P.gm1 * (ui[D + 1] - (0.5 / ui[0]) * si);
        const double p_out = P.gm1 * (uo[D + 1] - (0.5 / uo[0]) * so);
        const double c_in = sqrt(P.gamma * p_in / ui[0]);
        const double c_out = sqrt(P.gamma * p_out / uo[0]);
        const bool vin_max = fabs(vni) > fabs(vno);
        const bool cin_max = c_in > c_out;
        const double lmax = (vin_max ? fabs(vni) : fabs(vno)) + (cin_max ? c_in : c_out);
        const double a = P.half_lambda * lmax;
        double ab = 0.0;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            uib[e] += a * fb[e];
            uob[e] -= a * fb[e];
            ab += fb[e] * (ui[e] - uo[e]);
        }
        const double lb = ab * P.half_lambda;
        // velocity part: d|vn| = sign(vn) dvn, vn = sum_m m_m n_m / rho
        const double* us = vin_max ? ui : uo;
        double* usb = vin_max ? uib : uob;
        const double vn = vin_max ? vni : vno;
        const double vnb = lb * (vn < 0.0 ? -1.0 : 1.0);
#pragma unroll
        for (int m = 0; m < D; ++m) usb[m + 1] += vnb * n[m] / us[0];
        usb[0] += -vnb * vn / us[0];
        // sound speed part: c = sqrt(gamma p / rho), p = gm1 (E - |m|^2 / (2 rho))
        const double* uc = cin_max ? ui : uo;
        double* ucb = cin_max ? uib : uob;
        const double c = cin_max ? c_in : c_out;
        const double pc = cin_max ? p_in : p_out;
        const double sc = cin_max ? si : so;
        const double pb = lb * P.gamma / (2.0 * c * uc[0]);
        ucb[0] += -lb * c / (2.0 * uc[0]) + P.gm1 * pb * 0.5 * sc / (uc[0] * uc[0]);
#pragma unroll
        for (int m = 0; m < D; ++m) ucb[m + 1] += -P.gm1 * pb * uc[m + 1] / uc[0];
        ucb[D + 1] += P.gm1 * pb;
        (void)pc;
        return ab * lmax;
    }

    // w = c2e(u): wb -> ub (added)
    __device__ static void c2e_adj(const LawParams& P, const double* u, const double* wb, double* ub) {
        const double rho = u[0];
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) s += u[m + 1] * u[m + 1];
        const double k = (0.5 / rho) * s;
        const double p = P.gm1 * (u[D + 1] - k);
        const double ip = 1.0 / p;
        double kb = -wb[0] * ip;
        double ipb = -wb[0] * k - wb[D + 1] * rho;
#pragma unroll
        for (int m = 0; m < D; ++m) ipb += wb[m + 1] * u[m + 1];
        const double lpb = -wb[0] * P.inv_gm1, lrb = P.gamma * wb[0] * P.inv_gm1;
        double rb = lrb / rho - wb[D + 1] * ip;
        const double pb = lpb / p - ipb * ip * ip;
        ub[D + 1] += P.gm1 * pb;
        kb += -P.gm1 * pb;
#pragma unroll
        for (int m = 0; m < D; ++m) ub[m + 1] += wb[m + 1] * ip + kb * u[m + 1] / rho;
        rb += -kb * k / rho;
        ub[0] += rb;
    }

    // u = e2c(w): ub -> wb (added)
    __device__ static void e2c_adj(const LawParams& P, const double* win, const double* ub, double* wb) {
        double w[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) w[e] = win[e] * P.gm1;
        double s2 = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) s2 += w[m + 1] * w[m + 1];
        const double kk = s2 / (2.0 * w[D + 1]);
        const double s = P.gamma - w[0] + kk;
        const double rho_e = pow(P.gm1 / pow(-w[D + 1], P.gamma), P.inv_gm1) * exp(-s * P.inv_gm1);
        double rhoeb = -w[D + 1] * ub[0] + (1.0 - kk) * ub[D + 1];
        double wbl[NC];
#pragma unroll
        for (int m = 0; m < D; ++m) {
            rhoeb += w[m + 1] * ub[m + 1];
            wbl[m + 1] = rho_e * ub[m + 1];
        }
        double kkb = -rho_e * ub[D + 1];
        wbl[D + 1] = -rho_e * ub[0];
        // rho_e = exp(Lg), Lg = (log gm1 - gamma log(-w_E) - s) / gm1
        const double Lb = rhoeb * rho_e;
        const double sb = -Lb * P.inv_gm1;
        wbl[D + 1] += -P.gamma * Lb * P.inv_gm1 / w[D + 1];
        wbl[0] = -sb;
        kkb += sb;
        const double s2b = kkb / (2.0 * w[D + 1]);
        wbl[D + 1] += -kkb * s2 / (2.0 * w[D + 1] * w[D + 1]);
#pragma unroll
        for (int m = 0; m < D; ++m) wbl[m + 1] += 2.0 * w[m + 1] * s2b;
#pragma unroll
        for (int e = 0; e < NC; ++e) wb[e] += P.gm1 * wbl[e];
    }
};

}  // namespace sse
