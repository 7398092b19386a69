// sse_adjoint.h -- reverse-mode (adjoint) derivatives of the Euler physics of
// sse_physics.h, written by hand from the forward formulas:
//   logmean / inv_logmean           src/ConservationLaws/ConservationLaws.jl:132-156
//   EC two-point flux (contracted)  src/ConservationLaws/euler_navierstokes.jl:171-225
//   Lax-Friedrichs interface flux   src/ConservationLaws/ConservationLaws.jl:75-101,
//                                   wave speed euler_navierstokes.jl:133-150
//   conservative_to_entropy!        euler_navierstokes.jl:100-113
//   entropy_to_conservative!        euler_navierstokes.jl:115-131
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

// logmean(x, y) = N / Dn from x, y, their logarithms and reciprocals (no logarithm or
// division evaluated), with the numerators of its partial derivatives:
//   d logmean / dx = gx / Dn^2,  d logmean / dy = gy / Dn^2.
// The branch is the forward's (logmean_frac: |y - x| < 1e-2 (x + y)):
//   series: N = 105 s^7, Dn = 210 s^6 + 70 d^2 s^4 + 42 d^4 s^2 + 30 d^6 (s = x + y,
//           d = y - x; the reference's series scaled by s^6), g = N_s Dn - N (Dn_s -+ Dn_d);
//   log:    N = d, Dn = log y - log x, gx = -Dn + d / x, gy = Dn - d / y.
// The inverse logarithmic mean is Dn / N, with derivatives -gx / N^2 and -gy / N^2.
__device__ __forceinline__ void logmean_frac_vg(double x, double y, double lx, double ly, double ix, double iy,
                                                double& N, double& Dn, double& gx, double& gy) {
    const double s = x + y, d = y - x;
    const double s2 = s * s, d2 = d * d;
    if (d2 < 1.0e-4 * s2) {
        const double s4 = s2 * s2, d4 = d2 * d2;
        N = 105.0 * s * s4 * s2;
        Dn = ((210.0 * s2 + 70.0 * d2) * s2 + 42.0 * d4) * s2 + 30.0 * d4 * d2;
        const double Ns_s = 735.0 * s4 * s2;
        const double Ds_s = ((1260.0 * s2 + 280.0 * d2) * s2 + 84.0 * d4) * s;
        const double Ds_d = ((140.0 * s2 + 168.0 * d2) * s2 + 180.0 * d4) * d;
        gx = Ns_s * Dn - N * (Ds_s - Ds_d);
        gy = Ns_s * Dn - N * (Ds_s + Ds_d);
    } else {
        const double L = ly - lx;
        N = d;
        Dn = L;
        gx = -L + d * ix;
        gy = L - d * iy;
    }
}

template <int D>
struct EulerAdj {
    static constexpr int NC = D + 2;
    // hoisted per-node quantities: rho, v (D), p, log rho, beta = rho / p, log beta, 1 / rho, 1 / p
    static constexpr int NH = D + 7;
    enum { H_RHO = 0, H_V = 1, H_P = D + 1, H_LR = D + 2, H_B = D + 3, H_LB = D + 4, H_IR = D + 5, H_IP = D + 6 };

    __device__ static void hoist_full(const LawParams& P, const double* u, double* h) {
        const double ir = rcp_refined(u[0]);
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            h[H_V + m] = u[m + 1] * ir;
            s += h[H_V + m] * u[m + 1];
        }
        const double p = P.gm1 * (u[D + 1] - 0.5 * s);
        const double ip = rcp_refined(p);
        h[H_RHO] = u[0];
        h[H_P] = p;
        h[H_LR] = log(u[0]);
        h[H_B] = u[0] * ip;
        h[H_LB] = log(h[H_B]);
        h[H_IR] = ir;
        h[H_IP] = ip;
    }

    // The parts of the pull-back of Gb through G = sum_n a_n F_n(L, R) (hoisted
    // states) that both sides share: the averages, the log-means and their gradients.
    struct EcAdjShared {
        double f, rhoavgb, avb, Ccb, pavgb, drx, dry, dbx, dby, avL, avR;
    };
    __device__ static EcAdjShared ec_dot_adj_shared(const LawParams& P, const double* hL, const double* hR,
                                                    const double* a, const double* Gb, double* Vb) {
        EcAdjShared c;
        double vv = 0.0, av = 0.0;
        c.avL = 0.0;
        c.avR = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            const double vl = hL[H_V + m], vr = hR[H_V + m];
            vv += vl * vr;
            Vb[m] = 0.5 * (vl + vr);
            av += a[m] * Vb[m];
            c.avL += a[m] * vl;
            c.avR += a[m] * vr;
        }
        const double pL = hL[H_P], pR = hR[H_P];
        // rho_avg = logmean(rho) = N1 / D1 and ilm = 1 / logmean(beta) = D2 / N2 with their
        // derivatives, from one reciprocal of D1 N2 (as in the forward flux)
        double N1, D1, gx1, gy1, N2, D2, gx2, gy2;
        logmean_frac_vg(hL[H_RHO], hR[H_RHO], hL[H_LR], hR[H_LR], hL[H_IR], hR[H_IR], N1, D1, gx1, gy1);
        logmean_frac_vg(hL[H_B], hR[H_B], hL[H_LB], hR[H_LB], pL * hL[H_IR], pR * hR[H_IR], N2, D2, gx2, gy2);
        const double rr = rcp_refined(D1 * N2);
        const double iD1 = N2 * rr, iN2 = D1 * rr;
        const double rho_avg = N1 * iD1, ilm = D2 * iN2;
        c.drx = gx1 * iD1 * iD1;
        c.dry = gy1 * iD1 * iD1;
        c.dbx = -gx2 * iN2 * iN2;
        c.dby = -gy2 * iN2 * iN2;
        const double Cc = 0.5 * vv + P.inv_gm1 * ilm;
        c.f = rho_avg * av;
        double fb = Gb[0] + Gb[D + 1] * Cc;
        c.pavgb = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            fb += Gb[m + 1] * Vb[m];
            c.pavgb += Gb[m + 1] * a[m];
        }
        c.Ccb = Gb[D + 1] * c.f;
        c.rhoavgb = fb * av;
        c.avb = fb * rho_avg;
        return c;
    }

    // the side-specific rest in primitive variables: adds the pull-back into the (rho, v, p)
    // cotangent ubp = [rho, v_0..v_{D-1}, p] of the left (`left`) or the right state
    __device__ static void ec_dot_adj_side_prim(const LawParams& P, const double* hL, const double* hR,
                                                const double* a, const double* Gb, const EcAdjShared& c, bool left,
                                                double* ubp) {
        const double* hs = left ? hL : hR;
        const double* ho = left ? hR : hL;
        const double po = left ? hR[H_P] : hL[H_P];
        const double avo = left ? c.avR : c.avL;
#pragma unroll
        for (int m = 0; m < D; ++m)
            ubp[1 + m] += 0.5 * Gb[D + 1] * po * a[m] + 0.5 * (Gb[m + 1] * c.f + c.avb * a[m]) + 0.5 * c.Ccb * ho[H_V + m];
        const double bb = P.inv_gm1 * c.Ccb * (left ? c.dbx : c.dby);
        const double ips = hs[H_IP];
        ubp[D + 1] += 0.5 * Gb[D + 1] * avo + 0.5 * c.pavgb - bb * hs[H_B] * ips;
        ubp[0] += c.rhoavgb * (left ? c.drx : c.dry) + bb * ips;
    }

    // (rho, v, p) cotangent ubp of a state with hoisted quantities hs -> conservative
    // cotangent, added into ub: p = gm1 (E - 1/2 rho |v|^2), v = m / rho. Linear in ubp, so
    // a node's cotangents from several terms can be summed first and converted once.
    __device__ static void prim_cot_to_cons(const LawParams& P, const double* hs, const double* ubp, double* ub) {
        const double rho = hs[H_RHO], ir = hs[H_IR];
        const double pb = ubp[D + 1];
        double s = 0.0, vdot = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) s += hs[H_V + m] * hs[H_V + m];
        ub[D + 1] += P.gm1 * pb;
        const double rb = ubp[0] - 0.5 * P.gm1 * s * pb;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            const double vbt = ubp[1 + m] - P.gm1 * rho * hs[H_V + m] * pb;
            ub[m + 1] += vbt * ir;
            vdot += vbt * hs[H_V + m];
        }
        ub[0] += rb - vdot * ir;
    }

    // the side-specific rest: adds the pull-back into the conservative cotangent ub of
    // the left (`left`) or the right state
    __device__ static void ec_dot_adj_side_h(const LawParams& P, const double* hL, const double* hR, const double* a,
                                             const double* Gb, const EcAdjShared& c, bool left, double* ub) {
        double ubp[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) ubp[e] = 0.0;
        ec_dot_adj_side_prim(P, hL, hR, a, Gb, c, left, ubp);
        prim_cot_to_cons(P, left ? hL : hR, ubp, ub);
    }

    // ec_dot_adj_side from hoisted states: adds the pull-back of Gb through
    // G = sum_n a_n F_n(L, R) into the kept side's conservative cotangent ub.
    __device__ static void ec_dot_adj_h(const LawParams& P, const double* hL, const double* hR, const double* a,
                                        const double* Gb, bool left, double* ub) {
        double Vb[D];
        const EcAdjShared c = ec_dot_adj_shared(P, hL, hR, a, Gb, Vb);
        ec_dot_adj_side_h(P, hL, hR, a, Gb, c, left, ub);
    }

    // both sides from one evaluation of the shared part: the pull-back into the left
    // state's cotangent ubL and into the right state's ubR
    __device__ static void ec_dot_adj_h2(const LawParams& P, const double* hL, const double* hR, const double* a,
                                         const double* Gb, double* ubL, double* ubR) {
        double Vb[D];
        const EcAdjShared c = ec_dot_adj_shared(P, hL, hR, a, Gb, Vb);
        ec_dot_adj_side_h(P, hL, hR, a, Gb, c, true, ubL);
        ec_dot_adj_side_h(P, hL, hR, a, Gb, c, false, ubR);
    }

    // ec_dot_adj_h2 into (rho, v, p) cotangents (prim_cot_to_cons converts a node's sum)
    __device__ static void ec_dot_adj_h2p(const LawParams& P, const double* hL, const double* hR, const double* a,
                                          const double* Gb, double* ubpL, double* ubpR) {
        double Vb[D];
        const EcAdjShared c = ec_dot_adj_shared(P, hL, hR, a, Gb, Vb);
        ec_dot_adj_side_prim(P, hL, hR, a, Gb, c, true, ubpL);
        ec_dot_adj_side_prim(P, hL, hR, a, Gb, c, false, ubpR);
    }

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

    // As ec_dot_adj, but adds the pull-back of one side only (left: uL, else uR)
    // into ub: a single code path for every two-point term of a gathering node.
    __device__ static void ec_dot_adj_side(const LawParams& P, const double* uL, const double* uR, const double* a,
                                           const double* Gb, bool left, double* ub) {
        double vL[D], vR[D], Vb[D];
        double sL = 0.0, sR = 0.0, vv = 0.0, av = 0.0, avL = 0.0, avR = 0.0;
        const double iL = 1.0 / uL[0], iR = 1.0 / uR[0];
#pragma unroll
        for (int m = 0; m < D; ++m) {
            vL[m] = uL[m + 1] * iL;
            vR[m] = uR[m + 1] * iR;
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
        double fb = Gb[0] + Gb[D + 1] * Cc;
        double pavgb = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            fb += Gb[m + 1] * Vb[m];
            pavgb += Gb[m + 1] * a[m];
        }
        const double Ccb = Gb[D + 1] * f;
        const double rhoavgb = fb * av, avb = fb * rho_avg;
        // the kept side
        const double* us = left ? uL : uR;
        const double* vs = left ? vL : vR;      // this side's velocity
        const double* vo = left ? vR : vL;      // the other side's velocity
        const double po = left ? pR : pL;       // the other side's pressure
        const double ps = left ? pL : pR;
        const double avo = left ? avR : avL;    // a . v of the other side
        double vb[D];
#pragma unroll
        for (int m = 0; m < D; ++m)
            vb[m] = 0.5 * Gb[D + 1] * po * a[m] + 0.5 * (Gb[m + 1] * f + avb * a[m]) + 0.5 * Ccb * vo[m];
        double pb = 0.5 * Gb[D + 1] * avo + 0.5 * pavgb;
        double gx, gy;
        inv_logmean_grad(bL, bR, gx, gy);
        const double bb = P.inv_gm1 * Ccb * (left ? gx : gy);
        logmean_grad(uL[0], uR[0], gx, gy);
        const double rb = rhoavgb * (left ? gx : gy) + bb / ps;
        pb += -bb * us[0] / (ps * ps);
        (void)vs;
        prim_to_cons(P, us, rb, vb, pb, ub);
    }

    // Dissipation part of the Lax-Friedrichs flux, a (uin - uout) with
    // a = halfλ λmax(uin, uout, n): adds the pull-back of fb into the kept side's
    // cotangent ub (left: uin) and returns d/dhalfλ (fb . a (uin - uout)).
    __device__ static double lf_diss_adj(const LawParams& P, const double* ui, const double* uo, const double* n,
                                         const double* fb, bool left, double* ub) {
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
        const double sgn = left ? 1.0 : -1.0;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            ub[e] += sgn * a * fb[e];
            ab += fb[e] * (ui[e] - uo[e]);
        }
        const double lb = ab * P.half_lambda;
        const double* us = left ? ui : uo;
        if (vin_max == left) {  // the kept side sets the velocity part of λmax
            const double vn = left ? vni : vno;
            const double vnb = lb * (vn < 0.0 ? -1.0 : 1.0);
#pragma unroll
            for (int m = 0; m < D; ++m) ub[m + 1] += vnb * n[m] / us[0];
            ub[0] += -vnb * vn / us[0];
        }
        if (cin_max == left) {  // and the sound-speed part
            const double c = left ? c_in : c_out;
            const double sc = left ? si : so;
            const double pb = lb * P.gamma / (2.0 * c * us[0]);
            ub[0] += -lb * c / (2.0 * us[0]) + P.gm1 * pb * 0.5 * sc / (us[0] * us[0]);
#pragma unroll
            for (int m = 0; m < D; ++m) ub[m + 1] += -P.gm1 * pb * us[m + 1] / us[0];
            ub[D + 1] += P.gm1 * pb;
        }
        return ab * lmax;
    }

    // f* = F(uin, uout) . n + halfλ λmax (uin - uout) (Lax-Friedrichs, or only the
    // first term); fb -> uinb, uoutb, and returns d f* / d halfλ . fb
    __device__ static double lf_adj(const LawParams& P, bool lf, const double* ui, const double* uo, const double* n,
                                    const double* fb, double* uib, double* uob) {
        ec_dot_adj(P, ui, uo, n, fb, uib, uob);
        if (!lf) return 0.0;
        // wave speed (euler_navierstokes.jl:133-150)
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

    // u = e2c(w): ub -> wb (added), from the state's rho, v, p: du/dw is the symmetric
    // matrix A0 of the entropy variables (Hughes, Franca & Mallet, Comput. Methods Appl.
    // Mech. Engrg. 54, 1986), so wb += A0 ub with no logarithm or power.
    __device__ static void e2c_adj_prim(const LawParams& P, double rho, const double* v, double p,
                                        const double* ub, double* wb) {
        double vv = 0.0, vu = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            vv += v[m] * v[m];
            vu += v[m] * ub[m + 1];
        }
        const double E = p * P.inv_gm1 + 0.5 * rho * vv;
        const double rH = E + p;  // rho H
        const double c = rho * ub[0] + rho * vu + rH * ub[D + 1];
        wb[0] += rho * ub[0] + rho * vu + E * ub[D + 1];
#pragma unroll
        for (int m = 0; m < D; ++m) wb[m + 1] += v[m] * c + p * ub[m + 1];
        wb[D + 1] += E * ub[0] + rH * vu + (rH * rH / rho - P.gamma * p * p / rho * P.inv_gm1) * ub[D + 1];
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
