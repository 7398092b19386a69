// sse_physics.h -- per-node and per-pair physics of the residual, device side.
//
// Restates the reference's conservation-law functions:
//   logmean / inv_logmean           src/ConservationLaws/ConservationLaws.jl:132-156
//   conservative_to_entropy!        src/ConservationLaws/euler_navierstokes.jl:100-113
//   entropy_to_conservative!        src/ConservationLaws/euler_navierstokes.jl:115-131
//   wave_speed (Euler / advection)  src/ConservationLaws/euler_navierstokes.jl:133-150,
//                                   src/ConservationLaws/linear_advection_diffusion.jl:106-111
//   compute_two_point_flux (EC Ranocha, conservative, advection)
//                                   src/ConservationLaws/euler_navierstokes.jl:171-225,
//                                   src/ConservationLaws/linear_advection_diffusion.jl:113-128
// Flux tensors are stored F[n * NC + e] (direction n, variable e).
#pragma once

#include <hip/hip_runtime.h>

namespace sse {

enum { LAW_EULER = 1, LAW_ADVECTION = 2 };

struct LawParams {
    double gamma, gm1, inv_gm1, log_gm1;
    double a[3];
    double half_lambda;
    int inviscid_flux;   // sse_inviscid_flux
    int two_point_flux;  // sse_two_point_flux
};

__device__ __forceinline__ double logmean(double x, double y) {
    double f2 = (x * (x - 2.0 * y) + y * y) / (x * (x + 2.0 * y) + y * y);
    if (f2 < 1.0e-4) return (x + y) * 105.0 / (210.0 + f2 * (70.0 + f2 * (42.0 + f2 * 30.0)));
    return (y - x) / log(y / x);
}

__device__ __forceinline__ double inv_logmean(double x, double y) {
    double f2 = (x * (x - 2.0 * y) + y * y) / (x * (x + 2.0 * y) + y * y);
    if (f2 < 1.0e-4) return (210.0 + f2 * (70.0 + f2 * (42.0 + f2 * 30.0))) / ((x + y) * 105.0);
    return log(y / x) / (y - x);
}

template <int D, int LAW>
struct Physics;

template <int D>
struct Physics<D, LAW_EULER> {
    static constexpr int NC = D + 2;

    // entropy-conservative two-point flux of Ranocha (tensor form)
    __device__ static void two_point(const LawParams& P, const double* uL, const double* uR, double* F) {
        if (P.two_point_flux == 2) {  // ConservativeFlux: average of physical fluxes
            double FL[D * NC], FR[D * NC];
            physical_flux(P, uL, FL);
            physical_flux(P, uR, FR);
#pragma unroll
            for (int i = 0; i < D * NC; ++i) F[i] = 0.5 * (FL[i] + FR[i]);
            return;
        }
        double VL[D], VR[D];
        double sL = 0.0, sR = 0.0, vv = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            VL[m] = uL[m + 1] / uL[0];
            VR[m] = uR[m + 1] / uR[0];
        }
#pragma unroll
        for (int m = 0; m < D; ++m) {
            sL += VL[m] * VL[m];
            sR += VR[m] * VR[m];
            vv += VL[m] * VR[m];
        }
        const double pL = P.gm1 * (uL[D + 1] - 0.5 * uL[0] * sL);
        const double pR = P.gm1 * (uR[D + 1] - 0.5 * uR[0] * sR);
        const double rho_avg = logmean(uL[0], uR[0]);
        const double p_avg = 0.5 * (pL + pR);
        const double Cc = 0.5 * vv + P.inv_gm1 * inv_logmean(uL[0] / pL, uR[0] / pR);
#pragma unroll
        for (int n = 0; n < D; ++n) {
            const double vavg_n = 0.5 * (VL[n] + VR[n]);
            const double f_rho = rho_avg * vavg_n;
            F[n * NC + 0] = f_rho;
#pragma unroll
            for (int m = 0; m < D; ++m)
                F[n * NC + m + 1] = f_rho * (0.5 * (VL[m] + VR[m])) + (m == n ? p_avg : 0.0);
            F[n * NC + D + 1] = f_rho * Cc + 0.5 * (pL * VR[n] + pR * VL[n]);
        }
    }

    __device__ static void physical_flux(const LawParams& P, const double* u, double* F) {
        double V[D];
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) V[m] = u[m + 1] / u[0];
#pragma unroll
        for (int m = 0; m < D; ++m) s += u[m + 1] * V[m];
        const double p = P.gm1 * (u[D + 1] - 0.5 * s);
        const double h = u[D + 1] + p;
#pragma unroll
        for (int n = 0; n < D; ++n) {
            F[n * NC + 0] = u[n + 1];
#pragma unroll
            for (int m = 0; m < D; ++m) F[n * NC + m + 1] = u[m + 1] * V[n] + (m == n ? p : 0.0);
            F[n * NC + D + 1] = h * V[n];
        }
    }

    __device__ static double wave_speed(const LawParams& P, const double* ui, const double* uo, const double* n) {
        double si = 0.0, so = 0.0, vni = 0.0, vno = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            si += ui[m + 1] * ui[m + 1];
            so += uo[m + 1] * uo[m + 1];
        }
        const double p_in = P.gm1 * (ui[D + 1] - (0.5 / ui[0]) * si);
        const double p_out = P.gm1 * (uo[D + 1] - (0.5 / uo[0]) * so);
#pragma unroll
        for (int m = 0; m < D; ++m) {
            vni += (ui[m + 1] / ui[0]) * n[m];
            vno += (uo[m + 1] / uo[0]) * n[m];
        }
        const double c_in = sqrt(P.gamma * p_in / ui[0]);
        const double c_out = sqrt(P.gamma * p_out / uo[0]);
        return fmax(fabs(vni), fabs(vno)) + fmax(c_in, c_out);
    }

    // returns false for a non-physical state (rho <= 0 or p <= 0)
    __device__ static bool c2e(const LawParams& P, const double* u, double* w) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) s += u[m + 1] * u[m + 1];
        const double k = (0.5 / u[0]) * s;
        const double p = P.gm1 * (u[D + 1] - k);
        const double inv_p = 1.0 / p;
        w[0] = P.inv_gm1 * (P.gamma - log(p / pow(u[0], P.gamma))) - k * inv_p;
#pragma unroll
        for (int m = 0; m < D; ++m) w[m + 1] = u[m + 1] * inv_p;
        w[D + 1] = -u[0] * inv_p;
        return (u[0] > 0.0) && (p > 0.0);
    }

    __device__ static bool e2c(const LawParams& P, const double* win, double* u) {
        double w[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) w[e] = win[e] * P.gm1;
        double s2 = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) s2 += w[m + 1] * w[m + 1];
        const double k = s2 / (2.0 * w[D + 1]);
        const double s = P.gamma - w[0] + k;
        const double rho_e = pow(P.gm1 / pow(-w[D + 1], P.gamma), P.inv_gm1) * exp(-s * P.inv_gm1);
        u[0] = -w[D + 1] * rho_e;
#pragma unroll
        for (int m = 0; m < D; ++m) u[m + 1] = w[m + 1] * rho_e;
        u[D + 1] = rho_e * (1.0 - k);
        return win[D + 1] < 0.0;
    }
};

// ---------------------------------------------------------------------------
// Hoisted form of the entropy-conservative flux (Euler).
//
// Per node, once: h = [rho, v_0..v_{D-1}, p, log rho, beta = rho/p, log beta]
// (NH = D + 5 values). A pair then needs no logarithm: log(y/x) of the
// reference's logmean (ConservationLaws.jl:143) becomes log y - log x, which is
// only evaluated where |(y-x)/(y+x)| >= 1e-2, so its relative error stays
// below 1e-14. Both log-means of a flux come as fractions that share one reciprocal.
//
// ec_dot returns c_e = sum_n a_n F(uL,uR)[e,n] for a contraction vector a,
// without forming the (D+2) x D flux tensor:
//   c_0     = rho_avg (a . v_avg)
//   c_{m+1} = v_avg_m c_0 + a_m p_avg
//   c_{D+1} = C c_0 + (p_L (a . v_R) + p_R (a . v_L)) / 2
// which is the tensor of euler_navierstokes.jl:171-195 contracted with a.
// ---------------------------------------------------------------------------

// 1/x for a nonzero normal x: the hardware reciprocal estimate (v_rcp_f64) refined by
// two Newton-Raphson steps, e = 1 - x r, r <- r + r e, each of which squares the
// relative error (the iteration of the IEEE division sequence, without its scaling
// and rounding fix-up, which only matter for denormal, infinite or near-overflow
// operands). Five instructions instead of the eleven of a correctly rounded divide;
// within two ulps of 1/x (checked on the GPU against the divide).
__device__ __forceinline__ double rcp_refined(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

template <int D>
__device__ __forceinline__ void hoist(const LawParams& P, const double* u, double* h) {
    const double rho = u[0];
    const double irho = rcp_refined(rho);
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < D; ++m) {
        const double v = u[m + 1] * irho;
        h[1 + m] = v;
        s += v * v;
    }
    const double p = P.gm1 * (u[D + 1] - 0.5 * rho * s);
    const double beta = rho * rcp_refined(p);
    h[0] = rho;
    h[D + 1] = p;
    h[D + 2] = log(rho);
    h[D + 3] = beta;
    h[D + 4] = log(beta);
}

// The logarithmic mean as a fraction num/den (a branch, so that a wavefront none of whose
// pairs needs the series skips its polynomial):
//   |y - x| < 1e-2 (x + y): the reference's series 105 s / (210 + 70 f2 + 42 f2^2 + 30 f2^3),
//                           f2 = d^2 / s^2, with numerator and denominator scaled by s^6
//                           (s = x + y, d = y - x), so no divide is needed for f2;
//   otherwise:              (y - x) / (log y - log x).
// The inverse logarithmic mean is den/num of the same fraction.
__device__ __forceinline__ void logmean_frac(double x, double y, double lx, double ly, double& num, double& den) {
    const double s = x + y, d = y - x;
    const double s2 = s * s, d2 = d * d;
    num = d;
    den = ly - lx;
    if (d2 < 1.0e-4 * s2) {
        const double s6 = s2 * s2 * s2;
        num = 105.0 * s * s6;
        den = ((210.0 * s2 + 70.0 * d2) * s2 + 42.0 * d2 * d2) * s2 + 30.0 * d2 * d2 * d2;
    }
}

// ec_dot with the contraction vector given halved, ah = a / 2: the factors 1/2 of the
// averages then fall on ah (c_0 = rho_avg (ah . (v_L + v_R)), a_m p_avg = ah_m (p_L + p_R),
// (p_L (a . v_R) + p_R (a . v_L)) / 2 = p_L (ah . v_R) + p_R (ah . v_L)), five operations
// fewer per flux; scaling by 1/2 is exact, so the products are ec_dot's
template <int D>
__device__ __forceinline__ void ec_dot_h(const LawParams& P, const double* hL, const double* hR, const double* ah,
                                         double* c) {
    // rho_avg = n1/d1 (logmean of rho), ilm = d2/n2 (inverse logmean of rho/p): one reciprocal
    double n1, d1, n2, d2;
    logmean_frac(hL[0], hR[0], hL[D + 2], hR[D + 2], n1, d1);
    logmean_frac(hL[D + 3], hR[D + 3], hL[D + 4], hR[D + 4], n2, d2);
    const double rcp = rcp_refined(d1 * n2);
    const double rho_avg = n1 * n2 * rcp;
    const double ilm = d2 * d1 * rcp;
    double vv = 0.0, avL = 0.0, avR = 0.0;  // avL = (a . v_L) / 2
#pragma unroll
    for (int m = 0; m < D; ++m) {
        vv += hL[1 + m] * hR[1 + m];
        avL += ah[m] * hL[1 + m];
        avR += ah[m] * hR[1 + m];
    }
    const double pL = hL[D + 1], pR = hR[D + 1];
    const double ps = pL + pR;
    const double Cc = 0.5 * vv + P.inv_gm1 * ilm;
    const double c0 = rho_avg * (avL + avR);
    const double c0h = 0.5 * c0;
    c[0] = c0;
#pragma unroll
    for (int m = 0; m < D; ++m) c[1 + m] = c0h * (hL[1 + m] + hR[1 + m]) + ah[m] * ps;
    c[D + 1] = c0 * Cc + (pL * avR + pR * avL);
}

template <int D>
__device__ __forceinline__ void ec_dot(const LawParams& P, const double* hL, const double* hR, const double* a,
                                       double* c) {
    double ah[D];
#pragma unroll
    for (int m = 0; m < D; ++m) ah[m] = 0.5 * a[m];
    ec_dot_h<D>(P, hL, hR, ah, c);
}

// conservative_to_entropy! / entropy_to_conservative! (euler_navierstokes.jl:100-131)
// with the powers written through logarithms:
//   log(p / rho^gamma)                 = log p - gamma log rho
//   ((gamma-1)/(-w_E)^gamma)^(1/(gamma-1)) exp(-s/(gamma-1))
//                                      = exp((log(gamma-1) - gamma log(-w_E) - s) / (gamma-1))
// (two transcendental calls per node instead of five; same values up to rounding), and
// the divides as refined reciprocals.
template <int D>
__device__ __forceinline__ bool c2e_fast(const LawParams& P, const double* u, double* w) {
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < D; ++m) s += u[m + 1] * u[m + 1];
    const double k = (0.5 * rcp_refined(u[0])) * s;
    const double p = P.gm1 * (u[D + 1] - k);
    const double inv_p = rcp_refined(p);
    w[0] = P.inv_gm1 * (P.gamma - (log(p) - P.gamma * log(u[0]))) - k * inv_p;
#pragma unroll
    for (int m = 0; m < D; ++m) w[m + 1] = u[m + 1] * inv_p;
    w[D + 1] = -u[0] * inv_p;
    return (u[0] > 0.0) && (p > 0.0);
}

template <int D>
__device__ __forceinline__ bool e2c_fast(const LawParams& P, const double* win, double* u) {
    double w[D + 2];
#pragma unroll
    for (int e = 0; e < D + 2; ++e) w[e] = win[e] * P.gm1;
    double s2 = 0.0;
#pragma unroll
    for (int m = 0; m < D; ++m) s2 += w[m + 1] * w[m + 1];
    const double k = s2 * rcp_refined(2.0 * w[D + 1]);
    const double s = P.gamma - w[0] + k;
    const double rho_e = exp(P.inv_gm1 * (P.log_gm1 - P.gamma * log(-w[D + 1]) - s));
    u[0] = -w[D + 1] * rho_e;
#pragma unroll
    for (int m = 0; m < D; ++m) u[m + 1] = w[m + 1] * rho_e;
    u[D + 1] = rho_e * (1.0 - k);
    return win[D + 1] < 0.0;
}

// wave_speed from hoisted states (euler_navierstokes.jl:133-150). The larger sound speed
// is sqrt(gamma p / rho) of the side with the smaller beta = rho / p: one divide and one
// square root instead of two each (sqrt is monotone)
template <int D>
__device__ __forceinline__ double wave_speed_h(const LawParams& P, const double* hi, const double* ho,
                                               const double* n) {
    double vni = 0.0, vno = 0.0;
#pragma unroll
    for (int m = 0; m < D; ++m) {
        vni += hi[1 + m] * n[m];
        vno += ho[1 + m] * n[m];
    }
    return fmax(fabs(vni), fabs(vno)) + sqrt(P.gamma / fmin(hi[D + 3], ho[D + 3]));
}

template <int D>
struct Physics<D, LAW_ADVECTION> {
    static constexpr int NC = 1;
    __device__ static void two_point(const LawParams& P, const double* uL, const double* uR, double* F) {
        const double f1 = 0.5 * (uL[0] + uR[0]);
#pragma unroll
        for (int m = 0; m < D; ++m) F[m] = P.a[m] * f1;
    }
    __device__ static double wave_speed(const LawParams& P, const double*, const double*, const double* n) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) s += P.a[m] * n[m];
        return fabs(s);
    }
    __device__ static bool c2e(const LawParams&, const double* u, double* w) {
        w[0] = u[0];
        return true;
    }
    __device__ static bool e2c(const LawParams&, const double* w, double* u) {
        u[0] = w[0];
        return true;
    }
};

}  // namespace sse
