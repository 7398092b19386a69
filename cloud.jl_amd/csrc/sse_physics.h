// sse_physics.h -- per-node and per-pair physics of the residual, device side.
//
// Restates the reference's conservation-law functions:
//   logmean / inv_logmean           src/ConservationLaws/ConservationLaws.jl:132-156
//   conservative_to_entropy!        src/ConservationLaws/euler_navierstokes.jl:306-319
//   entropy_to_conservative!        src/ConservationLaws/euler_navierstokes.jl:321-337
//   wave_speed (Euler / advection)  src/ConservationLaws/euler_navierstokes.jl:339-356,
//                                   src/ConservationLaws/linear_advection_diffusion.jl:275-280
//   compute_two_point_flux (EC Ranocha, conservative, advection)
//                                   src/ConservationLaws/euler_navierstokes.jl:358-401,
//                                   src/ConservationLaws/linear_advection_diffusion.jl:282-288
// Flux tensors are stored F[n * NC + e] (direction n, variable e).
#pragma once

#include <hip/hip_runtime.h>

namespace sse {

enum { LAW_EULER = 1, LAW_ADVECTION = 2 };

struct LawParams {
    double gamma, gm1, inv_gm1;
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
