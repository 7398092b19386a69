// sse_oracle.cpp -- CPU oracle for the semi-discrete residual path.
//
// TEST INFRASTRUCTURE ONLY: this file restates the reference algorithm on the
// CPU so that tests, the smoke check and the CPU-baseline leg of the benchmark
// can check libsse_hip.so against it. The product never links, loads or calls
// it.
//
// It is a line-by-line restatement of StableSpectralElements.jl (the
// reference), consuming the same sse_desc the HIP library consumes:
//   semi_discrete_residual! Serial/Threaded   src/Solvers/Solvers.jl:474-514
//   nodal_values!, entropy_projection!       src/Solvers/flux_differencing_form.jl:171-292
//   time_derivative! (flux differencing)      src/Solvers/flux_differencing_form.jl:294-347
//   flux_difference! dense / sparse            src/Solvers/flux_differencing_form.jl:1-75
//   facet_correction! dense / sparse           src/Solvers/flux_differencing_form.jl:78-168
//   mass_matrix_solve! WA / diagonal           src/Solvers/mass_matrix.jl:169-196
//   numerical_flux! LF / central / EC          src/ConservationLaws/ConservationLaws.jl:75-128
//   logmean, inv_logmean                       src/ConservationLaws/ConservationLaws.jl:132-156
//   Euler entropy variables, wave speed, EC flux src/ConservationLaws/euler_navierstokes.jl:93-225
//   linear advection two-point flux            src/ConservationLaws/linear_advection_diffusion.jl:106-128
//   StandardForm nodal values / time deriv.   src/Solvers/standard_form_first_order.jl:1-94
//   BR1 auxiliary variable / time deriv.      src/Solvers/standard_form_second_order.jl:1-75
//   volume operators (skew / standard mapping) src/Solvers/operators.jl:1-37, 83-160
//   BR1 numerical fluxes                       src/ConservationLaws/linear_advection_diffusion.jl:54-102
//   CarpenterKennedy2N54 (OrdinaryDiffEq)      used at test/test_driver.jl:77-83
//
// The restatement is templated on the scalar type so that a dual-number
// instantiation can differentiate it.
//
// Pinning: the reference's golden L2 errors of its "Euler 1D Gauss collocation"
// test (test/runtests.jl:89-96) pin the EC flux, logmean, entropy variables,
// dense flux differencing, dense facet correction and CK2N54; closed-form
// invariants (entropy conservation, free-stream preservation, conservation,
// flux consistency/symmetry, Tadmor's condition) pin the sparse tensor-product
// paths (tests/test_oracle_*.py).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/sse_hip.h"

namespace {

using std::vector;

using std::exp;
using std::fabs;
using std::log;
using std::pow;
using std::sqrt;

// Forward-mode dual number (value, one tangent): instantiating the oracle on it
// differentiates the restated residual exactly (the role ForwardDiff plays for
// the Julia reference). Comparisons look at the value only.
struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double x) : v(x), d(0.0) {}
    Dual(double x, double t) : v(x), d(t) {}
    Dual& operator+=(const Dual& o) { v += o.v; d += o.d; return *this; }
    Dual& operator-=(const Dual& o) { v -= o.v; d -= o.d; return *this; }
    Dual& operator*=(const Dual& o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
    Dual& operator/=(const Dual& o) { d = (d * o.v - v * o.d) / (o.v * o.v); v /= o.v; return *this; }
};
inline Dual operator+(Dual a, const Dual& b) { return a += b; }
inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
inline Dual operator/(Dual a, const Dual& b) { return a /= b; }
inline Dual operator+(Dual a, double b) { return a += Dual(b); }
inline Dual operator+(double a, const Dual& b) { return Dual(a) += b; }
inline Dual operator-(Dual a, double b) { return a -= Dual(b); }
inline Dual operator-(double a, const Dual& b) { return Dual(a) -= b; }
inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
inline Dual operator*(double a, const Dual& b) { return Dual(a * b.v, a * b.d); }
inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
inline Dual operator/(double a, const Dual& b) { return Dual(a) /= b; }
inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<(const Dual& a, double b) { return a.v < b; }
inline bool operator>(const Dual& a, double b) { return a.v > b; }
inline Dual log(const Dual& a) { return Dual(std::log(a.v), a.d / a.v); }
inline Dual exp(const Dual& a) { const double e = std::exp(a.v); return Dual(e, e * a.d); }
inline Dual sqrt(const Dual& a) { const double r = std::sqrt(a.v); return Dual(r, 0.5 * a.d / r); }
inline Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }
inline Dual pow(const Dual& a, double b) {
    const double r = std::pow(a.v, b);
    return Dual(r, b * std::pow(a.v, b - 1.0) * a.d);
}

template <typename T>
struct Law {
    int kind, d, N_c;
    double gamma, gm1, inv_gm1;
    double a[3];

    // logmean (ConservationLaws.jl:132-145)
    static T logmean(const T& x, const T& y) {
        T f2 = (x * (x - 2.0 * y) + y * y) / (x * (x + 2.0 * y) + y * y);
        if (f2 < 1.0e-4) return (x + y) * 105.0 / (210.0 + f2 * (70.0 + f2 * (42.0 + f2 * 30.0)));
        return (y - x) / log(y / x);
    }
    // inv_logmean (ConservationLaws.jl:147-156)
    static T inv_logmean(const T& x, const T& y) {
        T f2 = (x * (x - 2.0 * y) + y * y) / (x * (x + 2.0 * y) + y * y);
        if (f2 < 1.0e-4) return (210.0 + f2 * (70.0 + f2 * (42.0 + f2 * 30.0))) / ((x + y) * 105.0);
        return log(y / x) / (y - x);
    }

    // compute_two_point_flux, tensor form F[e + N_c*n] (euler_navierstokes.jl:171-225;
    // linear_advection_diffusion.jl:113-128)
    void two_point(int flux, const T* uL, const T* uR, T* F) const {
        if (kind == SSE_EULER) {
            if (flux == SSE_TWO_POINT_CONSERVATIVE) {
                T FL[15], FR[15];
                physical_flux(uL, FL);
                physical_flux(uR, FR);
                for (int i = 0; i < N_c * d; ++i) F[i] = 0.5 * (FL[i] + FR[i]);
                return;
            }
            T VL[3], VR[3];
            T sL = 0.0, sR = 0.0;
            for (int m = 0; m < d; ++m) {
                VL[m] = uL[m + 1] / uL[0];
                VR[m] = uR[m + 1] / uR[0];
            }
            for (int m = 0; m < d; ++m) { sL += VL[m] * VL[m]; sR += VR[m] * VR[m]; }
            T pL = gm1 * (uL[d + 1] - 0.5 * uL[0] * sL);
            T pR = gm1 * (uR[d + 1] - 0.5 * uR[0] * sR);
            T rho_avg = logmean(uL[0], uR[0]);
            T V_avg[3];
            for (int m = 0; m < d; ++m) V_avg[m] = 0.5 * (VL[m] + VR[m]);
            T p_avg = 0.5 * (pL + pR);
            T vv = 0.0;
            for (int m = 0; m < d; ++m) vv += VL[m] * VR[m];
            T Cc = 0.5 * vv + inv_gm1 * inv_logmean(uL[0] / pL, uR[0] / pR);
            for (int n = 0; n < d; ++n) {
                T f_rho = rho_avg * V_avg[n];
                F[0 + N_c * n] = f_rho;
                for (int m = 0; m < d; ++m)
                    F[(m + 1) + N_c * n] = f_rho * V_avg[m] + (m == n ? p_avg : T(0.0));
                F[(d + 1) + N_c * n] = f_rho * Cc + 0.5 * (pL * VR[n] + pR * VL[n]);
            }
        } else {
            T flux_1d = 0.5 * (uL[0] + uR[0]);
            for (int m = 0; m < d; ++m) F[N_c * m] = a[m] * flux_1d;
        }
    }

    // physical_flux (euler_navierstokes.jl:264-274), tensor F[e + N_c*n]
    void physical_flux(const T* u, T* F) const {
        T V[3];
        for (int m = 0; m < d; ++m) V[m] = u[m + 1] / u[0];
        T s = 0.0;
        for (int m = 0; m < d; ++m) s += u[m + 1] * V[m];
        T p = gm1 * (u[d + 1] - 0.5 * s);
        T h = u[d + 1] + p;
        for (int n = 0; n < d; ++n) {
            F[0 + N_c * n] = u[n + 1];
            for (int m = 0; m < d; ++m) F[(m + 1) + N_c * n] = u[m + 1] * V[n] + (m == n ? p : T(0.0));
            F[(d + 1) + N_c * n] = h * V[n];
        }
    }

    // wave_speed (euler_navierstokes.jl:133-150; linear_advection_diffusion.jl:106-111)
    T wave_speed(const T* ui, const T* uo, const double* n) const {
        if (kind != SSE_EULER) {
            T s = 0.0;
            for (int m = 0; m < d; ++m) s += a[m] * n[m];
            return fabs(s);
        }
        T si = 0.0, so = 0.0, vni = 0.0, vno = 0.0;
        for (int m = 0; m < d; ++m) {
            si += ui[m + 1] * ui[m + 1];
            so += uo[m + 1] * uo[m + 1];
        }
        T p_in = gm1 * (ui[d + 1] - (0.5 / ui[0]) * si);
        T p_out = gm1 * (uo[d + 1] - (0.5 / uo[0]) * so);
        for (int m = 0; m < d; ++m) {
            vni += (ui[m + 1] / ui[0]) * n[m];
            vno += (uo[m + 1] / uo[0]) * n[m];
        }
        T c_in = sqrt(gamma * p_in / ui[0]);
        T c_out = sqrt(gamma * p_out / uo[0]);
        T av = fabs(vni) > fabs(vno) ? fabs(vni) : fabs(vno);
        T cm = c_in > c_out ? c_in : c_out;
        return av + cm;
    }

    // conservative_to_entropy! (euler_navierstokes.jl:100-113)
    void c2e(const T* u, T* w) const {
        T s = 0.0;
        for (int m = 0; m < d; ++m) s += u[m + 1] * u[m + 1];
        T k = (0.5 / u[0]) * s;
        T p = gm1 * (u[d + 1] - k);
        T inv_p = 1.0 / p;
        w[0] = inv_gm1 * (gamma - log(p / pow(u[0], gamma))) - k * inv_p;
        for (int m = 0; m < d; ++m) w[m + 1] = u[m + 1] * inv_p;
        w[d + 1] = -u[0] * inv_p;
    }

    // entropy_to_conservative! (euler_navierstokes.jl:115-131)
    void e2c(const T* w_in, T* u) const {
        T w[5];
        for (int e = 0; e < N_c; ++e) w[e] = w_in[e] * gm1;
        T s2 = 0.0;
        for (int m = 0; m < d; ++m) s2 += w[m + 1] * w[m + 1];
        T k = s2 / (2.0 * w[d + 1]);
        T s = gamma - w[0] + k;
        T rho_e = pow(gm1 / pow(-w[d + 1], gamma), inv_gm1) * exp(-s * inv_gm1);
        u[0] = -w[d + 1] * rho_e;
        for (int m = 0; m < d; ++m) u[m + 1] = w[m + 1] * rho_e;
        u[d + 1] = rho_e * (1.0 - k);
    }
};

struct CSC {
    vector<int> colptr, rowval;
    vector<double> nzval;
};

// `sparse(A)` of a dense column-major n_rows x n_cols matrix: structural
// nonzeros are the nonzero entries, rows ascending within each column.
CSC to_csc(const double* A, int n_rows, int n_cols) {
    CSC c;
    c.colptr.push_back(0);
    for (int j = 0; j < n_cols; ++j) {
        for (int i = 0; i < n_rows; ++i) {
            double v = A[i + (size_t)n_rows * j];
            if (v != 0.0) {
                c.rowval.push_back(i);
                c.nzval.push_back(v);
            }
        }
        c.colptr.push_back((int)c.rowval.size());
    }
    return c;
}

// The same nonzeros by rows: entry i of the result lists row i's nonzero columns,
// ascending (for y = A x, which then sums the nonzero terms in the order the dense
// loop would; the skipped products are exact zeros).
CSC to_csr(const double* A, int n_rows, int n_cols) {
    CSC c;
    c.colptr.push_back(0);
    for (int i = 0; i < n_rows; ++i) {
        for (int j = 0; j < n_cols; ++j) {
            double v = A[i + (size_t)n_rows * j];
            if (v != 0.0) {
                c.rowval.push_back(j);
                c.nzval.push_back(v);
            }
        }
        c.colptr.push_back((int)c.rowval.size());
    }
    return c;
}

template <typename T>
struct Oracle {
    sse_desc D;
    Law<T> law;
    int d, N_p, N_q, N_f, N_fac, N_c, N_e, npf;
    vector<CSC> S_csc;
    CSC C_csc;
    // nonzeros of D_m and R by rows (for D x, R x) and by columns (for D' x, R' x): the
    // reference applies these tensor-product operators without their zero entries
    vector<CSC> D_rows, D_cols;
    CSC R_rows, R_cols;
    vector<double> halfnJq;   // d x N_fac x N_q x N_e
    vector<double> n_f, BJf;  // d x N_f x N_e ; N_f x N_e
    vector<double> WJ, JinvW; // N_q x N_e
    // global intermediates, reference layout
    vector<T> u_q;            // N_q x N_c x N_e
    vector<T> u_f;            // N_f x N_e x N_c ("switched order")
    vector<int64_t> mapP0;    // 0-based
    bool second_order = false;
    T hl = 0.0;               // LaxFriedrichsNumericalFlux.halfλ (a T so that it can carry a tangent)
    T bd = 0.0;               // diffusion coefficient b (a T for the same reason)
    vector<T> q_q;            // N_q x N_c x d x N_e (auxiliary variable, SecondOrder)
    vector<T> q_f;            // N_f x N_e x N_c x d

    std::string init(const sse_desc* desc) {
        D = *desc;
        d = D.d; N_p = D.N_p; N_q = D.N_q; N_f = D.N_f; N_fac = D.N_fac; N_c = D.N_c; N_e = D.N_e;
        if (N_fac <= 0 || N_f % N_fac != 0) return "N_f must be a multiple of N_fac";
        npf = N_f / N_fac;
        hl = D.half_lambda;
        bd = D.b;
        law.kind = D.law; law.d = d; law.N_c = N_c;
        law.gamma = D.gamma; law.gm1 = D.gamma - 1.0; law.inv_gm1 = 1.0 / (D.gamma - 1.0);
        for (int m = 0; m < 3; ++m) law.a[m] = D.a[m];
        if (D.form == SSE_STANDARD) {
            if (!D.D) return "StandardForm needs the derivative operators D";
            second_order = (D.law == SSE_LINEAR_ADVECTION_DIFFUSION);
            if (second_order && D.viscous_flux != SSE_BR1) return "second-order laws need BR1";
        } else if (D.form != SSE_FLUX_DIFFERENCING) {
            return "unknown residual form";
        }
        if (D.S_sparse)
            for (int m = 0; m < d; ++m) S_csc.push_back(to_csc(D.S + (size_t)N_q * N_q * m, N_q, N_q));
        if (D.C_kind == SSE_C_SPARSE) C_csc = to_csc(D.C, N_q, N_f);
        if (D.D)
            for (int m = 0; m < d; ++m) {
                D_rows.push_back(to_csr(D.D + (size_t)N_q * N_q * m, N_q, N_q));
                D_cols.push_back(to_csc(D.D + (size_t)N_q * N_q * m, N_q, N_q));
            }
        if (D.R && !D.R_selection) {
            R_rows = to_csr(D.R, N_f, N_q);
            R_cols = to_csc(D.R, N_f, N_q);
        }
        halfnJq.assign((size_t)d * N_fac * N_q * N_e, 0.0);
        if (D.nJq) {
            for (size_t i = 0; i < halfnJq.size(); ++i) halfnJq[i] = 0.5 * D.nJq[i];
        } else {
            for (int k = 0; k < N_e; ++k)
                for (int i = 0; i < N_q; ++i)
                    for (int f = 0; f < N_fac; ++f)
                        for (int n = 0; n < d; ++n) {
                            double s = 0.0;
                            for (int m = 0; m < d; ++m)
                                s += D.Lambda_q[i + (size_t)N_q * (m + d * (n + (size_t)d * k))] *
                                     D.face_normals[m + d * f];
                            halfnJq[n + d * (f + (size_t)N_fac * (i + (size_t)N_q * k))] = 0.5 * s;
                        }
        }
        n_f.resize((size_t)d * N_f * N_e);
        BJf.resize((size_t)N_f * N_e);
        for (int k = 0; k < N_e; ++k)
            for (int j = 0; j < N_f; ++j) {
                double Jf = D.J_f[j + (size_t)N_f * k];
                BJf[j + (size_t)N_f * k] = D.B[j] * Jf;
                for (int m = 0; m < d; ++m)
                    n_f[m + d * (j + (size_t)N_f * k)] = D.nJf[m + d * (j + (size_t)N_f * k)] / Jf;
            }
        WJ.resize((size_t)N_q * N_e);
        JinvW.resize((size_t)N_q * N_e);
        for (int k = 0; k < N_e; ++k)
            for (int i = 0; i < N_q; ++i) {
                WJ[i + (size_t)N_q * k] = D.W[i] * D.J_q[i + (size_t)N_q * k];
                JinvW[i + (size_t)N_q * k] = D.W[i] / D.J_q[i + (size_t)N_q * k];
            }
        u_q.assign((size_t)N_q * N_c * N_e, T(0.0));
        u_f.assign((size_t)N_f * N_e * N_c, T(0.0));
        if (second_order) {
            q_q.assign((size_t)N_q * N_c * d * N_e, T(0.0));
            q_f.assign((size_t)N_f * N_e * N_c * d, T(0.0));
        }
        mapP0.resize((size_t)N_f * N_e);
        for (size_t i = 0; i < mapP0.size(); ++i) mapP0[i] = D.mapP[i] - 1;
        return "";
    }

    // y (n_out x N_c) = A (n_out x n_in, column-major) * x (n_in x N_c)
    void matmul(const double* A, int n_out, int n_in, const T* x, int ldx, T* y, int ldy) const {
        for (int e = 0; e < N_c; ++e)
            for (int i = 0; i < n_out; ++i) {
                T s = 0.0;
                for (int j = 0; j < n_in; ++j) s += A[i + (size_t)n_out * j] * x[j + (size_t)ldx * e];
                y[i + (size_t)ldy * e] = s;
            }
    }
    // y = A' x
    void matmul_t(const double* A, int n_rows_A, int n_cols_A, const T* x, int ldx, T* y, int ldy) const {
        for (int e = 0; e < N_c; ++e)
            for (int j = 0; j < n_cols_A; ++j) {
                T s = 0.0;
                for (int i = 0; i < n_rows_A; ++i) s += A[i + (size_t)n_rows_A * j] * x[i + (size_t)ldx * e];
                y[j + (size_t)ldy * e] = s;
            }
    }
    // y (n_out x N_c) = A x with A's nonzeros by rows (to_csr), or y = A' x with A's
    // nonzeros by columns (to_csc): the dense loops above without their zero products
    void matmul_sp(const CSC& A, int n_out, const T* x, int ldx, T* y, int ldy) const {
        for (int e = 0; e < N_c; ++e)
            for (int i = 0; i < n_out; ++i) {
                T s = 0.0;
                for (int q = A.colptr[i]; q < A.colptr[i + 1]; ++q) s += A.nzval[q] * x[A.rowval[q] + (size_t)ldx * e];
                y[i + (size_t)ldy * e] = s;
            }
    }
    void applyV(const T* x, T* y) const {  // N_p -> N_q
        if (D.V) matmul(D.V, N_q, N_p, x, N_p, y, N_q);
        else for (int e = 0; e < N_c; ++e) for (int i = 0; i < N_q; ++i) y[i + N_q * e] = x[i + N_p * e];
    }
    void applyVt(const T* x, T* y) const {  // N_q -> N_p
        if (D.V) matmul_t(D.V, N_q, N_p, x, N_q, y, N_p);
        else for (int e = 0; e < N_c; ++e) for (int i = 0; i < N_p; ++i) y[i + N_p * e] = x[i + N_q * e];
    }
    void applyMinv(T* x) const {  // in-place N_p x N_c
        if (!D.Minv) return;
        vector<T> t((size_t)N_p * N_c);
        matmul(D.Minv, N_p, N_p, x, N_p, t.data(), N_p);
        for (size_t i = 0; i < t.size(); ++i) x[i] = t[i];
    }
    // R applied to (N_q x N_c), result (N_f x N_c) with leading dim ldy
    void applyR(const T* x, T* y, int ldy) const {
        if (D.R_selection) {
            for (int e = 0; e < N_c; ++e)
                for (int j = 0; j < N_f; ++j) y[j + (size_t)ldy * e] = x[(D.R_selection[j] - 1) + (size_t)N_q * e];
        } else {
            matmul_sp(R_rows, N_f, x, N_q, y, ldy);
        }
    }
    void applyRt(const T* x, T* y) const {  // (N_f x N_c) -> (N_q x N_c)
        if (D.R_selection) {
            for (int e = 0; e < N_c; ++e) {
                for (int i = 0; i < N_q; ++i) y[i + (size_t)N_q * e] = 0.0;
                for (int j = 0; j < N_f; ++j) y[(D.R_selection[j] - 1) + (size_t)N_q * e] += x[j + (size_t)N_f * e];
            }
        } else {
            matmul_sp(R_cols, N_q, x, N_f, y, N_q);
        }
    }

    // mass_matrix_solve! (mass_matrix.jl:169-196); rhs N_p x N_c, temp N_q x N_c
    void mass_solve(int k, T* rhs, T* temp) const {
        if (D.mass_solver == SSE_MASS_DIAGONAL) {
            for (int e = 0; e < N_c; ++e)
                for (int i = 0; i < N_p; ++i) rhs[i + (size_t)N_p * e] /= WJ[i + (size_t)N_q * k];
            return;
        }
        applyMinv(rhs);
        applyV(rhs, temp);
        for (int e = 0; e < N_c; ++e)
            for (int i = 0; i < N_q; ++i) temp[i + (size_t)N_q * e] *= JinvW[i + (size_t)N_q * k];
        applyVt(temp, rhs);
        applyMinv(rhs);
    }

    // nodal_values! (flux_differencing_form.jl:253-292) with the three
    // entropy_projection! methods (:171-250)
    void nodal_values(const T* u, int k, T* w_q, T* w_f, T* w) {
        const T* uk = u + (size_t)N_p * N_c * k;
        T* uqk = u_q.data() + (size_t)N_q * N_c * k;
        T ub[5], wb[5];
        auto uf_at = [&](int j, int e) -> T& { return u_f[j + (size_t)N_f * (k + (size_t)N_e * e)]; };
        if (N_c == 1) {  // scalar: no entropy projection (:253-265)
            applyV(uk, uqk);
            applyR(uqk, w_f, N_f);
            for (int j = 0; j < N_f; ++j) uf_at(j, 0) = w_f[j];
            return;
        }
        if (!D.V && D.R_selection) {  // nodal diagonal-E (:171-187)
            applyV(uk, uqk);
            for (int e = 0; e < N_c; ++e)
                for (int j = 0; j < N_f; ++j) uf_at(j, e) = uqk[(D.R_selection[j] - 1) + (size_t)N_q * e];
            return;
        }
        if (!D.V) {  // nodal, general R (:190-211)
            applyV(uk, uqk);
            for (int i = 0; i < N_q; ++i) {
                for (int e = 0; e < N_c; ++e) ub[e] = uqk[i + (size_t)N_q * e];
                law.c2e(ub, wb);
                for (int e = 0; e < N_c; ++e) w_q[i + (size_t)N_q * e] = wb[e];
            }
            applyR(w_q, w_f, N_f);
            for (int j = 0; j < N_f; ++j) {
                for (int e = 0; e < N_c; ++e) wb[e] = w_f[j + (size_t)N_f * e];
                law.e2c(wb, ub);
                for (int e = 0; e < N_c; ++e) uf_at(j, e) = ub[e];
            }
            return;
        }
        // modal (:214-250)
        applyV(uk, uqk);
        for (int i = 0; i < N_q; ++i) {
            for (int e = 0; e < N_c; ++e) ub[e] = uqk[i + (size_t)N_q * e];
            law.c2e(ub, wb);
            for (int e = 0; e < N_c; ++e) w_q[i + (size_t)N_q * e] = wb[e];
        }
        for (int e = 0; e < N_c; ++e)
            for (int i = 0; i < N_q; ++i) w_q[i + (size_t)N_q * e] *= WJ[i + (size_t)N_q * k];
        applyVt(w_q, w);
        mass_solve(k, w, w_q);
        applyV(w, w_q);
        applyR(w_q, w_f, N_f);
        for (int i = 0; i < N_q; ++i) {
            for (int e = 0; e < N_c; ++e) wb[e] = w_q[i + (size_t)N_q * e];
            law.e2c(wb, ub);
            for (int e = 0; e < N_c; ++e) uqk[i + (size_t)N_q * e] = ub[e];
        }
        for (int j = 0; j < N_f; ++j) {
            for (int e = 0; e < N_c; ++e) wb[e] = w_f[j + (size_t)N_f * e];
            law.e2c(wb, ub);
            for (int e = 0; e < N_c; ++e) uf_at(j, e) = ub[e];
        }
    }

    void time_derivative(T* dudt, int k, T* f_f, T* r_q, T* temp) {
        const T* uqk = u_q.data() + (size_t)N_q * N_c * k;
        T uin[5], uout[5], F[15], ui[5];
        const double* Lk = D.Lambda_q + (size_t)N_q * d * d * k;
        // numerical_flux! (ConservationLaws.jl:75-128)
        for (int j = 0; j < N_f; ++j) {
            int64_t g = mapP0[j + (size_t)N_f * k];
            int jn = (int)(g % N_f), kn = (int)(g / N_f);
            for (int e = 0; e < N_c; ++e) {
                uin[e] = u_f[j + (size_t)N_f * (k + (size_t)N_e * e)];
                uout[e] = u_f[jn + (size_t)N_f * (kn + (size_t)N_e * e)];
            }
            law.two_point(D.two_point_flux, uin, uout, F);
            const double* nf = n_f.data() + d * (j + (size_t)N_f * k);
            T a = 0.0;
            if (D.inviscid_flux == SSE_LAX_FRIEDRICHS) a = hl * law.wave_speed(uin, uout, nf);
            for (int e = 0; e < N_c; ++e) {
                T fn = 0.0;
                for (int m = 0; m < d; ++m) fn += F[e + N_c * m] * nf[m];
                if (D.inviscid_flux == SSE_LAX_FRIEDRICHS) fn += a * (uin[e] - uout[e]);
                f_f[j + (size_t)N_f * e] = fn * BJf[j + (size_t)N_f * k];  // lmul!(BJf[k], f_f)
            }
        }
        // flux_difference!
        for (size_t i = 0; i < (size_t)N_q * N_c; ++i) r_q[i] = 0.0;
        if (!D.S_sparse) {  // dense (:1-35)
            for (int i = 0; i < N_q; ++i)
                for (int j = i + 1; j < N_q; ++j) {
                    for (int e = 0; e < N_c; ++e) { ui[e] = uqk[i + (size_t)N_q * e]; uout[e] = uqk[j + (size_t)N_q * e]; }
                    law.two_point(D.two_point_flux, ui, uout, F);
                    for (int e = 0; e < N_c; ++e) {
                        T diff = 0.0;
                        for (int m = 0; m < d; ++m) {
                            T Fm = 0.0;
                            for (int n = 0; n < d; ++n) {
                                double Lij = Lk[i + (size_t)N_q * (m + d * n)] + Lk[j + (size_t)N_q * (m + d * n)];
                                Fm += Lij * F[e + N_c * n];
                            }
                            diff += D.S[i + (size_t)N_q * (j + (size_t)N_q * m)] * Fm;
                        }
                        r_q[i + (size_t)N_q * e] -= diff;
                        r_q[j + (size_t)N_q * e] += diff;
                    }
                }
        } else {  // sparse (:37-75)
            for (int m = 0; m < d; ++m) {
                const CSC& Sm = S_csc[m];
                for (int j = 0; j < N_q; ++j)
                    for (int ii = Sm.colptr[j]; ii < Sm.colptr[j + 1]; ++ii) {
                        int i = Sm.rowval[ii];
                        if (i >= j) continue;
                        for (int e = 0; e < N_c; ++e) { ui[e] = uqk[i + (size_t)N_q * e]; uout[e] = uqk[j + (size_t)N_q * e]; }
                        law.two_point(D.two_point_flux, ui, uout, F);
                        double Sij = Sm.nzval[ii];
                        for (int e = 0; e < N_c; ++e) {
                            T Fm = 0.0;
                            for (int n = 0; n < d; ++n) {
                                double Lij = Lk[i + (size_t)N_q * (m + d * n)] + Lk[j + (size_t)N_q * (m + d * n)];
                                Fm += Lij * F[e + N_c * n];
                            }
                            T diff = Sij * Fm;
                            r_q[i + (size_t)N_q * e] -= diff;
                            r_q[j + (size_t)N_q * e] += diff;
                        }
                    }
            }
        }
        // facet_correction!
        const double* hJq = halfnJq.data() + (size_t)d * N_fac * N_q * k;
        auto fc_pair = [&](int i, int j, double Cij) {
            for (int e = 0; e < N_c; ++e) {
                ui[e] = uqk[i + (size_t)N_q * e];
                uout[e] = u_f[j + (size_t)N_f * (k + (size_t)N_e * e)];
            }
            law.two_point(D.two_point_flux, ui, uout, F);
            int f = j / npf;
            for (int e = 0; e < N_c; ++e) {
                T Fn = 0.0;
                for (int m = 0; m < d; ++m) {
                    double nJ = 0.5 * D.nJf[m + d * (j + (size_t)N_f * k)] + hJq[m + d * (f + (size_t)N_fac * i)];
                    Fn += nJ * F[e + N_c * m];
                }
                T diff = Cij * Fn;
                r_q[i + (size_t)N_q * e] -= diff;
                f_f[j + (size_t)N_f * e] -= diff;
            }
        };
        if (D.C_kind == SSE_C_DENSE) {  // (:91-124)
            for (int i = 0; i < N_q; ++i)
                for (int j = 0; j < N_f; ++j) fc_pair(i, j, D.C[i + (size_t)N_q * j]);
        } else if (D.C_kind == SSE_C_SPARSE) {  // (:126-168)
            for (int j = 0; j < N_f; ++j)
                for (int ii = C_csc.colptr[j]; ii < C_csc.colptr[j + 1]; ++ii) fc_pair(C_csc.rowval[ii], j, C_csc.nzval[ii]);
        }
        // lift, V' and mass solve (:340-346)
        applyRt(f_f, temp);
        for (size_t i = 0; i < (size_t)N_q * N_c; ++i) r_q[i] -= temp[i];
        T* dk = dudt + (size_t)N_p * N_c * k;
        applyVt(r_q, dk);
        mass_solve(k, dk, temp);
    }

    // ------------------------------------------------------------------ StandardForm
    // y = D_m x, y = D_m' x (N_q x N_c)
    void applyD(int m, const T* x, T* y) const { matmul_sp(D_rows[m], N_q, x, N_q, y, N_q); }
    void applyDt(int m, const T* x, T* y) const { matmul_sp(D_cols[m], N_q, x, N_q, y, N_q); }

    // r += (volume operator of direction n) x, before V' and the mass solve:
    //   SkewSymmetricMapping  sum_m D_m' (1/2 W Lambda_mn x) - 1/2 W Lambda_mn (D_m x)
    //                         + R' (1/2 B nJf_n (R x))        (operators.jl:132-160)
    //   StandardMapping       sum_m D_m' (W Lambda_mn x)        (operators.jl:83-130)
    void volume_term(int k, int n, const T* x, T* r, T* t1, T* t2, T* tf) const {
        const double* Lk = D.Lambda_q + (size_t)N_q * d * d * k;
        const bool skew = (D.mapping_form != SSE_STANDARD_MAPPING);
        for (int m = 0; m < d; ++m) {
            const double* Lmn = Lk + (size_t)N_q * (m + d * n);
            const double h = skew ? 0.5 : 1.0;
            for (int e = 0; e < N_c; ++e)
                for (int i = 0; i < N_q; ++i) t1[i + (size_t)N_q * e] = h * D.W[i] * Lmn[i] * x[i + (size_t)N_q * e];
            applyDt(m, t1, t2);
            for (size_t i = 0; i < (size_t)N_q * N_c; ++i) r[i] += t2[i];
            if (skew) {
                applyD(m, x, t2);
                for (int e = 0; e < N_c; ++e)
                    for (int i = 0; i < N_q; ++i)
                        r[i + (size_t)N_q * e] -= 0.5 * D.W[i] * Lmn[i] * t2[i + (size_t)N_q * e];
            }
        }
        if (skew) {
            applyR(x, tf, N_f);
            for (int e = 0; e < N_c; ++e)
                for (int j = 0; j < N_f; ++j)
                    tf[j + (size_t)N_f * e] *= 0.5 * D.B[j] * D.nJf[n + d * (j + (size_t)N_f * k)];
            applyRt(tf, t2);
            for (size_t i = 0; i < (size_t)N_q * N_c; ++i) r[i] += t2[i];
        }
    }

    // nodal_values! for StandardForm (standard_form_first_order.jl:1-14)
    void nodal_values_standard(const T* u, int k, T* w_f) {
        const T* uk = u + (size_t)N_p * N_c * k;
        T* uqk = u_q.data() + (size_t)N_q * N_c * k;
        applyV(uk, uqk);
        applyR(uqk, w_f, N_f);
        for (int e = 0; e < N_c; ++e)
            for (int j = 0; j < N_f; ++j) u_f[j + (size_t)N_f * (k + (size_t)N_e * e)] = w_f[j + (size_t)N_f * e];
    }

    // interface flux f* (N_f x N_c) at element k: numerical_flux! with the
    // inviscid flux and the conservative two-point flux (ConservationLaws.jl:75-128)
    void inviscid_interface(int k, T* f_f) const {
        T uin[5], uout[5], F[15];
        for (int j = 0; j < N_f; ++j) {
            int64_t g = mapP0[j + (size_t)N_f * k];
            int jn = (int)(g % N_f), kn = (int)(g / N_f);
            for (int e = 0; e < N_c; ++e) {
                uin[e] = u_f[j + (size_t)N_f * (k + (size_t)N_e * e)];
                uout[e] = u_f[jn + (size_t)N_f * (kn + (size_t)N_e * e)];
            }
            law.two_point(D.two_point_flux, uin, uout, F);
            const double* nf = n_f.data() + d * (j + (size_t)N_f * k);
            T a = 0.0;
            if (D.inviscid_flux == SSE_LAX_FRIEDRICHS) a = hl * law.wave_speed(uin, uout, nf);
            for (int e = 0; e < N_c; ++e) {
                T fn = 0.0;
                for (int m = 0; m < d; ++m) fn += F[e + N_c * m] * nf[m];
                if (D.inviscid_flux == SSE_LAX_FRIEDRICHS) fn += a * (uin[e] - uout[e]);
                f_f[j + (size_t)N_f * e] = fn;
            }
        }
    }

    // auxiliary_variable! (standard_form_second_order.jl:3-32): for each m,
    // q_m = M^{-1} V' [-(volume term m)(u_q) + R' (B J_f u*_m)], u*_m = {u} n_m (BR1,
    // linear_advection_diffusion.jl:75-86); q_q = V q_m, q_f = R q_q
    void auxiliary_variable(int k, T* r, T* t1, T* t2, T* tf, T* x) {
        const T* uqk = u_q.data() + (size_t)N_q * N_c * k;
        for (int m = 0; m < d; ++m) {
            for (size_t i = 0; i < (size_t)N_q * N_c; ++i) r[i] = 0.0;
            volume_term(k, m, uqk, r, t1, t2, tf);
            for (size_t i = 0; i < (size_t)N_q * N_c; ++i) r[i] = -r[i];
            for (int j = 0; j < N_f; ++j) {
                int64_t g = mapP0[j + (size_t)N_f * k];
                int jn = (int)(g % N_f), kn = (int)(g / N_f);
                const double nm = n_f[m + d * (j + (size_t)N_f * k)];
                for (int e = 0; e < N_c; ++e) {
                    T avg = 0.5 * (u_f[j + (size_t)N_f * (k + (size_t)N_e * e)] +
                                   u_f[jn + (size_t)N_f * (kn + (size_t)N_e * e)]);
                    tf[j + (size_t)N_f * e] = avg * nm * BJf[j + (size_t)N_f * k];
                }
            }
            applyRt(tf, t2);
            for (size_t i = 0; i < (size_t)N_q * N_c; ++i) r[i] += t2[i];
            applyVt(r, x);
            mass_solve(k, x, t1);
            T* qq = q_q.data() + (size_t)N_q * N_c * (m + (size_t)d * k);
            applyV(x, qq);
            applyR(qq, tf, N_f);
            for (int e = 0; e < N_c; ++e)
                for (int j = 0; j < N_f; ++j)
                    q_f[j + (size_t)N_f * (k + (size_t)N_e * (e + (size_t)N_c * m))] = tf[j + (size_t)N_f * e];
        }
    }

    // time_derivative! for StandardForm: first order (standard_form_first_order.jl:16-94)
    // and second order (standard_form_second_order.jl:36-75), as
    //   dudt = M^{-1} V' [ sum_n (volume term n)(f_n) - R' (B J_f f*) ]
    void time_derivative_standard(T* dudt, int k, T* f_f, T* r, T* t1, T* t2, T* tf, T* fq) {
        const T* uqk = u_q.data() + (size_t)N_q * N_c * k;
        inviscid_interface(k, f_f);
        if (second_order) {  // BR1 viscous flux (linear_advection_diffusion.jl:88-102)
            for (int j = 0; j < N_f; ++j) {
                int64_t g = mapP0[j + (size_t)N_f * k];
                int jn = (int)(g % N_f), kn = (int)(g / N_f);
                for (int e = 0; e < N_c; ++e) {
                    T s = 0.0;
                    for (int m = 0; m < d; ++m) {
                        T qin = q_f[j + (size_t)N_f * (k + (size_t)N_e * (e + (size_t)N_c * m))];
                        T qout = q_f[jn + (size_t)N_f * (kn + (size_t)N_e * (e + (size_t)N_c * m))];
                        s += bd * (-0.5 * (qin + qout)) * n_f[m + d * (j + (size_t)N_f * k)];
                    }
                    f_f[j + (size_t)N_f * e] += s;
                }
            }
        }
        for (size_t i = 0; i < (size_t)N_q * N_c; ++i) r[i] = 0.0;
        T ub[5], F[15];
        for (int n = 0; n < d; ++n) {
            // f_n at the volume nodes (physical_flux!)
            for (int i = 0; i < N_q; ++i) {
                if (D.law == SSE_EULER) {
                    for (int e = 0; e < N_c; ++e) ub[e] = uqk[i + (size_t)N_q * e];
                    law.physical_flux(ub, F);
                    for (int e = 0; e < N_c; ++e) fq[i + (size_t)N_q * e] = F[e + N_c * n];
                } else {
                    T f = D.a[n] * uqk[i];
                    if (second_order) f -= bd * q_q[i + (size_t)N_q * N_c * (n + (size_t)d * k)];
                    fq[i] = f;
                }
            }
            volume_term(k, n, fq, r, t1, t2, tf);
        }
        for (int e = 0; e < N_c; ++e)
            for (int j = 0; j < N_f; ++j) f_f[j + (size_t)N_f * e] *= BJf[j + (size_t)N_f * k];
        applyRt(f_f, t2);
        for (size_t i = 0; i < (size_t)N_q * N_c; ++i) r[i] -= t2[i];
        T* dk = dudt + (size_t)N_p * N_c * k;
        applyVt(r, dk);
        mass_solve(k, dk, t1);
    }

    void residual_standard(const T* u, T* dudt) {
#pragma omp parallel
        {
            vector<T> w_f((size_t)N_f * N_c);
#pragma omp for schedule(static)
            for (int k = 0; k < N_e; ++k) nodal_values_standard(u, k, w_f.data());
        }
        if (second_order) {
#pragma omp parallel
            {
                vector<T> r((size_t)N_q * N_c), t1((size_t)N_q * N_c), t2((size_t)N_q * N_c),
                    tf((size_t)N_f * N_c), x((size_t)N_p * N_c);
#pragma omp for schedule(static)
                for (int k = 0; k < N_e; ++k) auxiliary_variable(k, r.data(), t1.data(), t2.data(), tf.data(), x.data());
            }
        }
#pragma omp parallel
        {
            vector<T> f_f((size_t)N_f * N_c), r((size_t)N_q * N_c), t1((size_t)N_q * N_c), t2((size_t)N_q * N_c),
                tf((size_t)N_f * N_c), fq((size_t)N_q * N_c);
#pragma omp for schedule(static)
            for (int k = 0; k < N_e; ++k)
                time_derivative_standard(dudt, k, f_f.data(), r.data(), t1.data(), t2.data(), tf.data(), fq.data());
        }
    }

    void residual(const T* u, T* dudt, int nthreads) {
#ifdef _OPENMP
        if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
        if (D.form == SSE_STANDARD) {
            residual_standard(u, dudt);
            return;
        }
#pragma omp parallel
        {
            vector<T> w_q((size_t)N_q * N_c), w_f((size_t)N_f * N_c), w((size_t)N_p * N_c);
#pragma omp for schedule(static)
            for (int k = 0; k < N_e; ++k) nodal_values(u, k, w_q.data(), w_f.data(), w.data());
        }
#pragma omp parallel
        {
            vector<T> f_f((size_t)N_f * N_c), r_q((size_t)N_q * N_c), temp((size_t)N_q * N_c);
#pragma omp for schedule(static)
            for (int k = 0; k < N_e; ++k) time_derivative(dudt, k, f_f.data(), r_q.data(), temp.data());
        }
    }
};

// Carpenter & Kennedy (1994), NASA TM-109112, five-stage fourth-order 2N-storage
// coefficients, as used by OrdinaryDiffEq's CarpenterKennedy2N54.
const double CK_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                        -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double CK_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                        2277821191437.0 / 14882151754819.0};
const double CK_C[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                        2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

std::string g_err;
int g_threads = 0;

}  // namespace

extern "C" {

const char* sse_oracle_last_error(void) { return g_err.c_str(); }

void sse_oracle_set_threads(int n) { g_threads = n; }

int sse_oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// semi_discrete_residual!(dudt, u, solver, t) on the CPU.
int sse_oracle_residual(const sse_desc* desc, const double* u, double* dudt) {
    Oracle<double> o;
    g_err = o.init(desc);
    if (!g_err.empty()) return SSE_ERR_UNSUPPORTED;
    o.residual(u, dudt, g_threads);
    return 0;
}

// Residual repeated nrep times on one prepared oracle (for timing).
int sse_oracle_residual_repeat(const sse_desc* desc, const double* u, double* dudt, int nrep) {
    Oracle<double> o;
    g_err = o.init(desc);
    if (!g_err.empty()) return SSE_ERR_UNSUPPORTED;
    for (int r = 0; r < nrep; ++r) o.residual(u, dudt, g_threads);
    return 0;
}

// solve(ODEProblem, CarpenterKennedy2N54(), dt, adaptive=false) for nsteps steps;
// u is advanced in place. OrdinaryDiffEq's 2N form: tmp = A_s tmp + dt f(u); u += B_s tmp.
int sse_oracle_rk_steps(const sse_desc* desc, double* u, double dt, int nsteps) {
    Oracle<double> o;
    g_err = o.init(desc);
    if (!g_err.empty()) return SSE_ERR_UNSUPPORTED;
    size_t n = (size_t)desc->N_p * desc->N_c * desc->N_e;
    vector<double> k(n), tmp(n, 0.0);
    for (int step = 0; step < nsteps; ++step) {
        for (int s = 0; s < 5; ++s) {
            o.residual(u, k.data(), g_threads);
            for (size_t i = 0; i < n; ++i) {
                tmp[i] = (s == 0 ? 0.0 : CK_A[s] * tmp[i]) + dt * k[i];
                u[i] += CK_B[s] * tmp[i];
            }
        }
    }
    return 0;
}

// nsteps steps of an explicit Runge-Kutta method with s stages given by its Butcher
// tableau (A: s x s row-major, strictly lower triangular; b: s weights), fixed step
// dt, as OrdinaryDiffEq's explicit methods take with adaptive=false: stage i is
// k_i = f(u + dt sum_{j<i} A[i][j] k_j), then u += dt sum_i b_i k_i. With the
// Dormand-Prince 8(5,3) tableau this is DP8() (test/euler_3d.jl:49-56). The
// residual is autonomous, so the nodes c_i are not needed.
int sse_oracle_erk_steps(const sse_desc* desc, double* u, double dt, int nsteps, int s, const double* A,
                         const double* b) {
    if (s < 1 || s > 32) {
        g_err = "stages must be 1..32";
        return SSE_ERR_INVALID_ARGUMENT;
    }
    Oracle<double> o;
    g_err = o.init(desc);
    if (!g_err.empty()) return SSE_ERR_UNSUPPORTED;
    const size_t n = (size_t)desc->N_p * desc->N_c * desc->N_e;
    vector<double> k((size_t)s * n), y(n);
    for (int step = 0; step < nsteps; ++step) {
        for (int i = 0; i < s; ++i) {
            for (size_t t = 0; t < n; ++t) {
                double acc = 0.0;
                for (int j = 0; j < i; ++j) acc += A[i * s + j] * k[(size_t)j * n + t];
                y[t] = u[t] + dt * acc;
            }
            o.residual(y.data(), k.data() + (size_t)i * n, g_threads);
        }
        for (size_t t = 0; t < n; ++t) {
            double acc = 0.0;
            for (int i = 0; i < s; ++i) acc += b[i] * k[(size_t)i * n + t];
            u[t] += dt * acc;
        }
    }
    return 0;
}

// Directional derivative of the residual along (du, dlambda, db), lambda the
// LaxFriedrichsNumericalFlux coefficient (halfλ = λ / 2) and b the diffusion
// coefficient: dudt = R(u), ddudt = R'(u) du + dR/dλ dlambda + dR/db db.
// Forward-mode dual numbers through the same restatement.
int sse_oracle_residual_jvp(const sse_desc* desc, const double* u, const double* du, double dlambda, double db,
                            double* dudt, double* ddudt) {
    Oracle<Dual> o;
    g_err = o.init(desc);
    if (!g_err.empty()) return SSE_ERR_UNSUPPORTED;
    o.hl = Dual(desc->half_lambda, 0.5 * dlambda);
    o.bd = Dual(desc->b, db);
    const size_t n = (size_t)desc->N_p * desc->N_c * desc->N_e;
    vector<Dual> ud(n), rd(n);
    for (size_t i = 0; i < n; ++i) ud[i] = Dual(u[i], du[i]);
    o.residual(ud.data(), rd.data(), g_threads);
    for (size_t i = 0; i < n; ++i) { dudt[i] = rd[i].v; ddudt[i] = rd[i].d; }
    return 0;
}

// nsteps CK2N54 steps of (u, du) together: du becomes d u_N along (du_0, dlambda, db).
int sse_oracle_rk_jvp(const sse_desc* desc, double* u, double* du, double dlambda, double db, double dt,
                      int nsteps) {
    Oracle<Dual> o;
    g_err = o.init(desc);
    if (!g_err.empty()) return SSE_ERR_UNSUPPORTED;
    o.hl = Dual(desc->half_lambda, 0.5 * dlambda);
    o.bd = Dual(desc->b, db);
    const size_t n = (size_t)desc->N_p * desc->N_c * desc->N_e;
    vector<Dual> ud(n), k(n), tmp(n, Dual(0.0));
    for (size_t i = 0; i < n; ++i) ud[i] = Dual(u[i], du[i]);
    for (int step = 0; step < nsteps; ++step)
        for (int s = 0; s < 5; ++s) {
            o.residual(ud.data(), k.data(), g_threads);
            for (size_t i = 0; i < n; ++i) {
                tmp[i] = (s == 0 ? Dual(0.0) : CK_A[s] * tmp[i]) + dt * k[i];
                ud[i] += CK_B[s] * tmp[i];
            }
        }
    for (size_t i = 0; i < n; ++i) { u[i] = ud[i].v; du[i] = ud[i].d; }
    return 0;
}

// The tableau itself, for tests of its order conditions.
void sse_oracle_ck2n54(double* A, double* B, double* C) {
    for (int s = 0; s < 5; ++s) { A[s] = CK_A[s]; B[s] = CK_B[s]; C[s] = CK_C[s]; }
}

// Two-point flux, entropy variables and wave speed at single states, for the
// closed-form tests. F is N_c x d column-major.
int sse_oracle_two_point(int law, int d, double gamma, const double* a, int flux, const double* uL,
                         const double* uR, double* F) {
    Law<double> L;
    L.kind = law; L.d = d; L.N_c = (law == SSE_EULER ? d + 2 : 1);
    L.gamma = gamma; L.gm1 = gamma - 1.0; L.inv_gm1 = 1.0 / (gamma - 1.0);
    for (int m = 0; m < 3; ++m) L.a[m] = a ? a[m] : 0.0;
    L.two_point(flux, uL, uR, F);
    return 0;
}

void sse_oracle_c2e(int d, double gamma, const double* u, double* w) {
    Law<double> L;
    L.kind = SSE_EULER; L.d = d; L.N_c = d + 2; L.gamma = gamma; L.gm1 = gamma - 1.0; L.inv_gm1 = 1.0 / (gamma - 1.0);
    L.c2e(u, w);
}

void sse_oracle_e2c(int d, double gamma, const double* w, double* u) {
    Law<double> L;
    L.kind = SSE_EULER; L.d = d; L.N_c = d + 2; L.gamma = gamma; L.gm1 = gamma - 1.0; L.inv_gm1 = 1.0 / (gamma - 1.0);
    L.e2c(w, u);
}

double sse_oracle_logmean(double x, double y) { return Law<double>::logmean(x, y); }
double sse_oracle_inv_logmean(double x, double y) { return Law<double>::inv_logmean(x, y); }

}  // extern "C"
