// sse_standard.hip -- StandardForm residual kernels (first- and second-order laws),
// any element type and degree, one workgroup per element, working set in LDS.
//
// The reference's StandardForm time derivative with ReferenceOperators
// (src/Solvers/standard_form_first_order.jl:16-63) or PhysicalOperators
// (:65-94; src/Solvers/standard_form_second_order.jl:36-75) is
//   dudt = sum_n VOL_n f_n + FAC f*,
//   VOL_n = M^{-1} V' [ sum_m D_m' (1/2 W Lambda_mn) - (1/2 W Lambda_mn) D_m + R' (1/2 B nJf_n) R ]
//           (skew-symmetric mapping, src/Solvers/operators.jl:132-160), or
//   VOL_n = M^{-1} V' sum_m D_m' (W Lambda_mn)                   (standard mapping, :83-130),
//   FAC   = -M^{-1} V' R' (B J_f).
// Here the operators are applied matrix-free from the sparse rows of D_m, D_m',
// R and R': the dense per-element VOL/FAC matrices the reference builds for
// PhysicalOperators are never formed. The BR1 auxiliary variable
// (standard_form_second_order.jl:3-32) is q_m = -VOL_m u_q - FAC u*_m,
// u*_m = 1/2 (u^- + u^+) n_m, stored at volume and facet nodes for the second phase.
#include <hip/hip_runtime.h>

#include "sse_kernels.h"

namespace sse {

namespace {

struct StdLayout {
    int uq, x, r, t, ff, tf, L, q, o, size;
};

__host__ __device__ inline StdLayout std_layout(const DevParams& p) {
    StdLayout l;
    const int NC = p.N_c, Nq = p.N_q, Nf = p.N_f, d = p.d;
    int o = 0;
    l.uq = o; o += Nq * NC;
    l.x = o; o += Nq * NC;
    l.r = o; o += Nq * NC;
    l.t = o; o += Nq * NC;
    l.ff = o; o += Nf * NC;
    l.tf = o; o += Nf * NC;
    l.L = o; o += Nq * d * d;
    l.q = o; o += p.second_order ? Nq * NC * d : 0;
    l.o = o; o += p.N_p * NC;
    l.size = o;
    return l;
}

// r[i, e] += (volume term of direction n applied to x)[i, e] for the rows this
// thread owns; uses s_tf as facet scratch. Ends with a barrier.
template <int D>
__device__ void volume_term(const DevParams& p, int k, int n, const double* s_x, double* s_r, double* s_tf,
                            const double* s_L) {
    const int NC = p.N_c, Nq = p.N_q, Nf = p.N_f;
    const double h = p.skew ? 0.5 : 1.0;
    for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
        for (int e = 0; e < NC; ++e) {
            // sum_m D_m' (h W Lambda_mn x): row i gathers column i of each D_m
            double s = 0.0;
#pragma unroll 4
            for (int t = p.Dt_ptr[i]; t < p.Dt_ptr[i + 1]; ++t) {
                const int j = p.Dt_idx[t];
                double a = 0.0;
#pragma unroll
                for (int m = 0; m < D; ++m) a += p.Dt_val[t * D + m] * s_L[j + Nq * (m + D * n)];
                s += a * p.W[j] * s_x[j + Nq * e];
            }
            s *= h;
            if (p.skew) {  // - h W Lambda_mn (D_m x)
                double g = 0.0;
#pragma unroll 4
                for (int t = p.D_ptr[i]; t < p.D_ptr[i + 1]; ++t) {
                    const int j = p.D_idx[t];
                    double a = 0.0;
#pragma unroll
                    for (int m = 0; m < D; ++m) a += p.D_val[t * D + m] * s_L[i + Nq * (m + D * n)];
                    g += a * s_x[j + Nq * e];
                }
                s -= h * p.W[i] * g;
            }
            s_r[i + Nq * e] += s;
        }
    }
    if (p.skew) {  // + R' (1/2 B nJf_n (R x))
        const double* nJfk = p.nJf + (size_t)k * D * Nf;
        for (int j = threadIdx.x; j < Nf; j += blockDim.x) {
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
#pragma unroll 8
                for (int t = p.R_ptr[j]; t < p.R_ptr[j + 1]; ++t) s += p.R_val[t] * s_x[p.R_idx[t] + Nq * e];
                s_tf[j + Nf * e] = 0.5 * p.B[j] * nJfk[n + D * j] * s;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
#pragma unroll 8
                for (int t = p.Rt_ptr[i]; t < p.Rt_ptr[i + 1]; ++t) s += p.Rt_val[t] * s_tf[p.Rt_idx[t] + Nf * e];
                s_r[i + Nq * e] += s;
            }
        }
    }
    __syncthreads();
}

// y (N_p x NC) = V' x (x: N_q x NC), threads over (j, e), reading V' from global
// memory with consecutive j (coalesced), two accumulators per dot product
__device__ void std_vt(const DevParams& p, int NC, const double* x, double* y) {
    const int Nq = p.N_q, Np = p.N_p;
    for (int t = threadIdx.x; t < Np * NC; t += blockDim.x) {
        const int j = t % Np, e = t / Np;
        const double* xe = x + Nq * e;
        double s0 = 0.0, s1 = 0.0;
        int i = 0;
        for (; i + 1 < Nq; i += 2) {
            s0 += p.Vt[j + (size_t)Np * i] * xe[i];
            s1 += p.Vt[j + (size_t)Np * (i + 1)] * xe[i + 1];
        }
        if (i < Nq) s0 += p.Vt[j + (size_t)Np * i] * xe[i];
        y[j + Np * e] = s0 + s1;
    }
}

// y (N_q x NC) = scale V x (x: N_p x NC), threads over (i, e); scale = W / J_k if wj
__device__ void std_v(const DevParams& p, const double* Jk, bool wj, int NC, const double* x, double* y) {
    const int Nq = p.N_q, Np = p.N_p;
    for (int t = threadIdx.x; t < Nq * NC; t += blockDim.x) {
        const int i = t % Nq, e = t / Nq;
        const double* xe = x + Np * e;
        double s0 = 0.0, s1 = 0.0;
        int j = 0;
        for (; j + 1 < Np; j += 2) {
            s0 += p.V[i + (size_t)Nq * j] * xe[j];
            s1 += p.V[i + (size_t)Nq * (j + 1)] * xe[j + 1];
        }
        if (j < Np) s0 += p.V[i + (size_t)Nq * j] * xe[j];
        y[i + Nq * e] = (s0 + s1) * (wj ? p.W[i] / Jk[i] : 1.0);
    }
}

// s_o = M^{-1} V' s_r (weight-adjusted with M^{-1} = I, or diagonal W J); s_t is scratch.
__device__ void mass_solve_vt(const DevParams& p, int k, const double* s_r, double* s_t, double* s_o) {
    const int NC = p.N_c, Nq = p.N_q, Np = p.N_p;
    const double* Jk = p.Jq + (size_t)k * Nq;
    if (!p.has_V) {
        for (int j = threadIdx.x; j < Np; j += blockDim.x) {
            const double wj = p.W[j] * Jk[j];
            for (int e = 0; e < NC; ++e) s_o[j + Np * e] = s_r[j + Nq * e] / wj;
        }
        __syncthreads();
        return;
    }
    std_vt(p, NC, s_r, s_o);  // V' r
    __syncthreads();
    std_v(p, Jk, true, NC, s_o, s_t);  // J^{-1} W (V rhs)
    __syncthreads();
    std_vt(p, NC, s_t, s_o);
    __syncthreads();
}

}  // namespace

// BR1 auxiliary variable q_m (second-order laws, N_c = 1)
template <int D>
__global__ void __launch_bounds__(256) k_auxiliary_standard(DevParams p) {
    extern __shared__ double sm[];
    const StdLayout l = std_layout(p);
    const int k = blockIdx.x;
    const int Nq = p.N_q, Nf = p.N_f, Np = p.N_p;
    double *s_uq = sm + l.uq, *s_r = sm + l.r, *s_t = sm + l.t, *s_tf = sm + l.tf, *s_L = sm + l.L,
           *s_o = sm + l.o, *s_x = sm + l.x;
    const double* uqk = p.uq + (size_t)k * Nq;
    const double* Lk = p.Lam + (size_t)k * Nq * D * D;
    for (int t = threadIdx.x; t < Nq; t += blockDim.x) s_uq[t] = uqk[t];
    for (int t = threadIdx.x; t < Nq * D * D; t += blockDim.x) s_L[t] = Lk[t];
    __syncthreads();
    const double* nJfk = p.nJf + (size_t)k * D * Nf;
    const double* Jfk = p.Jf + (size_t)k * Nf;
    for (int m = 0; m < D; ++m) {
        for (int t = threadIdx.x; t < Nq; t += blockDim.x) s_r[t] = 0.0;
        __syncthreads();
        volume_term<D>(p, k, m, s_uq, s_r, s_tf, s_L);
        // r = -r + R' (B J_f u*_m),  u*_m = {u} n_m
        for (int j = threadIdx.x; j < Nf; j += blockDim.x) {
            const int g = p.mapP[j + (size_t)Nf * k];
            const double uo = p.uf[g];  // N_c = 1: u_f [k][j]
            const double ui = p.uf[j + (size_t)Nf * k];
            const double jf = Jfk[j];
            s_tf[j] = 0.5 * (ui + uo) * (nJfk[m + D * j] / jf) * (p.B[j] * jf);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
            double s = 0.0;
#pragma unroll 8
            for (int t = p.Rt_ptr[i]; t < p.Rt_ptr[i + 1]; ++t) s += p.Rt_val[t] * s_tf[p.Rt_idx[t]];
            s_r[i] = -s_r[i] + s;
        }
        __syncthreads();
        mass_solve_vt(p, k, s_r, s_t, s_o);
        // q_q = V q_m, q_f = R q_q
        double* qqk = p.qq + ((size_t)k * D + m) * Nq;
        double* qfk = p.qf + ((size_t)k * D + m) * Nf;
        if (p.has_V) {
            std_v(p, nullptr, false, 1, s_o, s_x);
            __syncthreads();
            for (int i = threadIdx.x; i < Nq; i += blockDim.x) qqk[i] = s_x[i];
        } else {
            for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
                s_x[i] = s_o[i];
                qqk[i] = s_o[i];
            }
        }
        __syncthreads();
        for (int j = threadIdx.x; j < Nf; j += blockDim.x) {
            double s = 0.0;
#pragma unroll 8
            for (int t = p.R_ptr[j]; t < p.R_ptr[j + 1]; ++t) s += p.R_val[t] * s_x[p.R_idx[t]];
            qfk[j] = s;
        }
        __syncthreads();
    }
}

template <int D, int LAW>
__global__ void __launch_bounds__(256) k_time_derivative_standard(DevParams p, double* __restrict__ dudt, RKStage rk) {
    using Ph = Physics<D, LAW>;
    constexpr int NC = Ph::NC;
    extern __shared__ double sm[];
    const StdLayout l = std_layout(p);
    const int k = blockIdx.x;
    const int Nq = p.N_q, Nf = p.N_f, Np = p.N_p;
    double *s_uq = sm + l.uq, *s_x = sm + l.x, *s_r = sm + l.r, *s_t = sm + l.t, *s_ff = sm + l.ff,
           *s_tf = sm + l.tf, *s_L = sm + l.L, *s_q = sm + l.q, *s_o = sm + l.o;
    const double* uqk = p.uq + (size_t)k * Nq * NC;
    const double* ufk = p.uf + (size_t)k * Nf * NC;
    const double* Lk = p.Lam + (size_t)k * Nq * D * D;
    for (int t = threadIdx.x; t < Nq * NC; t += blockDim.x) {
        s_uq[t] = uqk[t];
        s_r[t] = 0.0;
    }
    for (int t = threadIdx.x; t < Nq * D * D; t += blockDim.x) s_L[t] = Lk[t];
    if (p.second_order)
        for (int t = threadIdx.x; t < Nq * D; t += blockDim.x) s_q[t] = p.qq[(size_t)k * D * Nq + t];
    const double* nJfk = p.nJf + (size_t)k * D * Nf;
    const double* Jfk = p.Jf + (size_t)k * Nf;

    // interface flux f* (numerical_flux! with the conservative two-point flux;
    // BR1 adds -b {q} . n), scaled by B J_f
    for (int j = threadIdx.x; j < Nf; j += blockDim.x) {
        double uin[NC], uout[NC], F[D * NC], nrm[D];
        const int g = p.mapP[j + (size_t)Nf * k];
        const int kn = g / Nf, jn = g - kn * Nf;
        const double* ufn = p.uf + (size_t)kn * Nf * NC;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            uin[e] = ufk[j + Nf * e];
            uout[e] = ufn[jn + Nf * e];
        }
        const double jf = Jfk[j];
#pragma unroll
        for (int m = 0; m < D; ++m) nrm[m] = nJfk[m + D * j] / jf;
        Ph::two_point(p.law, uin, uout, F);
        const bool lf = (p.law.inviscid_flux == 1);
        const double a = lf ? p.law.half_lambda * Ph::wave_speed(p.law, uin, uout, nrm) : 0.0;
        const double bj = p.B[j] * jf;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            double fnv = 0.0;
#pragma unroll
            for (int m = 0; m < D; ++m) fnv += F[m * NC + e] * nrm[m];
            if (lf) fnv += a * (uin[e] - uout[e]);
            if (p.second_order) {
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    const double qi = p.qf[((size_t)k * D + m) * Nf + j];
                    const double qo = p.qf[((size_t)kn * D + m) * Nf + jn];
                    s += p.b * (-0.5 * (qi + qo)) * nrm[m];
                }
                fnv += s;
            }
            s_ff[j + Nf * e] = fnv * bj;
        }
    }
    __syncthreads();

    // sum_n VOL_n f_n (before V' and the mass solve)
    for (int n = 0; n < D; ++n) {
        for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
            double ui[NC], F[D * NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) ui[e] = s_uq[i + Nq * e];
            if constexpr (LAW == LAW_EULER) {
                Ph::physical_flux(p.law, ui, F);
#pragma unroll
                for (int e = 0; e < NC; ++e) s_x[i + Nq * e] = F[n * NC + e];
            } else {
                double f = p.law.a[n] * ui[0];
                if (p.second_order) f -= p.b * s_q[n * Nq + i];
                s_x[i] = f;
            }
        }
        __syncthreads();
        volume_term<D>(p, k, n, s_x, s_r, s_tf, s_L);
    }

    // r -= R' (B J_f f*)
    for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
#pragma unroll 8
            for (int t = p.Rt_ptr[i]; t < p.Rt_ptr[i + 1]; ++t) s += p.Rt_val[t] * s_ff[p.Rt_idx[t] + Nf * e];
            s_r[i + Nq * e] -= s;
        }
    }
    __syncthreads();
    mass_solve_vt(p, k, s_r, s_t, s_o);
    const size_t base = (size_t)k * Np * NC;
    if (!rk.active) {
        for (int t = threadIdx.x; t < Np * NC; t += blockDim.x) dudt[base + t] = s_o[t];
    } else {
        for (int t = threadIdx.x; t < Np * NC; t += blockDim.x) {
            const double kt = (rk.A == 0.0 ? 0.0 : rk.A * rk.tmp[base + t]) + rk.dt_of(k) * s_o[t];
            rk.tmp[base + t] = kt;
            rk.u[base + t] += rk.B * kt;
        }
    }
}

size_t lds_standard(const DevParams& p) { return sizeof(double) * (size_t)std_layout(p).size; }

template <int D, int LAW>
hipError_t launch_standard_t(const DevParams& p, double* dudt, const RKStage& rk, hipStream_t s, int which) {
    dim3 grid(p.N_e), block(256);
    const size_t lds = lds_standard(p);
    if ((which & 2) && p.second_order) {
        hipLaunchKernelGGL((k_auxiliary_standard<D>), grid, block, lds, s, p);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    if (which & 2) {
        hipLaunchKernelGGL((k_time_derivative_standard<D, LAW>), grid, block, lds, s, p, dudt, rk);
        return hipGetLastError();
    }
    return hipSuccess;
}

// Phase 2 of StandardForm (phase 1, u_q = V u and u_f = R u_q, is the generic
// nodal_values kernel without entropy projection).
hipError_t launch_standard(const DevParams& p, double* dudt, const RKStage& rk, hipStream_t s, int which) {
    const bool euler = (p.N_c != 1);
    switch (p.d) {
        case 1: return euler ? launch_standard_t<1, LAW_EULER>(p, dudt, rk, s, which)
                             : launch_standard_t<1, LAW_ADVECTION>(p, dudt, rk, s, which);
        case 2: return euler ? launch_standard_t<2, LAW_EULER>(p, dudt, rk, s, which)
                             : launch_standard_t<2, LAW_ADVECTION>(p, dudt, rk, s, which);
        case 3: return euler ? launch_standard_t<3, LAW_EULER>(p, dudt, rk, s, which)
                             : launch_standard_t<3, LAW_ADVECTION>(p, dudt, rk, s, which);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace sse
