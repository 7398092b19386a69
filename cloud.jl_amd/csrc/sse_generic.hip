// sse_generic.hip -- residual kernels for any element type and degree.
//
// One workgroup per element; sizes are runtime values and the per-element
// working set is staged in LDS. Phase 1 (nodal_values) and phase 2
// (time_derivative) are separate launches because phase 2 gathers the
// neighbours' facet states written by phase 1 (the reference's barrier between
// its two element loops, src/Solvers/Solvers.jl:505-511).
//
//   nodal_values     src/Solvers/flux_differencing_form.jl:171-292
//   time_derivative  src/Solvers/flux_differencing_form.jl:294-347
//   mass solve (WA)  src/Solvers/mass_matrix.jl:406-417
//
// The Hadamard sum is evaluated in gather form: the thread owning volume node
// i loops over every j with S_m[i,j] != 0 for some m and accumulates
//   r_q[i,e] -= sum_m S_m[i,j] sum_n (Lambda_i + Lambda_j)[m,n] F(u_i,u_j)[e,n],
// which equals the reference's scatter over i < j (S antisymmetric, F symmetric)
// without atomics; each pair is evaluated from both ends.
#include <hip/hip_runtime.h>

#include "sse_kernels.h"

namespace sse {

template <int D, int LAW>
__global__ void __launch_bounds__(256) k_nodal_values_generic(DevParams p, const double* __restrict__ u) {
    using Ph = Physics<D, LAW>;
    constexpr int NC = Ph::NC;
    extern __shared__ double sm[];
    const int k = blockIdx.x;
    const int N_p = p.N_p, N_q = p.N_q, N_f = p.N_f;
    double* s_u = sm;              // N_p x NC
    double* s_a = s_u + N_p * NC;  // N_q x NC
    double* s_w = s_a + N_q * NC;  // N_p x NC
    const double* uk = u + (size_t)k * N_p * NC;
    for (int t = threadIdx.x; t < N_p * NC; t += blockDim.x) s_u[t] = uk[t];
    __syncthreads();
    double* uqk = p.uq + (size_t)k * N_q * NC;
    double* ufk = p.uf + (size_t)k * N_f * NC;
    bool ok = true;

    if (!p.project) {
        // u_q = V u, u_f = R u_q (scalar laws and diagonal-E nodal schemes)
        for (int i = threadIdx.x; i < N_q; i += blockDim.x) {
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s;
                if (p.has_V) {
                    s = 0.0;
                    for (int j = 0; j < N_p; ++j) s += p.V[i + (size_t)N_q * j] * s_u[j + N_p * e];
                } else {
                    s = s_u[i + N_p * e];
                }
                s_a[i + N_q * e] = s;
                uqk[i + N_q * e] = s;
            }
        }
        __syncthreads();
        for (int j = threadIdx.x; j < N_f; j += blockDim.x) {
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int t = p.R_ptr[j]; t < p.R_ptr[j + 1]; ++t) s += p.R_val[t] * s_a[p.R_idx[t] + N_q * e];
                ufk[j + N_f * e] = s;
            }
        }
        return;
    }

    if (!p.has_V) {
        // nodal scheme, general R: u_q = u, w_q = c2e(u_q), u_f = e2c(R w_q)
        for (int i = threadIdx.x; i < N_q; i += blockDim.x) {
            double ub[NC], wb[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                ub[e] = s_u[i + N_p * e];
                uqk[i + N_q * e] = ub[e];
            }
            ok &= Ph::c2e(p.law, ub, wb);
#pragma unroll
            for (int e = 0; e < NC; ++e) s_a[i + N_q * e] = wb[e];
        }
    } else {
        // modal entropy projection
        const double* Jk = p.Jq + (size_t)k * N_q;
        for (int i = threadIdx.x; i < N_q; i += blockDim.x) {
            double ub[NC], wb[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int j = 0; j < N_p; ++j) s += p.V[i + (size_t)N_q * j] * s_u[j + N_p * e];
                ub[e] = s;
            }
            ok &= Ph::c2e(p.law, ub, wb);
            const double wj = p.W[i] * Jk[i];
#pragma unroll
            for (int e = 0; e < NC; ++e) s_a[i + N_q * e] = wb[e] * wj;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < N_p; j += blockDim.x) {  // w = V' (WJ w_q)
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int i = 0; i < N_q; ++i) s += p.V[i + (size_t)N_q * j] * s_a[i + N_q * e];
                s_w[j + N_p * e] = s;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < N_q; i += blockDim.x) {  // J^{-1}W (V w)
            const double jw = p.W[i] / Jk[i];
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int j = 0; j < N_p; ++j) s += p.V[i + (size_t)N_q * j] * s_w[j + N_p * e];
                s_a[i + N_q * e] = s * jw;
            }
        }
        __syncthreads();
        for (int j = threadIdx.x; j < N_p; j += blockDim.x) {  // w = V' (...)
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int i = 0; i < N_q; ++i) s += p.V[i + (size_t)N_q * j] * s_a[i + N_q * e];
                s_w[j + N_p * e] = s;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < N_q; i += blockDim.x) {  // w_q = V w
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int j = 0; j < N_p; ++j) s += p.V[i + (size_t)N_q * j] * s_w[j + N_p * e];
                s_a[i + N_q * e] = s;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < N_q; i += blockDim.x) {  // u_q = e2c(w_q)
            double ub[NC], wb[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) wb[e] = s_a[i + N_q * e];
            ok &= Ph::e2c(p.law, wb, ub);
#pragma unroll
            for (int e = 0; e < NC; ++e) uqk[i + N_q * e] = ub[e];
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < N_f; j += blockDim.x) {  // u_f = e2c(R w_q)
        double ub[NC], wb[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            for (int t = p.R_ptr[j]; t < p.R_ptr[j + 1]; ++t) s += p.R_val[t] * s_a[p.R_idx[t] + N_q * e];
            wb[e] = s;
        }
        ok &= Ph::e2c(p.law, wb, ub);
#pragma unroll
        for (int e = 0; e < NC; ++e) ufk[j + N_f * e] = ub[e];
    }
    if (!ok) atomicOr(p.status, 1);
}

// 0.5 * nJq[n, f, i] = 0.5 * sum_l Lambda_q[i, l, n] n_ref_f[l]
template <int D>
__device__ __forceinline__ double half_nJq(const double* s_L, int N_q, const double* fn, int n, int f, int i) {
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < D; ++l) s += s_L[i + N_q * (l + D * n)] * fn[l + D * f];
    return 0.5 * s;
}

template <int D, int LAW>
__global__ void __launch_bounds__(256) k_time_derivative_generic(DevParams p, double* __restrict__ dudt, RKStage rk) {
    using Ph = Physics<D, LAW>;
    constexpr int NC = Ph::NC;
    extern __shared__ double sm[];
    const int k = blockIdx.x;
    const int N_p = p.N_p, N_q = p.N_q, N_f = p.N_f, npf = p.npf;
    double* s_uq = sm;                 // N_q x NC
    double* s_uf = s_uq + N_q * NC;    // N_f x NC
    double* s_ff = s_uf + N_f * NC;    // N_f x NC
    double* s_rq = s_ff + N_f * NC;    // N_q x NC
    double* s_L = s_rq + N_q * NC;     // N_q x D x D
    double* s_t = s_L + N_q * D * D;   // N_q x NC
    double* s_o = s_t + N_q * NC;      // N_p x NC
    const double* uqk = p.uq + (size_t)k * N_q * NC;
    const double* ufk = p.uf + (size_t)k * N_f * NC;
    const double* Lk = p.Lam + (size_t)k * N_q * D * D;
    for (int t = threadIdx.x; t < N_q * NC; t += blockDim.x) s_uq[t] = uqk[t];
    for (int t = threadIdx.x; t < N_f * NC; t += blockDim.x) s_uf[t] = ufk[t];
    for (int t = threadIdx.x; t < N_q * D * D; t += blockDim.x) s_L[t] = Lk[t];
    __syncthreads();
    const double* nJfk = p.nJf + (size_t)k * D * N_f;
    const double* Jfk = p.Jf + (size_t)k * N_f;

    // interface flux and the facet-node side of the facet correction
    for (int j = threadIdx.x; j < N_f; j += blockDim.x) {
        double uin[NC], uout[NC], F[D * NC], nrm[D];
        const int g = p.mapP[j + (size_t)N_f * k];
        const int kn = g / N_f, jn = g - kn * N_f;
        const double* ufn = p.uf + (size_t)kn * N_f * NC;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            uin[e] = s_uf[j + N_f * e];
            uout[e] = ufn[jn + N_f * e];
        }
        const double jf = Jfk[j];
#pragma unroll
        for (int m = 0; m < D; ++m) nrm[m] = nJfk[m + D * j] / jf;
        Ph::two_point(p.law, uin, uout, F);
        double a = 0.0;
        const bool lf = (p.law.inviscid_flux == 1);
        if (lf) a = p.law.half_lambda * Ph::wave_speed(p.law, uin, uout, nrm);
        const double bj = p.B[j] * jf;
        double ff[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            double fnv = 0.0;
#pragma unroll
            for (int m = 0; m < D; ++m) fnv += F[m * NC + e] * nrm[m];
            if (lf) fnv += a * (uin[e] - uout[e]);
            ff[e] = fnv * bj;
        }
        if (p.C_kind) {
            const int f = j / npf;
            double hf[D];
#pragma unroll
            for (int m = 0; m < D; ++m) hf[m] = 0.5 * nJfk[m + D * j];
            for (int t = p.Cc_ptr[j]; t < p.Cc_ptr[j + 1]; ++t) {
                const int i = p.Cc_idx[t];
                const double cij = p.Cc_val[t];
                double ui[NC];
#pragma unroll
                for (int e = 0; e < NC; ++e) ui[e] = s_uq[i + N_q * e];
                Ph::two_point(p.law, ui, uin, F);
                double nJ[D];
#pragma unroll
                for (int m = 0; m < D; ++m) nJ[m] = hf[m] + half_nJq<D>(s_L, N_q, p.fn, m, f, i);
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    double fnv = 0.0;
#pragma unroll
                    for (int m = 0; m < D; ++m) fnv += nJ[m] * F[m * NC + e];
                    ff[e] -= cij * fnv;
                }
            }
        }
#pragma unroll
        for (int e = 0; e < NC; ++e) s_ff[j + N_f * e] = ff[e];
    }

    // volume Hadamard sum and the volume-node side of the facet correction
    for (int i = threadIdx.x; i < N_q; i += blockDim.x) {
        double ui[NC], uj[NC], F[D * NC], r[NC];
        double Li[D * D];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            ui[e] = s_uq[i + N_q * e];
            r[e] = 0.0;
        }
#pragma unroll
        for (int mn = 0; mn < D * D; ++mn) Li[mn] = s_L[i + N_q * mn];
        for (int t = p.S_ptr[i]; t < p.S_ptr[i + 1]; ++t) {
            const int j = p.S_idx[t];
#pragma unroll
            for (int e = 0; e < NC; ++e) uj[e] = s_uq[j + N_q * e];
            Ph::two_point(p.law, ui, uj, F);
            // a_n = sum_m S_m[i,j] (Lambda_i + Lambda_j)[m,n]
            double an[D];
#pragma unroll
            for (int n = 0; n < D; ++n) {
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < D; ++m) s += p.S_val[t * D + m] * (Li[m + D * n] + s_L[j + N_q * (m + D * n)]);
                an[n] = s;
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
#pragma unroll
                for (int n = 0; n < D; ++n) s += an[n] * F[n * NC + e];
                r[e] -= s;
            }
        }
        if (p.C_kind) {
            for (int t = p.Cr_ptr[i]; t < p.Cr_ptr[i + 1]; ++t) {
                const int j = p.Cr_idx[t];
                const double cij = p.Cr_val[t];
                const int f = j / npf;
#pragma unroll
                for (int e = 0; e < NC; ++e) uj[e] = s_uf[j + N_f * e];
                Ph::two_point(p.law, ui, uj, F);
                double nJ[D];
#pragma unroll
                for (int m = 0; m < D; ++m) nJ[m] = 0.5 * nJfk[m + D * j] + half_nJq<D>(s_L, N_q, p.fn, m, f, i);
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    double fnv = 0.0;
#pragma unroll
                    for (int m = 0; m < D; ++m) fnv += nJ[m] * F[m * NC + e];
                    r[e] -= cij * fnv;
                }
            }
        }
#pragma unroll
        for (int e = 0; e < NC; ++e) s_rq[i + N_q * e] = r[e];
    }
    __syncthreads();

    // r_q -= R' f_f
    for (int i = threadIdx.x; i < N_q; i += blockDim.x) {
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            for (int t = p.Rt_ptr[i]; t < p.Rt_ptr[i + 1]; ++t) s += p.Rt_val[t] * s_ff[p.Rt_idx[t] + N_f * e];
            s_rq[i + N_q * e] -= s;
        }
    }
    __syncthreads();

    const double* Jk = p.Jq + (size_t)k * N_q;
    if (p.has_V) {
        for (int j = threadIdx.x; j < N_p; j += blockDim.x) {  // V' r_q
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int i = 0; i < N_q; ++i) s += p.V[i + (size_t)N_q * j] * s_rq[i + N_q * e];
                s_o[j + N_p * e] = s;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < N_q; i += blockDim.x) {  // J^{-1}W (V rhs)
            const double jw = p.W[i] / Jk[i];
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int j = 0; j < N_p; ++j) s += p.V[i + (size_t)N_q * j] * s_o[j + N_p * e];
                s_t[i + N_q * e] = s * jw;
            }
        }
        __syncthreads();
        for (int j = threadIdx.x; j < N_p; j += blockDim.x) {
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int i = 0; i < N_q; ++i) s += p.V[i + (size_t)N_q * j] * s_t[i + N_q * e];
                s_o[j + N_p * e] = s;
            }
        }
    } else {
        for (int j = threadIdx.x; j < N_p; j += blockDim.x) {  // (W J)^{-1} r_q
            const double wj = p.W[j] * Jk[j];
#pragma unroll
            for (int e = 0; e < NC; ++e) s_o[j + N_p * e] = s_rq[j + N_q * e] / wj;
        }
    }
    __syncthreads();
    const size_t base = (size_t)k * N_p * NC;
    if (!rk.active) {
        for (int t = threadIdx.x; t < N_p * NC; t += blockDim.x) dudt[base + t] = s_o[t];
    } else {
        for (int t = threadIdx.x; t < N_p * NC; t += blockDim.x) {
            const double kt = (rk.A == 0.0 ? 0.0 : rk.A * rk.tmp[base + t]) + rk.dt_of(k) * s_o[t];
            rk.tmp[base + t] = kt;
            rk.u[base + t] += rk.B * kt;
        }
    }
}

size_t lds_nodal_values_generic(const DevParams& p) {
    return sizeof(double) * (size_t)(2 * p.N_p + p.N_q) * p.N_c;
}

size_t lds_time_derivative_generic(const DevParams& p) {
    return sizeof(double) * ((size_t)(3 * p.N_q + 2 * p.N_f + p.N_p) * p.N_c + (size_t)p.N_q * p.d * p.d);
}

template <int D, int LAW>
hipError_t launch_generic(const DevParams& p, const double* u, double* dudt, const RKStage& rk, hipStream_t s,
                          int which /* bit 0: phase 1, bit 1: phase 2 */) {
    dim3 grid(p.N_e), block(256);
    if (which & 1) {
        hipLaunchKernelGGL((k_nodal_values_generic<D, LAW>), grid, block, lds_nodal_values_generic(p), s, p, u);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    if (which & 2) {
        hipLaunchKernelGGL((k_time_derivative_generic<D, LAW>), grid, block, lds_time_derivative_generic(p), s, p,
                           dudt, rk);
        return hipGetLastError();
    }
    return hipSuccess;
}

hipError_t launch_generic_dispatch(const DevParams& p, const double* u, double* dudt, const RKStage& rk,
                                   hipStream_t s, int which) {
    const bool euler = (p.N_c != 1);
    switch (p.d) {
        case 1: return euler ? launch_generic<1, LAW_EULER>(p, u, dudt, rk, s, which)
                             : launch_generic<1, LAW_ADVECTION>(p, u, dudt, rk, s, which);
        case 2: return euler ? launch_generic<2, LAW_EULER>(p, u, dudt, rk, s, which)
                             : launch_generic<2, LAW_ADVECTION>(p, u, dudt, rk, s, which);
        case 3: return euler ? launch_generic<3, LAW_EULER>(p, u, dudt, rk, s, which)
                             : launch_generic<3, LAW_ADVECTION>(p, u, dudt, rk, s, which);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace sse
