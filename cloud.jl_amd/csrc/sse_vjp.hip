// sse_vjp.hip -- vector-Jacobian product of the flux-differencing Euler residual:
// given u and a cotangent ybar of dudt, ubar = (dR/du)' ybar and
// lambda_bar = (dR/dλ)' ybar, λ the Lax-Friedrichs coefficient (halfλ = λ / 2).
//
// The residual (src/Solvers/flux_differencing_form.jl:171-347) is reversed step by
// step; no operator is formed:
//   dudt = M^{-1} V' r_q                  ->  rbar = V M^{-1} ybar (M^{-1} symmetric)
//   r_q -= R' f_f                         ->  fbar = -R rbar
//   f_f = B J_f f*(u_f, u_f+, n)          ->  interface flux adjoint (own and neighbour side)
//   facet correction, Hadamard sum        ->  two-point flux adjoints, gathered per node
//   entropy projection (e2c, R, V, WADG, V', c2e, V) reversed in order.
// Every per-node sum is a gather with a fixed order (no atomics). Each two-point
// term is differentiated once from each end, where the gathering node keeps its
// own side of the pull-back.
//
// Phase V1 (per element): rbar and fbar. Phase V2 (per element): everything else,
// reading the neighbours' fbar for the exterior side of the interface flux.
#include <hip/hip_runtime.h>

#include "sse_adjoint.h"
#include "sse_kernels.h"

namespace sse {

namespace {

constexpr int VT = 256;

// s_o (N_p x NC) = M^{-1} x for x (N_p x NC) in s_o; s_t (N_q x NC) scratch.
__device__ void wadg_inplace(const DevParams& p, int k, double* s_o, double* s_t, int NC) {
    const int Nq = p.N_q, Np = p.N_p;
    const double* Jk = p.Jq + (size_t)k * Nq;
    for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
        const double jw = p.W[i] / Jk[i];
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            for (int j = 0; j < Np; ++j) s += p.V[i + (size_t)Nq * j] * s_o[j + Np * e];
            s_t[i + Nq * e] = s * jw;
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < Np; j += blockDim.x)
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            for (int i = 0; i < Nq; ++i) s += p.V[i + (size_t)Nq * j] * s_t[i + Nq * e];
            s_o[j + Np * e] = s;
        }
    __syncthreads();
}

__device__ double block_sum_vjp(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = VT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

}  // namespace

// V1: rbar (N_q x NC) = V M^{-1} ybar, fbar (N_f x NC) = -R rbar, per element
template <int D>
__global__ void __launch_bounds__(VT) k_vjp_v1(DevParams p, const double* __restrict__ ybar, double* __restrict__ rbar,
                                               double* __restrict__ fbar) {
    constexpr int NC = D + 2;
    extern __shared__ double sm[];
    const int k = blockIdx.x;
    const int Np = p.N_p, Nq = p.N_q, Nf = p.N_f;
    double* s_o = sm;               // Np x NC
    double* s_t = s_o + Np * NC;    // Nq x NC
    double* s_r = s_t + Nq * NC;    // Nq x NC
    const double* yk = ybar + (size_t)k * Np * NC;
    for (int t = threadIdx.x; t < Np * NC; t += blockDim.x) s_o[t] = yk[t];
    __syncthreads();
    const double* Jk = p.Jq + (size_t)k * Nq;
    if (p.has_V) {
        wadg_inplace(p, k, s_o, s_t, NC);
        for (int i = threadIdx.x; i < Nq; i += blockDim.x)
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int j = 0; j < Np; ++j) s += p.V[i + (size_t)Nq * j] * s_o[j + Np * e];
                s_r[i + Nq * e] = s;
            }
    } else {
        for (int i = threadIdx.x; i < Nq; i += blockDim.x)
            for (int e = 0; e < NC; ++e) s_r[i + Nq * e] = s_o[i + Np * e] / (p.W[i] * Jk[i]);
    }
    __syncthreads();
    double* rk = rbar + (size_t)k * Nq * NC;
    for (int t = threadIdx.x; t < Nq * NC; t += blockDim.x) rk[t] = s_r[t];
    double* fk = fbar + (size_t)k * Nf * NC;
    for (int j = threadIdx.x; j < Nf; j += blockDim.x)
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            for (int t = p.R_ptr[j]; t < p.R_ptr[j + 1]; ++t) s += p.R_val[t] * s_r[p.R_idx[t] + Nq * e];
            fk[j + Nf * e] = -s;
        }
}

template <int D>
__global__ void __launch_bounds__(VT) k_vjp_v2(DevParams p, const double* __restrict__ u,
                                               const double* __restrict__ rbar, const double* __restrict__ fbar,
                                               double* __restrict__ ubar, double* __restrict__ lam_partial) {
    constexpr int NC = D + 2;
    using A = EulerAdj<D>;
    using Ph = Physics<D, LAW_EULER>;
    extern __shared__ double sm[];
    const int k = blockIdx.x;
    const int Np = p.N_p, Nq = p.N_q, Nf = p.N_f, npf = p.npf;
    double* red = sm;                     // VT
    double* s_uq = red + VT;              // Nq x NC   projected volume states (u tilde)
    double* s_uf = s_uq + Nq * NC;        // Nf x NC   projected facet states
    double* s_rb = s_uf + Nf * NC;        // Nq x NC   rbar
    double* s_fb = s_rb + Nq * NC;        // Nf x NC   fbar
    double* s_ubq = s_fb + Nf * NC;       // Nq x NC   cotangent of u tilde (volume), then of w_q'
    double* s_ubf = s_ubq + Nq * NC;      // Nf x NC   cotangent of u tilde (facet), then of w_f
    double* s_L = s_ubf + Nf * NC;        // Nq x D x D
    double* s_u0 = s_L + Nq * D * D;      // Nq x NC   V u (un-projected)
    double* s_wq = s_u0 + Nq * NC;        // Nq x NC   projected entropy variables w_q'
    double* s_wf = s_wq + Nq * NC;        // Nf x NC   w_f = R w_q'
    double* s_t = s_wf + Nf * NC;         // Nq x NC   scratch
    double* s_o = s_t + Nq * NC;          // Np x NC   scratch
    const size_t kq = (size_t)k * Nq * NC, kf = (size_t)k * Nf * NC;
    for (int t = threadIdx.x; t < Nq * NC; t += blockDim.x) {
        s_uq[t] = p.uq[kq + t];
        s_rb[t] = rbar[kq + t];
        s_ubq[t] = 0.0;
    }
    for (int t = threadIdx.x; t < Nf * NC; t += blockDim.x) {
        s_uf[t] = p.uf[kf + t];
        s_fb[t] = fbar[kf + t];
        s_ubf[t] = 0.0;
    }
    const double* Lk = p.Lam + (size_t)k * Nq * D * D;
    for (int t = threadIdx.x; t < Nq * D * D; t += blockDim.x) s_L[t] = Lk[t];
    const double* uk = u + (size_t)k * Np * NC;
    for (int t = threadIdx.x; t < Np * NC; t += blockDim.x) s_o[t] = uk[t];
    __syncthreads();

    // recompute the entropy projection's intermediates
    const double* Jk = p.Jq + (size_t)k * Nq;
    if (p.has_V) {
        for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
            double ub[NC], wb[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int j = 0; j < Np; ++j) s += p.V[i + (size_t)Nq * j] * s_o[j + Np * e];
                ub[e] = s;
                s_u0[i + Nq * e] = s;
            }
            Ph::c2e(p.law, ub, wb);
            const double wj = p.W[i] * Jk[i];
#pragma unroll
            for (int e = 0; e < NC; ++e) s_t[i + Nq * e] = wb[e] * wj;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < Np; j += blockDim.x)
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int i = 0; i < Nq; ++i) s += p.V[i + (size_t)Nq * j] * s_t[i + Nq * e];
                s_o[j + Np * e] = s;
            }
        __syncthreads();
        wadg_inplace(p, k, s_o, s_t, NC);
        for (int i = threadIdx.x; i < Nq; i += blockDim.x)
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int j = 0; j < Np; ++j) s += p.V[i + (size_t)Nq * j] * s_o[j + Np * e];
                s_wq[i + Nq * e] = s;
            }
    } else {
        for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
            double ub[NC], wb[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                ub[e] = s_o[i + Np * e];
                s_u0[i + Nq * e] = ub[e];
            }
            Ph::c2e(p.law, ub, wb);
#pragma unroll
            for (int e = 0; e < NC; ++e) s_wq[i + Nq * e] = wb[e];
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < Nf; j += blockDim.x)
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            for (int t = p.R_ptr[j]; t < p.R_ptr[j + 1]; ++t) s += p.R_val[t] * s_wq[p.R_idx[t] + Nq * e];
            s_wf[j + Nf * e] = s;
        }

    // interface flux (own side and the neighbour's evaluation at the shared node),
    // and the facet side of the facet correction
    const double* nJfk = p.nJf + (size_t)k * D * Nf;
    const double* Jfk = p.Jf + (size_t)k * Nf;
    const bool lf = (p.law.inviscid_flux == 1);
    double lam = 0.0;
    for (int j = threadIdx.x; j < Nf; j += blockDim.x) {
        double uin[NC], uout[NC], nrm[D], nrm2[D], c[NC], ub[NC], junk[NC];
        const int g = p.mapP[j + (size_t)Nf * k];
        const int kn = g / Nf, jn = g - kn * Nf;
        const double* ufn = p.uf + (size_t)kn * Nf * NC;
        const double* fbn = fbar + (size_t)kn * Nf * NC;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            uin[e] = s_uf[j + Nf * e];
            uout[e] = ufn[jn + Nf * e];
            ub[e] = 0.0;
        }
        const double jf = Jfk[j];
#pragma unroll
        for (int m = 0; m < D; ++m) nrm[m] = nJfk[m + D * j] / jf;
        const double bj = p.B[j] * jf;
#pragma unroll
        for (int e = 0; e < NC; ++e) c[e] = bj * s_fb[j + Nf * e];
        const double dh = A::lf_adj(p.law, lf, uin, uout, nrm, c, ub, junk);
        lam += 0.5 * dh;  // halfλ = λ / 2
        // the neighbour evaluates f*(u_out, u_in, n') at its node jn; our state is its exterior
        const double* nJfn = p.nJf + (size_t)kn * D * Nf;
        const double jfn = p.Jf[jn + (size_t)Nf * kn];
#pragma unroll
        for (int m = 0; m < D; ++m) nrm2[m] = nJfn[m + D * jn] / jfn;
        const double bjn = p.B[jn] * jfn;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            c[e] = bjn * fbn[jn + Nf * e];
            junk[e] = 0.0;
        }
        A::lf_adj(p.law, lf, uout, uin, nrm2, c, junk, ub);
        // facet correction, facet side: -C_ij G(u_i, u_fj) enters r_q[i] and f_f[j]
        if (p.C_kind) {
            const int f = j / npf;
            for (int t = p.Cc_ptr[j]; t < p.Cc_ptr[j + 1]; ++t) {
                const int i = p.Cc_idx[t];
                const double cij = p.Cc_val[t];
                double ui[NC], nJ[D], gb[NC];
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    ui[e] = s_uq[i + Nq * e];
                    gb[e] = -cij * (s_rb[i + Nq * e] + s_fb[j + Nf * e]);
                    junk[e] = 0.0;
                }
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    double s = 0.0;
#pragma unroll
                    for (int l = 0; l < D; ++l) s += s_L[i + Nq * (l + D * m)] * p.fn[l + D * f];
                    nJ[m] = 0.5 * nJfk[m + D * j] + 0.5 * s;
                }
                A::ec_dot_adj(p.law, ui, uin, nJ, gb, junk, ub);
            }
        }
#pragma unroll
        for (int e = 0; e < NC; ++e) s_ubf[j + Nf * e] = ub[e];
    }

    // volume nodes: Hadamard sum and the volume side of the facet correction
    for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
        double ui[NC], uj[NC], ub[NC], junk[NC], gb[NC], an[D];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            ui[e] = s_uq[i + Nq * e];
            ub[e] = 0.0;
        }
        for (int t = p.S_ptr[i]; t < p.S_ptr[i + 1]; ++t) {
            const int j = p.S_idx[t];
#pragma unroll
            for (int n = 0; n < D; ++n) {
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < D; ++m)
                    s += p.S_val[t * D + m] * (s_L[i + Nq * (m + D * n)] + s_L[j + Nq * (m + D * n)]);
                an[n] = s;
            }
            // r_i -= T_ij, r_j -= T_ji = +T_ij: u_i's cotangent from both is (dT_ij/du_L)' (rbar_j - rbar_i)
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                uj[e] = s_uq[j + Nq * e];
                gb[e] = s_rb[j + Nq * e] - s_rb[i + Nq * e];
                junk[e] = 0.0;
            }
            A::ec_dot_adj(p.law, ui, uj, an, gb, ub, junk);
        }
        if (p.C_kind) {
            for (int t = p.Cr_ptr[i]; t < p.Cr_ptr[i + 1]; ++t) {
                const int j = p.Cr_idx[t];
                const double cij = p.Cr_val[t];
                const int f = j / npf;
                double nJ[D];
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    double s = 0.0;
#pragma unroll
                    for (int l = 0; l < D; ++l) s += s_L[i + Nq * (l + D * m)] * p.fn[l + D * f];
                    nJ[m] = 0.5 * nJfk[m + D * j] + 0.5 * s;
                }
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    uj[e] = s_uf[j + Nf * e];
                    gb[e] = -cij * (s_rb[i + Nq * e] + s_fb[j + Nf * e]);
                    junk[e] = 0.0;
                }
                A::ec_dot_adj(p.law, ui, uj, nJ, gb, ub, junk);
            }
        }
#pragma unroll
        for (int e = 0; e < NC; ++e) s_ubq[i + Nq * e] = ub[e];
    }
    __syncthreads();

    // reverse the entropy projection
    if (p.has_V || !p.R_sel) {
        // u tilde = e2c(w): cotangents of w_q' (volume; modal only) and w_f
        if (p.has_V)
            for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
                double w[NC], ubv[NC], wb[NC];
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    w[e] = s_wq[i + Nq * e];
                    ubv[e] = s_ubq[i + Nq * e];
                    wb[e] = 0.0;
                }
                A::e2c_adj(p.law, w, ubv, wb);
#pragma unroll
                for (int e = 0; e < NC; ++e) s_t[i + Nq * e] = wb[e];
            }
        for (int j = threadIdx.x; j < Nf; j += blockDim.x) {
            double w[NC], ubv[NC], wb[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                w[e] = s_wf[j + Nf * e];
                ubv[e] = s_ubf[j + Nf * e];
                wb[e] = 0.0;
            }
            A::e2c_adj(p.law, w, ubv, wb);
#pragma unroll
            for (int e = 0; e < NC; ++e) s_ubf[j + Nf * e] = wb[e];
        }
        __syncthreads();
        // w_q' cotangent += R' w_f cotangent (modal); nodal: w_q = c2e(u) cotangent = R' w_f cotangent
        for (int i = threadIdx.x; i < Nq; i += blockDim.x)
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int t = p.Rt_ptr[i]; t < p.Rt_ptr[i + 1]; ++t) s += p.Rt_val[t] * s_ubf[p.Rt_idx[t] + Nf * e];
                if (p.has_V) s_t[i + Nq * e] += s;
                else s_t[i + Nq * e] = s;
            }
        __syncthreads();
        if (p.has_V) {
            // w_q' = V w ; w = M^{-1} wt ; wt = V' (WJ w_q) ; w_q = c2e(V u)
            for (int j = threadIdx.x; j < Np; j += blockDim.x)
                for (int e = 0; e < NC; ++e) {
                    double s = 0.0;
                    for (int i = 0; i < Nq; ++i) s += p.V[i + (size_t)Nq * j] * s_t[i + Nq * e];
                    s_o[j + Np * e] = s;
                }
            __syncthreads();
            wadg_inplace(p, k, s_o, s_wq, NC);  // s_wq no longer needed
            for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
                const double wj = p.W[i] * Jk[i];
                double wb[NC], u0[NC], ubv[NC];
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    double s = 0.0;
                    for (int j = 0; j < Np; ++j) s += p.V[i + (size_t)Nq * j] * s_o[j + Np * e];
                    wb[e] = s * wj;
                    u0[e] = s_u0[i + Nq * e];
                    ubv[e] = 0.0;
                }
                A::c2e_adj(p.law, u0, wb, ubv);
#pragma unroll
                for (int e = 0; e < NC; ++e) s_t[i + Nq * e] = ubv[e];
            }
            __syncthreads();
            double* outk = ubar + (size_t)k * Np * NC;
            for (int j = threadIdx.x; j < Np; j += blockDim.x)
                for (int e = 0; e < NC; ++e) {
                    double s = 0.0;
                    for (int i = 0; i < Nq; ++i) s += p.V[i + (size_t)Nq * j] * s_t[i + Nq * e];
                    outk[j + Np * e] = s;
                }
        } else {
            // nodal, general R: u tilde_q = u, w_q = c2e(u)
            double* outk = ubar + (size_t)k * Np * NC;
            for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
                double wb[NC], u0[NC], ubv[NC];
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    wb[e] = s_t[i + Nq * e];
                    u0[e] = s_u0[i + Nq * e];
                    ubv[e] = s_ubq[i + Nq * e];
                }
                A::c2e_adj(p.law, u0, wb, ubv);
#pragma unroll
                for (int e = 0; e < NC; ++e) outk[i + Np * e] = ubv[e];
            }
        }
    } else {
        // nodal diagonal-E: u tilde_f = selection of u
        double* outk = ubar + (size_t)k * Np * NC;
        for (int i = threadIdx.x; i < Nq; i += blockDim.x)
            for (int e = 0; e < NC; ++e) {
                double s = s_ubq[i + Nq * e];
                for (int t = p.Rt_ptr[i]; t < p.Rt_ptr[i + 1]; ++t) s += p.Rt_val[t] * s_ubf[p.Rt_idx[t] + Nf * e];
                outk[i + Np * e] = s;
            }
    }
    const double tot = block_sum_vjp(lam, red);
    if (threadIdx.x == 0) lam_partial[k] = tot;
}

// out = scale * sum_k partial[k] (+ out when accumulating), in a fixed order
__global__ void __launch_bounds__(VT) k_vjp_sum(const double* __restrict__ partial, int n, double* __restrict__ out,
                                                double scale, int accumulate) {
    __shared__ double red[VT];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += VT) s += partial[k];
    const double tot = block_sum_vjp(s, red);
    if (threadIdx.x == 0) out[0] = (accumulate ? out[0] : 0.0) + scale * tot;
}

size_t lds_vjp_v1(const DevParams& p) {
    return sizeof(double) * (size_t)(p.N_p + 2 * p.N_q) * p.N_c;
}

size_t lds_vjp_v2(const DevParams& p) {
    return sizeof(double) * ((size_t)VT + (size_t)(6 * p.N_q + 5 * p.N_f) * p.N_c + (size_t)p.N_q * p.d * p.d +
                             (size_t)p.N_p * p.N_c);
}

template <int D>
hipError_t launch_vjp_t(const DevParams& p, const double* u, const double* ybar, double* ubar, double* rbar,
                        double* fbar, double* lam_partial, double* lam_out, double lam_scale, int accumulate,
                        hipStream_t s) {
    hipLaunchKernelGGL((k_vjp_v1<D>), dim3(p.N_e), dim3(VT), lds_vjp_v1(p), s, p, ybar, rbar, fbar);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_vjp_v2<D>), dim3(p.N_e), dim3(VT), lds_vjp_v2(p), s, p, u, rbar, fbar, ubar, lam_partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_vjp_sum, dim3(1), dim3(VT), 0, s, lam_partial, p.N_e, lam_out, lam_scale, accumulate);
    return hipGetLastError();
}

// Phases V1/V2 after the forward nodal_values (u tilde in p.uq / p.uf); the λ
// cotangent (times lam_scale) is written to, or added into, lam_out.
hipError_t launch_vjp(const DevParams& p, const double* u, const double* ybar, double* ubar, double* rbar,
                      double* fbar, double* lam_partial, double* lam_out, double lam_scale, int accumulate,
                      hipStream_t s) {
    switch (p.d) {
        case 1: return launch_vjp_t<1>(p, u, ybar, ubar, rbar, fbar, lam_partial, lam_out, lam_scale, accumulate, s);
        case 2: return launch_vjp_t<2>(p, u, ybar, ubar, rbar, fbar, lam_partial, lam_out, lam_scale, accumulate, s);
        case 3: return launch_vjp_t<3>(p, u, ybar, ubar, rbar, fbar, lam_partial, lam_out, lam_scale, accumulate, s);
        default: return hipErrorInvalidValue;
    }
}

// y = a x + b y over n entries
__global__ void k_axpby(size_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

hipError_t launch_axpby(size_t n, double a, const double* x, double b, double* y, hipStream_t s) {
    const int T = 256;
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + T - 1) / T)), dim3(T), 0, s, n, a, x, b, y);
    return hipGetLastError();
}

// d/du of the entropy integral sum_k sum_i WJ_k,i S((V u_k)_i): V'(WJ w(V u)) per element
template <int D>
__global__ void __launch_bounds__(VT) k_entropy_grad(DevParams p, const double* __restrict__ u, double* __restrict__ g) {
    constexpr int NC = D + 2;
    extern __shared__ double sm[];
    const int k = blockIdx.x;
    const int Np = p.N_p, Nq = p.N_q;
    double* s_u = sm;             // Np x NC
    double* s_w = s_u + Np * NC;  // Nq x NC
    const double* uk = u + (size_t)k * Np * NC;
    for (int t = threadIdx.x; t < Np * NC; t += blockDim.x) s_u[t] = uk[t];
    __syncthreads();
    const double* Jk = p.Jq + (size_t)k * Nq;
    for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
        double ub[NC], wb[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            if (p.has_V) {
                for (int j = 0; j < Np; ++j) s += p.V[i + (size_t)Nq * j] * s_u[j + Np * e];
            } else {
                s = s_u[i + Np * e];
            }
            ub[e] = s;
        }
        Physics<D, LAW_EULER>::c2e(p.law, ub, wb);
        const double wj = p.W[i] * Jk[i];
#pragma unroll
        for (int e = 0; e < NC; ++e) s_w[i + Nq * e] = wb[e] * wj;
    }
    __syncthreads();
    double* gk = g + (size_t)k * Np * NC;
    for (int j = threadIdx.x; j < Np; j += blockDim.x)
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            if (p.has_V) {
                for (int i = 0; i < Nq; ++i) s += p.V[i + (size_t)Nq * j] * s_w[i + Nq * e];
            } else {
                s = s_w[j + Nq * e];
            }
            gk[j + Np * e] = s;
        }
}

hipError_t launch_entropy_grad(const DevParams& p, const double* u, double* g, hipStream_t s) {
    const size_t lds = sizeof(double) * (size_t)(p.N_p + p.N_q) * p.N_c;
    switch (p.d) {
        case 1: hipLaunchKernelGGL((k_entropy_grad<1>), dim3(p.N_e), dim3(VT), lds, s, p, u, g); break;
        case 2: hipLaunchKernelGGL((k_entropy_grad<2>), dim3(p.N_e), dim3(VT), lds, s, p, u, g); break;
        case 3: hipLaunchKernelGGL((k_entropy_grad<3>), dim3(p.N_e), dim3(VT), lds, s, p, u, g); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sse
