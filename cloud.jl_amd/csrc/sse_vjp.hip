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
// LDS copy of V with an odd leading dimension, so that threads over i (V x) and
// threads over j (V' x) both read without bank conflicts
__host__ __device__ inline int vjp_ldv(const DevParams& p) { return p.N_q | 1; }

// y (N_q x NC) = diag(scale) V x for x (N_p x NC); threads over (i, e), two
// accumulators per dot product. scale: 0 none, 1 W / J_k, 2 W J_k.
__device__ void v_apply(const DevParams& p, const double* Vm, int ldv, const double* Jk, int mode, int NC,
                        const double* x, double* y) {
    const int Nq = p.N_q, Np = p.N_p;
    for (int t = threadIdx.x; t < Nq * NC; t += blockDim.x) {
        const int i = t % Nq, e = t / Nq;
        const double* xe = x + Np * e;
        double s0 = 0.0, s1 = 0.0;
        int j = 0;
        for (; j + 1 < Np; j += 2) {
            s0 += Vm[i + ldv * j] * xe[j];
            s1 += Vm[i + ldv * (j + 1)] * xe[j + 1];
        }
        if (j < Np) s0 += Vm[i + ldv * j] * xe[j];
        double sc = 1.0;
        if (mode == 1) sc = p.W[i] / Jk[i];
        else if (mode == 2) sc = p.W[i] * Jk[i];
        y[i + Nq * e] = (s0 + s1) * sc;
    }
}

// y (N_p x NC) = V' x for x (N_q x NC); threads over (j, e)
__device__ void vt_apply(const DevParams& p, const double* Vm, int ldv, int NC, const double* x, double* y) {
    const int Nq = p.N_q, Np = p.N_p;
    for (int t = threadIdx.x; t < Np * NC; t += blockDim.x) {
        const int j = t % Np, e = t / Np;
        const double* xe = x + Nq * e;
        const double* vj = Vm + ldv * j;
        double s0 = 0.0, s1 = 0.0;
        int i = 0;
        for (; i + 1 < Nq; i += 2) {
            s0 += vj[i] * xe[i];
            s1 += vj[i + 1] * xe[i + 1];
        }
        if (i < Nq) s0 += vj[i] * xe[i];
        y[j + Np * e] = s0 + s1;
    }
}

// s_o (N_p x NC) = M^{-1} s_o = V' (W / J_k (V s_o)) (weight-adjusted, M^{-1} = I);
// s_t (N_q x NC) scratch. Vm: V in LDS, Vm[i + ldv j] = V[i, j].
__device__ void wadg_inplace(const DevParams& p, const double* Vm, int ldv, int k, double* s_o, double* s_t,
                             int NC) {
    const double* Jk = p.Jq + (size_t)k * p.N_q;
    v_apply(p, Vm, ldv, Jk, 1, NC, s_o, s_t);
    __syncthreads();
    vt_apply(p, Vm, ldv, NC, s_t, s_o);
    __syncthreads();
}

// parts per node of the two-point term lists in phase V2
__host__ __device__ inline int vjp_split(const DevParams& p, int threads) {
    const int n = p.N_q > p.N_f ? p.N_q : p.N_f;
    const int s = threads / n;
    return s < 1 ? 1 : (s > 8 ? 8 : s);
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
    double* sV = s_r + Nq * NC;     // V (when modal), leading dimension ldv
    const int ldv = vjp_ldv(p);
    const double* yk = ybar + (size_t)k * Np * NC;
    for (int t = threadIdx.x; t < Np * NC; t += blockDim.x) s_o[t] = yk[t];
    if (p.has_V)
        for (int t = threadIdx.x; t < Nq * Np; t += blockDim.x) sV[(t % Nq) + ldv * (t / Nq)] = p.V[t];
    __syncthreads();
    const double* Jk = p.Jq + (size_t)k * Nq;
    if (p.has_V) {
        wadg_inplace(p, sV, ldv, k, s_o, s_t, NC);
        v_apply(p, sV, ldv, Jk, 0, NC, s_o, s_r);
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
#pragma unroll 8
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
    double* s_uq = sm;                    // Nq x NC   projected volume states (u tilde)
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
    double* sV = s_o + Np * NC;           // V (modal schemes), leading dimension ldv
    const int ldv = vjp_ldv(p);
    const int nV = p.has_V ? ldv * Np : 0;
    for (int t = threadIdx.x; t < (p.has_V ? Nq * Np : 0); t += blockDim.x) sV[(t % Nq) + ldv * (t / Nq)] = p.V[t];
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
        v_apply(p, sV, ldv, Jk, 0, NC, s_o, s_u0);  // V u
        __syncthreads();
        for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
            double ub[NC], wb[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) ub[e] = s_u0[i + Nq * e];
            Ph::c2e(p.law, ub, wb);
            const double wj = p.W[i] * Jk[i];
#pragma unroll
            for (int e = 0; e < NC; ++e) s_t[i + Nq * e] = wb[e] * wj;
        }
        __syncthreads();
        vt_apply(p, sV, ldv, NC, s_t, s_o);
        __syncthreads();
        wadg_inplace(p, sV, ldv, k, s_o, s_t, NC);
        v_apply(p, sV, ldv, Jk, 0, NC, s_o, s_wq);
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
#pragma unroll 8
            for (int t = p.R_ptr[j]; t < p.R_ptr[j + 1]; ++t) s += p.R_val[t] * s_wq[p.R_idx[t] + Nq * e];
            s_wf[j + Nf * e] = s;
        }

    // Cotangents of the projected states. Work items: facet node j or volume node
    // i, each split into `nsplit` parts over its list of two-point terms; parts
    // meet in s_part and are added in a fixed order.
    const double* nJfk = p.nJf + (size_t)k * D * Nf;
    const double* Jfk = p.Jf + (size_t)k * Nf;
    const bool lf = (p.law.inviscid_flux == 1);
    const int nsplit = vjp_split(p, blockDim.x);
    double* s_part = sV + nV;                              // nsplit x (Nf + Nq) x NC
    constexpr int NH = A::NH;
    double* s_h = s_part + (size_t)nsplit * (Nf + Nq) * NC;  // (Nf + Nq) x NH: facet nodes, then volume nodes
    for (int t = threadIdx.x; t < Nf + Nq; t += blockDim.x) {
        double ut[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) ut[e] = t < Nf ? s_uf[t + Nf * e] : s_uq[(t - Nf) + Nq * e];
        A::hoist_full(p.law, ut, s_h + (size_t)t * NH);
    }
    __syncthreads();
    double lam = 0.0;
    for (int w = threadIdx.x; w < (Nf + Nq) * nsplit; w += blockDim.x) {
        double ub[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) ub[e] = 0.0;
        const bool facet = w < Nf * nsplit;
        const int part = facet ? w / Nf : (w - Nf * nsplit) / Nq;
        const int node = facet ? w - part * Nf : w - Nf * nsplit - part * Nq;
        // the gathering node's state and its list of two-point terms
        double un[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) un[e] = facet ? s_uf[node + Nf * e] : s_uq[node + Nq * e];
        int nt;
        if (facet) nt = 2 + (p.C_kind ? p.Cc_ptr[node + 1] - p.Cc_ptr[node] : 0);
        else nt = (p.S_ptr[node + 1] - p.S_ptr[node]) + (p.C_kind ? p.Cr_ptr[node + 1] - p.Cr_ptr[node] : 0);
        const int nS = facet ? 0 : p.S_ptr[node + 1] - p.S_ptr[node];
        const double* hn = s_h + (size_t)(facet ? node : Nf + node) * NH;
        for (int q = part; q < nt; q += nsplit) {
            double xo[NC], a[D], gb[NC], hx[NH];
            const double* ho = hx;  // partner's hoisted state
            bool left = true;
            int lfmode = 0;  // 1: this node's interface flux, 2: the neighbour's at the shared node
            if (facet && q < 2) {
                // interface flux: f*(u_in, u_out, n) at this node (keep u_in), and the
                // neighbour's f*(u_out', u_in', n') at its node (keep its exterior state, ours)
                const int j = node;
                const int g = p.mapP[j + (size_t)Nf * k];
                const int kn = g / Nf, jn = g - kn * Nf;
                const double* ufn = p.uf + (size_t)kn * Nf * NC;
                const int kk = q == 0 ? k : kn, jj = q == 0 ? j : jn;
                const double jf = p.Jf[jj + (size_t)Nf * kk];
                const double* nJ = p.nJf + (size_t)kk * D * Nf;
#pragma unroll
                for (int m = 0; m < D; ++m) a[m] = nJ[m + D * jj] / jf;
                const double bj = p.B[jj] * jf;
                const double* fbs = q == 0 ? s_fb + j : fbar + (size_t)kn * Nf * NC + jn;
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    xo[e] = ufn[jn + Nf * e];
                    gb[e] = bj * fbs[Nf * e];
                }
                left = (q == 0);
                lfmode = q + 1;
                A::hoist_full(p.law, xo, hx);
            } else if (facet) {  // facet correction seen from facet node j (keep u_fj)
                const int j = node, t = p.Cc_ptr[j] + q - 2;
                const int i = p.Cc_idx[t];
                const double cij = p.Cc_val[t];
                const int f = j / npf;
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    double s = 0.0;
#pragma unroll
                    for (int l = 0; l < D; ++l) s += s_L[i + Nq * (l + D * m)] * p.fn[l + D * f];
                    a[m] = 0.5 * nJfk[m + D * j] + 0.5 * s;
                }
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    xo[e] = s_uq[i + Nq * e];
                    gb[e] = -cij * (s_rb[i + Nq * e] + s_fb[j + Nf * e]);
                }
                ho = s_h + (size_t)(Nf + i) * NH;
                left = false;
            } else if (q < nS) {  // Hadamard sum term (i, j): cotangent (dT_ij/du_L)' (rbar_j - rbar_i)
                const int i = node, t = p.S_ptr[i] + q;
                const int j = p.S_idx[t];
#pragma unroll
                for (int n = 0; n < D; ++n) {
                    double s = 0.0;
#pragma unroll
                    for (int m = 0; m < D; ++m)
                        s += p.S_val[t * D + m] * (s_L[i + Nq * (m + D * n)] + s_L[j + Nq * (m + D * n)]);
                    a[n] = s;
                }
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    xo[e] = s_uq[j + Nq * e];
                    gb[e] = s_rb[j + Nq * e] - s_rb[i + Nq * e];
                }
                ho = s_h + (size_t)(Nf + j) * NH;
            } else {  // facet correction seen from volume node i
                const int i = node, t = p.Cr_ptr[i] + q - nS;
                const int j = p.Cr_idx[t];
                const double cij = p.Cr_val[t];
                const int f = j / npf;
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    double s = 0.0;
#pragma unroll
                    for (int l = 0; l < D; ++l) s += s_L[i + Nq * (l + D * m)] * p.fn[l + D * f];
                    a[m] = 0.5 * nJfk[m + D * j] + 0.5 * s;
                }
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    xo[e] = s_uf[j + Nf * e];
                    gb[e] = -cij * (s_rb[i + Nq * e] + s_fb[j + Nf * e]);
                }
                ho = s_h + (size_t)j * NH;
            }
            const double* xL = left ? un : xo;
            const double* xR = left ? xo : un;
            A::ec_dot_adj_h(p.law, left ? hn : ho, left ? ho : hn, a, gb, left, ub);
            if (lfmode && lf) {
                const double dh = A::lf_diss_adj(p.law, xL, xR, a, gb, left, ub);
                if (lfmode == 1) lam += 0.5 * dh;  // halfλ = λ / 2
            }
        }
        const int slot = facet ? node : Nf + node;
#pragma unroll
        for (int e = 0; e < NC; ++e) s_part[(part * (Nf + Nq) + slot) * NC + e] = ub[e];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < Nf + Nq; t += blockDim.x) {
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            for (int part = 0; part < nsplit; ++part) s += s_part[(part * (Nf + Nq) + t) * NC + e];
            if (t < Nf) s_ubf[t + Nf * e] = s;
            else s_ubq[(t - Nf) + Nq * e] = s;
        }
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
#pragma unroll 8
                for (int t = p.Rt_ptr[i]; t < p.Rt_ptr[i + 1]; ++t) s += p.Rt_val[t] * s_ubf[p.Rt_idx[t] + Nf * e];
                if (p.has_V) s_t[i + Nq * e] += s;
                else s_t[i + Nq * e] = s;
            }
        __syncthreads();
        if (p.has_V) {
            // w_q' = V w ; w = M^{-1} wt ; wt = V' (WJ w_q) ; w_q = c2e(V u)
            vt_apply(p, sV, ldv, NC, s_t, s_o);
            __syncthreads();
            wadg_inplace(p, sV, ldv, k, s_o, s_wq, NC);  // s_wq no longer needed
            v_apply(p, sV, ldv, Jk, 2, NC, s_o, s_wq);   // WJ (V wt_bar)
            __syncthreads();
            for (int i = threadIdx.x; i < Nq; i += blockDim.x) {
                double wb[NC], u0[NC], ubv[NC];
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    wb[e] = s_wq[i + Nq * e];
                    u0[e] = s_u0[i + Nq * e];
                    ubv[e] = 0.0;
                }
                A::c2e_adj(p.law, u0, wb, ubv);
#pragma unroll
                for (int e = 0; e < NC; ++e) s_t[i + Nq * e] = ubv[e];
            }
            __syncthreads();
            vt_apply(p, sV, ldv, NC, s_t, ubar + (size_t)k * Np * NC);
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
#pragma unroll 8
                for (int t = p.Rt_ptr[i]; t < p.Rt_ptr[i + 1]; ++t) s += p.Rt_val[t] * s_ubf[p.Rt_idx[t] + Nf * e];
                outk[i + Np * e] = s;
            }
    }
    double* red = s_part;  // free again: VT doubles for the block sum
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
    return sizeof(double) * ((size_t)(p.N_p + 2 * p.N_q) * p.N_c + (p.has_V ? (size_t)vjp_ldv(p) * p.N_p : 0));
}

size_t lds_vjp_v2(const DevParams& p) {
    size_t part = (size_t)vjp_split(p, VT) * (p.N_q + p.N_f) * p.N_c;
    if (part < (size_t)VT) part = VT;  // also the block-sum buffer
    return sizeof(double) * ((size_t)(6 * p.N_q + 5 * p.N_f) * p.N_c + (size_t)p.N_q * p.d * p.d +
                             (size_t)p.N_p * p.N_c + part + (size_t)(p.N_q + p.N_f) * (p.d + 6) +
                             (p.has_V ? (size_t)vjp_ldv(p) * p.N_p : 0));
}

template <int D>
hipError_t launch_vjp_t(const DevParams& p, const double* u, const double* ybar, double* ubar, double* rbar,
                        double* fbar, double* lam_partial, double* lam_out, double lam_scale, int accumulate,
                        hipStream_t s) {
    static bool attr = false;  // allow the full LDS of a CU for phase V2
    if (!attr) {
        hipError_t ea = hipFuncSetAttribute((const void*)k_vjp_v2<D>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            160 * 1024);
        if (ea == hipSuccess)
            ea = hipFuncSetAttribute((const void*)k_vjp_v1<D>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (ea != hipSuccess) return ea;
        attr = true;
    }
    hipLaunchKernelGGL((k_vjp_v1<D>), dim3(p.N_e), dim3(VT), lds_vjp_v1(p), s, p, ybar, rbar, fbar);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_vjp_v2<D>), dim3(p.N_e), dim3(VT), lds_vjp_v2(p), s, p, u, rbar, fbar, ubar, lam_partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_vjp_sum, dim3(1), dim3(VT), 0, s, lam_partial, p.N_e, lam_out, lam_scale, accumulate);
    return hipGetLastError();
}

hipError_t launch_vjp_sum(const double* lam_partial, int n, double* lam_out, double lam_scale, int accumulate,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_vjp_sum, dim3(1), dim3(VT), 0, s, lam_partial, n, lam_out, lam_scale, accumulate);
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
    if (i < n) y[i] = b == 0.0 ? a * x[i] : a * x[i] + b * y[i];  // b = 0 does not read y
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
#pragma unroll 8
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
#pragma unroll 8
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
