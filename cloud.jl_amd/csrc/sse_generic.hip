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
//   mass solve (WA)  src/Solvers/mass_matrix.jl:185-196
//
// The time derivative evaluates every two-point flux once: the interface fluxes,
// the volume pairs i < j of the Hadamard sum and the facet-correction pairs form
// one flat list of work items shared evenly by the workgroup's threads (dense
// multidimensional operators have N_q (N_q - 1) / 2 volume and N_q N_f facet
// pairs per element), and each item adds its contribution to both ends through
// LDS atomic adds.
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
    double* s_b = s_w + N_p * NC;  // N_f x NC
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
#pragma unroll 8
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
        // modal entropy projection; the products with V and V' run over (node,
        // component) pairs so that a workgroup has N_q NC (or N_p NC) lanes of work
        const double* Jk = p.Jq + (size_t)k * N_q;
        for (int t = threadIdx.x; t < N_q * NC; t += blockDim.x) {  // u_q = V u
            const int i = t % N_q, e = t / N_q;
            double s = 0.0;
#pragma unroll 8
            for (int j = 0; j < N_p; ++j) s += p.V[i + (size_t)N_q * j] * s_u[j + N_p * e];
            s_a[t] = s;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < N_q; i += blockDim.x) {  // WJ c2e(u_q)
            double ub[NC], wb[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) ub[e] = s_a[i + N_q * e];
            ok &= Ph::c2e(p.law, ub, wb);
            const double wj = p.W[i] * Jk[i];
#pragma unroll
            for (int e = 0; e < NC; ++e) s_a[i + N_q * e] = wb[e] * wj;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < N_p * NC; t += blockDim.x) {  // w = V' (WJ w_q)
            const int j = t % N_p, e = t / N_p;
            double s = 0.0;
#pragma unroll 8
            for (int i = 0; i < N_q; ++i) s += p.Vt[j + (size_t)N_p * i] * s_a[i + N_q * e];
            s_w[t] = s;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < N_q * NC; t += blockDim.x) {  // J^{-1}W (V w)
            const int i = t % N_q, e = t / N_q;
            double s = 0.0;
#pragma unroll 8
            for (int j = 0; j < N_p; ++j) s += p.V[i + (size_t)N_q * j] * s_w[j + N_p * e];
            s_a[t] = s * (p.W[i] / Jk[i]);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < N_p * NC; t += blockDim.x) {  // w = V' (...)
            const int j = t % N_p, e = t / N_p;
            double s = 0.0;
#pragma unroll 8
            for (int i = 0; i < N_q; ++i) s += p.Vt[j + (size_t)N_p * i] * s_a[i + N_q * e];
            s_w[t] = s;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < N_q * NC; t += blockDim.x) {  // w_q = V w
            const int i = t % N_q, e = t / N_q;
            double s = 0.0;
#pragma unroll 8
            for (int j = 0; j < N_p; ++j) s += p.V[i + (size_t)N_q * j] * s_w[j + N_p * e];
            s_a[t] = s;
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
    for (int t = threadIdx.x; t < N_f * NC; t += blockDim.x) {  // w_f = R w_q
        const int j = t % N_f, e = t / N_f;
        double s = 0.0;
#pragma unroll 8
        for (int c = p.R_ptr[j]; c < p.R_ptr[j + 1]; ++c) s += p.R_val[c] * s_a[p.R_idx[c] + N_q * e];
        s_b[t] = s;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < N_f; j += blockDim.x) {  // u_f = e2c(w_f)
        double ub[NC], wb[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) wb[e] = s_b[j + N_f * e];
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

// DET: the deterministic variant. Each node sums its own terms in the order of its
// owner list (a volume node: its pairs, from either end, and its row of C; a facet
// node: its interface flux and its column of C), so every pair is evaluated twice and
// no sum depends on the order of atomics.
template <int D, int LAW, bool DET>
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
    if constexpr (!DET) {
        for (int t = threadIdx.x; t < N_q * NC; t += blockDim.x) s_rq[t] = 0.0;
        for (int t = threadIdx.x; t < N_f * NC; t += blockDim.x) s_ff[t] = 0.0;
    }
    __syncthreads();
    const double* nJfk = p.nJf + (size_t)k * D * N_f;
    const double* Jfk = p.Jf + (size_t)k * N_f;

    // The terms of the flux phase:
    //   the interface flux of facet node j, B_j J_f (F(u_f, u_f^+) n + LF), enters f_f[j];
    //   a volume pair i < j with S_m[i,j] != 0 for some m (the reference's scatter over
    //   the upper triangle, flux_differencing_form.jl:37-75):
    //   sv = sum_m S_m[i,j] sum_n (Lambda_i + Lambda_j)[m,n] F_n(u_i,u_j) is subtracted
    //   from r_q[i] and added to r_q[j] (S antisymmetric, F symmetric);
    //   a facet pair C_ij != 0: c_ij nJ(i,j).F(u_i, u_f_j) is subtracted from both r_q[i]
    //   and f_f[j] (flux_differencing_form.jl:91-168).
    auto interface_term = [&](int j, double* out) {
        double uin[NC], uout[NC], nrm[D], F[D * NC];
        const double jf = Jfk[j];
        const int g = p.mapP[j + (size_t)N_f * k];
        const int kn = g / N_f, jn = g - kn * N_f;
        const double* ufn = p.uf + (size_t)kn * N_f * NC;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            uin[e] = s_uf[j + N_f * e];
            uout[e] = ufn[jn + N_f * e];
        }
#pragma unroll
        for (int m = 0; m < D; ++m) nrm[m] = nJfk[m + D * j] / jf;
        Ph::two_point(p.law, uin, uout, F);
        double a = 0.0;
        const bool lf = (p.law.inviscid_flux == 1);
        if (lf) a = p.law.half_lambda * Ph::wave_speed(p.law, uin, uout, nrm);
        const double bj = p.B[j] * jf;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            double fnv = 0.0;
#pragma unroll
            for (int m = 0; m < D; ++m) fnv += F[m * NC + e] * nrm[m];
            if (lf) fnv += a * (uin[e] - uout[e]);
            out[e] = fnv * bj;
        }
    };
    // item c (volume pairs first, then facet pairs): its nodes and its term (sv of a
    // volume pair, c_ij nJ.F of a facet pair)
    auto item_term = [&](int c, int& i, int& j, double* out) {
        const int ij = p.item_ij[c];
        i = ij & 0xffff;
        j = ij >> 16;
        double ui[NC], uj[NC], F[D * NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) ui[e] = s_uq[i + N_q * e];
        if (c < p.n_vol_items) {
#pragma unroll
            for (int e = 0; e < NC; ++e) uj[e] = s_uq[j + N_q * e];
            Ph::two_point(p.law, ui, uj, F);
            double an[D];
#pragma unroll
            for (int n = 0; n < D; ++n) {
                double sv = 0.0;
#pragma unroll
                for (int m = 0; m < D; ++m)
                    sv += p.item_val[(size_t)c * D + m] * (s_L[i + N_q * (m + D * n)] + s_L[j + N_q * (m + D * n)]);
                an[n] = sv;
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                double sv = 0.0;
#pragma unroll
                for (int n = 0; n < D; ++n) sv += an[n] * F[n * NC + e];
                out[e] = sv;
            }
        } else {
            const double cij = p.item_val[(size_t)c * D];
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
                out[e] = fnv * cij;
            }
        }
    };
    if constexpr (!DET) {
        // one flat list of work items (interface fluxes, then the items), spread evenly
        // over the workgroup; each term is evaluated once and added to both ends with
        // LDS atomic adds, so the order of the sums (and the last bits of the result)
        // may vary between calls
        const int n_items = N_f + p.n_items;
        for (int t = threadIdx.x; t < n_items; t += blockDim.x) {
            double v[NC];
            if (t < N_f) {
                interface_term(t, v);
#pragma unroll
                for (int e = 0; e < NC; ++e) atomicAdd(&s_ff[t + N_f * e], v[e]);
                continue;
            }
            const int c = t - N_f;
            int i, j;
            item_term(c, i, j, v);
            if (c < p.n_vol_items) {
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    atomicAdd(&s_rq[i + N_q * e], -v[e]);
                    atomicAdd(&s_rq[j + N_q * e], v[e]);
                }
            } else {
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    atomicAdd(&s_rq[i + N_q * e], -v[e]);
                    atomicAdd(&s_ff[j + N_f * e], -v[e]);
                }
            }
        }
    } else {
        // node nd < N_q is volume node nd, else facet node nd - N_q; own_item[w] = c << 1 | end,
        // end = 1 when the node is the second node j of a volume pair
        for (int nd = threadIdx.x; nd < N_q + N_f; nd += blockDim.x) {
            double acc[NC], v[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) acc[e] = 0.0;
            if (nd >= N_q) interface_term(nd - N_q, acc);
            for (int w = p.own_ptr[nd]; w < p.own_ptr[nd + 1]; ++w) {
                const int c = p.own_item[w] >> 1;
                int i, j;
                item_term(c, i, j, v);
                const bool plus = c < p.n_vol_items && (p.own_item[w] & 1);
#pragma unroll
                for (int e = 0; e < NC; ++e) acc[e] += plus ? v[e] : -v[e];
            }
            if (nd < N_q) {
#pragma unroll
                for (int e = 0; e < NC; ++e) s_rq[nd + N_q * e] = acc[e];
            } else {
#pragma unroll
                for (int e = 0; e < NC; ++e) s_ff[nd - N_q + N_f * e] = acc[e];
            }
        }
    }
    __syncthreads();

    // r_q -= R' f_f
    for (int t = threadIdx.x; t < N_q * NC; t += blockDim.x) {
        const int i = t % N_q, e = t / N_q;
        double s = 0.0;
#pragma unroll 8
        for (int c = p.Rt_ptr[i]; c < p.Rt_ptr[i + 1]; ++c) s += p.Rt_val[c] * s_ff[p.Rt_idx[c] + N_f * e];
        s_rq[t] -= s;
    }
    __syncthreads();

    const double* Jk = p.Jq + (size_t)k * N_q;
    if (p.has_V) {  // over (mode or node, component) pairs, as in nodal_values
        for (int t = threadIdx.x; t < N_p * NC; t += blockDim.x) {  // V' r_q
            const int j = t % N_p, e = t / N_p;
            double s = 0.0;
#pragma unroll 8
            for (int i = 0; i < N_q; ++i) s += p.Vt[j + (size_t)N_p * i] * s_rq[i + N_q * e];
            s_o[t] = s;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < N_q * NC; t += blockDim.x) {  // J^{-1}W (V rhs)
            const int i = t % N_q, e = t / N_q;
            double s = 0.0;
#pragma unroll 8
            for (int j = 0; j < N_p; ++j) s += p.V[i + (size_t)N_q * j] * s_o[j + N_p * e];
            s_t[t] = s * (p.W[i] / Jk[i]);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < N_p * NC; t += blockDim.x) {
            const int j = t % N_p, e = t / N_p;
            double s = 0.0;
#pragma unroll 8
            for (int i = 0; i < N_q; ++i) s += p.Vt[j + (size_t)N_p * i] * s_t[i + N_q * e];
            s_o[t] = s;
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
    return sizeof(double) * (size_t)(2 * p.N_p + p.N_q + p.N_f) * p.N_c;
}

size_t lds_time_derivative_generic(const DevParams& p) {
    return sizeof(double) * ((size_t)(3 * p.N_q + 2 * p.N_f + p.N_p) * p.N_c + (size_t)p.N_q * p.d * p.d);
}

hipError_t launch_nodal_values_mfma(const DevParams& p, const double* u, hipStream_t s);

template <int D, int LAW>
hipError_t launch_generic(const DevParams& p, const double* u, double* dudt, const RKStage& rk, hipStream_t s,
                          int which /* bit 0: phase 1, bit 1: phase 2 */) {
    dim3 grid(p.N_e), block(256);
    if ((which & 1) && p.mfma) {
        hipError_t err = launch_nodal_values_mfma(p, u, s);
        if (err != hipSuccess) return err;
    } else if (which & 1) {
        hipLaunchKernelGGL((k_nodal_values_generic<D, LAW>), grid, block, lds_nodal_values_generic(p), s, p, u);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    if (which & 2) {
        if (p.det)
            hipLaunchKernelGGL((k_time_derivative_generic<D, LAW, true>), grid, block, lds_time_derivative_generic(p),
                               s, p, dudt, rk);
        else
            hipLaunchKernelGGL((k_time_derivative_generic<D, LAW, false>), grid, block, lds_time_derivative_generic(p),
                               s, p, dudt, rk);
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
