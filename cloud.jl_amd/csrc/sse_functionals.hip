// sse_functionals.hip -- conservation and invariant functionals on the device
// (the reference's src/Analysis/conservation.jl:113-189):
//   mass       sum_k 1' WJ_k V u_k                      (per variable)
//   energy     1/2 sum_k u_k' M_k u_k                   (per variable)
//   entropy    sum_k sum_i WJ_k,i S((V u_k)_i)
// and their rates along dudt:
//   mass       sum_k 1' WJ_k V dudt_k
//   energy     sum_k u_k' M_k dudt_k
//   entropy    sum_k (P_k w_q)' M_k dudt_k = sum_k w_q' WJ_k V dudt_k,  P_k = M_k^{-1} V' WJ_k
// and the squared L2 error against exact values at the volume quadrature nodes
// (the default ErrorAnalysis, src/Analysis/error.jl:13-57, 58-91):
//   error      sum_k sum_i W_i J_k,i (u_exact_k,i - (V u_k)_i)^2   (per variable)
//   volume     sum_k sum_i W_i J_k,i                                (for normalize)
// M_k is the mass matrix of the handle's solver: (V' diag(W / J_k) V)^{-1} for the
// weight-adjusted solver with M^{-1} = I (src/Solvers/mass_matrix.jl:150-153), W J_k
// for the diagonal one. The energy kinds form M_k in LDS; the caller may keep every
// element's M_k (`mass`) so that later calls read it instead. One workgroup per element
// writes its partial sums; a second kernel adds them in a fixed order (deterministic, no
// atomics).
#include <hip/hip_runtime.h>

#include "sse_kernels.h"

namespace sse {

namespace {

constexpr int FT = 128;  // threads per element
constexpr int FUNC_L2_ERROR = 6;

__device__ double block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = FT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// S(u) (euler_navierstokes.jl:93-98; linear_advection_diffusion.jl:104) and the
// entropy variables
template <int D, int LAW>
__device__ double entropy_of(const LawParams& P, const double* u) {
    if constexpr (LAW == LAW_EULER) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) s += u[m + 1] * u[m + 1];
        const double p = P.gm1 * (u[D + 1] - (0.5 / u[0]) * s);
        return -u[0] * log(p / pow(u[0], P.gamma)) * P.inv_gm1;
    } else {
        return 0.5 * u[0] * u[0];
    }
}

}  // namespace

// which: 0 mass, 1 energy, 2 entropy, 3 mass rate, 4 energy rate, 5 entropy rate,
// 6 squared L2 error and volume (dudt then holds u_exact, N_q x N_c per element).
// partial has N_e x ncomp entries, ncomp = N_c (N_c + 1 for the error).
template <int D, int LAW>
__global__ void __launch_bounds__(FT) k_functionals(DevParams p, const double* __restrict__ u,
                                                    const double* __restrict__ dudt, int which,
                                                    double* __restrict__ partial, double* __restrict__ mass,
                                                    int mass_ready) {
    using Ph = Physics<D, LAW>;
    constexpr int NC = Ph::NC;
    extern __shared__ double sm[];
    const int k = blockIdx.x;
    const int Np = p.N_p, Nq = p.N_q;
    double* red = sm;                // FT
    double* s_u = red + FT;          // Np x NC
    double* s_v = s_u + Np * NC;     // Np x NC (dudt)
    double* s_uq = s_v + Np * NC;    // Nq x NC
    double* s_vq = s_uq + Nq * NC;   // Nq x NC
    double* s_M = s_vq + Nq * NC;    // Np x Np (M_k^{-1}, then M_k), and 2 Np
    const double* Jk = p.Jq + (size_t)k * Nq;
    const size_t base = (size_t)k * Np * NC;
    const bool error = which == FUNC_L2_ERROR;
    const bool rate = which >= 3 && !error;
    const int kind = which % 3;
    for (int t = threadIdx.x; t < Np * NC; t += FT) {
        s_u[t] = u[base + t];
        s_v[t] = rate ? dudt[base + t] : 0.0;
    }
    __syncthreads();
    if (error) {
        for (int i = threadIdx.x; i < Nq; i += FT)
            for (int e = 0; e < NC; ++e) {
                double a = 0.0;
                if (p.has_V) {
                    for (int j = 0; j < Np; ++j) a += p.V[i + (size_t)Nq * j] * s_u[j + Np * e];
                } else {
                    a = s_u[i + Np * e];
                }
                s_uq[i + Nq * e] = a;
            }
        __syncthreads();
        const double* ex = dudt + (size_t)k * Nq * NC;
        double* out = partial + (size_t)k * (NC + 1);
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            for (int i = threadIdx.x; i < Nq; i += FT) {
                const double r = ex[i + Nq * e] - s_uq[i + Nq * e];
                s += p.W[i] * Jk[i] * r * r;
            }
            const double tot = block_sum(s, red);
            if (threadIdx.x == 0) out[e] = tot;
        }
        double s = 0.0;
        for (int i = threadIdx.x; i < Nq; i += FT) s += p.W[i] * Jk[i];
        const double tot = block_sum(s, red);
        if (threadIdx.x == 0) out[NC] = tot;
        return;
    }
    double* out = partial + (size_t)k * NC;
    if (kind != 1) {  // mass and entropy work at the volume nodes
        for (int i = threadIdx.x; i < Nq; i += FT)
            for (int e = 0; e < NC; ++e) {
                double a = 0.0, b = 0.0;
                if (p.has_V) {
                    for (int j = 0; j < Np; ++j) {
                        const double vij = p.V[i + (size_t)Nq * j];
                        a += vij * s_u[j + Np * e];
                        b += vij * s_v[j + Np * e];
                    }
                } else {
                    a = s_u[i + Np * e];
                    b = s_v[i + Np * e];
                }
                s_uq[i + Nq * e] = a;
                s_vq[i + Nq * e] = b;
            }
        __syncthreads();
        if (kind == 0) {
            for (int e = 0; e < NC; ++e) {
                double s = 0.0;
                for (int i = threadIdx.x; i < Nq; i += FT) s += p.W[i] * Jk[i] * (rate ? s_vq : s_uq)[i + Nq * e];
                const double tot = block_sum(s, red);
                if (threadIdx.x == 0) out[e] = tot;
            }
        } else {
            double s = 0.0;
            for (int i = threadIdx.x; i < Nq; i += FT) {
                double ui[NC];
#pragma unroll
                for (int e = 0; e < NC; ++e) ui[e] = s_uq[i + Nq * e];
                const double wj = p.W[i] * Jk[i];
                if (!rate) {
                    s += wj * entropy_of<D, LAW>(p.law, ui);
                } else {
                    double wi[NC];
                    if constexpr (LAW == LAW_EULER) {
                        Ph::c2e(p.law, ui, wi);
                    } else {
                        wi[0] = ui[0];
                    }
#pragma unroll
                    for (int e = 0; e < NC; ++e) s += wj * wi[e] * s_vq[i + Nq * e];
                }
            }
            const double tot = block_sum(s, red);
            if (threadIdx.x == 0) {
                out[0] = tot;
                for (int e = 1; e < NC; ++e) out[e] = 0.0;
            }
        }
        return;
    }
    if (!p.has_V) {  // nodal: M_k = diag(W J_k), no matrix to factor
        for (int e = 0; e < NC; ++e) {
            double s = 0.0;
            for (int i = threadIdx.x; i < Np; i += FT) s += p.W[i] * Jk[i] * s_u[i + Np * e] * (rate ? s_v : s_u)[i + Np * e];
            const double tot = block_sum(s, red);
            if (threadIdx.x == 0) out[e] = rate ? tot : 0.5 * tot;
        }
        return;
    }
    // energy: u' M_k x, x = u (energy) or dudt (rate). M_k is formed as the inverse of
    // M_k^{-1} (in LDS, by Gauss-Jordan elimination: M_k^{-1} is symmetric positive
    // definite, so no pivoting) or read from `mass`, where an earlier call kept it
    double* mk = mass ? mass + (size_t)k * Np * Np : nullptr;
    const double* Mk = mk;
    if (!(mk && mass_ready)) {
        if (p.has_V) {
            for (int t = threadIdx.x; t < Np * Np; t += FT) {  // M_k^{-1} = V' diag(W/J) V
                const int a = t % Np, b = t / Np;
                double s = 0.0;
                for (int i = 0; i < Nq; ++i) s += p.V[i + (size_t)Nq * a] * (p.W[i] / Jk[i]) * p.V[i + (size_t)Nq * b];
                s_M[t] = s;
            }
        } else {
            for (int t = threadIdx.x; t < Np * Np; t += FT) {
                const int a = t % Np, b = t / Np;
                s_M[t] = a == b ? 1.0 / (p.W[a] * Jk[a]) : 0.0;
            }
        }
        double* rowk = s_M + Np * Np;  // the pivot row and column of the step
        double* colk = rowk + Np;
        for (int kk = 0; kk < Np; ++kk) {
            __syncthreads();
            for (int t = threadIdx.x; t < Np; t += FT) {
                rowk[t] = s_M[kk + Np * t];
                colk[t] = s_M[t + Np * kk];
            }
            __syncthreads();
            const double ip = 1.0 / rowk[kk];
            for (int i = threadIdx.x; i < Np; i += FT) {
                const double ci = colk[i];
                for (int j = 0; j < Np; ++j) {
                    const double rkj = (j == kk ? 1.0 : rowk[j]) * ip;  // row kk after the step
                    s_M[i + Np * j] = i == kk ? rkj : (j == kk ? 0.0 : s_M[i + Np * j]) - ci * rkj;
                }
            }
        }
        __syncthreads();
        if (mk)
            for (int t = threadIdx.x; t < Np * Np; t += FT) mk[t] = s_M[t];
        Mk = s_M;
    }
    // u' M_k x: thread a takes row a of M_k (columns of consecutive words)
    const double* x = rate ? s_v : s_u;
    for (int e = 0; e < NC; ++e) {
        double s = 0.0;
        for (int a = threadIdx.x; a < Np; a += FT) {
            double y = 0.0;
            for (int b = 0; b < Np; ++b) y += Mk[a + Np * b] * x[b + Np * e];
            s += s_u[a + Np * e] * y;
        }
        const double tot = block_sum(s, red);
        if (threadIdx.x == 0) out[e] = rate ? tot : 0.5 * tot;
    }
}

// out[c] = sum_k partial[k, c] in a fixed order
__global__ void __launch_bounds__(FT) k_sum_partials(const double* __restrict__ partial, int N_e, int ncomp,
                                                     double* __restrict__ out) {
    __shared__ double red[FT];
    for (int c = 0; c < ncomp; ++c) {
        double s = 0.0;
        for (int k = threadIdx.x; k < N_e; k += FT) s += partial[(size_t)k * ncomp + c];
        const double tot = block_sum(s, red);
        if (threadIdx.x == 0) out[c] = tot;
    }
}

// LDS of one functional kind: the mass matrix and a pivot row and column only for
// the energy kinds of a modal (weight-adjusted) solver
size_t lds_functionals(const DevParams& p, int which) {
    const bool matrix = (which == 1 || which == 4) && p.has_V;
    return sizeof(double) * ((size_t)FT + 2 * (size_t)p.N_p * p.N_c + 2 * (size_t)p.N_q * p.N_c +
                             (matrix ? (size_t)p.N_p * p.N_p + 2 * (size_t)p.N_p : 0));
}

template <int D, int LAW>
hipError_t launch_functionals_t(const DevParams& p, const double* u, const double* dudt, int which, double* partial,
                                double* out, hipStream_t s, double* mass, int mass_ready) {
    // with M_k kept, the matrix is read from global memory: no LDS for it
    const bool kept = mass && mass_ready && (which == 1 || which == 4) && p.has_V;
    const size_t lds = lds_functionals(p, which) - (kept ? sizeof(double) * ((size_t)p.N_p * p.N_p + 2 * (size_t)p.N_p) : 0);
    hipLaunchKernelGGL((k_functionals<D, LAW>), dim3(p.N_e), dim3(FT), lds, s, p, u, dudt, which,
                       partial, mass, mass_ready);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FT), 0, s, partial, p.N_e,
                       which == FUNC_L2_ERROR ? p.N_c + 1 : p.N_c, out);
    return hipGetLastError();
}

hipError_t launch_functionals(const DevParams& p, const double* u, const double* dudt, int which, double* partial,
                              double* out, hipStream_t s, double* mass, int mass_ready) {
    const bool euler = (p.N_c != 1);
    switch (p.d) {
        case 1: return euler ? launch_functionals_t<1, LAW_EULER>(p, u, dudt, which, partial, out, s, mass, mass_ready)
                             : launch_functionals_t<1, LAW_ADVECTION>(p, u, dudt, which, partial, out, s, mass, mass_ready);
        case 2: return euler ? launch_functionals_t<2, LAW_EULER>(p, u, dudt, which, partial, out, s, mass, mass_ready)
                             : launch_functionals_t<2, LAW_ADVECTION>(p, u, dudt, which, partial, out, s, mass, mass_ready);
        case 3: return euler ? launch_functionals_t<3, LAW_EULER>(p, u, dudt, which, partial, out, s, mass, mass_ready)
                             : launch_functionals_t<3, LAW_ADVECTION>(p, u, dudt, which, partial, out, s, mass, mass_ready);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace sse
