// sse_tensor.hip -- residual kernels for ModalTensor triangles and tetrahedra
// with compile-time shapes (element type and degree P), for the Euler equations
// in FluxDifferencingForm with the entropy-conservative two-point flux.
//
// Same algorithm as the reference (src/Solvers/flux_differencing_form.jl), with
// the work organised for CDNA4:
//   * V and V' are applied sum-factorised (the WarpedTensorProductMap2D/3D
//     algorithm of src/MatrixFreeOperators/warped_product_{2d,3d}.jl), one
//     thread per output point, all variables at once, stages staged in LDS;
//   * the Hadamard sum and the facet correction are one list of unique pairs:
//     each volume pair (i < j) of the union sparsity of S_1..S_d and each
//     nonzero C_ij is evaluated exactly once per element; contributions go to
//     an LDS staging area and are summed into the owning node in a fixed,
//     host-built order (deterministic, no atomics);
//   * per-node quantities (velocity, pressure, log rho, rho/p, log(rho/p)) are
//     computed once, so a pair costs one divide and no logarithm;
//   * the flux tensor is never formed: it is contracted with the metric vector
//     inside the flux (ec_dot).
// A "team" of TS threads owns one element; a block holds EPB teams.
#include <hip/hip_runtime.h>

#include "sse_tensor.h"

namespace sse {

// ----------------------------------------------------------------------------- set-up
// The weight-adjusted inverse mass matrix of every element, V' diag(W / J_k) V
// (mass_matrix.jl:150-153 with the orthonormal modal basis: M^{-1} = I), column-major
// and symmetric: the upper triangle is summed, each entry written to both halves.
__global__ void __launch_bounds__(256) k_weight_adjusted_inverse(int N_p, int N_q, const double* __restrict__ V,
                                                                 const double* __restrict__ W,
                                                                 const double* __restrict__ Jq, double* __restrict__ Mw) {
    extern __shared__ double s_w[];  // N_q: W / J_k
    const size_t k = blockIdx.x;
    for (int i = threadIdx.x; i < N_q; i += blockDim.x) s_w[i] = W[i] / Jq[k * N_q + i];
    __syncthreads();
    double* Mk = Mw + k * N_p * N_p;
    for (int t = threadIdx.x; t < N_p * N_p; t += blockDim.x) {
        const int a = t % N_p, b = t / N_p;
        if (a > b) continue;
        double s = 0.0;
        for (int i = 0; i < N_q; ++i) s += V[i + (size_t)N_q * a] * s_w[i] * V[i + (size_t)N_q * b];
        Mk[a + (size_t)N_p * b] = s;
        Mk[b + (size_t)N_p * a] = s;
    }
}

hipError_t launch_weight_adjusted_inverse(const DevParams& p, double* Mw, hipStream_t s) {
    if (!p.has_V || !p.V) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_weight_adjusted_inverse, dim3(p.N_e), dim3(256), sizeof(double) * p.N_q, s, p.N_p, p.N_q,
                       p.V, p.W, p.Jq, Mw);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- nodal_values
// The sum-factorisation temporary Z shares the volume-node array: in V it is dead
// before the last stage writes the output, in V' the input is dead after the first
// stage. The modal input u and the modal w share one array (u is dead after u_q = V u).
template <int ELEM, int P>
struct K1Layout {
    using Sh = Shape<ELEM, P>;
    static constexpr int NC = Sh::NC;
    static_assert(Sh::NZ <= Sh::NQ, "Z inside the volume-node array");
    static constexpr int off_a = 0;                                  // NC x NQ (and Z)
    static constexpr int off_z = off_a;                              // NC x NZ
    static constexpr int off_W = off_a + NC * Sh::NQ;                // NC x NW
    static constexpr int off_u = off_W + NC * (Sh::NW > 0 ? Sh::NW : 1);  // NC x NP: u, then w
    static constexpr int off_w = off_u;
    static constexpr int size = off_u + NC * Sh::NP;
};

template <int ELEM, int P, int VAR>
__global__ void __launch_bounds__((Tune<ELEM, P, VAR>::TS1 * Tune<ELEM, P, VAR>::EPB1))
k_nodal_values_tensor(TensorParams tp, const double* __restrict__ u) {
    using Sh = Shape<ELEM, P>;
    using T = Tune<ELEM, P, VAR>;
    using SF = SumFact<ELEM, P, T::TS1>;
    using L = K1Layout<ELEM, P>;
    constexpr int D = Sh::D, NC = Sh::NC, NP = Sh::NP, NQ = Sh::NQ, NF = Sh::NF, TS = T::TS1;
    extern __shared__ double smem[];
    const DevParams& p = tp.base;
    const int team = threadIdx.x / TS, tl = threadIdx.x - team * TS;
    const int nb = gridDim.x;
    const int k = element_of_block(blockIdx.x, nb, tp.xcd_remap) * T::EPB1 + team;
    const bool valid = k < p.N_e;
    const int kk = valid ? k : 0;
    double* sm = smem + team * L::size;
    double* su = sm + L::off_u;
    double* sa = sm + L::off_a;
    double* sw = sm + L::off_w;
    double* sZ = sm + L::off_z;
    double* sW = sm + L::off_W;
    using Ph = Physics<D, LAW_EULER>;

    // the factors of V, shared by the block's teams, after the teams' arrays
    using VF = VFactors<ELEM, P>;
    double* sfac = smem + T::EPB1 * L::size;
    VF::stage(tp, sfac, threadIdx.x, TS * T::EPB1);
    const TensorParams tq = VF::in_lds(tp, sfac);
    const double* uk = u + (size_t)kk * NP * NC;
    for (int t = tl; t < NP * NC; t += TS) su[t] = uk[t];
    __syncthreads();
    SF::V(tq, tl, su, sa, sZ, sW);  // u_q = V u
    const double* Jk = p.Jq + (size_t)kk * NQ;
    bool ok = true;
    for (int i = tl; i < NQ; i += TS) {  // w_q = WJ c2e(u_q)
        double ub[NC], wb[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) ub[e] = sa[e * NQ + i];
        ok &= c2e_fast<D>(p.law, ub, wb);
        const double wj = p.W[i] * Jk[i];
#pragma unroll
        for (int e = 0; e < NC; ++e) sa[e * NQ + i] = wb[e] * wj;
    }
    __syncthreads();
    // the kept weight-adjusted inverse is staged where u_q and the temporary W were
    constexpr bool MW = NP * NP <= L::off_u - L::off_a;
    MassPrefetch<NP, TS> pre;
    if constexpr (MW) {
        if (tp.Mw) mass_prefetch<NP, TS>(tp.Mw + (size_t)kk * NP * NP, tl, pre);
    }
    SF::Vt(tq, tl, sa, sw, sZ, sW);  // w = V' (WJ w_q)
    if (MW && tp.Mw) {               // WADG: w = V' diag(W/J) V w with the kept matrix
        mass_apply_staged<NP, NC, TS>(pre, sa, sw, tl);
    } else {                         // WADG by sum factorisation: J^{-1} W (V w), then V'
        SF::V(tq, tl, sw, sa, sZ, sW);
        for (int i = tl; i < NQ; i += TS) {
            const double jw = p.W[i] / Jk[i];
#pragma unroll
            for (int e = 0; e < NC; ++e) sa[e * NQ + i] *= jw;
        }
        __syncthreads();
        SF::Vt(tq, tl, sa, sw, sZ, sW);
    }
    SF::V(tq, tl, sw, sa, sZ, sW);   // w_q = V w
    double* uqk = p.uq + (size_t)kk * NQ * NC;
    double* ufk = p.uf + (size_t)kk * NF * NC;
    for (int i = tl; i < NQ; i += TS) {  // u_q = e2c(w_q)
        double ub[NC], wb[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) wb[e] = sa[e * NQ + i];
        ok &= e2c_fast<D>(p.law, wb, ub);
        if (valid) {
#pragma unroll
            for (int e = 0; e < NC; ++e) uqk[e * NQ + i] = ub[e];
        }
    }
    // u_f = e2c(R w_q). A row's entries are taken four at a time, their loads issued
    // together (entries past the table width or the row's end are -1; a group whose first
    // entry is -1 ends the row)
    constexpr int G = 8;  // entries per load group
    auto row_sum = [&](int j, int w_begin, int w_end, double* wb) {
        for (int w0 = w_begin; w0 < w_end; w0 += G) {
            int ii[G];
            double rv[G];
#pragma unroll
            for (int q = 0; q < G; ++q) {
                const bool in = w0 + q < w_end;
                ii[q] = in ? tp.r_idx[(w0 + q) * NF + j] : -1;
                rv[q] = in ? tp.r_val[(w0 + q) * NF + j] : 0.0;
            }
            if (ii[0] < 0) break;
#pragma unroll
            for (int q = 0; q < G; ++q) {
                if (ii[q] < 0) continue;
#pragma unroll
                for (int e = 0; e < NC; ++e) wb[e] += rv[q] * sa[e * NQ + ii[q]];
            }
        }
    };
    auto store_facet = [&](int j, const double* wb) {
        double ub[NC];
        ok &= e2c_fast<D>(p.law, wb, ub);
        if (valid) {
#pragma unroll
            for (int e = 0; e < NC; ++e) ufk[e * NF + j] = ub[e];
            if (tp.uf_ext) {  // also the neighbour's exterior state: uf_ext at mapP(j, k) = u_f(j, k)
                const int g = p.mapP[j + (size_t)NF * k];
                const int kn = g / NF, jn = g - kn * NF;
                double* uxn = tp.uf_ext + (size_t)kn * NF * NC;
#pragma unroll
                for (int e = 0; e < NC; ++e) uxn[e * NF + jn] = ub[e];
            }
        }
    };
    // On a tet the last facet nodes (those of the face whose nodes each see a whole plane
    // of volume nodes) have the long rows: the team's otherwise idle threads take the
    // second half of the last NHELP rows and hand their partial sums over in LDS (the
    // modal array, dead by now)
    constexpr int NHELP = NF < TS ? TS - NF : 0;
    constexpr bool HELP = ELEM == ELEM_TET && NHELP > 0 && NHELP <= NP;
    if constexpr (HELP) {
        const int wh = (tp.wr / 2 + 3) / 4 * 4;  // split point (a multiple of four)
        const bool helper = tl >= NF;
        const int j = helper ? tl - NHELP : tl, h = j - (NF - NHELP);
        const bool split = h >= 0;
        double wb[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) wb[e] = 0.0;
        row_sum(j, helper ? wh : 0, (split && !helper) ? wh : tp.wr, wb);
        if (helper) {
#pragma unroll
            for (int e = 0; e < NC; ++e) sw[e * NHELP + h] = wb[e];
        }
        __syncthreads();
        if (!helper) {
            if (split) {
#pragma unroll
                for (int e = 0; e < NC; ++e) wb[e] += sw[e * NHELP + h];
            }
            store_facet(j, wb);
        }
    } else {
        for (int j = tl; j < NF; j += TS) {
            double wb[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) wb[e] = 0.0;
            row_sum(j, 0, tp.wr, wb);
            store_facet(j, wb);
        }
    }
    if (valid && !ok) atomicOr(p.status, 1);
}

// ----------------------------------------------------------------------------- time_derivative
template <int ELEM, int P, int VAR>
struct K2Layout {
    using Sh = Shape<ELEM, P>;
    using T = Tune<ELEM, P, VAR>;
    static constexpr int NC = Sh::NC, D = Sh::D, NH = Sh::NH;
    static constexpr int PS = T::TS * T::R;
    static constexpr int off_h = 0;                          // NH x NQ hoisted volume states
    static constexpr int off_L = off_h + NH * Sh::NQ;        // D*D x NQ metric terms
    static constexpr int off_hf = off_L + D * D * Sh::NQ;    // NH x NF hoisted facet states
    static constexpr int off_nJ = off_hf + NH * Sh::NF;      // D x NF scaled normals
    static constexpr int off_st = off_nJ + D * Sh::NF;       // PS x NC staging
    static constexpr int size = off_st + PS * NC;
    // phase-3 aliases
    static constexpr int off_r = off_h;    // NC x NQ (r_q, then V output)
    static constexpr int off_y = off_L;    // NC x NP
    static constexpr int off_z = off_hf;   // NC x NZ
    static constexpr int off_W = off_z + NC * Sh::NZ;  // NC x NW
    static_assert(NC * Sh::NQ <= NH * Sh::NQ, "r_q alias");
    static_assert(NC * Sh::NP <= D * D * Sh::NQ, "y alias");
    static_assert(off_W + NC * Sh::NW <= size, "sum-factorisation temporaries alias");
    static_assert(Sh::NF <= PS, "facet flux alias in staging");
};

template <int ELEM, int P, int VAR>
__global__ void __launch_bounds__((Tune<ELEM, P, VAR>::TS * Tune<ELEM, P, VAR>::EPB))
k_time_derivative_tensor(TensorParams tp, double* __restrict__ dudt, RKStage rk) {
    using Sh = Shape<ELEM, P>;
    using T = Tune<ELEM, P, VAR>;
    using SF = SumFact<ELEM, P, T::TS>;
    using L = K2Layout<ELEM, P, VAR>;
    constexpr int D = Sh::D, NC = Sh::NC, NH = Sh::NH, NP = Sh::NP, NQ = Sh::NQ, NF = Sh::NF, NPF = Sh::NPF;
    constexpr int TS = T::TS, PS = L::PS, MAXW = TENSOR_MAXW;
    static_assert(NQ <= TS && NF <= TS, "one node per thread");
    extern __shared__ double smem[];
    const DevParams& p = tp.base;
    const int team = threadIdx.x / TS, tl = threadIdx.x - team * TS;
    const int k = element_of_block(blockIdx.x, gridDim.x, tp.xcd_remap) * T::EPB + team;
    const bool valid = k < p.N_e;
    const int kk = valid ? k : 0;
    double* sm = smem + team * L::size;
    double* sh = sm + L::off_h;
    double* sL = sm + L::off_L;
    double* shf = sm + L::off_hf;
    double* snJ = sm + L::off_nJ;
    double* sst = sm + L::off_st;
    using Ph = Physics<D, LAW_EULER>;

    // ---- phase 0: hoist volume and facet states, interface flux
    double r[NC], ff[NC];
#pragma unroll
    for (int e = 0; e < NC; ++e) r[e] = ff[e] = 0.0;
    if (tl < NQ) {
        const double* uqk = p.uq + (size_t)kk * NQ * NC;
        double ub[NC], h[NH];
#pragma unroll
        for (int e = 0; e < NC; ++e) ub[e] = uqk[e * NQ + tl];
        hoist<D>(p.law, ub, h);
#pragma unroll
        for (int v = 0; v < NH; ++v) sh[v * NQ + tl] = h[v];
        const double* Lk = p.Lam + (size_t)kk * NQ * D * D;
#pragma unroll
        for (int mn = 0; mn < D * D; ++mn) sL[mn * NQ + tl] = Lk[mn * NQ + tl];
    }
    if (tl < NF) {
        const int j = tl;
        const double* ufk = p.uf + (size_t)kk * NF * NC;
        double uin[NC], uout[NC], hin[NH], hout[NH], nrm[D], c[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) uin[e] = ufk[e * NF + j];
        hoist<D>(p.law, uin, hin);
#pragma unroll
        for (int v = 0; v < NH; ++v) shf[v * NF + j] = hin[v];
        const int g = p.mapP[j + (size_t)NF * kk];
        const int kn = g / NF, jn = g - kn * NF;
        const double* ufn = p.uf + (size_t)kn * NF * NC;
#pragma unroll
        for (int e = 0; e < NC; ++e) uout[e] = ufn[e * NF + jn];
        hoist<D>(p.law, uout, hout);
        const double jf = p.Jf[j + (size_t)NF * kk];
        const double inv_jf = 1.0 / jf;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            const double nj = p.nJf[m + D * (j + (size_t)NF * kk)];
            snJ[m * NF + j] = nj;
            nrm[m] = nj * inv_jf;
        }
        ec_dot<D>(p.law, hin, hout, nrm, c);
        const double bj = p.B[j] * jf;
        if (p.law.inviscid_flux == 1) {
            const double a = p.law.half_lambda * wave_speed_h<D>(p.law, hin, hout, nrm);
#pragma unroll
            for (int e = 0; e < NC; ++e) ff[e] = (c[e] + a * (uin[e] - uout[e])) * bj;
        } else {
#pragma unroll
            for (int e = 0; e < NC; ++e) ff[e] = c[e] * bj;
        }
    }
    __syncthreads();

    // ---- phase 1: unique pairs, staged contributions, owner reduction
    const int npairs = tp.n_pairs;
    for (int pass = 0; pass < tp.n_pass; ++pass) {
#pragma unroll
        for (int rr = 0; rr < T::R; ++rr) {
            const int slot = rr * TS + tl;
            const int q = pass * PS + slot;
            if (q < npairs) {
                const int code = tp.pk[q];
                const int i = code & 0xfff, j = (code >> 12) & 0xfff;
                double hi[NH], hj[NH], a[D], c[NC];
#pragma unroll
                for (int v = 0; v < NH; ++v) hi[v] = sh[v * NQ + i];
                if (code >> 24) {  // facet pair: C_ij F(u_i, u_f_j) . (nJf_j + nJq_{f(j), i}) / 2
                    const double cij = tp.cf[q];
                    const int f = j / NPF;
#pragma unroll
                    for (int v = 0; v < NH; ++v) hj[v] = shf[v * NF + j];
#pragma unroll
                    for (int n = 0; n < D; ++n) {
                        double s = snJ[n * NF + j];
#pragma unroll
                        for (int l = 0; l < D; ++l) s += sL[(l + D * n) * NQ + i] * p.fn[l + D * f];
                        a[n] = 0.5 * cij * s;
                    }
                } else {  // volume pair: a_n = sum_m S_m[i,j] (Lambda_i + Lambda_j)[m,n]
#pragma unroll
                    for (int v = 0; v < NH; ++v) hj[v] = sh[v * NQ + j];
                    double sv[D];
#pragma unroll
                    for (int m = 0; m < D; ++m) sv[m] = tp.cf[m * npairs + q];
#pragma unroll
                    for (int n = 0; n < D; ++n) {
                        double s = 0.0;
#pragma unroll
                        for (int m = 0; m < D; ++m)
                            s += sv[m] * (sL[(m + D * n) * NQ + i] + sL[(m + D * n) * NQ + j]);
                        a[n] = s;
                    }
                }
                ec_dot<D>(p.law, hi, hj, a, c);
#pragma unroll
                for (int e = 0; e < NC; ++e) sst[e * PS + slot] = c[e];
            }
        }
        __syncthreads();
        if (tl < NQ) {
            int codes[MAXW];
#pragma unroll
            for (int w = 0; w < MAXW; ++w) codes[w] = w < tp.wv ? tp.vt[(pass * tp.wv + w) * NQ + tl] : -1;
#pragma unroll
            for (int w = 0; w < MAXW; ++w) {
                const int code = codes[w];
                if (code >= 0) {
                    const double sg = (code & 1) ? -1.0 : 1.0;
                    const double* src = sst + (code >> 1);
#pragma unroll
                    for (int e = 0; e < NC; ++e) r[e] += sg * src[e * PS];
                }
            }
        }
        if (tl < NF) {
            int codes[MAXW];
#pragma unroll
            for (int w = 0; w < MAXW; ++w) codes[w] = w < tp.wf ? tp.ft[(pass * tp.wf + w) * NF + tl] : -1;
#pragma unroll
            for (int w = 0; w < MAXW; ++w) {
                const int code = codes[w];
                if (code >= 0) {
                    const double* src = sst + (code >> 1);
#pragma unroll
                    for (int e = 0; e < NC; ++e) ff[e] -= src[e * PS];
                }
            }
        }
        __syncthreads();
    }

    // ---- phase 2: lift r_q -= R' f_f
    double* sff = sst;  // NC x NF
    if (tl < NF) {
#pragma unroll
        for (int e = 0; e < NC; ++e) sff[e * NF + tl] = ff[e];
    }
    __syncthreads();
    double* sr = sm + L::off_r;
    if (tl < NQ) {
        constexpr int MAXR = TENSOR_MAXR;
        int js[MAXR];
        double rvs[MAXR];
#pragma unroll
        for (int w = 0; w < MAXR; ++w) {
            js[w] = w < tp.wrt ? tp.rt_idx[w * NQ + tl] : -1;
            rvs[w] = w < tp.wrt ? tp.rt_val[w * NQ + tl] : 0.0;
        }
#pragma unroll
        for (int w = 0; w < MAXR; ++w) {
            if (js[w] >= 0) {
#pragma unroll
                for (int e = 0; e < NC; ++e) r[e] -= rvs[w] * sff[e * NF + js[w]];
            }
        }
#pragma unroll
        for (int e = 0; e < NC; ++e) sr[e * NQ + tl] = r[e];
    }
    __syncthreads();

    // ---- phase 3: dudt = V' r_q, then the weight-adjusted mass solve
    double* sy = sm + L::off_y;
    double* sZ = sm + L::off_z;
    double* sW = sm + L::off_W;
    SF::Vt(tp, tl, sr, sy, sZ, sW);
    SF::V(tp, tl, sy, sr, sZ, sW);
    const double* Jk = p.Jq + (size_t)kk * NQ;
    for (int i = tl; i < NQ; i += TS) {
        const double jw = p.W[i] / Jk[i];
#pragma unroll
        for (int e = 0; e < NC; ++e) sr[e * NQ + i] *= jw;
    }
    __syncthreads();
    SF::Vt(tp, tl, sr, sy, sZ, sW);
    if (!valid) return;
    const size_t base = (size_t)kk * NP * NC;
    if (!rk.active) {
        for (int t = tl; t < NP * NC; t += TS) dudt[base + t] = sy[t];
    } else {
        for (int t = tl; t < NP * NC; t += TS) {
            const double kt = (rk.A == 0.0 ? 0.0 : rk.A * rk.tmp[base + t]) + rk.dt_of(kk) * sy[t];
            rk.tmp[base + t] = kt;
            rk.u[base + t] += rk.B * kt;
        }
    }
}

// ----------------------------------------------------------------------------- launch
struct TensorConfig {
    int TS, EPB, R, lds1, lds2, NQ, NF, NP;
};

template <int ELEM, int P, int VAR>
TensorConfig tensor_config_of() {
    using Sh = Shape<ELEM, P>;
    using T = Tune<ELEM, P, VAR>;
    return TensorConfig{T::TS, T::EPB, T::R,
                        (int)(sizeof(double) * (K1Layout<ELEM, P>::size * T::EPB1 + VFactors<ELEM, P>::size)),
                        (int)(sizeof(double) * K2Layout<ELEM, P, VAR>::size * T::EPB), Sh::NQ, Sh::NF, Sh::NP};
}

template <int ELEM, int P, int VAR>
hipError_t launch_tensor_t(const TensorParams& tp, const double* u, double* dudt, const RKStage& rk, hipStream_t s,
                           int which) {
    using T = Tune<ELEM, P, VAR>;
    const TensorConfig c = tensor_config_of<ELEM, P, VAR>();
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e1 = hipFuncSetAttribute((const void*)k_nodal_values_tensor<ELEM, P, VAR>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, c.lds1);
        hipError_t e2 = hipFuncSetAttribute((const void*)k_time_derivative_tensor<ELEM, P, VAR>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, c.lds2);
        if (e1 != hipSuccess) return e1;
        if (e2 != hipSuccess) return e2;
        attr_set = true;
    }
    if (which & 1) {
        const int nb = (tp.base.N_e + T::EPB1 - 1) / T::EPB1;
        hipLaunchKernelGGL((k_nodal_values_tensor<ELEM, P, VAR>), dim3(nb), dim3(T::TS1 * T::EPB1), c.lds1, s, tp, u);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (which & 2) {
        const int nb = (tp.base.N_e + T::EPB - 1) / T::EPB;
        hipLaunchKernelGGL((k_time_derivative_tensor<ELEM, P, VAR>), dim3(nb), dim3(T::TS * T::EPB), c.lds2, s, tp,
                           dudt, rk);
        return hipGetLastError();
    }
    return hipSuccess;
}

#define SSE_ALL_CASES                                                                               \
    SSE_CASE(ELEM_TRI, 2, 0) SSE_CASE(ELEM_TRI, 3, 0) SSE_CASE(ELEM_TRI, 4, 0) SSE_CASE(ELEM_TRI, 5, 0) \
    SSE_CASE(ELEM_TRI, 6, 0) SSE_CASE(ELEM_TET, 2, 0) SSE_CASE(ELEM_TET, 3, 0) SSE_CASE(ELEM_TET, 4, 0) \
    SSE_CASE(ELEM_TET, 4, 1) SSE_CASE(ELEM_TET, 4, 2) SSE_CASE(ELEM_TET, 4, 3) SSE_CASE(ELEM_TET, 4, 4) \
    SSE_CASE(ELEM_TET, 5, 0) SSE_CASE(ELEM_TET, 5, 1) SSE_CASE(ELEM_TET, 3, 1) SSE_CASE(ELEM_TRI, 4, 1)

// returns false if (elem, P, variant) has no compiled kernels
bool tensor_config(int elem, int P, int var, TensorConfig* c) {
#define SSE_CASE(E, PP, VV) \
    if (elem == E && P == PP && var == VV) { *c = tensor_config_of<E, PP, VV>(); return true; }
    SSE_ALL_CASES
#undef SSE_CASE
    return false;
}

hipError_t launch_tensor(const TensorParams& tp, const double* u, double* dudt, const RKStage& rk, hipStream_t s,
                         int which) {
#define SSE_CASE(E, PP, VV) \
    if (tp.elem == E && tp.P == PP && tp.variant == VV) return launch_tensor_t<E, PP, VV>(tp, u, dudt, rk, s, which);
    SSE_ALL_CASES
#undef SSE_CASE
    return hipErrorInvalidValue;
}

}  // namespace sse
