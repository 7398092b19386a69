// sse_vjp_struct.hip -- vector-Jacobian product of the flux-differencing Euler
// residual for ModalTensor Tri/Tet with compile-time element shapes: the
// reverse of the line-structured time derivative (sse_struct.hip) and of the
// entropy projection, same algebra as the general kernels of sse_vjp.hip.
//
//   dudt = M^{-1} V' r_q        ->  rbar = V M^{-1} ybar         (phase V1, sum-factorised)
//   r_q -= R' f_f               ->  fbar = -R rbar
//   f_f = B J_f f*(u_f, u_f+)   ->  interface flux adjoint at each facet node, for its own
//                                   flux and for the neighbour's flux at the shared node
//   Hadamard rows along lines   ->  each volume node gathers (dT_ij/du_i)' (rbar_j - rbar_i)
//                                   over its line partners j (S_m from the line tables)
//   facet correction            ->  each C pair once, by column chunks: one evaluation of the
//                                   shared part of the flux adjoint pulled back to both the
//                                   volume and the facet state, the two sums accumulated in
//                                   LDS with atomic adds
//   u_q = e2c(w_q'), u_f = e2c(R w_q'), w_q' = V M^{-1} V' (WJ c2e(V u))  reversed.
// The pull-back through u = e2c(w) at the projected states uses du/dw in closed
// form from the states' (rho, v, p), so the projection is not recomputed. The
// facet-correction sums are the only ones whose order varies from run to run
// (LDS atomics); every other sum has a fixed order.
#include <hip/hip_runtime.h>

#include "sse_adjoint.h"
#include "sse_tensor.h"

namespace sse {

namespace {

// team size and split of a node's partner list (as the forward kernel)
template <int ELEM, int P>
struct VjpTune;
// WAVES: waves per SIMD V2's register allocation must allow (what its LDS allows)
template <int P> struct VjpTune<ELEM_TRI, P> { static constexpr int TS = 64, SPLIT = 1, WAVES = 3; };
template <> struct VjpTune<ELEM_TET, 2> { static constexpr int TS = 64, SPLIT = 1, WAVES = 3; };
template <> struct VjpTune<ELEM_TET, 3> { static constexpr int TS = 128, SPLIT = 2, WAVES = 3; };
template <> struct VjpTune<ELEM_TET, 4> { static constexpr int TS = 256, SPLIT = 2, WAVES = 2; };
template <> struct VjpTune<ELEM_TET, 5> { static constexpr int TS = 256, SPLIT = 1, WAVES = 2; };

template <int ELEM, int P>
struct VjpLayout {
    using Sh = Shape<ELEM, P>;
    using T = VjpTune<ELEM, P>;
    static constexpr int NC = Sh::NC, D = Sh::D, NH = EulerAdj<Sh::D>::NH;
    static constexpr int ITEMS = ELEM == ELEM_TET ? 3 * Sh::N1 * Sh::N1 + Sh::N1 * Sh::N1 * Sh::N1 : 3 * Sh::N1;
    static constexpr int MAXITEM = (ITEMS + T::TS - 1) / T::TS * T::TS;
    // the facet normals nJ and fbar at the facet nodes are read from global memory where
    // they are used, and two arrays live in space that is dead by then, so that the LDS
    // of one element allows WAVES waves per SIMD
    static constexpr int off_h = 0;                          // NH x NQ hoisted volume states
    static constexpr int off_L = off_h + NH * Sh::NQ;        // D*D x NQ
    static constexpr int off_hf = off_L + D * D * Sh::NQ;    // NH x NF hoisted facet states
    static constexpr int off_rb = off_hf + NH * Sh::NF;      // NC x NQ rbar
    static constexpr int off_st = off_rb + NC * Sh::NQ;      // NC x (NQ + NF) C-row and C-column sums,
                                                              // or NC x MAXITEM chunk slots (deterministic)
    static constexpr int ST = NC * (Sh::NQ + Sh::NF) > NC * MAXITEM ? NC * (Sh::NQ + Sh::NF) : NC * MAXITEM;
    static constexpr int off_wf = off_st + ST;               // NC x NF w_f cotangent
    static constexpr int off_fn = off_wf + NC * Sh::NF;      // D x NFAC reference face normals
    static constexpr int size = off_fn + D * Sh::NFAC;
    // after the facet-correction pairs: the split partial rows, then the w_q' cotangent, in
    // the metric's array; at the end the TS reduction buffer in the C sums' array
    static constexpr int off_ub = off_L;
    static constexpr int off_red = off_st;
    static_assert(NC <= D * D && T::TS <= ST, "aliases fit");
};

// LDS of the reverse projection (phase V3)
template <int ELEM, int P>
struct VjpProjLayout {
    using Sh = Shape<ELEM, P>;
    static constexpr int NC = Sh::NC;
    static constexpr int off_z = 0, off_W = off_z + NC * Sh::NZ, off_y = off_W + NC * (Sh::NW > 0 ? Sh::NW : 1);
    static constexpr int off_x = off_y + NC * Sh::NP, off_u = off_x + NC * Sh::NQ, off_u0 = off_u + NC * Sh::NP;
    static constexpr int off_w = off_u0 + NC * Sh::NQ;  // w_q' cotangent
    static constexpr int size = off_w + NC * Sh::NQ;
};

}  // namespace

// Input of phase V1: ybar := a ybar + b x (x optional, a = 0 does not read ybar),
// stored back to `store` when set (the RK adjoint's kbar = A_s kbar + B_{s-1} ubar).
struct VjpIn {
    const double* x;
    double a, b;
    double* store;
};

// Output mode of phase V2: ubar (and the element's λ cotangent) = scale * VJP,
// or += scale * VJP when accumulating (the RK adjoint's ubar += dt R'(u_s)' kbar).
struct VjpOut {
    double scale;
    int accumulate;
};

// V1: rbar = V M^{-1} ybar (M^{-1} = V' (W / J) V, weight-adjusted), fbar = -R rbar,
// fbar also into the neighbour's exterior slot [k][e][j] (mapP is an involution)
template <int ELEM, int P>
__global__ void __launch_bounds__((VjpTune<ELEM, P>::TS))
k_vjp_v1_tensor(TensorParams tp, const double* ybar, double* __restrict__ rbar, double* __restrict__ fbar,
                double* __restrict__ fbar_ext, VjpIn in) {
    using Sh = Shape<ELEM, P>;
    constexpr int TS = VjpTune<ELEM, P>::TS, NC = Sh::NC, NP = Sh::NP, NQ = Sh::NQ, NF = Sh::NF;
    using SF = SumFact<ELEM, P, TS>;
    extern __shared__ double sm[];
    const DevParams& p = tp.base;
    const int k = element_of_block(blockIdx.x, gridDim.x, tp.xcd_remap), tl = threadIdx.x;
    double* sy = sm;                 // NC x NP
    double* sx = sy + NC * NP;       // NC x NQ; also the sum-factorisation temporary Z
    double* sZ = sx;                 // (dead before V's last stage, and after V''s first)
    double* sW = sx + NC * NQ;       // NC x NW
    // the factors of V in LDS after the work arrays
    using VF = VFactors<ELEM, P>;
    double* sfac = sW + NC * (Sh::NW > 0 ? Sh::NW : 1);
    VF::stage(tp, sfac, tl, TS);
    const TensorParams tq = VF::in_lds(tp, sfac);
    // the kept weight-adjusted inverse (requested with ybar), staged in x and W
    constexpr bool MW = NP * NP <= NC * (NQ + Sh::NW);
    MassPrefetch<NP, TS> pre;
    if constexpr (MW) {
        if (tp.Mw) mass_prefetch<NP, TS>(tp.Mw + (size_t)k * NP * NP, tl, pre);
    }
    const size_t kb = (size_t)k * NP * NC;
    for (int t = tl; t < NP * NC; t += TS) {
        double y = in.a == 0.0 ? 0.0 : in.a * ybar[kb + t];
        if (in.x) y += in.b * in.x[kb + t];
        if (in.store) in.store[kb + t] = y;
        sy[t] = y;
    }
    __syncthreads();
    if (MW && tp.Mw) {  // M^{-1} ybar = V' diag(W/J) V ybar with the kept matrix
        mass_apply_staged<NP, NC, TS>(pre, sx, sy, tl);
    } else {
        SF::V(tq, tl, sy, sx, sZ, sW);
        const double* Jk = p.Jq + (size_t)k * NQ;
        for (int i = tl; i < NQ; i += TS) {
            const double jw = p.W[i] / Jk[i];
#pragma unroll
            for (int e = 0; e < NC; ++e) sx[e * NQ + i] *= jw;
        }
        __syncthreads();
        SF::Vt(tq, tl, sx, sy, sZ, sW);
    }
    SF::V(tq, tl, sy, sx, sZ, sW);
    double* rk = rbar + (size_t)k * NQ * NC;
    for (int t = tl; t < NQ * NC; t += TS) rk[t] = sx[t];
    for (int j = tl; j < NF; j += TS) {
        const int g = p.mapP[j + (size_t)NF * k];
        double acc[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) acc[e] = 0.0;
        // the row four entries at a time, their loads issued together (entries past the
        // table width count as padding, index -1; a group whose first entry is padding ends it)
        for (int w0 = 0; w0 < tp.wr; w0 += 4) {
            int ii[4];
            double rv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool in = w0 + q < tp.wr;
                ii[q] = in ? tp.r_idx[(w0 + q) * NF + j] : -1;
                rv[q] = in ? tp.r_val[(w0 + q) * NF + j] : 0.0;
            }
            if (ii[0] < 0) break;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (ii[q] < 0) continue;
#pragma unroll
                for (int e = 0; e < NC; ++e) acc[e] += rv[q] * sx[e * NQ + ii[q]];
            }
        }
        const int kn = g / NF, jn = g - kn * NF;
        double* fk = fbar + (size_t)k * NF * NC;
        double* fx = fbar_ext + (size_t)kn * NF * NC;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            fk[e * NF + j] = -acc[e];
            fx[e * NF + jn] = -acc[e];
        }
    }
}

// V2: the two-point terms and the entropy variables, one element per workgroup. It
// leaves the cotangent of the projected volume states w_q' in the element's slice of
// rbar (read into LDS at its start) for V3.
template <int V>
struct VjpHalf {
    static constexpr int value = V;
};

// DET: every sum in a fixed order, as the deterministic time derivative does: partner
// shares of the volume pairs added one offset and one half-team at a time, each volume node
// pulling back its own row of C, and the facet columns of C through per-chunk slots.
template <int ELEM, int P, bool DET>
__global__ void __launch_bounds__((VjpTune<ELEM, P>::TS), (VjpTune<ELEM, P>::WAVES))
k_vjp_v2_tensor(TensorParams tp, StructTables st, double* __restrict__ rbar, const double* __restrict__ fbar,
                const double* __restrict__ fbar_ext, double* __restrict__ lam_partial, VjpOut out) {
    using Sh = Shape<ELEM, P>;
    using T = VjpTune<ELEM, P>;
    using L = VjpLayout<ELEM, P>;
    using A = EulerAdj<Sh::D>;
    using SF = SumFact<ELEM, P, T::TS>;
    constexpr int D = Sh::D, NC = Sh::NC, NH = L::NH, NP = Sh::NP, NQ = Sh::NQ, NF = Sh::NF, NPF = Sh::NPF;
    constexpr int N1 = Sh::N1, TS = T::TS, MAXITEM = L::MAXITEM, SPLIT = T::SPLIT, NT = TS / SPLIT;
    // each unordered line pair once, as in the forward kernel: cyclic offsets
    // 1..DH from the node, and the offset N1/2 of an even line from its lower half
    constexpr int DH = (N1 - 1) / 2, EVEN = N1 % 2 == 0 ? 1 : 0, PER_F = DH + EVEN;
    constexpr int NPART = D * PER_F, QH = (NPART + SPLIT - 1) / SPLIT;
    static_assert(NQ <= NT && NF <= NT, "one node per thread");
    constexpr bool EARLY = !DET && SPLIT == 2 && NT == 64;
    extern __shared__ double sm[];
    const DevParams& p = tp.base;
    const int k = element_of_block(blockIdx.x, gridDim.x, tp.xcd_remap);
    const int tt = threadIdx.x, half = tt / NT, tl = tt - half * NT;
    double* sh = sm + L::off_h;
    double* sL = sm + L::off_L;
    double* shf = sm + L::off_hf;
    double* srb = sm + L::off_rb;
    double* srow = sm + L::off_st;       // NC x NQ: facet-correction sums of the volume nodes
    double* sfac = srow + NC * NQ;       // NC x NF: facet-correction sums of the facet nodes
    double* sub = sm + L::off_ub;
    double* swf = sm + L::off_wf;
    double* red = sm + L::off_red;
    double* sfn = sm + L::off_fn;

    // ---- phase A: loads, hoisting, interface flux adjoints; clear the C-pair sums
    // EARLY: where each half-team is one wavefront, the first sweeps all the volume pairs
    // right after its hoisting (a wavefront needs no workgroup barrier to see its own LDS
    // writes) while the second computes the interface adjoints; the default mode only
    if constexpr (EARLY) {
        if (half == 0) {
            for (int t = tl; t < NC * NQ; t += NT) srow[t] = 0.0;
        } else {
            for (int t = tl; t < NC * NF; t += NT) sfac[t] = 0.0;
        }
    } else {
        for (int t = tt; t < NC * (NQ + NF); t += TS) srow[t] = 0.0;
    }
    if (tt < D * Sh::NFAC) sfn[tt] = p.fn[tt];
    double hi[NH], ubq[NC];
    double lam = 0.0;
#pragma unroll
    for (int e = 0; e < NC; ++e) ubq[e] = 0.0;
    // the volume line pairs of node tl (phase B; in the early schedule run in phase A)
    const int a_last = tl % N1, a_mid = (tl / N1) % N1, a_first = tl / (N1 * N1);
    auto pair_of = [&](int q, int& f, int& b, int& j, bool& act) {
        f = q / PER_F;
        const int dlt = q - f * PER_F + 1;
        const int stride = D == 3 ? (f == 0 ? N1 * N1 : (f == 1 ? N1 : 1)) : (f == 0 ? N1 : 1);
        const int af = D == 3 ? (f == 0 ? a_first : (f == 1 ? a_mid : a_last)) : (f == 0 ? a_mid : a_last);
        act = !(EVEN && 2 * dlt == N1 && 2 * af >= N1);
        b = af + dlt < N1 ? af + dlt : af + dlt - N1;
        j = tl + (b - af) * stride;
    };
    // S_m[i, j] of pair q (zeros for m < f in the table), requested one pair ahead
    auto load_s = [&](int q, double* sv) {
        int f, b, j;
        bool act;
        pair_of(q, f, b, j, act);
        const double* sg = st.SG + ((f * N1 + b) * D) * NQ + tl;
#pragma unroll
        for (int m = 0; m < D; ++m) sv[m] = sg[m * NQ];
    };
    auto eval = [&](int q, const double* sv) {
        int f, b, j;
        bool act;
        pair_of(q, f, b, j, act);
        if (!act) return;
        double hj[NH], av[D], gb[NC], ubj[NC];
#pragma unroll
        for (int v = 0; v < NH; ++v) hj[v] = sh[v * NQ + j];
#pragma unroll
        for (int n = 0; n < D; ++n) av[n] = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            if (m < f) continue;
#pragma unroll
            for (int n = 0; n < D; ++n) av[n] += sv[m] * (sL[(m + D * n) * NQ + tl] + sL[(m + D * n) * NQ + j]);
        }
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            gb[e] = srb[e * NQ + j] - srb[e * NQ + tl];
            ubj[e] = 0.0;
        }
        // the pair's term enters r_q[i] as -c and r_q[j] as +c: one evaluation
        // pulls (rbar_j - rbar_i) back to both states; node j's share goes
        // through the row accumulator that phase D adds
        A::ec_dot_adj_h2p(p.law, hi, hj, av, gb, ubq, ubj);
#pragma unroll
        for (int e = 0; e < NC; ++e) atomicAdd(&srow[e * NQ + j], ubj[e]);
    };
    // with two threads per node, the half-team's pairs with compile-time pair indices
    // (half is uniform over a wavefront), the S values of each pair requested one pair ahead
    auto sweep = [&](auto hc) {
        constexpr int H = decltype(hc)::value;
        constexpr int Q0 = H * QH, Q1 = (H + 1) * QH < NPART ? (H + 1) * QH : NPART;
        double sv[2][D];
        load_s(Q0, sv[0]);
#pragma unroll
        for (int q = Q0; q < Q1; ++q) {
            if (q + 1 < Q1) load_s(q + 1, sv[(q + 1 - Q0) & 1]);
            eval(q, sv[(q - Q0) & 1]);
        }
    };
    if (half == 0 && tl < NQ) {
        const double* uqk = p.uq + (size_t)k * NQ * NC;
        const double* Lk = p.Lam + (size_t)k * NQ * D * D;
        const double* rk = rbar + (size_t)k * NQ * NC;
        double uq[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            uq[e] = uqk[e * NQ + tl];
            srb[e * NQ + tl] = rk[e * NQ + tl];
        }
        A::hoist_full(p.law, uq, hi);
#pragma unroll
        for (int v = 0; v < NH; ++v) sh[v * NQ + tl] = hi[v];
#pragma unroll
        for (int mn = 0; mn < D * D; ++mn) sL[mn * NQ + tl] = Lk[mn * NQ + tl];
    }
    if constexpr (EARLY) {
        if (half == 0) {  // the wavefront's own LDS writes (hoisted states, metric, rbar, zeros)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (tl < NQ) {
                sweep(VjpHalf<0>{});
                sweep(VjpHalf<1>{});
            }
        }
    }
    // the facet nodes' interface adjoints (the longer chain) on the last half-team, so
    // that with two threads per node they run beside the volume hoisting
    if (half == SPLIT - 1 && tl < NF) {
        const int g = p.mapP[tl + (size_t)NF * k];  // first: the neighbour's loads depend on it
        const double* ufk = p.uf + (size_t)k * NF * NC;
        const double* uxk = st.uf_ext + (size_t)k * NF * NC;
        const double* fk = fbar + (size_t)k * NF * NC;
        const double* fxk = fbar_ext + (size_t)k * NF * NC;
        double uf[NC], ux[NC], ubf[NC], hin[NH], hout[NH], fo[NC], fown[NC], nrm[D], gb[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) ubf[e] = 0.0;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            uf[e] = ufk[e * NF + tl];
            ux[e] = uxk[e * NF + tl];
            fown[e] = fk[e * NF + tl];
            fo[e] = fxk[e * NF + tl];
        }
        A::hoist_full(p.law, uf, hin);
        A::hoist_full(p.law, ux, hout);
#pragma unroll
        for (int v = 0; v < NH; ++v) shf[v * NF + tl] = hin[v];
        const double jf = p.Jf[tl + (size_t)NF * k];
#pragma unroll
        for (int m = 0; m < D; ++m) {
            nrm[m] = p.nJf[m + D * (tl + (size_t)NF * k)] / jf;
        }
        const bool lf = (p.law.inviscid_flux == 1);
        // this node's flux f*(u_in, u_out, n), cotangent B J_f fbar (keep u_in)
        const double bj = p.B[tl] * jf;
#pragma unroll
        for (int e = 0; e < NC; ++e) gb[e] = bj * fown[e];
        A::ec_dot_adj_h(p.law, hin, hout, nrm, gb, true, ubf);
        if (lf) lam += 0.5 * A::lf_diss_adj(p.law, uf, ux, nrm, gb, true, ubf);  // halfλ = λ / 2
        // the neighbour's flux at the shared node, f*(u_out, u_in, n'), cotangent
        // B' J_f' fbar' (keep u_in, its exterior state)
        const int kn = g / NF, jn = g - kn * NF;
        const double jfn = p.Jf[jn + (size_t)NF * kn];
#pragma unroll
        for (int m = 0; m < D; ++m) nrm[m] = p.nJf[m + D * (jn + (size_t)NF * kn)] / jfn;
        const double bjn = p.B[jn] * jfn;
#pragma unroll
        for (int e = 0; e < NC; ++e) gb[e] = bjn * fo[e];
        A::ec_dot_adj_h(p.law, hout, hin, nrm, gb, false, ubf);
        if (lf) A::lf_diss_adj(p.law, ux, uf, nrm, gb, false, ubf);
#pragma unroll
        for (int e = 0; e < NC; ++e) swf[e * NF + tl] = ubf[e];  // interface part of the u_f cotangent
    }
    __syncthreads();

    // ---- phase B: Hadamard row (line partners) and facet-correction row of node tl
    if (SPLIT > 1 && half > 0 && tl < NQ) {
#pragma unroll
        for (int v = 0; v < NH; ++v) hi[v] = sh[v * NQ + tl];
    }
    if (tl < NQ) {  // the node's metric and rbar are read from LDS (fewer live registers)
        if constexpr (!DET) {
            // with two threads per node, the half-team's pairs with compile-time pair indices
            // (half is uniform over a wavefront), the S values of each pair requested one
            // pair ahead
            if constexpr (EARLY) {  // swept in phase A
            } else if constexpr (SPLIT > 1) {
                if (half == 0)
                    sweep(VjpHalf<0>{});
                else
                    sweep(VjpHalf<1>{});
            } else {  // one thread per node: a runtime loop keeps the register count down
                double sa[D], sb[D];
                if (NPART > 0) load_s(0, sa);
#pragma unroll 1
                for (int q = 0; q < NPART; q += 2) {  // two pairs per trip: the S values alternate buffers
                    if (q + 1 < NPART) load_s(q + 1, sb);
                    eval(q, sa);
                    if (q + 1 < NPART) {
                        if (q + 2 < NPART) load_s(q + 2, sa);
                        eval(q + 1, sb);
                    }
                }
            }
        }
    }
    if constexpr (DET) {
        // the same pairs, every thread through QH trips; the half-teams add their partner
        // shares one after the other (the partners of one offset are distinct nodes)
#pragma unroll 1
        for (int t = 0; t < QH; ++t) {
            const int q = half * QH + t;
            const bool has = tl < NQ && q < NPART && q < (half + 1) * QH;
            double ubj[NC];
            int jd = 0;
            bool act = false;
#pragma unroll
            for (int e = 0; e < NC; ++e) ubj[e] = 0.0;
            if (has) {
                const int a_last = tl % N1, a_mid = (tl / N1) % N1, a_first = tl / (N1 * N1);
                const int f = q / PER_F, dlt = q - f * PER_F + 1;
                const int stride = D == 3 ? (f == 0 ? N1 * N1 : (f == 1 ? N1 : 1)) : (f == 0 ? N1 : 1);
                const int af = D == 3 ? (f == 0 ? a_first : (f == 1 ? a_mid : a_last)) : (f == 0 ? a_mid : a_last);
                act = !(EVEN && 2 * dlt == N1 && 2 * af >= N1);
                const int b = af + dlt < N1 ? af + dlt : af + dlt - N1;
                jd = tl + (b - af) * stride;
                if (act) {
                    double hj[NH], av[D], gb[NC];
#pragma unroll
                    for (int v = 0; v < NH; ++v) hj[v] = sh[v * NQ + jd];
#pragma unroll
                    for (int n = 0; n < D; ++n) av[n] = 0.0;
#pragma unroll
                    for (int m = 0; m < D; ++m) {
                        if (m < f) continue;
                        const double sv = st.SG[((f * N1 + b) * D + m) * NQ + tl];
#pragma unroll
                        for (int n = 0; n < D; ++n)
                            av[n] += sv * (sL[(m + D * n) * NQ + tl] + sL[(m + D * n) * NQ + jd]);
                    }
#pragma unroll
                    for (int e = 0; e < NC; ++e) gb[e] = srb[e * NQ + jd] - srb[e * NQ + tl];
                    A::ec_dot_adj_h2p(p.law, hi, hj, av, gb, ubq, ubj);
                }
            }
#pragma unroll 1
            for (int h = 0; h < SPLIT; ++h) {
                if (has && act && half == h) {
#pragma unroll
                    for (int e = 0; e < NC; ++e) srow[e * NQ + jd] += ubj[e];
                }
                __syncthreads();
            }
        }
        // the partner shares into the node's row; the accumulator area is then free
        if (half == 0 && tl < NQ) {
#pragma unroll
            for (int e = 0; e < NC; ++e) ubq[e] += srow[e * NQ + tl];
        }
        __syncthreads();
        // the node's own row of C (entries split between the half-teams)
        if (tl < NQ) {
#pragma unroll 1
            for (int w = half; w < st.rw; w += SPLIT) {
                const int j = st.cr_j[w * NQ + tl];
                if (j < 0) break;
                const double cv = st.cr_v[w * NQ + tl];
                const int fc = j / NPF;
                double hj[NH], av[D], gb[NC], dummy[NC];
#pragma unroll
                for (int v = 0; v < NH; ++v) hj[v] = shf[v * NF + j];
#pragma unroll
                for (int n = 0; n < D; ++n) {
                    double s = p.nJf[n + D * (j + (size_t)NF * k)];
#pragma unroll
                    for (int l = 0; l < D; ++l) s += sL[(l + D * n) * NQ + tl] * sfn[l + D * fc];
                    av[n] = 0.5 * cv * s;
                }
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    gb[e] = -(srb[e * NQ + tl] + fbar[(size_t)k * NF * NC + e * NF + j]);
                    dummy[e] = 0.0;
                }
                A::ec_dot_adj_h2p(p.law, hi, hj, av, gb, ubq, dummy);
            }
        }
    }
    // ---- phase C: facet-correction pairs by column chunks, both sides at once
#pragma unroll 1
    for (int it = 0; it < MAXITEM / TS; ++it) {
        const int item = it * TS + tt;
        if (item >= st.n_items) continue;  // the chunks are packed: items below n_items are real
        // the chunk (facet node, N1 volume nodes and C entries) requested at once
        const int j = st.ci_j[item];
        int ci[N1];
        double cv[N1];
#pragma unroll
        for (int w = 0; w < N1; ++w) {
            ci[w] = st.ci_i[w * MAXITEM + item];
            cv[w] = st.ci_v[w * MAXITEM + item];
        }
        double hj[NH], part[NC], fnl[D], njv[D], fbj[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            part[e] = 0.0;
            fbj[e] = fbar[(size_t)k * NF * NC + e * NF + j];
        }
#pragma unroll
        for (int n = 0; n < D; ++n) njv[n] = p.nJf[n + D * (j + (size_t)NF * k)];
#pragma unroll
        for (int v = 0; v < NH; ++v) hj[v] = shf[v * NF + j];
        const int fc = j / NPF;
#pragma unroll
        for (int l = 0; l < D; ++l) fnl[l] = sfn[l + D * fc];
#pragma unroll
        for (int w = 0; w < N1; ++w) {
            const int i = ci[w];
            if (i < 0) continue;
            double hv[NH], av[D], gb[NC], uv[NC];
#pragma unroll
            for (int v = 0; v < NH; ++v) hv[v] = sh[v * NQ + i];
#pragma unroll
            for (int n = 0; n < D; ++n) {
                double s = njv[n];
#pragma unroll
                for (int l = 0; l < D; ++l) s += sL[(l + D * n) * NQ + i] * fnl[l];
                av[n] = 0.5 * cv[w] * s;
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                gb[e] = -(srb[e * NQ + i] + fbj[e]);
                uv[e] = 0.0;
            }
            A::ec_dot_adj_h2p(p.law, hv, hj, av, gb, uv, part);
            if constexpr (!DET) {
#pragma unroll
                for (int e = 0; e < NC; ++e) atomicAdd(&srow[e * NQ + i], uv[e]);
            }
        }
        if constexpr (DET) {
#pragma unroll
            for (int e = 0; e < NC; ++e) srow[e * MAXITEM + item] = part[e];  // the chunk's slot
        } else {
#pragma unroll
            for (int e = 0; e < NC; ++e) atomicAdd(&sfac[e * NF + j], part[e]);
        }
    }
    // the lift's rows of R' (all MAXR entries, padding as a zero weight on node 0),
    // requested before the barriers so that their latency passes during phase D
    constexpr int MAXR = TENSOR_MAXR;
    int rj[MAXR];
    double rv[MAXR];
    if (half == 0 && tl < NQ) {
#pragma unroll
        for (int w = 0; w < MAXR; ++w) {
            rj[w] = w < tp.wrt ? tp.rt_idx[w * NQ + tl] : -1;
            rv[w] = w < tp.wrt ? tp.rt_val[w * NQ + tl] : 0.0;
        }
    }
    if constexpr (SPLIT > 1 && !EARLY) {
        __syncthreads();  // the pairs are done with the metric array, which takes the partial rows
        if (half == 1 && tl < NQ) {
#pragma unroll
            for (int e = 0; e < NC; ++e) sub[e * NQ + tl] = ubq[e];
        }
    }
    __syncthreads();
    // ---- phase D: cotangents of the entropy variables, w = c2e(u): volume and facet
    if (half == 0 && tl < NQ) {
        if (SPLIT > 1 && !EARLY) {
#pragma unroll
            for (int e = 0; e < NC; ++e) ubq[e] += sub[e * NQ + tl];
        }
        if constexpr (!DET) {
#pragma unroll
            for (int e = 0; e < NC; ++e) ubq[e] += srow[e * NQ + tl];
        }
    }
    if (half == 0 && tl < NF) {
        double ubf[NC], hf[NH];
        if constexpr (DET) {  // the node's chunk slots, in chunk order
#pragma unroll
            for (int e = 0; e < NC; ++e) ubf[e] = 0.0;
            const int c0 = st.ci_first[tl], c1 = st.ci_first[tl + 1];
            for (int c = c0; c < c1; ++c) {
#pragma unroll
                for (int e = 0; e < NC; ++e) ubf[e] += srow[e * MAXITEM + c];
            }
        } else {
#pragma unroll
            for (int e = 0; e < NC; ++e) ubf[e] = sfac[e * NF + tl];
        }
#pragma unroll
        for (int v = 0; v < NH; ++v) hf[v] = shf[v * NF + tl];
        // the C-pair sums are (rho, v, p) cotangents: to conservative, beside the interface part
        double ubc[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) ubc[e] = swf[e * NF + tl];
        A::prim_cot_to_cons(p.law, hf, ubf, ubc);
        double wb[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) wb[e] = 0.0;
        A::e2c_adj_prim(p.law, hf[A::H_RHO], hf + A::H_V, hf[A::H_P], ubc, wb);
#pragma unroll
        for (int e = 0; e < NC; ++e) swf[e * NF + tl] = wb[e];
    }
    __syncthreads();  // the w_f cotangents are complete
    if (half == 0 && tl < NQ) {
        double wb[NC];
#pragma unroll
        for (int v = 0; v < NH; ++v) hi[v] = sh[v * NQ + tl];
        // the pair sums are (rho, v, p) cotangents: to conservative once per node
        double ubc[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            ubc[e] = 0.0;
            wb[e] = 0.0;
        }
        A::prim_cot_to_cons(p.law, hi, ubq, ubc);
        A::e2c_adj_prim(p.law, hi[A::H_RHO], hi + A::H_V, hi[A::H_P], ubc, wb);
        // w_q' cotangent += R' (w_f cotangent)
#pragma unroll
        for (int w = 0; w < MAXR; ++w) {
            if (rj[w] < 0) continue;
#pragma unroll
            for (int e = 0; e < NC; ++e) wb[e] += rv[w] * swf[e * NF + rj[w]];
        }
#pragma unroll
        for (int e = 0; e < NC; ++e) sub[e * NQ + tl] = wb[e];
    }
    __syncthreads();  // the w_q' cotangent is complete
    double* wk = rbar + (size_t)k * NQ * NC;
    for (int t = tt; t < NQ * NC; t += TS) wk[t] = sub[t];
    // λ cotangent of this element, summed in a fixed order
    red[tt] = lam;
    __syncthreads();
    for (int s = TS / 2; s > 0; s >>= 1) {
        if (tt < s) red[tt] += red[tt + s];
        __syncthreads();
    }
    if (tt == 0) lam_partial[k] = (out.accumulate ? lam_partial[k] : 0.0) + out.scale * red[0];
}

// V3: the reverse of w_q' = V M^{-1} V' (WJ c2e(V u)), sum-factorised, one element per
// workgroup: ubar (+)= scale * V' c2e'(V u)' (WJ V M^{-1} V' wbar), M^{-1} = V' (W/J V .).
// Its own kernel so that these low-register stages run at higher occupancy than V2.
template <int ELEM, int P>
__global__ void __launch_bounds__((VjpTune<ELEM, P>::TS))
k_vjp_v3_tensor(TensorParams tp, const double* __restrict__ u, const double* __restrict__ wbar,
                double* __restrict__ ubar, VjpOut out) {
    using Sh = Shape<ELEM, P>;
    using T = VjpTune<ELEM, P>;
    using L = VjpProjLayout<ELEM, P>;
    using A = EulerAdj<Sh::D>;
    using SF = SumFact<ELEM, P, T::TS>;
    constexpr int NC = Sh::NC, NP = Sh::NP, NQ = Sh::NQ, TS = T::TS;
    extern __shared__ double sm[];
    const DevParams& p = tp.base;
    const int k = element_of_block(blockIdx.x, gridDim.x, tp.xcd_remap);
    const int tt = threadIdx.x;
    double* sZ = sm + L::off_z;
    double* sW = sm + L::off_W;
    double* sy = sm + L::off_y;
    double* sx = sm + L::off_x;
    double* su = sm + L::off_u;
    double* su0 = sm + L::off_u0;
    double* sw = sm + L::off_w;
    using VF = VFactors<ELEM, P>;
    double* sfac = sm + L::size;  // the factors of V in LDS after the work arrays
    VF::stage(tp, sfac, tt, TS);
    const TensorParams tq = VF::in_lds(tp, sfac);
    // the kept weight-adjusted inverse (requested with u and the cotangent), staged in the
    // temporaries Z and W
    constexpr bool MW = NP * NP <= L::off_y - L::off_z;
    MassPrefetch<NP, TS> pre;
    if constexpr (MW) {
        if (tp.Mw) mass_prefetch<NP, TS>(tp.Mw + (size_t)k * NP * NP, tt, pre);
    }
    const double* uk = u + (size_t)k * NP * NC;
    const double* wk = wbar + (size_t)k * NQ * NC;
    for (int t = tt; t < NP * NC; t += TS) su[t] = uk[t];
    for (int t = tt; t < NQ * NC; t += TS) sw[t] = wk[t];
    __syncthreads();
    SF::V(tq, tt, su, su0, sZ, sW);   // V u
    SF::Vt(tq, tt, sw, sy, sZ, sW);   // V' (w_q' cotangent)
    const double* Jk = p.Jq + (size_t)k * NQ;
    if (MW && tp.Mw) {                // M^{-1} = V' (W/J V .) with the kept matrix
        mass_apply_staged<NP, NC, TS>(pre, sZ, sy, tt);
    } else {
        SF::V(tq, tt, sy, sx, sZ, sW);
        for (int i = tt; i < NQ; i += TS) {
            const double jw = p.W[i] / Jk[i];
#pragma unroll
            for (int e = 0; e < NC; ++e) sx[e * NQ + i] *= jw;
        }
        __syncthreads();
        SF::Vt(tq, tt, sx, sy, sZ, sW);
    }
    SF::V(tq, tt, sy, sx, sZ, sW);    // V (M^{-1} V' w_q'bar), then times WJ and through c2e
    for (int i = tt; i < NQ; i += TS) {
        const double wj = p.W[i] * Jk[i];
        double u0[NC], wb[NC], ub[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            u0[e] = su0[e * NQ + i];
            wb[e] = wj * sx[e * NQ + i];
            ub[e] = 0.0;
        }
        A::c2e_adj(p.law, u0, wb, ub);
#pragma unroll
        for (int e = 0; e < NC; ++e) sw[e * NQ + i] = ub[e];
    }
    __syncthreads();
    SF::Vt(tq, tt, sw, sy, sZ, sW);
    double* outk = ubar + (size_t)k * NP * NC;
    for (int t = tt; t < NP * NC; t += TS) outk[t] = (out.accumulate ? outk[t] : 0.0) + out.scale * sy[t];
}

// ----------------------------------------------------------------------------- launch
template <int ELEM, int P>
size_t vjp_tensor_lds(int which) {
    using Sh = Shape<ELEM, P>;
    if (which == 1)
        return sizeof(double) * (Sh::NC * (Sh::NP + Sh::NQ + (Sh::NW > 0 ? Sh::NW : 1)) + VFactors<ELEM, P>::size);
    if (which == 3) return sizeof(double) * (VjpProjLayout<ELEM, P>::size + VFactors<ELEM, P>::size);
    return sizeof(double) * VjpLayout<ELEM, P>::size;
}

template <int ELEM, int P>
hipError_t launch_vjp_tensor_t(const TensorParams& tp, const StructTables& st, const double* u, const double* ybar,
                               double* ubar, double* rbar, double* fbar, double* fbar_ext, double* lam_partial,
                               double scale, int accumulate, const VjpIn& in, hipStream_t s) {
    using T = VjpTune<ELEM, P>;
    static bool attr = false;
    const size_t l1 = vjp_tensor_lds<ELEM, P>(1), l2 = vjp_tensor_lds<ELEM, P>(2), l3 = vjp_tensor_lds<ELEM, P>(3);
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)k_vjp_v1_tensor<ELEM, P>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)l1);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_vjp_v2_tensor<ELEM, P, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)l2);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_vjp_v2_tensor<ELEM, P, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)l2);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_vjp_v3_tensor<ELEM, P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)l3);
        if (e != hipSuccess) return e;
        attr = true;
    }
    const int nb = tp.base.N_e;
    hipLaunchKernelGGL((k_vjp_v1_tensor<ELEM, P>), dim3(nb), dim3(T::TS), l1, s, tp, ybar, rbar, fbar, fbar_ext, in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (st.det)
        hipLaunchKernelGGL((k_vjp_v2_tensor<ELEM, P, true>), dim3(nb), dim3(T::TS), l2, s, tp, st, rbar, fbar,
                           fbar_ext, lam_partial, VjpOut{scale, accumulate});
    else
        hipLaunchKernelGGL((k_vjp_v2_tensor<ELEM, P, false>), dim3(nb), dim3(T::TS), l2, s, tp, st, rbar, fbar,
                           fbar_ext, lam_partial, VjpOut{scale, accumulate});
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_vjp_v3_tensor<ELEM, P>), dim3(nb), dim3(T::TS), l3, s, tp, u, rbar, ubar,
                       VjpOut{scale, accumulate});
    return hipGetLastError();
}

#define SSE_VJP_CASES                                                                                      \
    SSE_CASE(ELEM_TRI, 2) SSE_CASE(ELEM_TRI, 3) SSE_CASE(ELEM_TRI, 4) SSE_CASE(ELEM_TRI, 5) SSE_CASE(ELEM_TRI, 6) \
    SSE_CASE(ELEM_TET, 2) SSE_CASE(ELEM_TET, 3) SSE_CASE(ELEM_TET, 4) SSE_CASE(ELEM_TET, 5)

// LDS of the larger phase, or 0 when (elem, P) has no compiled kernels
size_t vjp_tensor_config(int elem, int P) {
#define SSE_CASE(E_, PP)                                                                         \
    if (elem == E_ && P == PP) {                                                                 \
        const size_t a = vjp_tensor_lds<E_, PP>(1), b = vjp_tensor_lds<E_, PP>(2);               \
        const size_t c = vjp_tensor_lds<E_, PP>(3), ab = a > b ? a : b;                          \
        return ab > c ? ab : c;                                                                  \
    }
    SSE_VJP_CASES
#undef SSE_CASE
    return 0;
}

// ubar = scale * (dR/du)' y and lam_partial[k] = scale * element k's (dR/dλ)' y, or both
// added (accumulate != 0), with y = in_a ybar + in_b in_x (in_x optional; y also stored to
// in_store when set); phase 1 must have run on u with tp/st's buffers.
hipError_t launch_vjp_tensor(const TensorParams& tp, const StructTables& st, const double* u, const double* ybar,
                             double* ubar, double* rbar, double* fbar, double* fbar_ext, double* lam_partial,
                             double scale, int accumulate, const double* in_x, double in_a, double in_b,
                             double* in_store, hipStream_t s) {
    const VjpIn in{in_x, in_a, in_b, in_store};
#define SSE_CASE(E_, PP)                                                                                   \
    if (tp.elem == E_ && tp.P == PP)                                                                       \
        return launch_vjp_tensor_t<E_, PP>(tp, st, u, ybar, ubar, rbar, fbar, fbar_ext, lam_partial, scale,  \
                                           accumulate, in, s);
    SSE_VJP_CASES
#undef SSE_CASE
    return hipErrorInvalidValue;
}

}  // namespace sse
