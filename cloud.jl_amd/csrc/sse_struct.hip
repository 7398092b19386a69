// sse_struct.hip -- time_derivative for ModalTensor Tri/Tet (Euler, FluxDifferencingForm,
// entropy-conservative two-point flux) driven by the tensor-line structure of
// the operators instead of pair tables.
//
// The union sparsity of S_1..S_d on collapsed-coordinate tensor elements is a
// set of 1D lines (src/Solvers/operators.jl:182-200 with D_xi of
// src/SpatialDiscretizations/SpatialDiscretizations.jl:414-418): node i pairs
// with every node j that differs from i in exactly one tensor index (family f).
// Each unordered volume pair is evaluated once, by the thread of one of its nodes
// (cyclic offsets along the line), and its term
//   c_e = sum_n a_n F(u_i, u_j)[e,n],  a_n = sum_m S_m[i,j] (Lambda_i + Lambda_j)[m,n]
// enters r_q[i] as -c (in registers) and r_q[j] as +c (S antisymmetric, F
// symmetric; an LDS atomic add): the reference's scatter over i < j,
// src/Solvers/flux_differencing_form.jl:37-75. Each facet-correction pair (C
// nonzero) is evaluated once too, by column chunks balanced across the team: the
// same two-point term enters the volume node's row and the facet node's column,
// both accumulated in LDS with atomic adds. These atomic sums are the only ones
// whose order varies from run to run.
//
// A team of TS threads owns one element. The exterior facet state is read
// contiguously: phase 1 writes each facet state also into the neighbour's
// exterior slot (the reference's u_f[CI[mapP[:,k]], :] gather).
#include <hip/hip_runtime.h>

#include "sse_tensor.h"

namespace sse {

// team size, elements per block, threads per node (SPLIT threads share a
// node's partner list; their partial sums meet in LDS)
template <int ELEM, int P>
struct StructTune;
// WAVES: waves per SIMD the register allocation must allow (the LDS of Tet p = 4
// allows 4; one more register would drop it to 3)
template <int P> struct StructTune<ELEM_TRI, P> { static constexpr int TS = 64, EPB = 2, SPLIT = 1, WAVES = 1; };
template <> struct StructTune<ELEM_TET, 2> { static constexpr int TS = 64, EPB = 1, SPLIT = 1, WAVES = 1; };
template <> struct StructTune<ELEM_TET, 3> { static constexpr int TS = 128, EPB = 1, SPLIT = 2, WAVES = 4; };
template <> struct StructTune<ELEM_TET, 4> { static constexpr int TS = 256, EPB = 1, SPLIT = 2, WAVES = 4; };
template <> struct StructTune<ELEM_TET, 5> { static constexpr int TS = 256, EPB = 1, SPLIT = 1, WAVES = 1; };

template <int ELEM, int P>
struct StructLayout {
    using Sh = Shape<ELEM, P>;
    using T = StructTune<ELEM, P>;
    static constexpr int NC = Sh::NC, D = Sh::D, NH = Sh::NH;
    // facet column chunks per element (N1 entries each): faces with one volume
    // line per facet node, plus N1 chunks per node of a face coupled to N1^2 nodes
    static constexpr int ITEMS = ELEM == ELEM_TET ? 3 * Sh::N1 * Sh::N1 + Sh::N1 * Sh::N1 * Sh::N1 : 3 * Sh::N1;
    static constexpr int MAXITEM = (ITEMS + T::TS - 1) / T::TS * T::TS;
    static constexpr int off_h = 0;                            // NH x NQ
    static constexpr int off_L = off_h + NH * Sh::NQ;          // D*D x NQ
    static constexpr int off_hf = off_L + D * D * Sh::NQ;      // NH x NF
    static constexpr int off_nJ = off_hf + NH * Sh::NF;        // D x NF
    static constexpr int off_st = off_nJ + D * Sh::NF;         // row and facet accumulators of C
    static constexpr int ST = NC * (Sh::NQ + Sh::NF) > NC * MAXITEM ? NC * (Sh::NQ + Sh::NF) : NC * MAXITEM;
    static constexpr int off_fn = off_st + ST;                 // D x NFAC reference face normals
    static constexpr int size = off_fn + D * Sh::NFAC;
    static constexpr int off_ff = off_h;                       // NC x NF (after the pair phase)
    static constexpr int off_r = off_hf;                       // NC x NQ
    static constexpr int off_rb = off_L;                       // NC x NQ partial rows of split threads
    static constexpr int off_y = off_L;                        // NC x NP
    static constexpr int off_z = off_st;                       // NC x NZ
    static constexpr int off_W = off_z + NC * Sh::NZ;          // NC x NW
    static_assert(NC * Sh::NF <= NH * Sh::NQ, "f_f alias");
    static_assert(NC * Sh::NQ <= NH * Sh::NF + D * Sh::NF + ST, "r_q alias");
    static_assert(NC * Sh::NP <= D * D * Sh::NQ, "y alias");
    static_assert(NC * Sh::NQ <= D * D * Sh::NQ, "split partial-row alias");
    static_assert(off_W + NC * Sh::NW <= off_fn, "sum-factorisation temporaries alias");
};


template <int V>
struct IntConst {
    static constexpr int value = V;
};

// DET: the deterministic variant. The partner shares of the volume pairs are added in a
// fixed order (the pairs of one offset at a time, one half-team after the other: the
// targets of one offset are distinct), each volume node evaluates its own row of C, and
// the column sums of C go through per-chunk slots summed in chunk order. The C pairs are
// then evaluated twice (row and column), and no sum depends on the order of atomics.
template <int ELEM, int P, bool DET>
__global__ void __launch_bounds__((StructTune<ELEM, P>::TS * StructTune<ELEM, P>::EPB), (StructTune<ELEM, P>::WAVES))
k_time_derivative_struct(TensorParams tp, StructTables st, double* __restrict__ dudt, RKStage rk, int n_teams) {
    using Sh = Shape<ELEM, P>;
    using T = StructTune<ELEM, P>;
    using SF = SumFact<ELEM, P, T::TS>;
    using L = StructLayout<ELEM, P>;
    constexpr int D = Sh::D, NC = Sh::NC, NH = Sh::NH, NP = Sh::NP, NQ = Sh::NQ, NF = Sh::NF, NPF = Sh::NPF;
    constexpr int N1 = Sh::N1, TS = T::TS, MAXITEM = L::MAXITEM, SPLIT = T::SPLIT, NT = TS / SPLIT;
    constexpr int MAXR = TENSOR_MAXR;
    // volume pairs: on each line node a evaluates its pairs with the nodes at cyclic
    // offsets 1..DH from it, which covers every unordered pair of an odd-length line
    // once; on an even-length line the offset N1/2 pairs nodes both ways and only
    // the node in the lower half of the line evaluates it
    constexpr int DH = (N1 - 1) / 2, EVEN = N1 % 2 == 0 ? 1 : 0, PER_F = DH + EVEN;
    constexpr int NPART = D * PER_F, QH = (NPART + SPLIT - 1) / SPLIT;  // volume pairs per thread
    static_assert(NQ <= NT && NF <= NT, "one node per thread");
    extern __shared__ double smem[];
    const DevParams& p = tp.base;
    const int team = threadIdx.x / TS, tt = threadIdx.x - team * TS;
    const int half = tt / NT, tl = tt - half * NT;  // node index tl, partner share `half`
    double* sm = smem + team * L::size;
    double* sh = sm + L::off_h;
    double* sL = sm + L::off_L;
    double* shf = sm + L::off_hf;
    double* snJ = sm + L::off_nJ;
    double* sst = sm + L::off_st;
    using Ph = Physics<D, LAW_EULER>;

    const double Wi = tl < NQ ? p.W[tl] : 0.0;
    const double Bj = tl < NF ? p.B[tl] : 0.0;

    // ---- prefetch registers. With two threads per node, the first half-team loads and
    // hoists the volume nodes and the second the facet nodes (the interface flux is the
    // longer chain of the two)
    constexpr int FH = SPLIT - 1;
    double pu[NC], pL[D * D], pJ, pf[NC], px[NC], pn[D], pJf;
    auto prefetch = [&](int k) {
        if (half == 0 && tl < NQ) {
            const double* uqk = p.uq + (size_t)k * NQ * NC;
            const double* Lk = p.Lam + (size_t)k * NQ * D * D;
#pragma unroll
            for (int e = 0; e < NC; ++e) pu[e] = uqk[e * NQ + tl];
#pragma unroll
            for (int mn = 0; mn < D * D; ++mn) pL[mn] = Lk[mn * NQ + tl];
            pJ = p.Jq[(size_t)k * NQ + tl];
        }
        if (FH == 0 && tl < NF) {
            const double* ufk = p.uf + (size_t)k * NF * NC;
            const double* uxk = st.uf_ext + (size_t)k * NF * NC;
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                pf[e] = ufk[e * NF + tl];
                px[e] = uxk[e * NF + tl];
            }
#pragma unroll
            for (int m = 0; m < D; ++m) pn[m] = p.nJf[m + D * (tl + (size_t)NF * k)];
            pJf = p.Jf[tl + (size_t)NF * k];
        }
    };

    // one element per team; a team past the last element computes on a valid
    // element (block-wide barriers) and stores nothing
    const int last = p.N_e - 1;
    const int kb = element_of_block(blockIdx.x, gridDim.x, tp.xcd_remap) * T::EPB;
    {
        const int k = min(kb + team, last);
        const bool valid = kb + team < p.N_e;
        prefetch(k);
        // ---- phase A: hoist, interface flux; clear the facet-correction accumulators
        double* srow = sst;            // NC x NQ: C-row sums of the volume nodes
        double* sfac = sst + NC * NQ;  // NC x NF: C-column sums of the facet nodes
        double* sfn = sm + L::off_fn;
        // the row accumulators start at zero; each facet node's column accumulator starts at
        // minus its interface flux (written below by the node's thread), so that phase D reads
        // f_f - (column sums of C) from one place and the flux needs no register meanwhile
        for (int t = tt; t < NC * NQ; t += TS) sst[t] = 0.0;
        if (tt < D * Sh::NFAC) sfn[tt] = p.fn[tt];
        double hi[NH], Li[D * D], r[NC], ff[NC], Ji = 1.0;
#pragma unroll
        for (int e = 0; e < NC; ++e) r[e] = ff[e] = 0.0;
        if (half == 0 && tl < NQ) {
            hoist<D>(p.law, pu, hi);
#pragma unroll
            for (int v = 0; v < NH; ++v) sh[v * NQ + tl] = hi[v];
#pragma unroll
            for (int mn = 0; mn < D * D; ++mn) {
                Li[mn] = pL[mn];
                sL[mn * NQ + tl] = pL[mn];
            }
            Ji = pJ;
        }
        if (half == FH && tl < NF) {
            if constexpr (FH > 0) {  // the facet loads of the second half-team
                const double* ufk = p.uf + (size_t)k * NF * NC;
                const double* uxk = st.uf_ext + (size_t)k * NF * NC;
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    pf[e] = ufk[e * NF + tl];
                    px[e] = uxk[e * NF + tl];
                }
#pragma unroll
                for (int m = 0; m < D; ++m) pn[m] = p.nJf[m + D * (tl + (size_t)NF * k)];
                pJf = p.Jf[tl + (size_t)NF * k];
            }
            double hin[NH], hout[NH], nrm[D], c[NC];
            hoist<D>(p.law, pf, hin);
            hoist<D>(p.law, px, hout);
#pragma unroll
            for (int v = 0; v < NH; ++v) shf[v * NF + tl] = hin[v];
            const double inv_jf = 1.0 / pJf;
#pragma unroll
            for (int m = 0; m < D; ++m) {
                snJ[m * NF + tl] = pn[m];
                nrm[m] = pn[m] * inv_jf;
            }
            ec_dot<D>(p.law, hin, hout, nrm, c);
            const double bj = Bj * pJf;
            if (p.law.inviscid_flux == 1) {
                const double al = p.law.half_lambda * wave_speed_h<D>(p.law, hin, hout, nrm);
#pragma unroll
                for (int e = 0; e < NC; ++e) ff[e] = (c[e] + al * (pf[e] - px[e])) * bj;
            } else {
#pragma unroll
                for (int e = 0; e < NC; ++e) ff[e] = c[e] * bj;
            }
            if constexpr (!DET) {
#pragma unroll
                for (int e = 0; e < NC; ++e) sfac[e * NF + tl] = -ff[e];
            }
        }
        __syncthreads();

        // ---- phase B: volume Hadamard row of node tl
        if (SPLIT > 1 && half > 0 && tl < NQ) {
#pragma unroll
            for (int v = 0; v < NH; ++v) hi[v] = sh[v * NQ + tl];
#pragma unroll
            for (int mn = 0; mn < D * D; ++mn) Li[mn] = sL[mn * NQ + tl];
        }
        if (tl < NQ) {
            // tensor indices of node tl (last index fastest) and the family strides
            const int a_last = tl % N1, a_mid = (tl / N1) % N1, a_first = tl / (N1 * N1);
            // pair q of this node: family f, partner index b on the line, partner node j;
            // `act` is false for the offset N1/2 of an even line taken by the other node
            auto pair_of = [&](int q, int& f, int& b, int& j, bool& act) {
                f = q / PER_F;
                const int dlt = q - f * PER_F + 1;
                const int stride = D == 3 ? (f == 0 ? N1 * N1 : (f == 1 ? N1 : 1)) : (f == 0 ? N1 : 1);
                const int af = D == 3 ? (f == 0 ? a_first : (f == 1 ? a_mid : a_last)) : (f == 0 ? a_mid : a_last);
                act = !(EVEN && 2 * dlt == N1 && 2 * af >= N1);
                b = af + dlt < N1 ? af + dlt : af + dlt - N1;
                j = tl + (b - af) * stride;
            };
            // S_m[i, j] of pair q (the table holds the zeros of m < f too); requested one
            // pair ahead and unconditionally, so that the loads overlap the flux
            // evaluation of the pair before
            auto load_s = [&](int q, double* sv) {
                int f, b, j;
                bool act;
                pair_of(q, f, b, j, act);
                const double* sg = st.SGh + ((f * N1 + b) * D) * NQ + tl;
#pragma unroll
                for (int m = 0; m < D; ++m) sv[m] = sg[m * NQ];
            };
            auto eval = [&](int q, const double* sv) {
                int f, b, j;
                bool act;
                pair_of(q, f, b, j, act);
                if (!act) return;
                double hj[NH], av[D], c[NC];
#pragma unroll
                for (int v = 0; v < NH; ++v) hj[v] = sh[v * NQ + j];
                // -a_n / 2 = sum_{m >= f} (-S_m[i,j] / 2) (Lambda_i + Lambda_j)[m,n]  (S_m = 0 for
                // m < f): the flux is linear in a, so c comes out negated, as both rows take it
#pragma unroll
                for (int n = 0; n < D; ++n) av[n] = 0.0;
#pragma unroll
                for (int m = 0; m < D; ++m) {  // static indices into Li
                    if (m < f) continue;
#pragma unroll
                    for (int n = 0; n < D; ++n) av[n] += sv[m] * (Li[m + D * n] + sL[(m + D * n) * NQ + j]);
                }
                ec_dot_h<D>(p.law, hi, hj, av, c);
                // S antisymmetric, F symmetric: node j's row receives +c, through
                // the row accumulator that phase D subtracts
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    r[e] += c[e];
                    atomicAdd(&srow[e * NQ + j], c[e]);
                }
            };
            if constexpr (!DET) {
                // the half-team's pairs with compile-time pair indices (half is uniform over
                // a wavefront): the family of each pair, and with it the S_m that vanish, is
                // then known, and the straight-line code lets the next pair's loads overlap
                auto sweep = [&](auto hc) {
                    constexpr int H = decltype(hc)::value;
                    constexpr int Q0 = H * QH, Q1 = (H + 1) * QH < NPART ? (H + 1) * QH : NPART;
                    double sv[2][D];  // the S values of a pair are requested one pair ahead
                    load_s(Q0, sv[0]);
#pragma unroll
                    for (int q = Q0; q < Q1; ++q) {
                        if (q + 1 < Q1) load_s(q + 1, sv[(q + 1 - Q0) & 1]);
                        eval(q, sv[(q - Q0) & 1]);
                    }
                };
                if (half == 0) {
                    sweep(IntConst<0>{});
                } else if constexpr (SPLIT > 1) {
                    sweep(IntConst<1>{});
                }
            }
        }
        if constexpr (DET) {
            // the same pairs, every thread through QH trips: at each trip the half-teams
            // add their partner shares one after the other (plain read-add-write: within a
            // half-team, the partners of one offset are distinct nodes)
#pragma unroll 1
            for (int t = 0; t < QH; ++t) {
                const int q = half * QH + t;
                const bool has = tl < NQ && q < NPART && q < (half + 1) * QH;
                double c[NC];
                int jd = 0;
                bool act = false;
                if (has) {
                    const int a_last = tl % N1, a_mid = (tl / N1) % N1, a_first = tl / (N1 * N1);
                    const int f = q / PER_F, dlt = q - f * PER_F + 1;
                    const int stride = D == 3 ? (f == 0 ? N1 * N1 : (f == 1 ? N1 : 1)) : (f == 0 ? N1 : 1);
                    const int af = D == 3 ? (f == 0 ? a_first : (f == 1 ? a_mid : a_last)) : (f == 0 ? a_mid : a_last);
                    act = !(EVEN && 2 * dlt == N1 && 2 * af >= N1);
                    const int b = af + dlt < N1 ? af + dlt : af + dlt - N1;
                    jd = tl + (b - af) * stride;
                    if (act) {
                        double hj[NH], av[D];
#pragma unroll
                        for (int v = 0; v < NH; ++v) hj[v] = sh[v * NQ + jd];
#pragma unroll
                        for (int n = 0; n < D; ++n) av[n] = 0.0;
#pragma unroll
                        for (int m = 0; m < D; ++m) {
                            if (m < f) continue;
                            const double sv = st.SGh[((f * N1 + b) * D + m) * NQ + tl];
#pragma unroll
                            for (int n = 0; n < D; ++n) av[n] += sv * (Li[m + D * n] + sL[(m + D * n) * NQ + jd]);
                        }
                        ec_dot_h<D>(p.law, hi, hj, av, c);  // negated, as above
#pragma unroll
                        for (int e = 0; e < NC; ++e) r[e] += c[e];
                    }
                }
#pragma unroll 1
                for (int h = 0; h < SPLIT; ++h) {
                    if (has && act && half == h) {
#pragma unroll
                        for (int e = 0; e < NC; ++e) srow[e * NQ + jd] += c[e];
                    }
                    __syncthreads();
                }
            }
            // the partner shares into the node's row; the accumulator area is then free
            if (half == 0 && tl < NQ) {
#pragma unroll
                for (int e = 0; e < NC; ++e) r[e] -= srow[e * NQ + tl];
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
                    double hj[NH], av[D], c[NC];
#pragma unroll
                    for (int v = 0; v < NH; ++v) hj[v] = shf[v * NF + j];
#pragma unroll
                    for (int n = 0; n < D; ++n) {
                        double s = snJ[n * NF + j];
#pragma unroll
                        for (int l = 0; l < D; ++l) s += sL[(l + D * n) * NQ + tl] * sfn[l + D * fc];
                        av[n] = 0.25 * cv * s;  // half the contraction vector
                    }
                    ec_dot_h<D>(p.law, hi, hj, av, c);
#pragma unroll
                    for (int e = 0; e < NC; ++e) r[e] -= c[e];
                }
            }
        }
        // ---- phase C: facet-correction pairs by column chunks, both sums by LDS atomics
#pragma unroll 1
        for (int it = 0; it < MAXITEM / TS; ++it) {
            const int item = it * TS + tt;
            if (item >= st.n_items) continue;  // the chunks are packed: items below n_items are real
            // the chunk (facet node, N1 volume nodes and C entries) requested at once
            const int j = st.ci_j[item];
            int ci[N1];
#pragma unroll
            for (int w = 0; w < N1; ++w) ci[w] = st.ci_i[w * MAXITEM + item];
            const int fc = j / NPF;
            double hj[NH], fnl[D], nj[D], part[NC];
#pragma unroll
            for (int l = 0; l < D; ++l) fnl[l] = sfn[l + D * fc];
#pragma unroll
            for (int e = 0; e < NC; ++e) part[e] = 0.0;
#pragma unroll
            for (int v = 0; v < NH; ++v) hj[v] = shf[v * NF + j];
#pragma unroll
            for (int n = 0; n < D; ++n) nj[n] = snJ[n * NF + j];
#pragma unroll
            for (int w = 0; w < N1; ++w) {
                const int i = ci[w];
                if (i < 0) continue;
                const double cvw = st.ci_v[w * MAXITEM + item];
                double hv[NH], av[D], c[NC];
#pragma unroll
                for (int v = 0; v < NH; ++v) hv[v] = sh[v * NQ + i];
#pragma unroll
                for (int n = 0; n < D; ++n) {
                    double s = nj[n];
#pragma unroll
                    for (int l = 0; l < D; ++l) s += sL[(l + D * n) * NQ + i] * fnl[l];
                    av[n] = 0.25 * cvw * s;  // half the contraction vector
                }
                ec_dot_h<D>(p.law, hv, hj, av, c);
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    part[e] += c[e];
                    if constexpr (!DET) atomicAdd(&srow[e * NQ + i], c[e]);
                }
            }
            if constexpr (DET) {
#pragma unroll
                for (int e = 0; e < NC; ++e) sst[e * MAXITEM + item] = part[e];  // the chunk's slot
            } else {
#pragma unroll
                for (int e = 0; e < NC; ++e) atomicAdd(&sfac[e * NF + j], part[e]);
            }
        }
        // the lift's rows of R' (all MAXR entries at once, padding as a zero weight on node 0)
        // and this thread's share of the factors of V, requested here so that their latency
        // passes during phases D and E
        int rj[MAXR];
        double rv[MAXR];
        if (half == 0 && tl < NQ) {
#pragma unroll
            for (int w = 0; w < MAXR; ++w) {
                rj[w] = w < tp.wrt ? tp.rt_idx[w * NQ + tl] : -1;
                rv[w] = w < tp.wrt ? tp.rt_val[w * NQ + tl] : 0.0;
            }
        }
        using VF = VFactors<ELEM, P>;
        constexpr bool STAGE = VF::size <= NH * NQ;
        constexpr int NVF = STAGE ? (VF::size + TS - 1) / TS : 1;
        double vf[NVF];
        if constexpr (STAGE) VF::fetch(tp, vf, tt, TS);
        __syncthreads();
        // ---- phase D: the C sums into r (volume nodes) and f_f (facet nodes); f_f to LDS
        double* sff = sm + L::off_ff;
        double* srb = sm + L::off_rb;
        if (SPLIT > 1 && half == 1 && tl < NQ) {
#pragma unroll
            for (int e = 0; e < NC; ++e) srb[e * NQ + tl] = r[e];
        }
        if constexpr (!DET) {
            if (half == 0 && tl < NQ) {  // C-row sums, read before r_q may overwrite them
#pragma unroll
                for (int e = 0; e < NC; ++e) r[e] -= srow[e * NQ + tl];
            }
            if (half == 0 && tl < NF) {
#pragma unroll
                for (int e = 0; e < NC; ++e) sff[e * NF + tl] = -sfac[e * NF + tl];
            }
        } else if (half == FH && tl < NF) {  // the node's chunk slots, in chunk order
            double col[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) col[e] = 0.0;
            const int c0 = st.ci_first[tl], c1 = st.ci_first[tl + 1];
            for (int c = c0; c < c1; ++c) {
#pragma unroll
                for (int e = 0; e < NC; ++e) col[e] += sst[e * MAXITEM + c];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) sff[e * NF + tl] = ff[e] - col[e];
        }
        __syncthreads();
        // ---- phase E: lift r_q -= R' f_f
        double* sr = sm + L::off_r;
        if (half == 0 && tl < NQ) {
            if (SPLIT > 1) {
#pragma unroll
                for (int e = 0; e < NC; ++e) r[e] += srb[e * NQ + tl];
            }
#pragma unroll
            for (int w = 0; w < MAXR; ++w) {
                if (rj[w] < 0) continue;
#pragma unroll
                for (int e = 0; e < NC; ++e) r[e] -= rv[w] * sff[e * NF + rj[w]];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) sr[e * NQ + tl] = r[e];
        }
        __syncthreads();
        // ---- phase F: dudt = V' r_q, weight-adjusted mass solve; the factors of V are
        // copied into the (now dead) hoisted-state array when they fit there
        double* sy = sm + L::off_y;
        double* sZ = sm + L::off_z;
        double* sW = sm + L::off_W;
        TensorParams tq = tp;
        if constexpr (STAGE) {
            VF::put(vf, sm + L::off_h, tt, TS);
            tq = VF::in_lds(tp, sm + L::off_h);
            __syncthreads();
        }
        // the weight-adjusted inverse V' diag(W/J) V kept per element, requested before
        // the transform and staged where r_q and the temporaries were
        constexpr bool MW = NP * NP <= L::off_fn - L::off_hf;
        MassPrefetch<NP, TS> pre;
        if constexpr (MW) {
            if (tp.Mw) mass_prefetch<NP, TS>(tp.Mw + (size_t)k * NP * NP, tt, pre);
        }
        SF::Vt(tq, tt, sr, sy, sZ, sW);
        if (MW && tp.Mw) {
            mass_apply_staged<NP, NC, TS>(pre, sm + L::off_hf, sy, tt);
        } else {
            SF::V(tq, tt, sy, sr, sZ, sW);
            if (half == 0 && tl < NQ) {
                const double jw = Wi / Ji;
#pragma unroll
                for (int e = 0; e < NC; ++e) sr[e * NQ + tl] *= jw;
            }
            __syncthreads();
            SF::Vt(tq, tt, sr, sy, sZ, sW);
        }
        const size_t base = (size_t)k * NP * NC;
        if (!valid) {
        } else if (!rk.active) {
            for (int t = tt; t < NP * NC; t += TS) dudt[base + t] = sy[t];
        } else {
            for (int t = tt; t < NP * NC; t += TS) {
                const double kt = (rk.A == 0.0 ? 0.0 : rk.A * rk.tmp[base + t]) + rk.dt_of(k) * sy[t];
                rk.tmp[base + t] = kt;
                if (rk.keep) rk.keep[base + t] = rk.u[base + t];
                rk.u[base + t] += rk.B * kt;
            }
        }
    }
}

struct StructConfig {
    int TS, EPB, lds, max_item;
};

template <int ELEM, int P>
StructConfig struct_config_of() {
    using T = StructTune<ELEM, P>;
    return StructConfig{T::TS, T::EPB, (int)(sizeof(double) * StructLayout<ELEM, P>::size * T::EPB),
                        StructLayout<ELEM, P>::MAXITEM};
}

template <int ELEM, int P, bool DET>
hipError_t launch_struct_v(const TensorParams& tp, const StructTables& st, double* dudt, const RKStage& rk,
                           hipStream_t s, int n_blocks) {
    using T = StructTune<ELEM, P>;
    const StructConfig c = struct_config_of<ELEM, P>();
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_time_derivative_struct<ELEM, P, DET>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, c.lds);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((k_time_derivative_struct<ELEM, P, DET>), dim3(n_blocks), dim3(T::TS * T::EPB), c.lds, s, tp,
                       st, dudt, rk, n_blocks * T::EPB);
    return hipGetLastError();
}

template <int ELEM, int P>
hipError_t launch_struct_t(const TensorParams& tp, const StructTables& st, double* dudt, const RKStage& rk,
                           hipStream_t s, int n_blocks) {
    return st.det ? launch_struct_v<ELEM, P, true>(tp, st, dudt, rk, s, n_blocks)
                  : launch_struct_v<ELEM, P, false>(tp, st, dudt, rk, s, n_blocks);
}

#define SSE_STRUCT_CASES                                                                       \
    SSE_CASE(ELEM_TRI, 2) SSE_CASE(ELEM_TRI, 3) SSE_CASE(ELEM_TRI, 4) SSE_CASE(ELEM_TRI, 5) \
    SSE_CASE(ELEM_TRI, 6) SSE_CASE(ELEM_TET, 2) SSE_CASE(ELEM_TET, 3) SSE_CASE(ELEM_TET, 4) SSE_CASE(ELEM_TET, 5)

template <int ELEM, int P>
int struct_blocks_per_cu() {
    using T = StructTune<ELEM, P>;
    const StructConfig c = struct_config_of<ELEM, P>();
    (void)hipFuncSetAttribute((const void*)k_time_derivative_struct<ELEM, P, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, c.lds);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_time_derivative_struct<ELEM, P, false>, T::TS * T::EPB,
                                                     c.lds) !=
        hipSuccess)
        return 0;
    return nb;
}

int struct_occupancy(int elem, int P) {
#define SSE_CASE(E, PP) \
    if (elem == E && P == PP) return struct_blocks_per_cu<E, PP>();
    SSE_STRUCT_CASES
#undef SSE_CASE
    return 0;
}

bool struct_config(int elem, int P, StructConfig* c) {
#define SSE_CASE(E, PP) \
    if (elem == E && P == PP) { *c = struct_config_of<E, PP>(); return true; }
    SSE_STRUCT_CASES
#undef SSE_CASE
    return false;
}

hipError_t launch_struct(const TensorParams& tp, const StructTables& st, double* dudt, const RKStage& rk,
                         hipStream_t s, int n_blocks) {
#define SSE_CASE(E, PP) \
    if (tp.elem == E && tp.P == PP) return launch_struct_t<E, PP>(tp, st, dudt, rk, s, n_blocks);
    SSE_STRUCT_CASES
#undef SSE_CASE
    return hipErrorInvalidValue;
}

}  // namespace sse
