// sse_std_tensor.hip -- StandardForm residual of scalar linear laws (advection,
// advection-diffusion with BR1) on ModalTensor triangles and tetrahedra, with
// compile-time element shapes.
//
// Same algebra as the general StandardForm kernels (the reference's
// src/Solvers/standard_form_first_order.jl:16-94 and
// src/Solvers/standard_form_second_order.jl:3-75 with the operators of
// src/Solvers/operators.jl:83-160), arranged for throughput:
//  * a workgroup advances E "slots" at once. A slot is one element of one
//    problem; with an ensemble of R problems that share the mesh and differ
//    only in the diffusion coefficient, consecutive slots are the members of
//    one element, so they also share its geometric factors. Every operator
//    entry loaded (V factors, rows of D_m, D_m', R, R') serves all E slots
//    from registers;
//  * V and V' are sum-factorised (warped-product factors), the mass solve is
//    the weight-adjusted one with M^{-1} = I (src/Solvers/mass_matrix.jl:59-75);
//  * the reference derivative operators are applied along lines: with
//    D_m = sum_l diag(c_lm) D1_l (D1_l the 1D derivative in collapsed direction l
//    on every line of the tensor grid, src/SpatialDiscretizations/
//    SpatialDiscretizations.jl:414-418) and the per-element metric folded into
//    Lt_ln = sum_m c_lm Lambda_mn once at set-up,
//      sum_m D_m' (W Lambda_mn x) = sum_l D1_l' (W Lt_ln x),
//      sum_m Lambda_mn (D_m x)    = sum_l Lt_ln (D1_l x),
//    so a node's derivative reads N1 neighbours per direction and no table;
//  * the BR1 auxiliary variable of all d directions goes through one
//    three-direction mass solve, and R q_q reuses the stored facet values;
//  * with the skew-symmetric mapping, R x at the facets is never recomputed:
//    R u_q = u_f and R f_n = a_n u_f - b q_f,n (R is linear), so the interface
//    and surface terms combine pointwise at the facet nodes.
//
// Lanes: lane g = k * R + r is member r of element k. Intermediate arrays
// (u_q, u_f, q_q, q_f) are indexed by lane; states and time steps by
// r * N_e + k (each member a reference-layout (N_p, 1, N_e) array).
//
// Tangent pairs: the residual is linear in u and in b (viscous flux -b q,
// BR1 flux -b {q}.n, q linear in u), so v = du/db obeys the semi-discrete system
// v' = R(v; b) + dR/db(u) with dR/db(u) the viscous terms of u at unit b. A
// member pair (u, v) advanced together by the same low-storage RK stages gives
// the exact derivative of the discrete solution; only the time derivative
// kernel knows the pairing (it adds -q(u) to the tangent's viscous flux).
#include <hip/hip_runtime.h>

#include "sse_tensor.h"

namespace sse {

struct Ensemble {
    int R;              // members per element (1: a single problem)
    const double* b;    // diffusion coefficient of each member, or null for the law's b
    int tangent;        // members come in pairs (u, v = du/db): the odd member's residual
                        // is R(v; b) + dR/db(u), dR/db(u) = -(viscous terms of u at b = 1)
};

// threads per workgroup and slots per workgroup of the three kernels (few slots
// per workgroup measured fastest: more workgroups in flight hide the latency of
// the dependent stages better than wider reuse of the small tables)
template <int ELEM, int P>
struct StdTune;
#define SSE_STD_TUNE(E_, PP, ts, e1, e2, e3) \
    template <>                              \
    struct StdTune<E_, PP> {                 \
        static constexpr int TS = ts, E1 = e1, E2 = e2, E3 = e3; \
    };
SSE_STD_TUNE(ELEM_TRI, 2, 64, 8, 2, 2)
SSE_STD_TUNE(ELEM_TRI, 3, 64, 8, 2, 2)
SSE_STD_TUNE(ELEM_TRI, 4, 64, 8, 2, 2)
SSE_STD_TUNE(ELEM_TRI, 5, 64, 8, 2, 2)
SSE_STD_TUNE(ELEM_TRI, 6, 64, 8, 2, 2)
SSE_STD_TUNE(ELEM_TET, 2, 64, 8, 2, 2)
SSE_STD_TUNE(ELEM_TET, 3, 64, 8, 2, 2)
SSE_STD_TUNE(ELEM_TET, 4, 128, 8, 1, 2)
SSE_STD_TUNE(ELEM_TET, 5, 256, 8, 1, 2)
#undef SSE_STD_TUNE

namespace {

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// waves per SIMD that `lds_doubles` of LDS per workgroup of `threads` allow (160 KB
// per CU, 4 SIMDs), capped at 7: the register budget a kernel is held to, so that
// the compiler does not trade the occupancy the LDS allows for early loads
constexpr int lds_waves(int lds_doubles, int threads) {
    const int wg = (160 * 1024) / (8 * lds_doubles);
    const int w = wg * (threads / 64) / 4;
    return w < 1 ? 1 : (w > 7 ? 7 : w);
}

// Node i of the tensor grid has coordinate (i / stride_l) % N1 in direction l
// (the last direction fastest).
template <int ELEM, int P>
struct Lines {
    static constexpr int D = Shape<ELEM, P>::D, N1 = P + 1;
    __host__ __device__ static constexpr int stride(int l) {
        return D == 3 ? (l == 0 ? N1 * N1 : (l == 1 ? N1 : 1)) : (l == 0 ? N1 : 1);
    }
    __device__ static int coord(int i, int l) { return (i / stride(l)) % N1; }
};

__device__ __forceinline__ size_t state_row(long long g, int R, int N_e) {
    const long long k = g / R, r = g - k * R;
    return (size_t)r * N_e + k;
}

__device__ __forceinline__ double member_b(const DevParams& p, const Ensemble& en, long long g) {
    return en.b ? en.b[g % en.R] : p.b;
}

// lane of the exterior state of facet node j of lane g
__device__ __forceinline__ long long neighbour(const DevParams& p, int R, long long g, int NF, int j, int& jn) {
    const long long k = g / R, r = g - k * R;
    const int gl = p.mapP[j + (size_t)NF * k];
    const int kn = gl / NF;
    jn = gl - kn * NF;
    return (long long)kn * R + r;
}

// acc[c] += sum over row `row` of an ELL table (width a multiple of 4, -1 pads
// with value 0) of val * x[c * XS + idx]: four entries per step, so that a
// step's loads are in flight together
template <int NCV, int XS>
__device__ __forceinline__ void ell_apply(const int* __restrict__ idx, const double* __restrict__ val, int width,
                                          int nrows, int row, const double* x, double* acc) {
    for (int w = 0; w < width; w += 4) {
        int ii[4];
        double vv[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            ii[t] = idx[(w + t) * nrows + row];
            vv[t] = val[(w + t) * nrows + row];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int i = ii[t] < 0 ? 0 : ii[t];
#pragma unroll
            for (int c = 0; c < NCV; ++c) acc[c] += vv[t] * x[c * XS + i];
        }
        if (ii[3] < 0) break;
    }
}

}  // namespace

// ----------------------------------------------------------------------------- phase 1
// u_q = V u, u_f = R u_q
template <int ELEM, int P>
struct StdNodalLayout {
    using Sh = Shape<ELEM, P>;
    static constexpr int E = StdTune<ELEM, P>::E1;
    // the sum-factorisation temporary Z lives in u_q's array (dead before V's last stage)
    static constexpr int u = 0, q = u + E * Sh::NP, z = q, w = q + E * Sh::NQ;
    static constexpr int size = w + E * cmax(Sh::NW, 1);
};

template <int ELEM, int P>
__global__ void __launch_bounds__((StdTune<ELEM, P>::TS))
k_std_nodal_tensor(TensorParams tp, Ensemble en, const double* __restrict__ u) {
    using Sh = Shape<ELEM, P>;
    using T = StdTune<ELEM, P>;
    using L = StdNodalLayout<ELEM, P>;
    constexpr int E = T::E1, TS = T::TS, NP = Sh::NP, NQ = Sh::NQ, NF = Sh::NF;
    using SF = SumFact<ELEM, P, TS, E>;
    extern __shared__ double sm[];
    const DevParams& p = tp.base;
    const int tl = threadIdx.x;
    const long long nl = (long long)p.N_e * en.R;
    const long long g0 = (long long)element_of_block(blockIdx.x, gridDim.x, tp.xcd_remap) * E;
    double *su = sm + L::u, *sq = sm + L::q;
    for (int t = tl; t < E * NP; t += TS) {
        const int s = t / NP, j = t - s * NP;
        const long long g = g0 + s;
        su[t] = g < nl ? u[state_row(g, en.R, p.N_e) * NP + j] : 0.0;
    }
    __syncthreads();
    SF::V(tp, tl, su, sq, sm + L::z, sm + L::w);
    const int nvalid = (int)(nl - g0 < E ? nl - g0 : E);
    for (int t = tl; t < nvalid * NQ; t += TS) p.uq[(size_t)g0 * NQ + t] = sq[t];
    for (int j = tl; j < NF; j += TS) {
        double acc[E];
#pragma unroll
        for (int s = 0; s < E; ++s) acc[s] = 0.0;
        ell_apply<E, NQ>(tp.r_idx, tp.r_val, tp.wr, NF, j, sq, acc);
#pragma unroll
        for (int s = 0; s < E; ++s)
            if (s < nvalid) p.uf[(size_t)(g0 + s) * NF + j] = acc[s];
    }
}

// ----------------------------------------------------------------------------- BR1 auxiliary variable
// q_m = M^{-1} V' [ -VT_m(u_q) + R' (B J_f u*_m) ],  u*_m = 1/2 (u^- + u^+) n_m,
// VT_n(x) = h sum_m D_m' (W Lambda_mn x) - h W sum_m Lambda_mn (D_m x) + R' (1/2 B nJf_n R x)
// (h = 1/2, skew-symmetric mapping) or sum_m D_m' (W Lambda_mn x) (standard mapping).
// With the skew form the facet terms combine to R' (1/2 B nJf_n u^+). All d
// directions go through each pass together.
template <int ELEM, int P, bool SKEW>
struct StdAuxLayout {
    using Sh = Shape<ELEM, P>;
    static constexpr int E = StdTune<ELEM, P>::E2, D = Sh::D;
    static constexpr int A = cmax(D * Sh::NQ, D * (Sh::NZ + Sh::NP));  // D1_l u, then SumFact Z and y
    static constexpr int B = cmax(D * Sh::NQ, D * Sh::NW);             // g_ln of one n, then SumFact W
    // u^- only for the standard mapping
    static constexpr int uq = 0, up = uq + E * Sh::NQ, um = up + E * Sh::NF, a = um + (SKEW ? 0 : E * Sh::NF);
    static constexpr int b = a + E * A, r = b + E * B, tf = r + E * D * Sh::NQ;
    static constexpr int d1 = tf + E * D * Sh::NF;
    static constexpr int size = d1 + D * Sh::N1 * Sh::N1;
};

template <int ELEM, int P, bool SKEW>
__global__ void __launch_bounds__((StdTune<ELEM, P>::TS), (lds_waves(StdAuxLayout<ELEM, P, SKEW>::size, StdTune<ELEM, P>::TS)))
    k_std_aux_tensor(TensorParams tp, Ensemble en) {
    using Sh = Shape<ELEM, P>;
    using T = StdTune<ELEM, P>;
    using L = StdAuxLayout<ELEM, P, SKEW>;
    using LN = Lines<ELEM, P>;
    constexpr int E = T::E2, TS = T::TS, D = Sh::D, NQ = Sh::NQ, NF = Sh::NF, N1 = Sh::N1;
    using SF = SumFact<ELEM, P, TS, D * E>;
    extern __shared__ double sm[];
    const DevParams& p = tp.base;
    const int tl = threadIdx.x;
    const int R = en.R;
    const long long nl = (long long)p.N_e * R;
    const long long g0 = (long long)element_of_block(blockIdx.x, gridDim.x, tp.xcd_remap) * E;
    const int nvalid = (int)(nl - g0 < E ? nl - g0 : E);
    double *suq = sm + L::uq, *sup = sm + L::up, *sum = sm + L::um, *sDu = sm + L::a, *sg = sm + L::b,
           *sr = sm + L::r, *stf = sm + L::tf, *sD1 = sm + L::d1;
    constexpr double h = SKEW ? 0.5 : 1.0;

    for (int t = tl; t < D * N1 * N1; t += TS) sD1[t] = tp.D1[t];
    for (int t = tl; t < E * NQ; t += TS) suq[t] = t < nvalid * NQ ? p.uq[(size_t)g0 * NQ + t] : 0.0;
    for (int t = tl; t < E * NF; t += TS) {
        const int s = t / NF, j = t - s * NF;
        const long long g = g0 + s;
        double up = 0.0, um = 0.0;
        if (s < nvalid) {
            int jn;
            const long long gn = neighbour(p, R, g, NF, j, jn);
            up = p.uf[(size_t)gn * NF + jn];
            if (!SKEW) um = p.uf[(size_t)g * NF + j];
        }
        sup[t] = up;
        if (!SKEW) sum[t] = um;
    }
    __syncthreads();
    if constexpr (SKEW) {  // D1_l u_q along the lines of each direction l
        for (int i = tl; i < NQ; i += TS) {
#pragma unroll
            for (int l = 0; l < D; ++l) {
                const int sl = LN::stride(l), a = LN::coord(i, l), base = i - a * sl;
                const double* drow = sD1 + (l * N1 + a) * N1;
                double acc[E];
#pragma unroll
                for (int s = 0; s < E; ++s) acc[s] = 0.0;
#pragma unroll
                for (int bb = 0; bb < N1; ++bb) {
                    const double dv = drow[bb];
#pragma unroll
                    for (int s = 0; s < E; ++s) acc[s] += dv * suq[s * NQ + base + bb * sl];
                }
#pragma unroll
                for (int s = 0; s < E; ++s) sDu[(s * D + l) * NQ + i] = acc[s];
            }
        }
        __syncthreads();
    }
    // facet: 1/2 B nJf_n u^+ (skew) or 1/2 B nJf_n (u^- + u^+)
    for (int t = tl; t < E * NF; t += TS) {
        const int s = t / NF, j = t - s * NF;
        const long long k = (g0 + (s < nvalid ? s : 0)) / R;
        const double* nJ = p.nJf + (size_t)k * D * NF + D * j;
        const double x = 0.5 * p.B[j] * (SKEW ? sup[t] : sup[t] + sum[t]);  // sum unread when SKEW
#pragma unroll
        for (int n = 0; n < D; ++n) stf[(s * D + n) * NF + j] = nJ[n] * x;
    }
    // one direction n at a time (three g_ln arrays in LDS instead of nine):
    //   g_ln = h W Lt_ln u_q;  r_n = h W sum_l Lt_ln (D1_l u_q) - sum_l D1_l' g_ln + R' tf_n
    // node i's sums stay in registers across the passes
    static_assert(NQ <= TS, "one volume node per thread");
    double acc[E * D];
#pragma unroll
    for (int c = 0; c < E * D; ++c) acc[c] = 0.0;
#pragma unroll 1
    for (int n = 0; n < D; ++n) {
        for (int t = tl; t < E * NQ; t += TS) {
            const int s = t / NQ, i = t - s * NQ;
            const long long k = (g0 + (s < nvalid ? s : 0)) / R;
            const double* Lk = p.Lam + (size_t)k * NQ * D * D;
            const double hw = h * p.W[i];
            const double x = suq[t];
            double dterm = 0.0;
#pragma unroll
            for (int l = 0; l < D; ++l) {
                const double lam = Lk[i + NQ * (l + D * n)];
                sg[(s * D + l) * NQ + i] = hw * lam * x;
                if constexpr (SKEW) dterm += lam * sDu[(s * D + l) * NQ + i];
            }
            sr[(s * D + n) * NQ + i] = hw * dterm;
        }
        __syncthreads();
        if (tl < NQ) {
            const int i = tl;
#pragma unroll 1
            for (int l = 0; l < D; ++l) {
                const int sl = LN::stride(l), a = LN::coord(i, l), base = i - a * sl;
#pragma unroll
                for (int bb = 0; bb < N1; ++bb) {
                    const double dv = sD1[(l * N1 + bb) * N1 + a];
#pragma unroll
                    for (int s = 0; s < E; ++s) acc[s * D + n] -= dv * sg[(s * D + l) * NQ + base + bb * sl];
                }
            }
        }
        __syncthreads();
    }
    if (tl < NQ) {
        const int i = tl;
        ell_apply<E * D, NF>(tp.rt_idx, tp.rt_val, tp.wrt, NQ, i, stf, acc);
#pragma unroll
        for (int c = 0; c < E * D; ++c) sr[c * NQ + i] += acc[c];
    }
    __syncthreads();
    // three-direction mass solve: q = V' (W/J o V (V' r)); then q_q = V q
    double* sZ = sDu;
    double* sy = sDu + E * D * Sh::NZ;
    double* sW = sg;
    SF::Vt(tp, tl, sr, sy, sZ, sW);
    SF::V(tp, tl, sy, sr, sZ, sW);
    for (int t = tl; t < E * D * NQ; t += TS) {
        const int c = t / NQ, i = t - c * NQ, s = c / D;
        const long long k = (g0 + (s < nvalid ? s : 0)) / R;
        sr[t] *= p.W[i] / p.Jq[(size_t)k * NQ + i];
    }
    __syncthreads();
    SF::Vt(tp, tl, sr, sy, sZ, sW);
    SF::V(tp, tl, sy, sr, sZ, sW);
    for (int t = tl; t < nvalid * D * NQ; t += TS) p.qq[(size_t)g0 * D * NQ + t] = sr[t];
    for (int j = tl; j < NF; j += TS) {  // q_f = R q_q
        double acc[E * D];
#pragma unroll
        for (int c = 0; c < E * D; ++c) acc[c] = 0.0;
        ell_apply<E * D, NQ>(tp.r_idx, tp.r_val, tp.wr, NF, j, sr, acc);
#pragma unroll
        for (int c = 0; c < E * D; ++c)
            if (c < nvalid * D) p.qf[((size_t)g0 * D + c) * NF + j] = acc[c];
    }
}

// ----------------------------------------------------------------------------- time derivative
// dudt = M^{-1} V' [ sum_n VT_n(f_n) - R' (B J_f f*) ],  f_n = a_n u - b q_n,
// f* = a.n {u} + halfλ |a.n| [u] - b {q}.n (LF; BR1 when second order).
template <int ELEM, int P>
struct StdTdLayout {
    using Sh = Shape<ELEM, P>;
    static constexpr int E = StdTune<ELEM, P>::E3, D = Sh::D;
    static constexpr int F = 2 * D * Sh::NQ;  // f_n and G_m, then SumFact Z, W, y, tmp
    static_assert(F >= Sh::NZ + cmax(Sh::NW, 1) + Sh::NP + Sh::NQ, "mass-solve scratch");
    static constexpr int f = 0, r = f + E * F, t = r + E * Sh::NQ, d1 = t + E * Sh::NF;
    static constexpr int size = d1 + D * Sh::N1 * Sh::N1;
};

template <int ELEM, int P, bool SKEW, bool SECOND>
__global__ void __launch_bounds__((StdTune<ELEM, P>::TS))
k_std_td_tensor(TensorParams tp, Ensemble en, double* __restrict__ dudt, RKStage rk) {
    using Sh = Shape<ELEM, P>;
    using T = StdTune<ELEM, P>;
    using L = StdTdLayout<ELEM, P>;
    constexpr int E = T::E3, TS = T::TS, D = Sh::D, NP = Sh::NP, NQ = Sh::NQ, NF = Sh::NF;
    using SF = SumFact<ELEM, P, TS, E>;
    extern __shared__ double sm[];
    const DevParams& p = tp.base;
    const int tl = threadIdx.x;
    const int R = en.R;
    const long long nl = (long long)p.N_e * R;
    const long long g0 = (long long)element_of_block(blockIdx.x, gridDim.x, tp.xcd_remap) * E;
    const int nvalid = (int)(nl - g0 < E ? nl - g0 : E);
    double *sf = sm + L::f, *sG = sf + E * D * NQ, *sr = sm + L::r, *st = sm + L::t, *sD1 = sm + L::d1;
    constexpr double h = SKEW ? 0.5 : 1.0;
    using LN = Lines<ELEM, P>;
    constexpr int N1 = Sh::N1;
    for (int t = tl; t < D * N1 * N1; t += TS) sD1[t] = tp.D1[t];

    // volume fluxes f_n and G_l = h W sum_n Lt_ln f_n
    for (int t = tl; t < E * NQ; t += TS) {
        const int s = t / NQ, i = t - s * NQ;
        const long long g = g0 + (s < nvalid ? s : 0);
        const long long k = g / R;
        const double bm = SECOND ? member_b(p, en, g) : 0.0;
        const bool tan = SECOND && en.tangent && ((g % R) & 1) && s < nvalid;
        const double x = s < nvalid ? p.uq[(size_t)g * NQ + i] : 0.0;
        double f[D];
#pragma unroll
        for (int n = 0; n < D; ++n) {
            f[n] = p.law.a[n] * x;
            if constexpr (SECOND) {
                f[n] -= bm * (s < nvalid ? p.qq[((size_t)g * D + n) * NQ + i] : 0.0);
                if (tan) f[n] -= p.qq[((size_t)(g - 1) * D + n) * NQ + i];  // dR/db(u) of the pair's u
            }
            sf[(s * D + n) * NQ + i] = f[n];
        }
        const double* Lk = p.Lam + (size_t)k * NQ * D * D;
        const double hw = h * p.W[i];
#pragma unroll
        for (int m = 0; m < D; ++m) {
            double a = 0.0;
#pragma unroll
            for (int n = 0; n < D; ++n) a += Lk[i + NQ * (m + D * n)] * f[n];
            sG[(s * D + m) * NQ + i] = hw * a;
        }
    }
    // facet values: 1/2 B sum_n nJf_n (R f_n) - B J_f f*
    for (int t = tl; t < E * NF; t += TS) {
        const int s = t / NF, j = t - s * NF;
        double val = 0.0;
        if (s < nvalid) {
            const long long g = g0 + s;
            const long long k = g / R;
            int jn;
            const long long gn = neighbour(p, R, g, NF, j, jn);
            const double um = p.uf[(size_t)g * NF + j], up = p.uf[(size_t)gn * NF + jn];
            const double jf = p.Jf[(size_t)k * NF + j];
            const double* nJ = p.nJf + (size_t)k * D * NF + D * j;
            double an = 0.0;
#pragma unroll
            for (int m = 0; m < D; ++m) an += p.law.a[m] * nJ[m];
            an /= jf;  // a . n
            double fs = 0.5 * an * (um + up);
            if (p.law.inviscid_flux == 1) fs += p.law.half_lambda * fabs(an) * (um - up);
            double surf = SKEW ? an * jf * um : 0.0;  // sum_n nJf_n a_n u^-
            if constexpr (SECOND) {
                const double bm = member_b(p, en, g);
                double qn = 0.0, qm = 0.0;
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    const double qi = p.qf[((size_t)g * D + m) * NF + j];
                    const double qo = p.qf[((size_t)gn * D + m) * NF + jn];
                    qn += (qi + qo) * nJ[m];
                    qm += qi * nJ[m];
                }
                qn *= bm;
                qm *= bm;
                if (en.tangent && ((g % R) & 1)) {  // + the pair's u at unit b
#pragma unroll
                    for (int m = 0; m < D; ++m) {
                        const double qi = p.qf[((size_t)(g - 1) * D + m) * NF + j];
                        const double qo = p.qf[((size_t)(gn - 1) * D + m) * NF + jn];
                        qn += (qi + qo) * nJ[m];
                        qm += qi * nJ[m];
                    }
                }
                fs -= 0.5 * qn / jf;
                if (SKEW) surf -= qm;
            }
            val = 0.5 * p.B[j] * surf - p.B[j] * jf * fs;
        }
        st[t] = val;
    }
    __syncthreads();
    // r = sum_l D1_l' G_l - h W sum_{l,n} Lt_ln (D1_l f_n) + R' t
    for (int i = tl; i < NQ; i += TS) {
        double acc[E];
#pragma unroll
        for (int s = 0; s < E; ++s) acc[s] = 0.0;
#pragma unroll 1
        for (int l = 0; l < D; ++l) {
            const int sl = LN::stride(l), a = LN::coord(i, l), base = i - a * sl;
#pragma unroll
            for (int bb = 0; bb < N1; ++bb) {
                const double dv = sD1[(l * N1 + bb) * N1 + a];
#pragma unroll
                for (int s = 0; s < E; ++s) acc[s] += dv * sG[(s * D + l) * NQ + base + bb * sl];
            }
            if constexpr (SKEW) {
                double y[E][D];
#pragma unroll
                for (int s = 0; s < E; ++s)
#pragma unroll
                    for (int n = 0; n < D; ++n) y[s][n] = 0.0;
                const double* drow = sD1 + (l * N1 + a) * N1;
#pragma unroll
                for (int bb = 0; bb < N1; ++bb) {
                    const double dv = drow[bb];
#pragma unroll
                    for (int s = 0; s < E; ++s)
#pragma unroll
                        for (int n = 0; n < D; ++n) y[s][n] += dv * sf[(s * D + n) * NQ + base + bb * sl];
                }
                const double hw = h * p.W[i];
#pragma unroll
                for (int s = 0; s < E; ++s) {
                    const long long k = (g0 + (s < nvalid ? s : 0)) / R;
                    const double* Lk = p.Lam + (size_t)k * NQ * D * D;
                    double c = 0.0;
#pragma unroll
                    for (int n = 0; n < D; ++n) c += Lk[i + NQ * (l + D * n)] * y[s][n];
                    acc[s] -= hw * c;
                }
            }
        }
        ell_apply<E, NF>(tp.rt_idx, tp.rt_val, tp.wrt, NQ, i, st, acc);
#pragma unroll
        for (int s = 0; s < E; ++s) sr[s * NQ + i] = acc[s];
    }
    __syncthreads();
    // mass solve: o = V' (W/J o V (V' r))
    double* sZ = sf;
    double* sW = sZ + E * Sh::NZ;
    double* sy = sW + E * cmax(Sh::NW, 1);
    double* sx = sy + E * NP;
    SF::Vt(tp, tl, sr, sy, sZ, sW);
    SF::V(tp, tl, sy, sx, sZ, sW);
    for (int t = tl; t < E * NQ; t += TS) {
        const int s = t / NQ, i = t - s * NQ;
        const long long k = (g0 + (s < nvalid ? s : 0)) / R;
        sx[t] *= p.W[i] / p.Jq[(size_t)k * NQ + i];
    }
    __syncthreads();
    SF::Vt(tp, tl, sx, sy, sZ, sW);
    for (int t = tl; t < nvalid * NP; t += TS) {
        const int s = t / NP, j = t - s * NP;
        const size_t o = state_row(g0 + s, R, p.N_e) * NP + j;
        if (!rk.active) {
            dudt[o] = sy[t];
        } else {
            const double kt = (rk.A == 0.0 ? 0.0 : rk.A * rk.tmp[o]) + (rk.dt_e ? rk.dt_e[o / NP] : rk.dt) * sy[t];
            rk.tmp[o] = kt;
            rk.u[o] += rk.B * kt;
        }
    }
}

// ----------------------------------------------------------------------------- set-up
// In place: Lambda_mn -> Lt_ln = sum_m c_lm Lambda_mn at every node of every
// element (c: N_q x d x d, c[(i d + l) d + m]).
__global__ void k_collapse_metric(double* __restrict__ Lam, const double* __restrict__ c, int NQ, int D,
                                  long long n_nodes) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_nodes) return;
    const long long k = t / NQ;
    const int i = (int)(t - k * NQ);
    double* Lk = Lam + (size_t)k * NQ * D * D;
    double in[9], out[9];
    for (int m = 0; m < D; ++m)
        for (int n = 0; n < D; ++n) in[m + 3 * n] = Lk[i + NQ * (m + D * n)];
    for (int l = 0; l < D; ++l)
        for (int n = 0; n < D; ++n) {
            double s = 0.0;
            for (int m = 0; m < D; ++m) s += c[(i * D + l) * D + m] * in[m + 3 * n];
            out[l + 3 * n] = s;
        }
    for (int l = 0; l < D; ++l)
        for (int n = 0; n < D; ++n) Lk[i + NQ * (l + D * n)] = out[l + 3 * n];
}

hipError_t collapse_metric(double* Lam, const double* c, int NQ, int D, int N_e, hipStream_t s) {
    const long long n = (long long)NQ * N_e;
    hipLaunchKernelGGL(k_collapse_metric, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Lam, c, NQ, D, n);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- launch
template <int ELEM, int P>
size_t std_tensor_lds(int which) {
    if (which == 0) return sizeof(double) * StdNodalLayout<ELEM, P>::size;
    if (which == 1) return sizeof(double) * cmax(StdAuxLayout<ELEM, P, true>::size, StdAuxLayout<ELEM, P, false>::size);
    if (which == 3) return sizeof(double) * StdAuxLayout<ELEM, P, true>::size;
    return sizeof(double) * StdTdLayout<ELEM, P>::size;
}

template <int ELEM, int P, bool SKEW, bool SECOND>
hipError_t launch_std_tensor_t(const TensorParams& tp, const Ensemble& en, const double* u, double* dudt,
                               const RKStage& rk, hipStream_t s, int which) {
    using T = StdTune<ELEM, P>;
    static bool attr_set = false;
    const size_t l0 = std_tensor_lds<ELEM, P>(0), l1 = std_tensor_lds<ELEM, P>(SKEW ? 3 : 1),
                 l2 = std_tensor_lds<ELEM, P>(2);
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_std_nodal_tensor<ELEM, P>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)l0);
        if (e == hipSuccess && SECOND)
            e = hipFuncSetAttribute((const void*)k_std_aux_tensor<ELEM, P, SKEW>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)l1);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_std_td_tensor<ELEM, P, SKEW, SECOND>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)l2);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const long long nl = (long long)tp.base.N_e * en.R;
    if (which & 1) {
        const int nb = (int)((nl + T::E1 - 1) / T::E1);
        hipLaunchKernelGGL((k_std_nodal_tensor<ELEM, P>), dim3(nb), dim3(T::TS), l0, s, tp, en, u);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (which & 2) {
        if (SECOND) {
            const int nb = (int)((nl + T::E2 - 1) / T::E2);
            hipLaunchKernelGGL((k_std_aux_tensor<ELEM, P, SKEW>), dim3(nb), dim3(T::TS), l1, s, tp, en);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        const int nb = (int)((nl + T::E3 - 1) / T::E3);
        hipLaunchKernelGGL((k_std_td_tensor<ELEM, P, SKEW, SECOND>), dim3(nb), dim3(T::TS), l2, s, tp, en, dudt, rk);
        return hipGetLastError();
    }
    return hipSuccess;
}

template <int ELEM, int P>
hipError_t launch_std_tensor_p(const TensorParams& tp, const Ensemble& en, const double* u, double* dudt,
                               const RKStage& rk, hipStream_t s, int which) {
    const bool skew = tp.base.skew, second = tp.base.second_order;
    if (skew && second) return launch_std_tensor_t<ELEM, P, true, true>(tp, en, u, dudt, rk, s, which);
    if (skew) return launch_std_tensor_t<ELEM, P, true, false>(tp, en, u, dudt, rk, s, which);
    if (second) return launch_std_tensor_t<ELEM, P, false, true>(tp, en, u, dudt, rk, s, which);
    return launch_std_tensor_t<ELEM, P, false, false>(tp, en, u, dudt, rk, s, which);
}

#define SSE_STD_CASES                                                                               \
    SSE_CASE(ELEM_TRI, 2) SSE_CASE(ELEM_TRI, 3) SSE_CASE(ELEM_TRI, 4) SSE_CASE(ELEM_TRI, 5)          \
    SSE_CASE(ELEM_TRI, 6) SSE_CASE(ELEM_TET, 2) SSE_CASE(ELEM_TET, 3) SSE_CASE(ELEM_TET, 4)          \
    SSE_CASE(ELEM_TET, 5)

// largest LDS of the three kernels, or 0 if (elem, P) has no compiled kernels
size_t std_tensor_config(int elem, int P) {
#define SSE_CASE(E_, PP)                                                                              \
    if (elem == E_ && P == PP)                                                                        \
        return cmax(cmax((int)std_tensor_lds<E_, PP>(0), (int)std_tensor_lds<E_, PP>(1)),             \
                    (int)std_tensor_lds<E_, PP>(2));
    SSE_STD_CASES
#undef SSE_CASE
    return 0;
}

hipError_t launch_std_tensor(const TensorParams& tp, int R, const double* b_member, int tangent, const double* u,
                             double* dudt, const RKStage& rk, hipStream_t s, int which) {
    if (tangent && ((R & 1) || !tp.base.second_order)) return hipErrorInvalidValue;
    const Ensemble en{R, b_member, tangent};
#define SSE_CASE(E_, PP) \
    if (tp.elem == E_ && tp.P == PP) return launch_std_tensor_p<E_, PP>(tp, en, u, dudt, rk, s, which);
    SSE_STD_CASES
#undef SSE_CASE
    return hipErrorInvalidValue;
}

}  // namespace sse
