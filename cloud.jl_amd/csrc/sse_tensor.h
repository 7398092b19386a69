// sse_tensor.h -- compile-time element shapes, launch shapes and sum-factorised
// V / V' shared by the tensor-product kernels (ModalTensor Tri/Tet).
#pragma once

#include <hip/hip_runtime.h>

#include "sse_kernels.h"

namespace sse {

enum { ELEM_TRI = 2, ELEM_TET = 3 };

// widest owner-reduction table and R' table the tensor kernels accept (checked on the host)
constexpr int TENSOR_MAXW = 16;
constexpr int TENSOR_MAXR = 10;

template <int ELEM, int P>
struct Shape;

template <int P>
struct Shape<ELEM_TRI, P> {
    static constexpr int D = 2, N1 = P + 1, NQ = N1 * N1, NFAC = 3, NPF = N1, NF = NFAC * NPF;
    static constexpr int NP = (P + 1) * (P + 2) / 2, NC = 4, NH = D + 5;
    static constexpr int NZ = (P + 1) * N1;  // Z[b1][a2]
    static constexpr int NW = 0;
};

template <int P>
struct Shape<ELEM_TET, P> {
    static constexpr int D = 3, N1 = P + 1, NQ = N1 * N1 * N1, NFAC = 4, NPF = N1 * N1, NF = NFAC * NPF;
    static constexpr int NP = (P + 1) * (P + 2) * (P + 3) / 6, NC = 5, NH = D + 5;
    static constexpr int NPAIR12 = (P + 1) * (P + 2) / 2;
    static constexpr int NZ = NPAIR12 * N1;  // Z[q(b1,b2)][a3]
    static constexpr int NW = (P + 1) * N1 * N1;  // W[b1][a2][a3]
};

// Launch shape of the two kernels. K2 (time_derivative): team size TS, EPB
// elements per block, R pairs per thread per pass. K1 (nodal_values): TS1, EPB1.
// V selects among variants of the same (element, degree) compiled for tuning.
template <int ELEM, int P, int V>
struct Tune;
#define SSE_TUNE(E, PP, VV, ts, epb, r, ts1, epb1)                                       \
    template <>                                                                        \
    struct Tune<E, PP, VV> {                                                           \
        static constexpr int TS = ts, EPB = epb, R = r, TS1 = ts1, EPB1 = epb1;         \
    };
SSE_TUNE(ELEM_TRI, 2, 0, 64, 2, 1, 64, 2)
SSE_TUNE(ELEM_TRI, 3, 0, 64, 2, 2, 64, 2)
SSE_TUNE(ELEM_TRI, 4, 0, 64, 2, 3, 64, 2)
SSE_TUNE(ELEM_TRI, 5, 0, 64, 2, 5, 64, 2)
SSE_TUNE(ELEM_TRI, 6, 0, 64, 2, 7, 64, 2)
SSE_TUNE(ELEM_TET, 2, 0, 64, 2, 3, 64, 2)
SSE_TUNE(ELEM_TET, 3, 0, 128, 1, 3, 64, 2)
SSE_TUNE(ELEM_TET, 4, 0, 128, 1, 2, 128, 2)
SSE_TUNE(ELEM_TET, 4, 1, 128, 1, 2, 64, 1)
SSE_TUNE(ELEM_TET, 4, 2, 128, 1, 2, 128, 2)
SSE_TUNE(ELEM_TET, 4, 3, 128, 2, 2, 128, 4)
SSE_TUNE(ELEM_TET, 4, 4, 128, 1, 1, 64, 2)
SSE_TUNE(ELEM_TET, 5, 0, 256, 1, 2, 256, 1)
SSE_TUNE(ELEM_TET, 5, 1, 256, 1, 2, 128, 2)
SSE_TUNE(ELEM_TET, 3, 1, 64, 2, 3, 64, 4)
SSE_TUNE(ELEM_TRI, 4, 1, 64, 2, 3, 64, 4)
#undef SSE_TUNE

// Line-structured operator tables (host-built from the descriptor's S, C and R).
struct StructTables {
    const double* SG;    // D x N1 x D x NQ: S_m[i, j(i, f, b)], index ((f * N1 + b) * D + m) * NQ + i
    const double* SGh;   // -SG / 2: the forward's volume pairs contract with -a / 2
    const int* ci_i;     // CH x MAXITEM: volume nodes of each column chunk (-1 pad)
    const double* ci_v;  // CH x MAXITEM
    const int* ci_j;     // MAXITEM: facet node of each chunk (-1 = no chunk)
    int ch, n_items;
    const double* uf_ext;  // exterior facet states [k][e][j]
    // deterministic mode: the rows of C (RW x NQ, facet node and entry, -1 pad) and the
    // first chunk of each facet node (NF + 1 prefix offsets into the chunks)
    const int* cr_j;
    const double* cr_v;
    const int* ci_first;
    int rw;
    int det;  // 1: every sum in a fixed order (no atomic accumulation)
};

// q index of (b1, b2), b1 + b2 <= P, enumerated with b1 outer: also the mode
// number of (b1, b2) in 2D.
template <int P>
__device__ __forceinline__ constexpr int q_index(int b1, int b2) {
    return b1 * (P + 1) - (b1 * (b1 - 1)) / 2 + b2;
}

// number of 3D modes before (b1, b2, 0)
template <int P>
__device__ __forceinline__ constexpr int mode_offset3(int b1, int b2) {
    int off = 0;
    for (int i = 0; i < b1; ++i) off += (P - i + 1) * (P - i + 2) / 2;
    for (int j = 0; j < b2; ++j) off += P - b1 - j + 1;
    return off;
}

__device__ __forceinline__ int element_of_block(int b, int nb, int remap) {
    if (!remap || (nb & 7)) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// ----------------------------------------------------------------------------- factors of V in LDS
// The sum-factorisation factors A, B (and C on the Tet) of V are a few KB read by every
// stage of every application; a kernel copies them to LDS once and hands SumFact a
// TensorParams whose VA, VB, VC point there (LDS latency instead of cache latency in the
// factor loops).
template <int ELEM, int P>
struct VFactors {
    static constexpr int N1 = P + 1;
    static constexpr int NA = N1 * (P + 1), NB = N1 * (P + 1) * (P + 1);
    static constexpr int NCF = ELEM == ELEM_TET ? N1 * (P + 1) * (P + 1) * (P + 1) : 0;
    static constexpr int size = NA + NB + NCF;
    // all `nt` threads of the caller copy; a barrier must follow before use
    __device__ static void stage(const TensorParams& tp, double* dst, int t, int nt) {
        for (int i = t; i < NA; i += nt) dst[i] = tp.VA[i];
        for (int i = t; i < NB; i += nt) dst[NA + i] = tp.VB[i];
        if constexpr (NCF > 0)
            for (int i = t; i < NCF; i += nt) dst[NA + NB + i] = tp.VC[i];
    }
    // the same copy in two steps: thread t's entries t, t + nt, ... (of the concatenated
    // factors) into registers, then into LDS; v holds ceil(size / nt) values
    __device__ static void fetch(const TensorParams& tp, double* v, int t, int nt) {
#pragma unroll
        for (int r = 0; r * nt < size; ++r) {
            const int i = t + r * nt;
            v[r] = i < NA ? tp.VA[i] : i < NA + NB ? tp.VB[i - NA] : i < size ? tp.VC[i - NA - NB] : 0.0;
        }
    }
    __device__ static void put(const double* v, double* dst, int t, int nt) {
#pragma unroll
        for (int r = 0; r * nt < size; ++r)
            if (t + r * nt < size) dst[t + r * nt] = v[r];
    }
    __device__ static TensorParams in_lds(const TensorParams& tp, const double* src) {
        TensorParams r = tp;
        r.VA = src;
        r.VB = src + NA;
        r.VC = NCF > 0 ? src + NA + NB : tp.VC;
        return r;
    }
};

// ----------------------------------------------------------------------------- sum factorisation
// All arrays are component-major LDS arrays: x[e * stride + point]. NCV
// components (conserved variables, or several elements or directions side by
// side) go through together, so each factor entry loaded serves all of them.
template <int ELEM, int P, int TS, int NCV = Shape<ELEM, P>::NC>
struct SumFact;

template <int P, int TS, int NCV>
struct SumFact<ELEM_TRI, P, TS, NCV> {
    using Sh = Shape<ELEM_TRI, P>;
    static constexpr int N1 = Sh::N1, NC = NCV, NP = Sh::NP, NQ = Sh::NQ, NZ = Sh::NZ;

    // y[a1,a2] = sum_b1 A[a1,b1] sum_{b2 <= P-b1} B[a2,b1,b2] x[q(b1,b2)]
    __device__ static void V(const TensorParams& tp, int tl, const double* x, double* y, double* Z, double*) {
        for (int o = tl; o < NZ; o += TS) {
            const int b1 = o / N1, a2 = o - b1 * N1;
            double s[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) s[e] = 0.0;
            const int off = q_index<P>(b1, 0);
            for (int b2 = 0; b2 <= P - b1; ++b2) {
                const double bv = tp.VB[a2 + N1 * (b1 + (P + 1) * b2)];
#pragma unroll
                for (int e = 0; e < NC; ++e) s[e] += bv * x[e * NP + off + b2];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) Z[e * NZ + o] = s[e];
        }
        __syncthreads();
        for (int o = tl; o < NQ; o += TS) {
            const int a1 = o / N1, a2 = o - a1 * N1;
            double s[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) s[e] = 0.0;
#pragma unroll
            for (int b1 = 0; b1 <= P; ++b1) {
                const double av = tp.VA[a1 + N1 * b1];
#pragma unroll
                for (int e = 0; e < NC; ++e) s[e] += av * Z[e * NZ + b1 * N1 + a2];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) y[e * NQ + o] = s[e];
        }
        __syncthreads();
    }

    // y[q(b1,b2)] = sum_a2 B[a2,b1,b2] sum_a1 A[a1,b1] x[a1,a2]
    __device__ static void Vt(const TensorParams& tp, int tl, const double* x, double* y, double* Z, double*) {
        for (int o = tl; o < NZ; o += TS) {
            const int b1 = o / N1, a2 = o - b1 * N1;
            double s[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) s[e] = 0.0;
#pragma unroll
            for (int a1 = 0; a1 < N1; ++a1) {
                const double av = tp.VA[a1 + N1 * b1];
#pragma unroll
                for (int e = 0; e < NC; ++e) s[e] += av * x[e * NQ + a1 * N1 + a2];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) Z[e * NZ + o] = s[e];
        }
        __syncthreads();
        for (int m = tl; m < NP; m += TS) {
            int b1 = 0;
            while (q_index<P>(b1 + 1, 0) <= m && b1 < P) ++b1;
            const int b2 = m - q_index<P>(b1, 0);
            double s[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) s[e] = 0.0;
#pragma unroll
            for (int a2 = 0; a2 < N1; ++a2) {
                const double bv = tp.VB[a2 + N1 * (b1 + (P + 1) * b2)];
#pragma unroll
                for (int e = 0; e < NC; ++e) s[e] += bv * Z[e * NZ + b1 * N1 + a2];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) y[e * NP + m] = s[e];
        }
        __syncthreads();
    }
};

template <int P, int TS, int NCV>
struct SumFact<ELEM_TET, P, TS, NCV> {
    using Sh = Shape<ELEM_TET, P>;
    static constexpr int N1 = Sh::N1, NC = NCV, NP = Sh::NP, NQ = Sh::NQ, NZ = Sh::NZ, NW = Sh::NW;
    static constexpr int NN = N1 * N1;

    __device__ static void decode_q(int q, int& b1, int& b2) {
        b1 = 0;
        while (b1 < P && q_index<P>(b1 + 1, 0) <= q) ++b1;
        b2 = q - q_index<P>(b1, 0);
    }

    // y[a1,a2,a3] = sum A[a1,b1] B[a2,b1,b2] C[a3,b1,b2,b3] x[sigma(b1,b2,b3)]
    __device__ static void V(const TensorParams& tp, int tl, const double* x, double* y, double* Z, double* Wt) {
        for (int o = tl; o < NZ; o += TS) {
            const int q = o / N1, a3 = o - q * N1;
            int b1, b2;
            decode_q(q, b1, b2);
            const int off = mode_offset3<P>(b1, b2);
            double s[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) s[e] = 0.0;
            for (int b3 = 0; b3 <= P - b1 - b2; ++b3) {
                const double cv = tp.VC[a3 + N1 * (b1 + (P + 1) * (b2 + (P + 1) * b3))];
#pragma unroll
                for (int e = 0; e < NC; ++e) s[e] += cv * x[e * NP + off + b3];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) Z[e * NZ + o] = s[e];
        }
        __syncthreads();
        for (int o = tl; o < NW; o += TS) {
            const int b1 = o / NN, a23 = o - b1 * NN, a2 = a23 / N1, a3 = a23 - a2 * N1;
            double s[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) s[e] = 0.0;
            const int q0 = q_index<P>(b1, 0);
            for (int b2 = 0; b2 <= P - b1; ++b2) {
                const double bv = tp.VB[a2 + N1 * (b1 + (P + 1) * b2)];
#pragma unroll
                for (int e = 0; e < NC; ++e) s[e] += bv * Z[e * NZ + (q0 + b2) * N1 + a3];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) Wt[e * NW + o] = s[e];
        }
        __syncthreads();
        for (int o = tl; o < NQ; o += TS) {
            const int a1 = o / NN, a23 = o - a1 * NN;
            double s[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) s[e] = 0.0;
#pragma unroll
            for (int b1 = 0; b1 <= P; ++b1) {
                const double av = tp.VA[a1 + N1 * b1];
#pragma unroll
                for (int e = 0; e < NC; ++e) s[e] += av * Wt[e * NW + b1 * NN + a23];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) y[e * NQ + o] = s[e];
        }
        __syncthreads();
    }

    __device__ static void Vt(const TensorParams& tp, int tl, const double* x, double* y, double* Z, double* Wt) {
        for (int o = tl; o < NW; o += TS) {
            const int b1 = o / NN, a23 = o - b1 * NN;
            double s[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) s[e] = 0.0;
#pragma unroll
            for (int a1 = 0; a1 < N1; ++a1) {
                const double av = tp.VA[a1 + N1 * b1];
#pragma unroll
                for (int e = 0; e < NC; ++e) s[e] += av * x[e * NQ + a1 * NN + a23];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) Wt[e * NW + o] = s[e];
        }
        __syncthreads();
        for (int o = tl; o < NZ; o += TS) {
            const int q = o / N1, a3 = o - q * N1;
            int b1, b2;
            decode_q(q, b1, b2);
            double s[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) s[e] = 0.0;
#pragma unroll
            for (int a2 = 0; a2 < N1; ++a2) {
                const double bv = tp.VB[a2 + N1 * (b1 + (P + 1) * b2)];
#pragma unroll
                for (int e = 0; e < NC; ++e) s[e] += bv * Wt[e * NW + b1 * NN + a2 * N1 + a3];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) Z[e * NZ + o] = s[e];
        }
        __syncthreads();
        for (int m = tl; m < NP; m += TS) {  // one thread per mode (b1, b2, b3)
            int b1, b2, b3;
            decode_mode(m, b1, b2, b3);
            const int q = q_index<P>(b1, b2);
            double s[NC];
#pragma unroll
            for (int e = 0; e < NC; ++e) s[e] = 0.0;
#pragma unroll
            for (int a3 = 0; a3 < N1; ++a3) {
                const double cv = tp.VC[a3 + N1 * (b1 + (P + 1) * (b2 + (P + 1) * b3))];
#pragma unroll
                for (int e = 0; e < NC; ++e) s[e] += cv * Z[e * NZ + q * N1 + a3];
            }
#pragma unroll
            for (int e = 0; e < NC; ++e) y[e * NP + m] = s[e];
        }
        __syncthreads();
    }

    // mode number m -> (b1, b2, b3), modes nested b1, b2, b3 (b3 fastest)
    __device__ static void decode_mode(int m, int& b1, int& b2, int& b3) {
        b1 = 0;
        while (b1 < P && m >= (P - b1 + 1) * (P - b1 + 2) / 2) {
            m -= (P - b1 + 1) * (P - b1 + 2) / 2;
            ++b1;
        }
        b2 = 0;
        while (b2 < P - b1 && m >= P - b1 - b2 + 1) {
            m -= P - b1 - b2 + 1;
            ++b2;
        }
        b3 = m;
    }
};

// The weight-adjusted mass solve with the element's kept inverse Mk = V' diag(W/J_k) V
// (NP x NP column-major in global memory): mass_prefetch requests a team's share of Mk
// into registers (issued before a transform, whose time hides the latency), then
// mass_apply_staged writes it to LDS (sM, NP x NP, not aliasing x) and computes
// x = Mk x in place for NC columns of NP modes in LDS: thread t of a team of TS takes
// the items (e, j) = t, t + TS, ..., into registers, then writes them back. Ends with a
// barrier.
template <int NP, int TS>
struct MassPrefetch {
    static constexpr int MI = (NP * NP + TS - 1) / TS;
    double v[MI];
};

template <int NP, int TS>
__device__ __forceinline__ void mass_prefetch(const double* __restrict__ Mk, int t, MassPrefetch<NP, TS>& pre) {
#pragma unroll
    for (int r = 0; r < MassPrefetch<NP, TS>::MI; ++r) {
        const int c = t + r * TS;
        pre.v[r] = c < NP * NP ? Mk[c] : 0.0;
    }
}

template <int NP, int NC, int TS>
__device__ __forceinline__ void mass_apply_staged(const MassPrefetch<NP, TS>& pre, double* sM, double* x, int t) {
#pragma unroll
    for (int r = 0; r < MassPrefetch<NP, TS>::MI; ++r) {
        const int c = t + r * TS;
        if (c < NP * NP) sM[c] = pre.v[r];
    }
    __syncthreads();
    constexpr int IT = (NP * NC + TS - 1) / TS;
    double y[IT];
#pragma unroll
    for (int r = 0; r < IT; ++r) {
        const int it = t + r * TS;
        double s0 = 0.0, s1 = 0.0;
        if (it < NP * NC) {
            const int e = it / NP, j = it - e * NP;
            const double* xe = x + e * NP;
#pragma unroll 4
            for (int i = 0; i + 1 < NP; i += 2) {
                s0 += sM[j + NP * i] * xe[i];
                s1 += sM[j + NP * (i + 1)] * xe[i + 1];
            }
            if constexpr (NP % 2 == 1) s0 += sM[j + NP * (NP - 1)] * xe[NP - 1];
        }
        y[r] = s0 + s1;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < IT; ++r) {
        const int it = t + r * TS;
        if (it < NP * NC) x[it] = y[r];
    }
    __syncthreads();
}

}  // namespace sse
