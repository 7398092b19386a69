// Nodal values (entropy projection) of the dense multidimensional schemes on the matrix
// cores: the V, V' and R products of 16 elements at a time as FP64 MFMA tiles.
//
// The products of flux_differencing_form.jl:253-292 (u_q = V u, w = V'(WJ c2e(u_q)),
// the weight-adjusted solve M^{-1} = V' (W/J V .), w_q = V w, u_q = e2c(w_q),
// u_f = e2c(R w_q)) are, over a workgroup's 16 elements, matrix products
// C (M x 80) = A (M x K) B (K x 80) whose 80 columns are the 5 components of the 16
// elements (64 columns in 2D). Each wavefront computes 16 x 16 tiles of C with v_mfma_f64_16x16x4f64
// (one K step of 4 per instruction), two tiles at a time so that two accumulation
// chains interleave. The operator (A) fragments are read from global memory (the same
// small operator for every workgroup, so they stay in L2), a group of four K steps ahead;
// the element data (B) and the results live in LDS (the facet rows reuse the space of the
// modal coefficients, a chunk at a time, so that two workgroups fit a compute unit). The
// entropy-variable maps run per (element, node) between the products, as in
// k_nodal_values_generic.
//
// Lane layout of v_mfma_f64_16x16x4f64 (checked on the device): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16], and receives D[4 r + l / 16][l % 16], r = 0..3.
#include <hip/hip_runtime.h>

#include "sse_kernels.h"
#include "sse_physics.h"

namespace sse {

namespace {

typedef double dvec4 __attribute__((ext_vector_type(4)));

constexpr int MF_E = 16;  // elements per workgroup

__host__ __device__ constexpr int round16(int n) { return (n + 15) / 16 * 16; }
__host__ __device__ constexpr int round4(int n) { return (n + 3) / 4 * 4; }

// C[0, Mp) x [0, NCOL) = A B, A (M x K) through the accessor a(i, k) (zero outside the
// operator), B (Kp x NCOL) and C row-major in LDS with leading dimension NCOL; Mp a
// multiple of 16, Kp of 4. Wave w takes the tiles w, w + nw, ... in pairs; the K steps go
// in groups of four, the operator fragments of the next group requested before the
// MFMAs of the current one.
template <int NCOL, class AFn>
__device__ __forceinline__ void mm_tiles(const AFn& a, int Mp, int Kp, const double* B, double* C) {
    constexpr int NT = NCOL / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int units = (Mp / 16) * NT, steps = Kp / 4, groups = (steps + 3) / 4;
    for (int u0 = wave; u0 < units; u0 += 2 * nw) {
        const int u1 = u0 + nw;
        const bool two = u1 < units;  // uniform over the wavefront
        const int rt0 = u0 / NT, ct0 = u0 - rt0 * NT;
        const int rt1 = two ? u1 / NT : rt0, ct1 = two ? u1 - (u1 / NT) * NT : ct0;
        const int i0 = 16 * rt0 + li, i1 = 16 * rt1 + li;
        auto fetch = [&](int g, double* f0, double* f1) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = 4 * (4 * g + q) + lk;
                f0[q] = a(i0, k);
                f1[q] = a(i1, k);
            }
        };
        dvec4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
        double c0[4], c1[4], n0[4], n1[4];
        fetch(0, c0, c1);
        for (int g = 0; g < groups; ++g) {
            if (g + 1 < groups) fetch(g + 1, n0, n1);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int st = 4 * g + q;
                if (st < steps) {
                    const int k = 4 * st + lk;
                    const double b0 = B[k * NCOL + 16 * ct0 + li], b1 = B[k * NCOL + 16 * ct1 + li];
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(c0[q], b0, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(c1[q], b1, acc1, 0, 0, 0);
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                c0[q] = n0[q];
                c1[q] = n1[q];
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) C[(16 * rt0 + 4 * r + lk) * NCOL + 16 * ct0 + li] = acc0[r];
        if (two) {
#pragma unroll
            for (int r = 0; r < 4; ++r) C[(16 * rt1 + 4 * r + lk) * NCOL + 16 * ct1 + li] = acc1[r];
        }
    }
}

template <int D, int LAW>
__global__ void __launch_bounds__(256) k_nodal_values_mfma(DevParams p, const double* __restrict__ u) {
    using Ph = Physics<D, LAW>;
    constexpr int NC = Ph::NC, NCOL = MF_E * NC;
    static_assert(NCOL % 16 == 0, "16 elements of NC components fill whole column tiles");
    extern __shared__ double sm[];
    const int N_p = p.N_p, N_q = p.N_q, N_f = p.N_f, N_e = p.N_e;
    const int NPp = round16(N_p), NQp = round16(N_q);
    double* X = sm;                // NPp x NCOL: modal coefficients, then NPp facet rows at a time
    double* Y = X + NPp * NCOL;    // NQp x NCOL: volume nodes
    const int k0 = blockIdx.x * MF_E;
    const int tid = threadIdx.x, nt = blockDim.x;

    // u of the 16 elements, read contiguously ([k][c][j]), into X[j][e NC + c]; rows past
    // N_p and elements past N_e are zero
    for (int t = tid; t < NPp * NCOL; t += nt) X[t] = 0.0;
    __syncthreads();
    for (int t = tid; t < MF_E * NC * N_p; t += nt) {
        const int e = t / (NC * N_p), r = t - e * NC * N_p, c = r / N_p, j = r - c * N_p;
        if (k0 + e < N_e) X[j * NCOL + e * NC + c] = u[(size_t)(k0 + e) * NC * N_p + r];
    }
    __syncthreads();

    auto aV = [&](int i, int k) { return (i < N_q && k < N_p) ? p.V[i + (size_t)N_q * k] : 0.0; };
    auto aVt = [&](int i, int k) { return (i < N_p && k < N_q) ? p.Vt[i + (size_t)N_p * k] : 0.0; };

    bool ok = true;

    mm_tiles<NCOL>(aV, NQp, round4(N_p), X, Y);  // u_q = V u
    __syncthreads();
    for (int t = tid; t < MF_E * N_q; t += nt) {  // WJ c2e(u_q)
        const int e = t / N_q, i = t - e * N_q, k = k0 + e;
        double ub[NC], wb[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) ub[c] = Y[i * NCOL + e * NC + c];
        const bool good = Ph::c2e(p.law, ub, wb);
        if (k < N_e) ok &= good;
        const double wj = k < N_e ? p.W[i] * p.Jq[(size_t)k * N_q + i] : 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) Y[i * NCOL + e * NC + c] = k < N_e ? wb[c] * wj : 0.0;
    }
    __syncthreads();
    mm_tiles<NCOL>(aVt, NPp, round4(N_q), Y, X);  // w = V' (WJ w_q)
    __syncthreads();
    mm_tiles<NCOL>(aV, NQp, round4(N_p), X, Y);  // V w
    __syncthreads();
    for (int t = tid; t < MF_E * N_q; t += nt) {  // J^{-1} W (V w)
        const int e = t / N_q, i = t - e * N_q, k = k0 + e;
        const double jw = k < N_e ? p.W[i] / p.Jq[(size_t)k * N_q + i] : 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) Y[i * NCOL + e * NC + c] *= jw;
    }
    __syncthreads();
    mm_tiles<NCOL>(aVt, NPp, round4(N_q), Y, X);  // w = V' (...)
    __syncthreads();
    mm_tiles<NCOL>(aV, NQp, round4(N_p), X, Y);  // w_q = V w
    __syncthreads();
    // u_q = e2c(w_q); then w_f = R w_q, NPp facet rows at a time in the space of X
    for (int t = tid; t < MF_E * N_q; t += nt) {
        const int e = t / N_q, i = t - e * N_q, k = k0 + e;
        if (k >= N_e) continue;
        double ub[NC], wb[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) wb[c] = Y[i * NCOL + e * NC + c];
        ok &= Ph::e2c(p.law, wb, ub);
        double* uqk = p.uq + (size_t)k * N_q * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) uqk[i + N_q * c] = ub[c];
    }
    for (int f0 = 0; f0 < N_f; f0 += NPp) {
        const int rows = min(NPp, round16(N_f - f0));
        auto aR = [&](int i, int k) {
            return (f0 + i < N_f && k < N_q) ? p.Rd[f0 + i + (size_t)N_f * k] : 0.0;
        };
        mm_tiles<NCOL>(aR, rows, round4(N_q), Y, X);
        __syncthreads();
        const int nf = min(NPp, N_f - f0);
        for (int t = tid; t < MF_E * nf; t += nt) {  // u_f = e2c(w_f)
            const int e = t / nf, jl = t - e * nf, k = k0 + e;
            if (k >= N_e) continue;
            double ub[NC], wb[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) wb[c] = X[jl * NCOL + e * NC + c];
            ok &= Ph::e2c(p.law, wb, ub);
            double* ufk = p.uf + (size_t)k * N_f * NC;
#pragma unroll
            for (int c = 0; c < NC; ++c) ufk[f0 + jl + N_f * c] = ub[c];
        }
        __syncthreads();
    }
    if (!ok) atomicOr(p.status, 1);
}

}  // namespace

size_t lds_nodal_values_mfma(const DevParams& p) {
    return sizeof(double) * (size_t)(round16(p.N_p) + round16(p.N_q)) * MF_E * p.N_c;
}

// Euler in 2D or 3D with the entropy projection and a modal V only
hipError_t launch_nodal_values_mfma(const DevParams& p, const double* u, hipStream_t s) {
    if (p.N_c != p.d + 2 || !p.project || !p.has_V || !p.Rd) return hipErrorInvalidValue;
    const size_t lds = lds_nodal_values_mfma(p);
    const dim3 grid((p.N_e + MF_E - 1) / MF_E), block(256);
    switch (p.d) {
        case 2: {
            const hipError_t e = hipFuncSetAttribute((const void*)k_nodal_values_mfma<2, LAW_EULER>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL((k_nodal_values_mfma<2, LAW_EULER>), grid, block, lds, s, p, u);
            break;
        }
        case 3: {
            const hipError_t e = hipFuncSetAttribute((const void*)k_nodal_values_mfma<3, LAW_EULER>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL((k_nodal_values_mfma<3, LAW_EULER>), grid, block, lds, s, p, u);
            break;
        }
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sse
