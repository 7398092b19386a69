// sse_setup.hip -- set-up on the device: geometric factors of curved elements, the
// reference mapping of StandardForm, and the L2 projection of initial data.
//
// These are the steps before the residual path (they run once per problem), done
// per element with one workgroup per element: the element's mapping nodes are
// staged in LDS, the small dense reference operators (column-major, read through
// the L1/L2 caches, one operator entry feeding every component of a node) are
// applied as row gathers, and the pointwise metric algebra runs in registers.
//
//   k_geometric_factors   GeometricFactors (src/SpatialDiscretizations/mesh.jl:229-282
//                         exact metrics, :284-339 conservative-curl metrics on
//                         triangles, :410-506 on tetrahedra) and project_jacobian!
//                         (src/SpatialDiscretizations/SpatialDiscretizations.jl:310-317)
//   k_reference_mapping   apply_reference_mapping (SpatialDiscretizations.jl:395-411)
//   k_project_l2          project_function with a modal V (src/Solvers/Solvers.jl:396-417):
//                         Cholesky of V' W J V in LDS, two triangular solves
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/sse_hip.h"
#include "sse_physics.h"

namespace sse {
int set_global_error(int code, const std::string& msg);
}

namespace {

constexpr int SETUP_THREADS = 256;

struct GeoArgs {
    int d, curl, N_p, N_q, N_f, N_fac, N_p1, N_e;
    const double *D, *Vq, *Vf, *nrstJ, *PJ, *I1, *D1, *Vq1, *Vf1;
    const double *x, *y, *z;
    double *J_q, *Lam, *J_f, *nJf, *nJq;
};

// LDS regions of k_geometric_factors (doubles)
struct GeoLayout {
    int X, dX, Jn, X1, dX1, F, Cc, Jq, size;
    __host__ __device__ GeoLayout(const GeoArgs& a) {
        const int d = a.d;
        X = 0;
        dX = X + d * a.N_p;                 // dX[(m d + n) N_p + j] = d x_m / d r_n at mapping node j
        Jn = dX + d * d * a.N_p;            // J at the mapping nodes (curl metrics)
        const int nc = (a.curl && d == 3) ? a.N_p1 : a.N_p;  // nodes the metric terms live on
        X1 = Jn + a.N_p;                    // x, y, z at the degree-(N+1) nodes (curl, 3D)
        dX1 = X1 + ((a.curl && d == 3) ? 3 * a.N_p1 : 0);     // D1_n x, D1_n y
        F = dX1 + ((a.curl && d == 3) ? 6 * a.N_p1 : 0);      // (D1_n a) b of one product pair
        Cc = F + ((a.curl && d == 3) ? 3 * a.N_p1 : 0);       // Cc[(l d + m) nc + j] = J dxi_l/dx_m
        Jq = Cc + (a.curl ? d * d * nc : 0);                  // J_q before projection
        size = Jq + a.N_q;
    }
};

__device__ inline void adjugate(int d, const double* A, double* L, double* J) {
    // A[m d + n] = d x_m / d r_n ; L[l d + m] = J d r_l / d x_m (metrics, mesh.jl:211-227)
    if (d == 1) {
        *J = A[0];
        L[0] = 1.0;
    } else if (d == 2) {
        *J = A[0] * A[3] - A[1] * A[2];
        L[0] = A[3];
        L[1] = -A[1];
        L[2] = -A[2];
        L[3] = A[0];
    } else {
        L[0] = A[4] * A[8] - A[5] * A[7];
        L[1] = A[2] * A[7] - A[1] * A[8];
        L[2] = A[1] * A[5] - A[2] * A[4];
        L[3] = A[5] * A[6] - A[3] * A[8];
        L[4] = A[0] * A[8] - A[2] * A[6];
        L[5] = A[2] * A[3] - A[0] * A[5];
        L[6] = A[3] * A[7] - A[4] * A[6];
        L[7] = A[1] * A[6] - A[0] * A[7];
        L[8] = A[0] * A[4] - A[1] * A[3];
        *J = A[0] * L[0] + A[1] * L[3] + A[2] * L[6];
    }
}

__global__ void __launch_bounds__(SETUP_THREADS) k_geometric_factors(GeoArgs a) {
    extern __shared__ double sm[];
    const GeoLayout L(a);
    const int k = blockIdx.x, tt = threadIdx.x, d = a.d, dd = d * d;
    const int N_p = a.N_p, N_q = a.N_q, N_f = a.N_f;
    const double* xs[3] = {a.x, a.y, a.z};
    // mapping nodes of element k
    for (int t = tt; t < d * N_p; t += SETUP_THREADS) {
        const int m = t / N_p, j = t - m * N_p;
        sm[L.X + t] = xs[m][(size_t)k * N_p + j];
    }
    __syncthreads();
    // nodal derivatives d x_m / d r_n = D_n x_m at the mapping nodes
    for (int t = tt; t < d * N_p; t += SETUP_THREADS) {
        const int m = t / N_p, i = t - m * N_p;
        double acc[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < N_p; ++j) {
            const double xj = sm[L.X + m * N_p + j];
            for (int n = 0; n < d; ++n) acc[n] += a.D[(size_t)(n * N_p + j) * N_p + i] * xj;
        }
        for (int n = 0; n < d; ++n) sm[L.dX + (m * d + n) * N_p + i] = acc[n];
    }
    __syncthreads();
    if (a.curl) {
        // J at the mapping nodes; for triangles also the metric terms (geometric_factors,
        // mesh.jl:306: rxJ = ys, sxJ = -yr, ryJ = -xs, syJ = xr)
        for (int i = tt; i < N_p; i += SETUP_THREADS) {
            double A[9], Lm[9], J;
            for (int q = 0; q < dd; ++q) A[q] = sm[L.dX + q * N_p + i];
            adjugate(d, A, Lm, &J);
            sm[L.Jn + i] = J;
            if (d == 2)
                for (int q = 0; q < 4; ++q) sm[L.Cc + q * N_p + i] = Lm[q];
        }
        if (d == 3) {
            // conservative curl form on the degree-(N+1) nodes (mesh.jl:452-487):
            // J dr_l/dx = curl terms of (D y) z, J dr_l/dy of (D x) z, J dr_l/dz of (D y) x
            const int N1 = a.N_p1;
            for (int t = tt; t < 3 * N1; t += SETUP_THREADS) {
                const int m = t / N1, i = t - m * N1;
                double acc = 0.0;
                for (int j = 0; j < N_p; ++j) acc += a.I1[(size_t)j * N1 + i] * sm[L.X + m * N_p + j];
                sm[L.X1 + t] = acc;
            }
            __syncthreads();
            for (int t = tt; t < 2 * N1; t += SETUP_THREADS) {  // D1_n x and D1_n y
                const int m = t / N1, i = t - m * N1;
                double acc[3] = {0.0, 0.0, 0.0};
                for (int j = 0; j < N1; ++j) {
                    const double xj = sm[L.X1 + m * N1 + j];
                    for (int n = 0; n < 3; ++n) acc[n] += a.D1[(size_t)(n * N1 + j) * N1 + i] * xj;
                }
                for (int n = 0; n < 3; ++n) sm[L.dX1 + (m * 3 + n) * N1 + i] = acc[n];
            }
            // the three products (a, b) and the metric column m they give, with sign
            const int pa[3] = {1, 0, 1}, pb[3] = {2, 2, 0};
            const double sg[3] = {1.0, -1.0, -1.0};
            for (int c = 0; c < 3; ++c) {
                __syncthreads();
                for (int t = tt; t < 3 * N1; t += SETUP_THREADS) {
                    const int n = t / N1, i = t - n * N1;
                    sm[L.F + t] = sm[L.dX1 + (pa[c] * 3 + n) * N1 + i] * sm[L.X1 + pb[c] * N1 + i];
                }
                __syncthreads();
                for (int i = tt; i < N1; i += SETUP_THREADS) {
                    // r: Dt Fs - Ds Ft ; s: Dr Ft - Dt Fr ; t: Ds Fr - Dr Fs
                    double dr[3] = {0, 0, 0}, ds[3] = {0, 0, 0}, dt[3] = {0, 0, 0};
                    for (int j = 0; j < N1; ++j) {
                        const double r = a.D1[(size_t)(0 * N1 + j) * N1 + i];
                        const double s = a.D1[(size_t)(1 * N1 + j) * N1 + i];
                        const double tq = a.D1[(size_t)(2 * N1 + j) * N1 + i];
                        for (int q = 0; q < 3; ++q) {
                            const double f = sm[L.F + q * N1 + j];
                            dr[q] += r * f;
                            ds[q] += s * f;
                            dt[q] += tq * f;
                        }
                    }
                    const int m = c;  // metric column: x, y, z
                    sm[L.Cc + (0 * 3 + m) * N1 + i] = sg[c] * (dt[1] - ds[2]);
                    sm[L.Cc + (1 * 3 + m) * N1 + i] = sg[c] * (dr[2] - dt[0]);
                    sm[L.Cc + (2 * 3 + m) * N1 + i] = sg[c] * (ds[0] - dr[1]);
                }
            }
        }
        __syncthreads();
    }
    // volume quadrature nodes: Lambda_q, J_q, nJq
    const int npf = N_f / a.N_fac;
    const int nc = (a.curl && d == 3) ? a.N_p1 : N_p;
    const double* Vc = (a.curl && d == 3) ? a.Vq1 : a.Vq;
    const double* src = a.curl ? sm + L.Cc : sm + L.dX;
    for (int i = tt; i < N_q; i += SETUP_THREADS) {
        double acc[9], Lm[9], J = 0.0;
        for (int q = 0; q < dd; ++q) acc[q] = 0.0;
        for (int j = 0; j < nc; ++j) {
            const double v = Vc[(size_t)j * N_q + i];
            for (int q = 0; q < dd; ++q) acc[q] += v * src[q * nc + j];
        }
        if (a.curl) {
            for (int j = 0; j < N_p; ++j) J += a.Vq[(size_t)j * N_q + i] * sm[L.Jn + j];
            for (int q = 0; q < dd; ++q) Lm[q] = acc[q];
        } else {
            adjugate(d, acc, Lm, &J);
        }
        double* Lk = a.Lam + (size_t)k * N_q * dd;  // Lambda_q[i, l, m, k], column-major
        for (int l = 0; l < d; ++l)
            for (int m = 0; m < d; ++m) Lk[(size_t)(m * d + l) * N_q + i] = Lm[l * d + m];
        sm[L.Jq + i] = J;
        // nJq[n, f, i, k] = sum_m Lambda_q[i, m, n, k] nrstJ[m] at the face's first node
        double* nk = a.nJq + ((size_t)k * N_q + i) * a.N_fac * d;
        for (int f = 0; f < a.N_fac; ++f)
            for (int n = 0; n < d; ++n) {
                double s = 0.0;
                for (int m = 0; m < d; ++m) s += Lm[m * d + n] * a.nrstJ[(size_t)m * N_f + f * npf];
                nk[f * d + n] = s;
            }
    }
    // facet quadrature nodes: nJf, J_f
    const double* Vfc = (a.curl && d == 3) ? a.Vf1 : a.Vf;
    for (int i = tt; i < N_f; i += SETUP_THREADS) {
        double acc[9], Lm[9], J;
        for (int q = 0; q < dd; ++q) acc[q] = 0.0;
        for (int j = 0; j < nc; ++j) {
            const double v = Vfc[(size_t)j * N_f + i];
            for (int q = 0; q < dd; ++q) acc[q] += v * src[q * nc + j];
        }
        if (a.curl) {
            for (int q = 0; q < dd; ++q) Lm[q] = acc[q];
        } else {
            adjugate(d, acc, Lm, &J);
        }
        double n2 = 0.0;
        for (int m = 0; m < d; ++m) {
            double s = 0.0;
            for (int n = 0; n < d; ++n) s += Lm[n * d + m] * a.nrstJ[(size_t)n * N_f + i];
            a.nJf[((size_t)k * N_f + i) * d + m] = s;
            n2 += s * s;
        }
        a.J_f[(size_t)k * N_f + i] = sqrt(n2);
    }
    __syncthreads();
    // J_q, projected onto the solution space when a projector is given
    for (int i = tt; i < N_q; i += SETUP_THREADS) {
        double J = sm[L.Jq + i];
        if (a.PJ) {
            J = 0.0;
            for (int j = 0; j < N_q; ++j) J += a.PJ[(size_t)j * N_q + i] * sm[L.Jq + j];
        }
        a.J_q[(size_t)k * N_q + i] = J;
    }
}

__global__ void k_reference_mapping(int N_q, int d, size_t N_e, const double* __restrict__ Lr,
                                    const double* __restrict__ Jr, const double* __restrict__ Lam,
                                    double* __restrict__ out) {
    // out[i, m, n, k] = sum_l Lr[i, m, l] Lam[i, l, n, k] / Jr[i]
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N_q * N_e) return;
    const size_t k = t / N_q;
    const int i = (int)(t - k * N_q);
    const int dd = d * d;
    const double* Lk = Lam + k * N_q * dd;
    double lam[9], lr[9];
    for (int q = 0; q < dd; ++q) {
        lam[q] = Lk[(size_t)q * N_q + i];
        lr[q] = Lr[(size_t)q * N_q + i];
    }
    const double inv = 1.0 / Jr[i];
    double* ok = out + k * N_q * dd;
    for (int m = 0; m < d; ++m)
        for (int n = 0; n < d; ++n) {
            double s = 0.0;
            for (int l = 0; l < d; ++l) s += lr[l * d + m] * lam[n * d + l];
            ok[(size_t)(n * d + m) * N_q + i] = s * inv;
        }
}

// L2 projection of one element: u = (V' W J V)^{-1} V' W J u_q, Cholesky in LDS.
// status[0] receives 1 + (the first element whose matrix is not positive definite).
__global__ void __launch_bounds__(SETUP_THREADS)
    k_project_l2(int N_q, int N_p, int N_c, const double* __restrict__ V, const double* __restrict__ W,
                 const double* __restrict__ J_q, const double* __restrict__ u_q, double* __restrict__ u,
                 int* status) {
    extern __shared__ double sm[];
    const int k = blockIdx.x, tt = threadIdx.x;
    double* M = sm;                      // N_p x N_p, row i at M[i N_p]
    double* y = M + N_p * N_p;           // N_p x N_c right-hand sides, y[a N_c + c]
    double* wj = y + N_p * N_c;          // W J at the quadrature nodes
    double* uq = wj + N_q;               // N_q x N_c, W J u_q
    __shared__ int bad;
    if (tt == 0) bad = 0;
    for (int i = tt; i < N_q; i += SETUP_THREADS) wj[i] = W[i] * J_q[(size_t)k * N_q + i];
    __syncthreads();
    for (int t = tt; t < N_q * N_c; t += SETUP_THREADS) {
        const int c = t / N_q, i = t - c * N_q;
        uq[i * N_c + c] = wj[i] * u_q[((size_t)k * N_c + c) * N_q + i];
    }
    // lower triangle of M = V' diag(W J) V
    const int ntri = N_p * (N_p + 1) / 2;
    for (int t = tt; t < ntri; t += SETUP_THREADS) {
        int a = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while (a * (a + 1) / 2 > t) --a;
        while ((a + 1) * (a + 2) / 2 <= t) ++a;
        const int b = t - a * (a + 1) / 2;
        const double* va = V + (size_t)a * N_q;
        const double* vb = V + (size_t)b * N_q;
        double s = 0.0;
        for (int i = 0; i < N_q; ++i) s += va[i] * wj[i] * vb[i];
        M[a * N_p + b] = s;
    }
    __syncthreads();
    for (int t = tt; t < N_p * N_c; t += SETUP_THREADS) {
        const int a = t / N_c, c = t - a * N_c;
        const double* va = V + (size_t)a * N_q;
        double s = 0.0;
        for (int i = 0; i < N_q; ++i) s += va[i] * uq[i * N_c + c];
        y[t] = s;
    }
    // Cholesky M = L L' (right-looking, column j at a time)
    for (int j = 0; j < N_p; ++j) {
        __syncthreads();
        const double piv = M[j * N_p + j];
        if (!(piv > 0.0)) {  // every thread reads the same pivot and leaves together
            if (tt == 0) bad = 1;
            break;
        }
        const double ljj = sqrt(piv);
        const double inv = 1.0 / ljj;
        const int m = N_p - j - 1;
        // column below the diagonal, then the trailing lower triangle
        for (int t = tt; t < m; t += SETUP_THREADS) M[(j + 1 + t) * N_p + j] *= inv;
        __syncthreads();
        if (tt == 0) M[j * N_p + j] = ljj;
        for (int t = tt; t < m * m; t += SETUP_THREADS) {
            const int r = t / m, c = t - r * m;
            if (c <= r) M[(j + 1 + r) * N_p + (j + 1 + c)] -= M[(j + 1 + r) * N_p + j] * M[(j + 1 + c) * N_p + j];
        }
    }
    __syncthreads();
    if (bad) {
        if (tt == 0) atomicCAS(status, 0, k + 1);
        return;
    }
    // L z = y, then L' x = z
    for (int j = 0; j < N_p; ++j) {
        if (tt < N_c) y[j * N_c + tt] /= M[j * N_p + j];
        __syncthreads();
        for (int t = tt; t < (N_p - j - 1) * N_c; t += SETUP_THREADS) {
            const int r = j + 1 + t / N_c, c = t % N_c;
            y[r * N_c + c] -= M[r * N_p + j] * y[j * N_c + c];
        }
        __syncthreads();
    }
    for (int j = N_p - 1; j >= 0; --j) {
        if (tt < N_c) y[j * N_c + tt] /= M[j * N_p + j];
        __syncthreads();
        for (int t = tt; t < j * N_c; t += SETUP_THREADS) {
            const int r = t / N_c, c = t - r * N_c;
            y[r * N_c + c] -= M[j * N_p + r] * y[j * N_c + c];
        }
        __syncthreads();
    }
    for (int t = tt; t < N_p * N_c; t += SETUP_THREADS) {
        const int c = t / N_p, a = t - c * N_p;
        u[((size_t)k * N_c + c) * N_p + a] = y[a * N_c + c];
    }
}

// ----------------------------------------------------------------------------- host side
struct DevBuf {
    std::vector<void*> ptrs;
    ~DevBuf() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    int alloc(size_t n, T** out) {
        void* p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess)
            return sse::set_global_error(SSE_ERR_OUT_OF_MEMORY, "hipMalloc failed in set-up");
        ptrs.push_back(p);
        *out = (T*)p;
        return 0;
    }
    // device copy of a host operator
    int put(const double* host, size_t n, const double** out) {
        if (!host) {
            *out = nullptr;
            return 0;
        }
        double* p;
        if (int rc = alloc(n, &p)) return rc;
        if (hipMemcpy(p, host, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return sse::set_global_error(SSE_ERR_HIP, "hipMemcpy failed in set-up");
        *out = p;
        return 0;
    }
};

int hip_fail(hipError_t e, const char* what) {
    return sse::set_global_error(SSE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

size_t max_lds() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 65536;
    return (size_t)v;
}

}  // namespace

extern "C" int sse_geometric_factors(const sse_geometry_ops* ops, int32_t N_e, const double* x, const double* y,
                                     const double* z, double* J_q, double* Lambda_q, double* J_f, double* nJf,
                                     double* nJq, int32_t where) {
    using sse::set_global_error;
    if (!ops) return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_geometric_factors: ops is NULL");
    const int d = ops->d;
    const bool curl = ops->metrics == SSE_METRICS_CURL;
    if (d < 1 || d > 3) return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_geometric_factors: d must be 1, 2 or 3");
    if (ops->metrics != SSE_METRICS_EXACT && !curl)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_geometric_factors: unknown metric type");
    if (curl && d == 1) return set_global_error(SSE_ERR_UNSUPPORTED, "sse_geometric_factors: curl metrics need d = 2 or 3");
    if (N_e < 0 || ops->N_p < 1 || ops->N_q < 1 || ops->N_f < 1 || ops->N_fac < 1 || ops->N_f % ops->N_fac)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_geometric_factors: bad sizes (N_f must be a multiple of N_fac)");
    if (!ops->D || !ops->Vq || !ops->Vf || !ops->nrstJ)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_geometric_factors: D, Vq, Vf and nrstJ are required");
    if (curl && d == 3 && (ops->N_p1 < 1 || !ops->I1 || !ops->D1 || !ops->Vq1 || !ops->Vf1))
        return set_global_error(SSE_ERR_INVALID_ARGUMENT,
                                "sse_geometric_factors: curl metrics on tetrahedra need N_p1, I1, D1, Vq1 and Vf1");
    if (!x || (d >= 2 && !y) || (d == 3 && !z) || !J_q || !Lambda_q || !J_f || !nJf || !nJq)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_geometric_factors: NULL coordinate or output array");
    if (where != SSE_DEVICE && where != SSE_HOST)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_geometric_factors: where must be SSE_DEVICE or SSE_HOST");
    if (N_e == 0) return 0;
    GeoArgs a{};
    a.d = d;
    a.curl = curl ? 1 : 0;
    a.N_p = ops->N_p;
    a.N_q = ops->N_q;
    a.N_f = ops->N_f;
    a.N_fac = ops->N_fac;
    a.N_p1 = (curl && d == 3) ? ops->N_p1 : 0;
    a.N_e = N_e;
    const GeoLayout L(a);
    const size_t lds = sizeof(double) * (size_t)L.size;
    if (lds > max_lds())
        return set_global_error(SSE_ERR_UNSUPPORTED, "sse_geometric_factors: element too large for one workgroup's LDS");
    DevBuf buf;
    int rc;
    const size_t Np = a.N_p, Nq = a.N_q, Nf = a.N_f, N1 = a.N_p1;
    if ((rc = buf.put(ops->D, Np * Np * d, &a.D)) || (rc = buf.put(ops->Vq, Nq * Np, &a.Vq)) ||
        (rc = buf.put(ops->Vf, Nf * Np, &a.Vf)) || (rc = buf.put(ops->nrstJ, Nf * d, &a.nrstJ)) ||
        (rc = buf.put(ops->P_J, ops->P_J ? Nq * Nq : 0, &a.PJ)))
        return rc;
    if (curl && d == 3) {
        if ((rc = buf.put(ops->I1, N1 * Np, &a.I1)) || (rc = buf.put(ops->D1, N1 * N1 * 3, &a.D1)) ||
            (rc = buf.put(ops->Vq1, Nq * N1, &a.Vq1)) || (rc = buf.put(ops->Vf1, Nf * N1, &a.Vf1)))
            return rc;
    }
    const size_t ne = (size_t)N_e, dd = (size_t)d * d;
    const size_t n_x = Np * ne, n_Jq = Nq * ne, n_L = Nq * dd * ne, n_Jf = Nf * ne, n_nJf = d * Nf * ne,
                 n_nJq = d * (size_t)a.N_fac * Nq * ne;
    if (where == SSE_HOST) {
        double *dx, *dy = nullptr, *dz = nullptr, *o1, *o2, *o3, *o4, *o5;
        if ((rc = buf.alloc(n_x, &dx)) || (d >= 2 && (rc = buf.alloc(n_x, &dy))) ||
            (d == 3 && (rc = buf.alloc(n_x, &dz))) || (rc = buf.alloc(n_Jq, &o1)) || (rc = buf.alloc(n_L, &o2)) ||
            (rc = buf.alloc(n_Jf, &o3)) || (rc = buf.alloc(n_nJf, &o4)) || (rc = buf.alloc(n_nJq, &o5)))
            return rc;
        hipError_t e;
        if ((e = hipMemcpy(dx, x, n_x * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess ||
            (d >= 2 && (e = hipMemcpy(dy, y, n_x * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) ||
            (d == 3 && (e = hipMemcpy(dz, z, n_x * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess))
            return hip_fail(e, "sse_geometric_factors: copy of the mapping nodes");
        a.x = dx, a.y = dy, a.z = dz;
        a.J_q = o1, a.Lam = o2, a.J_f = o3, a.nJf = o4, a.nJq = o5;
    } else {
        a.x = x, a.y = y, a.z = z;
        a.J_q = J_q, a.Lam = Lambda_q, a.J_f = J_f, a.nJf = nJf, a.nJq = nJq;
    }
    hipLaunchKernelGGL(k_geometric_factors, dim3(N_e), dim3(SETUP_THREADS), lds, 0, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "k_geometric_factors launch");
    if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, "k_geometric_factors");
    if (where == SSE_HOST) {
        const std::pair<double*, std::pair<const double*, size_t>> outs[5] = {
            {J_q, {a.J_q, n_Jq}}, {Lambda_q, {a.Lam, n_L}}, {J_f, {a.J_f, n_Jf}}, {nJf, {a.nJf, n_nJf}},
            {nJq, {a.nJq, n_nJq}}};
        for (const auto& o : outs)
            if ((e = hipMemcpy(o.first, o.second.first, o.second.second * sizeof(double), hipMemcpyDeviceToHost)) !=
                hipSuccess)
                return hip_fail(e, "sse_geometric_factors: copy of the geometric factors");
    }
    return 0;
}

extern "C" int sse_apply_reference_mapping(int32_t N_q, int32_t d, int32_t N_e, const double* Lambda_ref,
                                           const double* J_ref, const double* Lambda_q, double* out, int32_t where) {
    using sse::set_global_error;
    if (N_q < 1 || d < 1 || d > 3 || N_e < 0 || !Lambda_ref || !J_ref || !Lambda_q || !out)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_apply_reference_mapping: bad arguments");
    if (where != SSE_DEVICE && where != SSE_HOST)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_apply_reference_mapping: where must be SSE_DEVICE or SSE_HOST");
    if (N_e == 0) return 0;
    DevBuf buf;
    int rc;
    const double *Lr, *Jr;
    const size_t n = (size_t)N_q * d * d * N_e;
    if ((rc = buf.put(Lambda_ref, (size_t)N_q * d * d, &Lr)) || (rc = buf.put(J_ref, N_q, &Jr))) return rc;
    const double* src = Lambda_q;
    double* dst = out;
    hipError_t e;
    if (where == SSE_HOST) {
        double *s2, *d2;
        if ((rc = buf.alloc(n, &s2)) || (rc = buf.alloc(n, &d2))) return rc;
        if ((e = hipMemcpy(s2, Lambda_q, n * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, "sse_apply_reference_mapping: copy in");
        src = s2, dst = d2;
    }
    const size_t total = (size_t)N_q * N_e;
    hipLaunchKernelGGL(k_reference_mapping, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, 0, N_q, d,
                       (size_t)N_e, Lr, Jr, src, dst);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "k_reference_mapping launch");
    if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, "k_reference_mapping");
    if (where == SSE_HOST && (e = hipMemcpy(out, dst, n * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "sse_apply_reference_mapping: copy out");
    return 0;
}

extern "C" int sse_project_l2(int32_t N_q, int32_t N_p, int32_t N_c, int32_t N_e, const double* V, const double* W,
                              const double* J_q, const double* u_q, double* u, int32_t where) {
    using sse::set_global_error;
    if (N_q < 1 || N_p < 1 || N_c < 1 || N_e < 0 || !V || !W || !J_q || !u_q || !u)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_project_l2: bad arguments");
    if (N_c > SETUP_THREADS)
        return set_global_error(SSE_ERR_UNSUPPORTED, "sse_project_l2: more than 256 components");
    if (where != SSE_DEVICE && where != SSE_HOST)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_project_l2: where must be SSE_DEVICE or SSE_HOST");
    if (N_e == 0) return 0;
    const size_t lds = sizeof(double) * ((size_t)N_p * N_p + (size_t)N_p * N_c + N_q + (size_t)N_q * N_c);
    if (lds > max_lds())
        return set_global_error(SSE_ERR_UNSUPPORTED, "sse_project_l2: element too large for one workgroup's LDS");
    DevBuf buf;
    int rc;
    const double *dV, *dW;
    if ((rc = buf.put(V, (size_t)N_q * N_p, &dV)) || (rc = buf.put(W, N_q, &dW))) return rc;
    const size_t ne = (size_t)N_e, n_J = (size_t)N_q * ne, n_uq = (size_t)N_q * N_c * ne,
                 n_u = (size_t)N_p * N_c * ne;
    const double *jq = J_q, *uq = u_q;
    double* uo = u;
    int* status;
    if ((rc = buf.alloc(1, &status))) return rc;
    hipError_t e;
    if ((e = hipMemset(status, 0, sizeof(int))) != hipSuccess) return hip_fail(e, "sse_project_l2: status");
    if (where == SSE_HOST) {
        double *a1, *a2, *a3;
        if ((rc = buf.alloc(n_J, &a1)) || (rc = buf.alloc(n_uq, &a2)) || (rc = buf.alloc(n_u, &a3))) return rc;
        if ((e = hipMemcpy(a1, J_q, n_J * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(a2, u_q, n_uq * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, "sse_project_l2: copy in");
        jq = a1, uq = a2, uo = a3;
    }
    hipLaunchKernelGGL(k_project_l2, dim3(N_e), dim3(SETUP_THREADS), lds, 0, N_q, N_p, N_c, dV, dW, jq, uq, uo, status);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "k_project_l2 launch");
    if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, "k_project_l2");
    int st = 0;
    if ((e = hipMemcpy(&st, status, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "sse_project_l2: status");
    if (st)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_project_l2: V' W J V is not positive definite on element " +
                                                              std::to_string(st - 1) + " (J <= 0 at a quadrature node?)");
    if (where == SSE_HOST && (e = hipMemcpy(u, uo, n_u * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "sse_project_l2: copy out");
    return 0;
}

// ----------------------------------------------------------------------------- flux evaluator
namespace {

template <int D>
__global__ void __launch_bounds__(256) k_ec_flux_batch(sse::LawParams law, int n, const double* __restrict__ uL,
                                                       const double* __restrict__ uR,
                                                       const double* __restrict__ a, double* __restrict__ c) {
    constexpr int NC = D + 2, NH = D + 5;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double ul[NC], ur[NC], hl[NH], hr[NH], av[D], cv[NC];
#pragma unroll
    for (int e = 0; e < NC; ++e) {
        ul[e] = uL[(size_t)t * NC + e];
        ur[e] = uR[(size_t)t * NC + e];
    }
#pragma unroll
    for (int m = 0; m < D; ++m) av[m] = a[(size_t)t * D + m];
    sse::hoist<D>(law, ul, hl);
    sse::hoist<D>(law, ur, hr);
    sse::ec_dot<D>(law, hl, hr, av, cv);
#pragma unroll
    for (int e = 0; e < NC; ++e) c[(size_t)t * NC + e] = cv[e];
}

}  // namespace

extern "C" int sse_ec_flux_batch(int32_t d, double gamma, int32_t n, const double* uL, const double* uR,
                                 const double* a, double* c, int32_t where) {
    using sse::set_global_error;
    if (d < 1 || d > 3 || !(gamma > 1.0) || n < 0 || !uL || !uR || !a || !c)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_ec_flux_batch: bad arguments");
    if (where != SSE_DEVICE && where != SSE_HOST)
        return set_global_error(SSE_ERR_INVALID_ARGUMENT, "sse_ec_flux_batch: where must be SSE_DEVICE or SSE_HOST");
    if (n == 0) return 0;
    const size_t nc = (size_t)n * (d + 2), na = (size_t)n * d;
    DevBuf buf;
    int rc;
    const double *l = uL, *r = uR, *av = a;
    double* out = c;
    if (where == SSE_HOST) {
        if ((rc = buf.put(uL, nc, &l)) || (rc = buf.put(uR, nc, &r)) || (rc = buf.put(a, na, &av)) ||
            (rc = buf.alloc(nc, &out)))
            return rc;
    }
    sse::LawParams law{};
    law.gamma = gamma;
    law.gm1 = gamma - 1.0;
    law.inv_gm1 = 1.0 / (gamma - 1.0);
    law.log_gm1 = std::log(gamma - 1.0);
    const dim3 grid((unsigned)((n + 255) / 256));
    switch (d) {
        case 1: hipLaunchKernelGGL(k_ec_flux_batch<1>, grid, dim3(256), 0, 0, law, n, l, r, av, out); break;
        case 2: hipLaunchKernelGGL(k_ec_flux_batch<2>, grid, dim3(256), 0, 0, law, n, l, r, av, out); break;
        default: hipLaunchKernelGGL(k_ec_flux_batch<3>, grid, dim3(256), 0, 0, law, n, l, r, av, out); break;
    }
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "k_ec_flux_batch launch");
    if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, "k_ec_flux_batch");
    if (where == SSE_HOST && (e = hipMemcpy(c, out, nc * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "sse_ec_flux_batch: copy out");
    return 0;
}
