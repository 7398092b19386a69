// sse_api.hip -- the C-ABI of libsse_hip.so (declared in include/sse_hip.h).
//
// sse_create validates the descriptor (the reference's Solver fields), turns
// the dense reference operators into compressed-row tables, uploads operators
// and geometry, and allocates the phase-1 intermediates. sse_residual and
// sse_rk_steps launch the kernels on the handle's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/sse_hip.h"
#include "sse_kernels.h"
#include "sse_tensor.h"

namespace sse {
hipError_t launch_generic_dispatch(const DevParams& p, const double* u, double* dudt, const RKStage& rk,
                                   hipStream_t s, int which);
size_t lds_nodal_values_generic(const DevParams& p);
size_t lds_nodal_values_mfma(const DevParams& p);
size_t lds_time_derivative_generic(const DevParams& p);
hipError_t launch_standard(const DevParams& p, double* dudt, const RKStage& rk, hipStream_t s, int which);
size_t lds_standard(const DevParams& p);
hipError_t launch_functionals(const DevParams& p, const double* u, const double* dudt, int which, double* partial,
                              double* out, hipStream_t s, double* mass = nullptr, int mass_ready = 0);
size_t lds_functionals(const DevParams& p, int which);
hipError_t launch_vjp(const DevParams& p, const double* u, const double* ybar, double* ubar, double* rbar,
                      double* fbar, double* lam_partial, double* lam_out, double lam_scale, int accumulate,
                      hipStream_t s);
size_t lds_vjp_v2(const DevParams& p);
hipError_t launch_axpby(size_t n, double a, const double* x, double b, double* y, hipStream_t s);
hipError_t launch_entropy_grad(const DevParams& p, const double* u, double* g, hipStream_t s);
struct TensorConfig {
    int TS, EPB, R, lds1, lds2, NQ, NF, NP;
};
bool tensor_config(int elem, int P, int var, TensorConfig* c);
hipError_t launch_weight_adjusted_inverse(const DevParams& p, double* Mw, hipStream_t s);
hipError_t launch_tensor(const TensorParams& tp, const double* u, double* dudt, const RKStage& rk, hipStream_t s,
                         int which);
struct StructConfig {
    int TS, EPB, lds, max_item;
};
bool struct_config(int elem, int P, StructConfig* c);
int struct_occupancy(int elem, int P);
hipError_t launch_struct(const TensorParams& tp, const StructTables& st, double* dudt, const RKStage& rk,
                         hipStream_t s, int n_blocks);
size_t std_tensor_config(int elem, int P);
size_t vjp_tensor_config(int elem, int P);
hipError_t launch_vjp_tensor(const TensorParams& tp, const StructTables& st, const double* u, const double* ybar,
                             double* ubar, double* rbar, double* fbar, double* fbar_ext, double* lam_partial,
                             double scale, int accumulate, const double* in_x, double in_a, double in_b,
                             double* in_store, hipStream_t s);
hipError_t launch_vjp_sum(const double* lam_partial, int n, double* lam_out, double lam_scale, int accumulate,
                          hipStream_t s);
hipError_t collapse_metric(double* Lam, const double* c, int NQ, int D, int N_e, hipStream_t s);
hipError_t launch_std_tensor(const TensorParams& tp, int R, const double* b_member, int tangent, const double* u,
                             double* dudt, const RKStage& rk, hipStream_t s, int which);
}  // namespace sse

using sse::DevParams;
using sse::RKStage;

struct sse_handle {
    DevParams P{};
    sse::TensorParams TP{};
    bool use_tensor = false;
    sse::StructTables ST{};
    bool use_struct = false;
    int struct_blocks = 0;
    bool use_std_tensor = false;  // StandardForm scalar-law kernels with compile-time shapes
    int members = 1;              // lanes per element the intermediates hold (ensemble capacity)
    int element = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    double* rk_tmp = nullptr;
    double* erk_k = nullptr;     // sse_erk_steps: erk_stages stage derivatives, then the stage state
    // sse_rk_steps: one CK2N54 step captured as a HIP graph, for the state, step and
    // stream it was captured with
    hipGraph_t rk_graph = nullptr;
    hipGraphExec_t rk_exec = nullptr;
    const double* rk_graph_u = nullptr;
    double rk_graph_dt = 0.0;
    hipStream_t rk_graph_stream = nullptr;
    int erk_stages = 0;
    double* stage_u = nullptr;
    double* stage_dudt = nullptr;
    double* fpartial = nullptr;  // functionals: per-element partial sums and the result
    double* fout = nullptr;
    double* ferr = nullptr;      // staging of u_exact for sse_l2_error with host arrays
    double* fmass = nullptr;     // every element's M_k, kept by the first energy functional
    bool fmass_ready = false;
    double *vjp_rbar = nullptr, *vjp_fbar = nullptr, *vjp_lam = nullptr, *vjp_lamout = nullptr;  // VJP scratch
    double* vjp_fext = nullptr;  // fbar in the neighbour's exterior slot (tensor VJP kernels)
    bool vjp_tensor = false;
    hipStream_t adj_stream = nullptr;  // sse_rk_adjoint: the segment recomputation, beside the reverse sweep
    double* noalloc_ws = nullptr;  // this device's address of the file-scope workspace `scratch`
    std::string noalloc_err;       // why that address is missing, if it is
    std::vector<void*> allocations;
    std::string err;
};

namespace {

thread_local std::string g_create_error;

// Workspace of sse_functional_noalloc: the per-element partial sums and their totals.
// It is a file-scope device array, so the entry point allocates nothing; every GPU the
// library runs on holds its own copy (8 MiB) until the process ends. Host code never
// passes the name `scratch` to a kernel (on the host it is only a stand-in symbol): it
// uses the device address hipGetSymbolAddress returns, looked up once per device.
static __device__ __align__(256) double scratch[SSE_FUNCTIONAL_NOALLOC_MAX_DOUBLES];

constexpr int kMaxDevices = 64;
std::mutex g_ws_mutex;
double* g_ws_ptr[kMaxDevices] = {};
hipError_t g_ws_status[kMaxDevices] = {};
bool g_ws_done[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first call per
// device, with that device current). Returns the lookup's status.
hipError_t workspace_address(double** out) {
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if (!g_ws_done[dev]) {
        void* p = nullptr;
        g_ws_status[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ws_ptr[dev] = g_ws_status[dev] == hipSuccess ? static_cast<double*>(p) : nullptr;
        g_ws_done[dev] = true;
    }
    *out = g_ws_ptr[dev];
    return g_ws_status[dev];
}

// Carpenter & Kennedy (1994), NASA TM-109112: five-stage fourth-order 2N-storage
// Runge-Kutta, the coefficients of OrdinaryDiffEq's CarpenterKennedy2N54.
const double CK_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                        -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double CK_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                        2277821191437.0 / 14882151754819.0};

int fail(sse_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    else g_create_error = msg;
    return code;
}

#define SSE_HIP_TRY(h, expr)                                                             \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return fail((h), SSE_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

template <typename T>
int alloc(sse_handle* h, size_t n, T** out);

// per-element partial sums and the result of sse_functional / sse_l2_error
int functional_buffers(sse_handle* h) {
    int rc;
    if (!h->fpartial && ((rc = alloc(h, (size_t)h->P.N_e * (h->P.N_c + 1), &h->fpartial)) ||
                         (rc = alloc(h, (size_t)h->P.N_c + 1, &h->fout))))
        return rc;
    return 0;
}

template <typename T>
int upload(sse_handle* h, const std::vector<T>& host, const T** dev) {
    if (host.empty()) {
        *dev = nullptr;
        return 0;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, host.size() * sizeof(T));
    if (e != hipSuccess) return fail(h, SSE_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    h->allocations.push_back(p);
    e = hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    *dev = static_cast<const T*>(p);
    return 0;
}

template <typename T>
int alloc(sse_handle* h, size_t n, T** dev) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n * sizeof(T) + 16);
    if (e != hipSuccess) return fail(h, SSE_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    h->allocations.push_back(p);
    *dev = static_cast<T*>(p);
    return 0;
}

struct Rows {
    std::vector<int> ptr, idx;
    std::vector<double> val;
};

// rows of a dense column-major n_rows x n_cols matrix (nonzero entries)
Rows rows_of(const double* A, int n_rows, int n_cols) {
    Rows r;
    r.ptr.push_back(0);
    for (int i = 0; i < n_rows; ++i) {
        for (int j = 0; j < n_cols; ++j) {
            double v = A[i + (size_t)n_rows * j];
            if (v != 0.0) {
                r.idx.push_back(j);
                r.val.push_back(v);
            }
        }
        r.ptr.push_back((int)r.idx.size());
    }
    return r;
}

// rows of the transpose
Rows cols_of(const double* A, int n_rows, int n_cols) {
    Rows r;
    r.ptr.push_back(0);
    for (int j = 0; j < n_cols; ++j) {
        for (int i = 0; i < n_rows; ++i) {
            double v = A[i + (size_t)n_rows * j];
            if (v != 0.0) {
                r.idx.push_back(i);
                r.val.push_back(v);
            }
        }
        r.ptr.push_back((int)r.idx.size());
    }
    return r;
}

int validate(const sse_desc* D, std::string& msg) {
    if (!D) { msg = "descriptor is NULL"; return SSE_ERR_INVALID_ARGUMENT; }
    if (D->abi_version != SSE_ABI_VERSION) { msg = "ABI version mismatch"; return SSE_ERR_INVALID_ARGUMENT; }
    if (D->d < 1 || D->d > 3) { msg = "d must be 1, 2 or 3"; return SSE_ERR_INVALID_ARGUMENT; }
    if (D->N_p <= 0 || D->N_q <= 0 || D->N_f <= 0 || D->N_fac <= 0 || D->N_e <= 0 || D->N_c <= 0) {
        msg = "sizes must be positive"; return SSE_ERR_INVALID_ARGUMENT;
    }
    if (D->N_f % D->N_fac != 0) { msg = "N_f must be a multiple of N_fac"; return SSE_ERR_INVALID_ARGUMENT; }
    const bool standard = (D->form == SSE_STANDARD);
    if (!standard && D->form != SSE_FLUX_DIFFERENCING) { msg = "unknown residual form"; return SSE_ERR_INVALID_ARGUMENT; }
    if (D->law == SSE_EULER) {
        if (D->N_c != D->d + 2) { msg = "Euler needs N_c = d + 2"; return SSE_ERR_INVALID_ARGUMENT; }
        if (!(D->gamma > 1.0)) { msg = "gamma must exceed 1"; return SSE_ERR_INVALID_ARGUMENT; }
    } else if (D->law == SSE_LINEAR_ADVECTION || D->law == SSE_LINEAR_ADVECTION_DIFFUSION) {
        if (D->N_c != 1) { msg = "linear advection(-diffusion) needs N_c = 1"; return SSE_ERR_INVALID_ARGUMENT; }
        if (D->law == SSE_LINEAR_ADVECTION_DIFFUSION) {
            if (!standard) {
                msg = "second-order laws are implemented for StandardForm (as in the reference)";
                return SSE_ERR_UNSUPPORTED;
            }
            if (D->viscous_flux != SSE_BR1) { msg = "second-order laws need the BR1 viscous flux"; return SSE_ERR_UNSUPPORTED; }
            if (!(D->b >= 0.0)) { msg = "diffusion coefficient must be non-negative"; return SSE_ERR_INVALID_ARGUMENT; }
        }
    } else {
        msg = "unknown conservation law"; return SSE_ERR_INVALID_ARGUMENT;
    }
    if (standard) {
        if (!D->D) { msg = "StandardForm needs the derivative operators D"; return SSE_ERR_INVALID_ARGUMENT; }
        if (D->mapping_form != SSE_SKEW_SYMMETRIC_MAPPING && D->mapping_form != SSE_STANDARD_MAPPING) {
            msg = "unknown mapping form"; return SSE_ERR_INVALID_ARGUMENT;
        }
        if (D->two_point_flux != SSE_TWO_POINT_CONSERVATIVE) {
            msg = "StandardForm's interface flux uses the conservative two-point flux"; return SSE_ERR_INVALID_ARGUMENT;
        }
    }
    if (D->inviscid_flux < 1 || D->inviscid_flux > 3) { msg = "unknown inviscid flux"; return SSE_ERR_INVALID_ARGUMENT; }
    if (D->two_point_flux < 1 || D->two_point_flux > 2) { msg = "unknown two-point flux"; return SSE_ERR_INVALID_ARGUMENT; }
    if (!D->V && D->N_p != D->N_q) { msg = "nodal scheme (V = NULL) needs N_p == N_q"; return SSE_ERR_INVALID_ARGUMENT; }
    if (D->Minv) { msg = "a non-identity weight-adjusted M^{-1} is not supported"; return SSE_ERR_UNSUPPORTED; }
    if (D->V && D->mass_solver != SSE_MASS_WEIGHT_ADJUSTED) {
        msg = "modal schemes use the weight-adjusted mass solver"; return SSE_ERR_UNSUPPORTED;
    }
    if (!D->V && D->mass_solver != SSE_MASS_DIAGONAL) {
        msg = "nodal schemes use the diagonal mass solver"; return SSE_ERR_UNSUPPORTED;
    }
    if (!D->R || !D->W || !D->B || (!standard && !D->S) || !D->J_q || !D->Lambda_q || !D->nJf || !D->J_f || !D->face_normals ||
        !D->mapP) {
        msg = "missing operator or geometry array"; return SSE_ERR_INVALID_ARGUMENT;
    }
    if (D->C_kind != SSE_C_NONE && !D->C) { msg = "C is NULL but C_kind says present"; return SSE_ERR_INVALID_ARGUMENT; }
    const long long nf = (long long)D->N_f * D->N_e;
    for (long long t = 0; t < nf; ++t) {
        if (D->mapP[t] < 1 || D->mapP[t] > nf) { msg = "mapP entry out of range"; return SSE_ERR_INVALID_ARGUMENT; }
    }
    if (nf > 2147483647LL) { msg = "N_f * N_e exceeds int32 range"; return SSE_ERR_UNSUPPORTED; }
    if (D->R_selection) {
        for (int j = 0; j < D->N_f; ++j)
            if (D->R_selection[j] < 1 || D->R_selection[j] > D->N_q) {
                msg = "R_selection out of range"; return SSE_ERR_INVALID_ARGUMENT;
            }
    }
    for (long long t = 0; t < (long long)D->N_q * D->N_e; ++t)
        if (!(D->J_q[t] > 0.0)) { msg = "J_q must be positive"; return SSE_ERR_INVALID_ARGUMENT; }
    for (long long t = 0; t < nf; ++t)
        if (!(D->J_f[t] > 0.0)) { msg = "J_f must be positive"; return SSE_ERR_INVALID_ARGUMENT; }
    return 0;
}

// The kernels derive 0.5 nJq from Lambda_q and the face normals (the identity
// of src/SpatialDiscretizations/mesh.jl:262-268); a supplied nJq must agree.
int check_nJq(const sse_desc* D, std::string& msg) {
    if (!D->nJq) return 0;
    const int d = D->d;
    double scale = 0.0, worst = 0.0;
    for (long long k = 0; k < D->N_e; ++k)
        for (int i = 0; i < D->N_q; ++i)
            for (int f = 0; f < D->N_fac; ++f)
                for (int n = 0; n < d; ++n) {
                    double s = 0.0;
                    for (int m = 0; m < d; ++m)
                        s += D->Lambda_q[i + (size_t)D->N_q * (m + d * (n + (size_t)d * k))] * D->face_normals[m + d * f];
                    double g = D->nJq[n + d * (f + (size_t)D->N_fac * (i + (size_t)D->N_q * k))];
                    scale = fmax(scale, fabs(g));
                    worst = fmax(worst, fabs(g - s));
                }
    if (worst > 1e-12 * fmax(scale, 1.0)) {
        msg = "nJq is not Lambda_q contracted with the face normals";
        return SSE_ERR_UNSUPPORTED;
    }
    return 0;
}

// ELL table (width = longest row) of a list of rows; -1 pads
template <typename V>
void ell(const std::vector<std::vector<std::pair<int, V>>>& rows, int& width, std::vector<int>& idx,
         std::vector<V>& val) {
    width = 0;
    for (auto& r : rows) width = std::max(width, (int)r.size());
    const int n = (int)rows.size();
    idx.assign((size_t)std::max(width, 1) * n, -1);
    val.assign((size_t)std::max(width, 1) * n, V(0));
    for (int o = 0; o < n; ++o)
        for (int w = 0; w < (int)rows[o].size(); ++w) {
            idx[(size_t)w * n + o] = rows[o][w].first;
            val[(size_t)w * n + o] = rows[o][w].second;
        }
}

// Product of the warped-product factors, column-major N_q x N_p, for checking
// the dense V the caller passed.
std::vector<double> warped_dense(const sse_desc* D) {
    const int P = D->V_p, d = D->d;
    const int M1 = D->V_M[0], M2 = D->V_M[1], M3 = d == 3 ? D->V_M[2] : 1;
    const int NQ = M1 * M2 * M3;
    std::vector<double> V;
    std::vector<std::vector<double>> cols;
    for (int i = 0; i <= P; ++i)
        for (int j = 0; j <= P - i; ++j) {
            if (d == 2) {
                std::vector<double> c(NQ);
                for (int a1 = 0; a1 < M1; ++a1)
                    for (int a2 = 0; a2 < M2; ++a2)
                        c[a1 * M2 + a2] = D->V_A[a1 + M1 * i] * D->V_B[a2 + M2 * (i + (P + 1) * j)];
                cols.push_back(c);
            } else {
                for (int kk = 0; kk <= P - i - j; ++kk) {
                    std::vector<double> c(NQ);
                    for (int a1 = 0; a1 < M1; ++a1)
                        for (int a2 = 0; a2 < M2; ++a2)
                            for (int a3 = 0; a3 < M3; ++a3)
                                c[(a1 * M2 + a2) * M3 + a3] = D->V_A[a1 + M1 * i] *
                                                              D->V_B[a2 + M2 * (i + (P + 1) * j)] *
                                                              D->V_C[a3 + M3 * (i + (P + 1) * (j + (P + 1) * kk))];
                    cols.push_back(c);
                }
            }
        }
    for (auto& c : cols) V.insert(V.end(), c.begin(), c.end());
    return V;
}

// Set up the compile-time-shaped tensor kernels when the discretisation is one
// of the compiled (element, degree) shapes; returns 0 (and leaves use_tensor
// false) when it is not.
int build_tensor(sse_handle* h, const sse_desc* D) {
    const DevParams& P = h->P;
    if (D->law != SSE_EULER || D->two_point_flux != SSE_TWO_POINT_EC) return 0;
    if (!(D->element == SSE_TRI || D->element == SSE_TET) || !D->V || D->V_p <= 0 || !D->V_A || !D->V_B) return 0;
    if (D->element == SSE_TET && !D->V_C) return 0;
    if (D->C_kind == SSE_C_NONE || D->mass_solver != SSE_MASS_WEIGHT_ADJUSTED || D->Minv) return 0;
    sse::TensorConfig cfg;
    int variant = 0;
    if (const char* env = getenv("SSE_TENSOR_VARIANT")) variant = atoi(env);
    if (!sse::tensor_config(D->element, D->V_p, variant, &cfg)) {
        variant = 0;
        if (!sse::tensor_config(D->element, D->V_p, variant, &cfg)) return 0;
    }
    for (int m = 0; m < D->d; ++m)
        if (D->V_M[m] != D->V_p + 1) return 0;
    if (cfg.NQ != D->N_q || cfg.NF != D->N_f || cfg.NP != D->N_p) return 0;
    // the dense V must be the product of the factors
    {
        std::vector<double> Vw = warped_dense(D);
        if ((int)Vw.size() != D->N_q * D->N_p) return 0;
        double scale = 0.0, worst = 0.0;
        for (size_t t = 0; t < Vw.size(); ++t) {
            scale = std::max(scale, std::fabs(D->V[t]));
            worst = std::max(worst, std::fabs(D->V[t] - Vw[t]));
        }
        if (worst > 1e-12 * std::max(scale, 1.0)) return 0;
    }
    const int d = D->d, N_q = D->N_q, N_f = D->N_f;
    sse::TensorParams& T = h->TP;
    T.base = P;
    T.elem = D->element;
    T.P = D->V_p;
    T.variant = variant;
    // pair list: volume pairs (union sparsity of S, i < j), then facet pairs (nonzeros of C)
    std::vector<int> pk;
    std::vector<std::vector<double>> cfm(d);
    for (int i = 0; i < N_q; ++i)
        for (int j = i + 1; j < N_q; ++j) {
            bool nz = false;
            for (int m = 0; m < d; ++m) nz |= D->S[i + (size_t)N_q * (j + (size_t)N_q * m)] != 0.0;
            if (!nz) continue;
            pk.push_back(i | (j << 12));
            for (int m = 0; m < d; ++m) cfm[m].push_back(D->S[i + (size_t)N_q * (j + (size_t)N_q * m)]);
        }
    T.n_vol_pairs = (int)pk.size();
    // order volume pairs by (j - i, i): lanes of a wave then walk neighbouring
    // nodes of one line family at one distance (few LDS bank conflicts)
    {
        std::vector<int> ord(pk.size());
        for (size_t t = 0; t < ord.size(); ++t) ord[t] = (int)t;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
            const int ia = pk[a] & 0xfff, ja = (pk[a] >> 12) & 0xfff, ib = pk[b] & 0xfff, jb = (pk[b] >> 12) & 0xfff;
            if (ja - ia != jb - ib) return ja - ia < jb - ib;
            return ia < ib;
        });
        std::vector<int> pk2;
        std::vector<std::vector<double>> cf2(d);
        for (int t : ord) {
            pk2.push_back(pk[t]);
            for (int m = 0; m < d; ++m) cf2[m].push_back(cfm[m][t]);
        }
        pk.swap(pk2);
        cfm.swap(cf2);
    }
    // facet pairs ordered by (face, rank of i within its column, facet node j)
    {
        const int npf = N_f / D->N_fac;
        std::vector<std::array<int, 4>> fp;  // face, rank, j, i
        for (int j = 0; j < N_f; ++j) {
            int rank = 0;
            for (int i = 0; i < N_q; ++i) {
                if (D->C[i + (size_t)N_q * j] == 0.0) continue;
                fp.push_back({j / npf, rank++, j, i});
            }
        }
        std::stable_sort(fp.begin(), fp.end());
        for (auto& e : fp) {
            const int j = e[2], i = e[3];
            pk.push_back(i | (j << 12) | (1 << 24));
            cfm[0].push_back(D->C[i + (size_t)N_q * j]);
            for (int m = 1; m < d; ++m) cfm[m].push_back(0.0);
        }
    }
    T.n_pairs = (int)pk.size();
    T.pass_size = cfg.TS * cfg.R;
    T.n_pass = (T.n_pairs + T.pass_size - 1) / T.pass_size;
    std::vector<double> cf;
    for (int m = 0; m < d; ++m) cf.insert(cf.end(), cfm[m].begin(), cfm[m].end());
    // owner reduction tables
    std::vector<std::vector<std::vector<int>>> vin(T.n_pass, std::vector<std::vector<int>>(N_q));
    std::vector<std::vector<std::vector<int>>> fin(T.n_pass, std::vector<std::vector<int>>(N_f));
    for (int q = 0; q < T.n_pairs; ++q) {
        const int pass = q / T.pass_size, slot = q % T.pass_size;
        const int i = pk[q] & 0xfff, j = (pk[q] >> 12) & 0xfff;
        if (q < T.n_vol_pairs) {
            vin[pass][i].push_back(2 * slot + 1);  // r_q[i] -= c
            vin[pass][j].push_back(2 * slot);      // r_q[j] += c
        } else {
            vin[pass][i].push_back(2 * slot + 1);  // r_q[i] -= c
            fin[pass][j].push_back(2 * slot + 1);  // f_f[j] -= c
        }
    }
    T.wv = T.wf = 0;
    for (int ps = 0; ps < T.n_pass; ++ps) {
        for (auto& v : vin[ps]) T.wv = std::max(T.wv, (int)v.size());
        for (auto& v : fin[ps]) T.wf = std::max(T.wf, (int)v.size());
    }
    std::vector<int> vt((size_t)T.n_pass * std::max(T.wv, 1) * N_q, -1), ft((size_t)T.n_pass * std::max(T.wf, 1) * N_f, -1);
    for (int ps = 0; ps < T.n_pass; ++ps) {
        for (int o = 0; o < N_q; ++o)
            for (int w = 0; w < (int)vin[ps][o].size(); ++w) vt[((size_t)ps * T.wv + w) * N_q + o] = vin[ps][o][w];
        for (int o = 0; o < N_f; ++o)
            for (int w = 0; w < (int)fin[ps][o].size(); ++w) ft[((size_t)ps * T.wf + w) * N_f + o] = fin[ps][o][w];
    }
    // R and R' as ELL tables
    std::vector<std::vector<std::pair<int, double>>> rrows(N_f), rtrows(N_q);
    for (int j = 0; j < N_f; ++j)
        for (int i = 0; i < N_q; ++i) {
            const double v = D->R[j + (size_t)N_f * i];
            if (v == 0.0) continue;
            rrows[j].push_back({i, v});
            rtrows[i].push_back({j, v});
        }
    std::vector<int> r_idx, rt_idx;
    std::vector<double> r_val, rt_val;
    ell(rrows, T.wr, r_idx, r_val);
    ell(rtrows, T.wrt, rt_idx, rt_val);
    const int P1 = D->V_p + 1;
    std::vector<double> VA(D->V_A, D->V_A + (size_t)D->V_M[0] * P1);
    std::vector<double> VB(D->V_B, D->V_B + (size_t)D->V_M[1] * P1 * P1);
    std::vector<double> VC;
    if (d == 3) VC.assign(D->V_C, D->V_C + (size_t)D->V_M[2] * P1 * P1 * P1);
    int rc;
    if ((rc = upload(h, pk, &T.pk)) || (rc = upload(h, cf, &T.cf)) || (rc = upload(h, vt, &T.vt)) ||
        (rc = upload(h, ft, &T.ft)) || (rc = upload(h, r_idx, &T.r_idx)) || (rc = upload(h, r_val, &T.r_val)) ||
        (rc = upload(h, rt_idx, &T.rt_idx)) || (rc = upload(h, rt_val, &T.rt_val)) || (rc = upload(h, VA, &T.VA)) ||
        (rc = upload(h, VB, &T.VB)))
        return rc;
    if (d == 3 && (rc = upload(h, VC, &T.VC))) return rc;
    T.xcd_remap = 1;
    if (cfg.lds1 > 160 * 1024 || cfg.lds2 > 160 * 1024) return 0;
    if (T.wv > 16 || T.wf > 16 || T.wrt > 10) return 0;  // TENSOR_MAXW, TENSOR_MAXR
    h->use_tensor = true;
    return 0;
}

// Set up the compile-time-shaped StandardForm kernels of scalar linear laws
// (ModalTensor Tri/Tet, weight-adjusted mass solve); returns 0 and leaves
// use_std_tensor false when the discretisation is not one of the compiled shapes.
int build_std_tensor(sse_handle* h, const sse_desc* D) {
    const DevParams& P = h->P;
    if (!P.standard || D->N_c != 1) return 0;
    if (D->law != SSE_LINEAR_ADVECTION && D->law != SSE_LINEAR_ADVECTION_DIFFUSION) return 0;
    if (!(D->element == SSE_TRI || D->element == SSE_TET) || !D->V || D->V_p <= 0 || !D->V_A || !D->V_B) return 0;
    if (D->element == SSE_TET && !D->V_C) return 0;
    if (D->mass_solver != SSE_MASS_WEIGHT_ADJUSTED || D->Minv || D->R_selection || !D->R) return 0;
    const int p = D->V_p, n1 = p + 1;
    for (int m = 0; m < D->d; ++m)
        if (D->V_M[m] != n1) return 0;
    const bool tet = D->element == SSE_TET;
    const int nq = tet ? n1 * n1 * n1 : n1 * n1, nf = tet ? 4 * n1 * n1 : 3 * n1;
    const int np = tet ? n1 * (p + 2) * (p + 3) / 6 : n1 * (p + 2) / 2;
    if (D->N_q != nq || D->N_f != nf || D->N_p != np) return 0;
    const size_t lds = sse::std_tensor_config(D->element, p);
    if (lds == 0 || lds > 160 * 1024) return 0;
    {  // the dense V must be the product of the factors
        std::vector<double> Vw = warped_dense(D);
        if ((int)Vw.size() != nq * np) return 0;
        double scale = 0.0, worst = 0.0;
        for (size_t t = 0; t < Vw.size(); ++t) {
            scale = std::max(scale, std::fabs(D->V[t]));
            worst = std::max(worst, std::fabs(D->V[t] - Vw[t]));
        }
        if (worst > 1e-12 * std::max(scale, 1.0)) return 0;
    }
    sse::TensorParams& T = h->TP;
    T.base = P;
    T.elem = D->element;
    T.P = p;
    T.variant = 0;
    std::vector<std::vector<std::pair<int, double>>> rrows(nf), rtrows(nq);
    for (int j = 0; j < nf; ++j)
        for (int i = 0; i < nq; ++i) {
            const double v = D->R[j + (size_t)nf * i];
            if (v == 0.0) continue;
            rrows[j].push_back({i, v});
            rtrows[i].push_back({j, v});
        }
    std::vector<int> r_idx, rt_idx;
    std::vector<double> r_val, rt_val;
    ell(rrows, T.wr, r_idx, r_val);
    ell(rtrows, T.wrt, rt_idx, rt_val);
    // the kernels read ELL rows four entries at a time
    auto pad4 = [](int& width, int rows, std::vector<int>& idx, std::vector<double>& val) {
        const int w4 = (std::max(width, 1) + 3) / 4 * 4;
        idx.resize((size_t)w4 * rows, -1);
        val.resize((size_t)w4 * rows, 0.0);
        width = w4;
    };
    pad4(T.wr, nf, r_idx, r_val);
    pad4(T.wrt, nq, rt_idx, rt_val);
    std::vector<double> VA(D->V_A, D->V_A + (size_t)n1 * n1);
    std::vector<double> VB(D->V_B, D->V_B + (size_t)n1 * n1 * n1);
    std::vector<double> VC;
    if (tet) VC.assign(D->V_C, D->V_C + (size_t)n1 * n1 * n1 * n1);
    // D_m = sum_l diag(c_lm) D1_l along the lines of the tensor grid: D1_l row a
    // from the line through node a * stride_l (row sums of a derivative vanish),
    // c by least squares, then the whole of every D_m is checked against it
    const int d = D->d;
    int stride[3];
    for (int l = 0; l < d; ++l) {
        stride[l] = 1;
        for (int t = l + 1; t < d; ++t) stride[l] *= n1;
    }
    auto coord = [&](int i, int l) { return (i / stride[l]) % n1; };
    auto Dm = [&](int i, int j, int m) { return D->D[i + (size_t)nq * (j + (size_t)nq * m)]; };
    std::vector<double> D1((size_t)d * n1 * n1, 0.0), c((size_t)nq * d * d, 0.0);
    for (int l = 0; l < d; ++l)
        for (int a = 0; a < n1; ++a) {
            const int i0 = a * stride[l];
            int mbest = -1;
            double best = 0.0;
            for (int m = 0; m < d; ++m) {
                double w = 0.0;
                for (int b2 = 0; b2 < n1; ++b2)
                    if (b2 != a) w += std::fabs(Dm(i0, i0 + (b2 - a) * stride[l], m));
                if (w > best) { best = w; mbest = m; }
            }
            if (mbest < 0) return 0;
            double rs = 0.0;
            for (int b2 = 0; b2 < n1; ++b2)
                if (b2 != a) {
                    const double v = Dm(i0, i0 + (b2 - a) * stride[l], mbest);
                    D1[((size_t)l * n1 + a) * n1 + b2] = v;
                    rs += v;
                }
            D1[((size_t)l * n1 + a) * n1 + a] = -rs;
        }
    double dmax = 0.0;
    for (size_t t = 0; t < (size_t)nq * nq * d; ++t) dmax = std::max(dmax, std::fabs(D->D[t]));
    for (int i = 0; i < nq; ++i)
        for (int l = 0; l < d; ++l) {
            const int a = coord(i, l);
            double den = 0.0;
            for (int b2 = 0; b2 < n1; ++b2)
                if (b2 != a) den += D1[((size_t)l * n1 + a) * n1 + b2] * D1[((size_t)l * n1 + a) * n1 + b2];
            for (int m = 0; m < d; ++m) {
                double num = 0.0;
                for (int b2 = 0; b2 < n1; ++b2)
                    if (b2 != a) num += Dm(i, i + (b2 - a) * stride[l], m) * D1[((size_t)l * n1 + a) * n1 + b2];
                c[((size_t)i * d + l) * d + m] = den > 0.0 ? num / den : 0.0;
            }
        }
    for (int m = 0; m < d; ++m)
        for (int i = 0; i < nq; ++i)
            for (int j = 0; j < nq; ++j) {
                double rec = 0.0;
                for (int l = 0; l < d; ++l) {
                    bool on_line = true;
                    for (int t = 0; t < d; ++t)
                        if (t != l && coord(i, t) != coord(j, t)) on_line = false;
                    if (on_line) rec += c[((size_t)i * d + l) * d + m] * D1[((size_t)l * n1 + coord(i, l)) * n1 + coord(j, l)];
                }
                if (std::fabs(rec - Dm(i, j, m)) > 1e-12 * std::max(dmax, 1.0)) return 0;
            }
    int rc;
    const double* c_dev = nullptr;
    if ((rc = upload(h, r_idx, &T.r_idx)) || (rc = upload(h, r_val, &T.r_val)) || (rc = upload(h, rt_idx, &T.rt_idx)) ||
        (rc = upload(h, rt_val, &T.rt_val)) || (rc = upload(h, VA, &T.VA)) || (rc = upload(h, VB, &T.VB)) ||
        (rc = upload(h, D1, &T.D1)) || (rc = upload(h, c, &c_dev)))
        return rc;
    if (tet && (rc = upload(h, VC, &T.VC))) return rc;
    // the kernels read Lt_ln = sum_m c_lm Lambda_mn in place of Lambda (this handle's copy)
    hipError_t e = sse::collapse_metric(const_cast<double*>(P.Lam), c_dev, nq, d, P.N_e, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("metric set-up: ") + hipGetErrorString(e));
    h->TP.base = h->P;
    T.xcd_remap = 1;
    h->use_std_tensor = true;
    return 0;
}

// Tables of the line-structured time_derivative kernel (requires the tensor
// kernels to be set up). Returns 0 and leaves use_struct false when the
// operators do not have the tensor-line structure.
int build_struct(sse_handle* h, const sse_desc* D) {
    if (!h->use_tensor) return 0;
    sse::StructConfig cfg;
    if (!sse::struct_config(D->element, D->V_p, &cfg)) return 0;
    const int d = D->d, N_q = D->N_q, N_f = D->N_f, N1 = D->V_p + 1, N_e = D->N_e;
    // mapP must pair facet nodes (an involution) for the exterior-state scatter
    for (long long t = 0; t < (long long)N_f * N_e; ++t) {
        const long long g = D->mapP[t] - 1;
        if (D->mapP[g] - 1 != t) return 0;
    }
    int stride[3] = {1, 1, 1};
    for (int f = d - 2; f >= 0; --f) stride[f] = stride[f + 1] * N1;
    // The kernels evaluate each line pair once, at the node that owns it, and give the
    // partner -S_m[i, j]: S must be antisymmetric (it is 1/2 (W D - D' W) in the reference);
    // a descriptor whose S is not is left to the general kernels.
    double smax = 0.0;
    for (size_t t = 0; t < (size_t)N_q * N_q * d; ++t) smax = std::max(smax, std::fabs(D->S[t]));
    for (int m = 0; m < d; ++m)
        for (int i = 0; i < N_q; ++i)
            for (int j = 0; j < i; ++j) {
                const double a = D->S[i + (size_t)N_q * (j + (size_t)N_q * m)];
                const double b = D->S[j + (size_t)N_q * (i + (size_t)N_q * m)];
                if (std::fabs(a + b) > 1e-12 * smax) return 0;
            }
    // S on tensor lines, S_m[i, j] = 0 for m < f (family f); checked against every nonzero
    std::vector<double> SG((size_t)d * N1 * d * N_q, 0.0);
    std::vector<char> covered((size_t)N_q * N_q, 0);
    for (int i = 0; i < N_q; ++i) {
        int a[3], t = i;
        for (int f = d - 1; f >= 0; --f) {
            a[f] = t % N1;
            t /= N1;
        }
        for (int f = 0; f < d; ++f)
            for (int b = 0; b < N1; ++b) {
                if (b == a[f]) continue;
                const int j = i + (b - a[f]) * stride[f];
                covered[(size_t)i * N_q + j] = 1;
                for (int m = 0; m < d; ++m) {
                    const double v = D->S[i + (size_t)N_q * (j + (size_t)N_q * m)];
                    if (m < f && v != 0.0) return 0;
                    SG[(((size_t)f * N1 + b) * d + m) * N_q + i] = v;
                }
            }
    }
    for (int i = 0; i < N_q; ++i)
        for (int j = 0; j < N_q; ++j) {
            if (i == j || covered[(size_t)i * N_q + j]) continue;
            for (int m = 0; m < d; ++m)
                if (D->S[i + (size_t)N_q * (j + (size_t)N_q * m)] != 0.0) return 0;
        }
    // C by column chunks: every nonzero once, grouped by facet node
    std::vector<std::vector<std::pair<int, double>>> ccols(N_f);
    for (int j = 0; j < N_f; ++j)
        for (int i = 0; i < N_q; ++i) {
            const double c = D->C[i + (size_t)N_q * j];
            if (c == 0.0) continue;
            ccols[j].push_back({i, c});
        }
    sse::StructTables& S = h->ST;
    const int ch = N1;
    const int max_item = cfg.max_item;
    std::vector<int> ci_i((size_t)ch * max_item, -1), ci_j(max_item, -1);
    std::vector<double> ci_v((size_t)ch * max_item, 0.0);
    int n_items = 0;
    for (int j = 0; j < N_f; ++j) {
        for (size_t s0 = 0; s0 < ccols[j].size(); s0 += ch) {
            if (n_items >= max_item) return 0;
            for (int w = 0; w < ch && s0 + w < ccols[j].size(); ++w) {
                ci_i[(size_t)w * max_item + n_items] = ccols[j][s0 + w].first;
                ci_v[(size_t)w * max_item + n_items] = ccols[j][s0 + w].second;
            }
            ci_j[n_items] = j;
            ++n_items;
        }
    }
    S.ch = ch;
    S.n_items = n_items;
    // deterministic mode: the rows of C (each volume node evaluates its own) and the first
    // chunk of each facet node (chunks were made in facet-node order)
    std::vector<int> ci_first(N_f + 1, 0);
    for (int j = 0; j < N_f; ++j) ci_first[j + 1] = ci_first[j] + (int)((ccols[j].size() + ch - 1) / ch);
    std::vector<std::vector<std::pair<int, double>>> crows(N_q);
    for (int j = 0; j < N_f; ++j)
        for (const auto& e : ccols[j]) crows[e.first].push_back({j, e.second});
    int rw = 1;
    for (int i = 0; i < N_q; ++i) rw = std::max(rw, (int)crows[i].size());
    std::vector<int> cr_j((size_t)rw * N_q, -1);
    std::vector<double> cr_v((size_t)rw * N_q, 0.0);
    for (int i = 0; i < N_q; ++i)
        for (size_t w = 0; w < crows[i].size(); ++w) {
            cr_j[w * N_q + i] = crows[i][w].first;
            cr_v[w * N_q + i] = crows[i][w].second;
        }
    S.rw = rw;
    S.det = D->deterministic ? 1 : 0;
    int rc;
    std::vector<double> SGh(SG.size());
    for (size_t t = 0; t < SG.size(); ++t) SGh[t] = -0.5 * SG[t];
    if ((rc = upload(h, SG, &S.SG)) || (rc = upload(h, SGh, &S.SGh)) || (rc = upload(h, ci_i, &S.ci_i)) || (rc = upload(h, ci_v, &S.ci_v)) ||
        (rc = upload(h, ci_j, &S.ci_j)) || (rc = upload(h, ci_first, &S.ci_first)) || (rc = upload(h, cr_j, &S.cr_j)) ||
        (rc = upload(h, cr_v, &S.cr_v)))
        return rc;
    double* ext = nullptr;
    if ((rc = alloc(h, (size_t)N_f * D->N_c * N_e, &ext))) return rc;
    S.uf_ext = ext;
    h->TP.uf_ext = ext;
    int dev = 0, n_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return fail(h, SSE_ERR_HIP, "device query failed");
    const int per_cu = sse::struct_occupancy(D->element, D->V_p);
    if (per_cu <= 0) {
        h->TP.uf_ext = nullptr;
        return 0;
    }
    h->struct_blocks = (N_e + cfg.EPB - 1) / cfg.EPB;
    h->use_struct = true;
    return 0;
}

int launch(sse_handle* h, const double* u, double* dudt, const RKStage& rk, int which) {
    if (h->use_struct) {
        hipError_t e = hipSuccess;
        if (which & 1) e = sse::launch_tensor(h->TP, u, dudt, rk, h->stream, 1);
        if (e == hipSuccess && (which & 2)) e = sse::launch_struct(h->TP, h->ST, dudt, rk, h->stream, h->struct_blocks);
        if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
        return 0;
    }
    hipError_t e;
    if (h->use_std_tensor) {
        e = sse::launch_std_tensor(h->TP, 1, nullptr, 0, u, dudt, rk, h->stream, which);
    } else if (h->P.standard) {  // phase 1 without entropy projection, then the StandardForm kernels
        e = (which & 1) ? sse::launch_generic_dispatch(h->P, u, dudt, rk, h->stream, 1) : hipSuccess;
        if (e == hipSuccess) e = sse::launch_standard(h->P, dudt, rk, h->stream, which);
    } else {
        e = h->use_tensor ? sse::launch_tensor(h->TP, u, dudt, rk, h->stream, which)
                          : sse::launch_generic_dispatch(h->P, u, dudt, rk, h->stream, which);
    }
    if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return 0;
}

// One launch over R members of an ensemble handle (R <= h->members); u, dudt and
// the RK buffers hold the R members back to back.
int launch_ens(sse_handle* h, const double* u, double* dudt, const RKStage& rk, int which, int R, const double* b,
               int tangent) {
    if (!h->use_std_tensor || R > h->members) return fail(h, SSE_ERR_UNSUPPORTED, "not an ensemble handle");
    hipError_t e = sse::launch_std_tensor(h->TP, R, b, tangent, u, dudt, rk, h->stream, which);
    if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return 0;
}

int check_status(sse_handle* h) {
    int st = 0;
    SSE_HIP_TRY(h, hipMemcpyAsync(&st, h->P.status, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    SSE_HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (st) {
        SSE_HIP_TRY(h, hipMemsetAsync(h->P.status, 0, sizeof(int), h->stream));
        SSE_HIP_TRY(h, hipStreamSynchronize(h->stream));
        return fail(h, SSE_ERR_NONPHYSICAL_STATE, "non-physical state (density or pressure <= 0) in the residual");
    }
    return 0;
}

size_t state_size(const sse_handle* h) { return (size_t)h->P.N_p * h->P.N_c * h->P.N_e; }

}  // namespace

namespace sse {
// error of a call that has no handle (set-up entries, failed creation)
int set_global_error(int code, const std::string& msg) { return fail(nullptr, code, msg); }
}  // namespace sse

extern "C" {

int32_t sse_abi_version(void) { return SSE_ABI_VERSION; }

const char* sse_last_error(const sse_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

}  // extern "C"

namespace {

// A handle for `members` problems that share the descriptor's discretisation and
// mesh (members > 1 needs the compile-time-shaped StandardForm kernels).
int create_handle(const sse_desc* D, int members, sse_handle** out) {
    if (!out) return fail(nullptr, SSE_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    std::string msg;
    int code = validate(D, msg);
    if (!code && D->form == SSE_FLUX_DIFFERENCING) code = check_nJq(D, msg);  // StandardForm does not use nJq
    if (code) return fail(nullptr, code, msg);

    sse_handle* h = new sse_handle();
    DevParams& P = h->P;
    const int d = D->d;
    P.d = d; P.N_p = D->N_p; P.N_q = D->N_q; P.N_f = D->N_f; P.N_fac = D->N_fac;
    P.npf = D->N_f / D->N_fac; P.N_c = D->N_c; P.N_e = D->N_e;
    P.has_V = D->V != nullptr;
    P.R_sel = D->R_selection != nullptr;
    P.standard = D->form == SSE_STANDARD;
    P.skew = D->mapping_form != SSE_STANDARD_MAPPING;
    P.second_order = D->law == SSE_LINEAR_ADVECTION_DIFFUSION;
    P.b = D->b;
    // entropy projection only in the flux-differencing form (flux_differencing_form.jl:253-292)
    P.project = !P.standard && (D->N_c > 1) && !(!P.has_V && P.R_sel);
    P.C_kind = D->C_kind;
    P.mass_diag = D->mass_solver == SSE_MASS_DIAGONAL;
    P.law.gamma = D->gamma;
    P.law.gm1 = D->gamma - 1.0;
    P.law.inv_gm1 = D->law == SSE_EULER ? 1.0 / (D->gamma - 1.0) : 0.0;
    P.law.log_gm1 = D->law == SSE_EULER ? std::log(D->gamma - 1.0) : 0.0;
    for (int m = 0; m < 3; ++m) P.law.a[m] = D->a[m];
    P.law.half_lambda = D->half_lambda;
    P.law.inviscid_flux = D->inviscid_flux;
    P.law.two_point_flux = D->two_point_flux;
    h->element = D->element;

    int rc = 0;
    auto cleanup_fail = [&](int c) {
        g_create_error = h->err;
        sse_destroy(h);
        return c;
    };
    const int N_q = D->N_q, N_f = D->N_f, N_p = D->N_p, N_e = D->N_e;
    // operators
    if (P.has_V) {
        std::vector<double> V(D->V, D->V + (size_t)N_q * N_p), Vt((size_t)N_q * N_p);
        for (int i = 0; i < N_q; ++i)
            for (int j = 0; j < N_p; ++j) Vt[j + (size_t)N_p * i] = V[i + (size_t)N_q * j];
        if ((rc = upload(h, V, &P.V)) || (rc = upload(h, Vt, &P.Vt))) return cleanup_fail(rc);
    }
    Rows R, Rt;
    if (D->R_selection) {
        std::vector<double> Rd((size_t)N_f * N_q, 0.0);
        for (int j = 0; j < N_f; ++j) Rd[j + (size_t)N_f * (D->R_selection[j] - 1)] = 1.0;
        R = rows_of(Rd.data(), N_f, N_q);
        Rt = cols_of(Rd.data(), N_f, N_q);
    } else {
        R = rows_of(D->R, N_f, N_q);
        Rt = cols_of(D->R, N_f, N_q);
    }
    if ((rc = upload(h, R.ptr, &P.R_ptr)) || (rc = upload(h, R.idx, &P.R_idx)) || (rc = upload(h, R.val, &P.R_val)) ||
        (rc = upload(h, Rt.ptr, &P.Rt_ptr)) || (rc = upload(h, Rt.idx, &P.Rt_idx)) ||
        (rc = upload(h, Rt.val, &P.Rt_val)))
        return cleanup_fail(rc);
    // the matrix-core nodal values of the general path (Euler with a modal V) when the
    // element data of 16 elements fit the LDS of a compute unit, for meshes of at least
    // 4,096 elements (one workgroup of 16 per CU and more; on fewer elements the vector
    // kernel, one element per workgroup, is faster). SSE_NODAL_MFMA=0 or 1 overrides.
    if (P.project && P.has_V && D->law == SSE_EULER && D->d >= 2 && sse::lds_nodal_values_mfma(P) <= 160 * 1024) {
        const char* env = getenv("SSE_NODAL_MFMA");
        const bool on = env && env[0] ? env[0] != '0' : N_e >= 4096;
        if (on) {
            std::vector<double> Rd((size_t)N_f * N_q, 0.0);
            if (D->R_selection) {
                for (int j = 0; j < N_f; ++j) Rd[j + (size_t)N_f * (D->R_selection[j] - 1)] = 1.0;
            } else {
                std::copy(D->R, D->R + (size_t)N_f * N_q, Rd.begin());
            }
            if ((rc = upload(h, Rd, &P.Rd))) return cleanup_fail(rc);
            P.mfma = 1;
        }
    }
    // union sparsity by rows of d stacked N_q x N_q operators A_m (column-major,
    // A[i + N_q (j + N_q m)]), optionally transposed, with or without the diagonal
    auto union_rows = [&](const double* A, bool transpose, bool skip_diag, const int** p_ptr, const int** p_idx,
                          const double** p_val) {
        std::vector<int> ptr(1, 0), idx;
        std::vector<double> val;
        auto at = [&](int i, int j, int m) {
            return transpose ? A[j + (size_t)N_q * (i + (size_t)N_q * m)] : A[i + (size_t)N_q * (j + (size_t)N_q * m)];
        };
        for (int i = 0; i < N_q; ++i) {
            for (int j = 0; j < N_q; ++j) {
                if (skip_diag && j == i) continue;
                bool nz = false;
                for (int m = 0; m < d; ++m) nz |= at(i, j, m) != 0.0;
                if (!nz) continue;
                idx.push_back(j);
                for (int m = 0; m < d; ++m) val.push_back(at(i, j, m));
            }
            ptr.push_back((int)idx.size());
        }
        int r;
        if ((r = upload(h, ptr, p_ptr)) || (r = upload(h, idx, p_idx)) || (r = upload(h, val, p_val))) return r;
        return 0;
    };
    if (P.standard) {  // rows of D_m and of D_m' (StandardForm volume operators)
        if ((rc = union_rows(D->D, false, false, &P.D_ptr, &P.D_idx, &P.D_val)) ||
            (rc = union_rows(D->D, true, false, &P.Dt_ptr, &P.Dt_idx, &P.Dt_val)))
            return cleanup_fail(rc);
    } else if ((rc = union_rows(D->S, false, true, &P.S_ptr, &P.S_idx, &P.S_val))) {
        return cleanup_fail(rc);
    }
    if (!P.standard) {  // flux work items: volume pairs i < j, then the nonzeros of C
        if (N_q >= 32768 || N_f >= 32768) {  // item_ij packs i | j << 16 into an int
            h->err = "too many nodes per element";
            return cleanup_fail(SSE_ERR_UNSUPPORTED);
        }
        std::vector<int> ij;
        std::vector<double> val;
        for (int i = 0; i < N_q; ++i)
            for (int j = i + 1; j < N_q; ++j) {
                bool nz = false;
                for (int m = 0; m < d; ++m) nz |= D->S[i + (size_t)N_q * (j + (size_t)N_q * m)] != 0.0;
                if (!nz) continue;
                ij.push_back(i | j << 16);
                for (int m = 0; m < d; ++m) val.push_back(D->S[i + (size_t)N_q * (j + (size_t)N_q * m)]);
            }
        P.n_vol_items = (int)ij.size();
        if (D->C_kind != SSE_C_NONE)
            for (int i = 0; i < N_q; ++i)
                for (int j = 0; j < N_f; ++j) {
                    const double c = D->C[i + (size_t)N_q * j];
                    if (c == 0.0) continue;
                    ij.push_back(i | j << 16);
                    val.push_back(c);
                    for (int m = 1; m < d; ++m) val.push_back(0.0);
                }
        P.n_items = (int)ij.size();
        // Order the items so that consecutive ones (the lanes of a wavefront) touch
        // different nodes: by the cyclic offset between their two nodes, then by
        // node. Row-major order would send the atomic adds of a whole wavefront to
        // one address.
        {
            std::vector<int> perm(ij.size());
            for (size_t t = 0; t < perm.size(); ++t) perm[t] = (int)t;
            auto key = [&](int t) {
                const int i = ij[t] & 0xffff, j = ij[t] >> 16;
                const bool vol = t < P.n_vol_items;
                const long long off = vol ? ((j - i) % N_q + N_q) % N_q : ((i - j) % N_q + N_q) % N_q;
                return std::make_tuple(vol ? 0 : 1, off, vol ? i : j);
            };
            std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key(a) < key(b); });
            std::vector<int> ij2(ij.size());
            std::vector<double> val2(val.size());
            for (size_t t = 0; t < perm.size(); ++t) {
                ij2[t] = ij[perm[t]];
                for (int m = 0; m < d; ++m) val2[t * d + m] = val[(size_t)perm[t] * d + m];
            }
            ij.swap(ij2);
            val.swap(val2);
        }
        if (!ij.empty() && ((rc = upload(h, ij, &P.item_ij)) || (rc = upload(h, val, &P.item_val))))
            return cleanup_fail(rc);
        // deterministic mode: each node's items, in item order (volume nodes, then facet nodes)
        P.det = D->deterministic ? 1 : 0;
        if (P.det) {
            std::vector<std::vector<int>> own((size_t)N_q + N_f);
            for (int c = 0; c < (int)ij.size(); ++c) {
                const int i = ij[c] & 0xffff, j = ij[c] >> 16;
                own[i].push_back(c << 1);
                if (c < P.n_vol_items)
                    own[j].push_back(c << 1 | 1);
                else
                    own[(size_t)N_q + j].push_back(c << 1);
            }
            std::vector<int> optr((size_t)N_q + N_f + 1, 0), oitem;
            for (size_t n = 0; n < own.size(); ++n) {
                optr[n + 1] = optr[n] + (int)own[n].size();
                oitem.insert(oitem.end(), own[n].begin(), own[n].end());
            }
            if ((rc = upload(h, optr, &P.own_ptr)) || (!oitem.empty() && (rc = upload(h, oitem, &P.own_item))))
                return cleanup_fail(rc);
        }
    }
    if (D->C_kind != SSE_C_NONE) {
        Rows Cr = rows_of(D->C, N_q, N_f), Cc = cols_of(D->C, N_q, N_f);
        if ((rc = upload(h, Cr.ptr, &P.Cr_ptr)) || (rc = upload(h, Cr.idx, &P.Cr_idx)) ||
            (rc = upload(h, Cr.val, &P.Cr_val)) || (rc = upload(h, Cc.ptr, &P.Cc_ptr)) ||
            (rc = upload(h, Cc.idx, &P.Cc_idx)) || (rc = upload(h, Cc.val, &P.Cc_val)))
            return cleanup_fail(rc);
    }
    {
        std::vector<double> W(D->W, D->W + N_q), B(D->B, D->B + N_f);
        std::vector<double> fn(D->face_normals, D->face_normals + (size_t)d * D->N_fac);
        std::vector<double> Jq(D->J_q, D->J_q + (size_t)N_q * N_e);
        std::vector<double> Lam(D->Lambda_q, D->Lambda_q + (size_t)N_q * d * d * N_e);
        std::vector<double> nJf(D->nJf, D->nJf + (size_t)d * N_f * N_e);
        std::vector<double> Jf(D->J_f, D->J_f + (size_t)N_f * N_e);
        std::vector<int> mapP((size_t)N_f * N_e);
        for (size_t t = 0; t < mapP.size(); ++t) mapP[t] = (int)(D->mapP[t] - 1);
        if ((rc = upload(h, W, &P.W)) || (rc = upload(h, B, &P.B)) || (rc = upload(h, fn, &P.fn)) ||
            (rc = upload(h, Jq, &P.Jq)) || (rc = upload(h, Lam, &P.Lam)) || (rc = upload(h, nJf, &P.nJf)) ||
            (rc = upload(h, Jf, &P.Jf)) || (rc = upload(h, mapP, &P.mapP)))
            return cleanup_fail(rc);
    }
    const size_t ens = (size_t)members;
    h->members = members;
    if ((rc = alloc(h, (size_t)N_q * D->N_c * N_e * ens, &P.uq)) ||
        (rc = alloc(h, (size_t)N_f * D->N_c * N_e * ens, &P.uf)) || (rc = alloc(h, 1, &P.status)))
        return cleanup_fail(rc);
    if (hipMemset(P.status, 0, sizeof(int)) != hipSuccess) {
        h->err = "hipMemset failed";
        return cleanup_fail(SSE_ERR_HIP);
    }
    if (P.second_order &&
        ((rc = alloc(h, (size_t)N_q * d * N_e * ens, &P.qq)) || (rc = alloc(h, (size_t)N_f * d * N_e * ens, &P.qf))))
        return cleanup_fail(rc);
    if (sse::lds_time_derivative_generic(P) > 64 * 1024 || sse::lds_nodal_values_generic(P) > 64 * 1024 ||
        (P.standard && sse::lds_standard(P) > 64 * 1024)) {
        h->err = "element too large for the LDS-staged kernels";
        return cleanup_fail(SSE_ERR_UNSUPPORTED);
    }
    if (!P.standard && (rc = build_tensor(h, D))) return cleanup_fail(rc);
    if (P.standard && (rc = build_std_tensor(h, D))) return cleanup_fail(rc);
    const char* kenv = getenv("SSE_KERNELS");  // "generic", "tensor" (table-driven pairs) or unset
    if (!(kenv && (!strcmp(kenv, "generic") || !strcmp(kenv, "tensor"))) && (rc = build_struct(h, D)))
        return cleanup_fail(rc);
    if (const char* env = getenv("SSE_FORCE_GENERIC")) {
        if (env[0] == '1') h->use_tensor = false;
    }
    if (kenv && !strcmp(kenv, "generic")) h->use_tensor = h->use_std_tensor = false;
    if (!h->use_tensor) h->use_struct = false;
    // the tensor kernels' weight-adjusted mass solve with a kept matrix per element
    // (N_p^2 doubles each; up to 16 GiB) in place of V, W/J and V' (SSE_MASS_MATRICES=0: off)
    if (h->use_tensor && P.has_V && (size_t)N_p * N_p * N_e * sizeof(double) <= ((size_t)16 << 30)) {
        const char* menv = getenv("SSE_MASS_MATRICES");
        if (!(menv && menv[0] == '0')) {
            double* Mw = nullptr;
            if ((rc = alloc(h, (size_t)N_p * N_p * N_e, &Mw))) return cleanup_fail(rc);
            hipError_t e = sse::launch_weight_adjusted_inverse(P, Mw, nullptr);
            if (e == hipSuccess) e = hipDeviceSynchronize();
            if (e != hipSuccess) {
                h->err = std::string("weight-adjusted inverse: ") + hipGetErrorString(e);
                return cleanup_fail(SSE_ERR_HIP);
            }
            h->TP.Mw = Mw;
        }
    }
    if (members > 1 && !h->use_std_tensor) {
        h->err = "an ensemble of problems needs the compiled StandardForm shapes of scalar laws";
        return cleanup_fail(SSE_ERR_UNSUPPORTED);
    }
    if (hipStreamCreate(&h->own_stream) != hipSuccess) {
        h->err = "hipStreamCreate failed";
        return cleanup_fail(SSE_ERR_HIP);
    }
    h->stream = h->own_stream;
    // set-up of sse_functional_noalloc: this device's address of its workspace
    if (hipError_t e = workspace_address(&h->noalloc_ws); e != hipSuccess)
        h->noalloc_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    *out = h;
    return 0;
}

}  // namespace

extern "C" {

int sse_create(const sse_desc* D, sse_handle** out) { return create_handle(D, 1, out); }

int sse_describe(const sse_handle* h, char* buf, int32_t len) {
    if (!h || !buf || len <= 0) return SSE_ERR_INVALID_ARGUMENT;
    const DevParams& P = h->P;
    if (h->use_struct) {
        const sse::TensorParams& T = h->TP;
        const sse::StructTables& S = h->ST;
        snprintf(buf, len, "struct elem=%d P=%d variant=%d blocks=%d items=%d ch=%d wrt=%d%s", T.elem, T.P, T.variant,
                 h->struct_blocks, S.n_items, S.ch, T.wrt, S.det ? " deterministic" : "");
    } else if (h->use_tensor) {
        const sse::TensorParams& T = h->TP;
        snprintf(buf, len, "tensor elem=%d P=%d variant=%d pairs=%d vol_pairs=%d passes=%d pass_size=%d wv=%d wf=%d wrt=%d wr=%d",
                 T.elem, T.P, T.variant, T.n_pairs, T.n_vol_pairs, T.n_pass, T.pass_size, T.wv, T.wf, T.wrt, T.wr);
    } else if (h->use_std_tensor) {
        snprintf(buf, len, "standard tensor elem=%d P=%d N_e=%d members=%d mapping=%s%s wr=%d wrt=%d", h->TP.elem,
                 h->TP.P, P.N_e, h->members, P.skew ? "skew" : "standard", P.second_order ? " br1" : "", h->TP.wr,
                 h->TP.wrt);
    } else if (P.standard) {
        snprintf(buf, len, "standard d=%d N_p=%d N_q=%d N_f=%d N_c=%d N_e=%d mapping=%s%s", P.d, P.N_p, P.N_q, P.N_f,
                 P.N_c, P.N_e, P.skew ? "skew" : "standard", P.second_order ? " br1" : "");
    } else {
        snprintf(buf, len, "generic d=%d N_p=%d N_q=%d N_f=%d N_c=%d N_e=%d%s%s", P.d, P.N_p, P.N_q, P.N_f, P.N_c,
                 P.N_e, P.mfma ? " nodal_values=mfma" : "", P.det ? " deterministic" : "");
    }
    return 0;
}

int sse_l2_error(sse_handle* h, const double* u, const double* u_exact, int32_t normalize, int32_t where,
                 double* out) {
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    if (!u || !u_exact || !out) return fail(h, SSE_ERR_INVALID_ARGUMENT, "u, u_exact or out is NULL");
    if (where != SSE_DEVICE && where != SSE_HOST) return fail(h, SSE_ERR_INVALID_ARGUMENT, "bad `where`");
    if (h->members != 1) return fail(h, SSE_ERR_UNSUPPORTED, "the error of an ensemble handle");
    if (sse::lds_functionals(h->P, 6) > 64 * 1024) return fail(h, SSE_ERR_UNSUPPORTED, "element too large for the functionals");
    const size_t n = state_size(h), nq = (size_t)h->P.N_q * h->P.N_c * h->P.N_e;
    int rc;
    if ((rc = functional_buffers(h))) return rc;
    const double* ud = u;
    const double* xd = u_exact;
    if (where == SSE_HOST) {
        if (!h->stage_u && ((rc = alloc(h, n, &h->stage_u)) || (rc = alloc(h, n, &h->stage_dudt)))) return rc;
        if (!h->ferr && (rc = alloc(h, nq, &h->ferr))) return rc;
        SSE_HIP_TRY(h, hipMemcpyAsync(h->stage_u, u, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        SSE_HIP_TRY(h, hipMemcpyAsync(h->ferr, u_exact, nq * sizeof(double), hipMemcpyHostToDevice, h->stream));
        ud = h->stage_u;
        xd = h->ferr;
    }
    hipError_t e = sse::launch_functionals(h->P, ud, xd, 6, h->fpartial, h->fout, h->stream);
    if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    const int nc = h->P.N_c;
    std::vector<double> sums(nc + 1);
    SSE_HIP_TRY(h, hipMemcpyAsync(sums.data(), h->fout, (nc + 1) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SSE_HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (int c = 0; c < nc; ++c) out[c] = std::sqrt(normalize ? sums[c] / sums[nc] : sums[c]);
    return 0;
}

int sse_functional(sse_handle* h, const double* u, const double* dudt, int32_t which, int32_t where, double* out) {
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    if (!u || !out) return fail(h, SSE_ERR_INVALID_ARGUMENT, "u or out is NULL");
    if (which < SSE_MASS || which > SSE_ENTROPY) return fail(h, SSE_ERR_INVALID_ARGUMENT, "unknown functional");
    if (where != SSE_DEVICE && where != SSE_HOST) return fail(h, SSE_ERR_INVALID_ARGUMENT, "bad `where`");
    if (sse::lds_functionals(h->P, which + (dudt ? 3 : 0)) > 64 * 1024) return fail(h, SSE_ERR_UNSUPPORTED, "element too large for the functionals");
    const size_t n = state_size(h);
    int rc;
    if ((rc = functional_buffers(h))) return rc;
    const double* ud = u;
    const double* vd = dudt;
    if (where == SSE_HOST) {
        if (!h->stage_u && ((rc = alloc(h, n, &h->stage_u)) || (rc = alloc(h, n, &h->stage_dudt)))) return rc;
        SSE_HIP_TRY(h, hipMemcpyAsync(h->stage_u, u, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        ud = h->stage_u;
        if (dudt) {
            SSE_HIP_TRY(h, hipMemcpyAsync(h->stage_dudt, dudt, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
            vd = h->stage_dudt;
        }
    }
    const int kind = which + (dudt ? 3 : 0);
    // the mass-matrix factors depend only on the handle's geometry: kept after the first energy call
    if (which == SSE_ENERGY && h->P.has_V && !h->fmass &&
        (rc = alloc(h, (size_t)h->P.N_p * h->P.N_p * h->P.N_e, &h->fmass)))
        return rc;
    double* mass = which == SSE_ENERGY ? h->fmass : nullptr;
    hipError_t e = sse::launch_functionals(h->P, ud, vd, kind, h->fpartial, h->fout, h->stream, mass,
                                           mass && h->fmass_ready ? 1 : 0);
    if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    if (mass) h->fmass_ready = true;  // stream order: later calls read M_k after this one wrote it
    const int nout = which == SSE_ENTROPY ? 1 : h->P.N_c;
    SSE_HIP_TRY(h, hipMemcpyAsync(out, h->fout, nout * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SSE_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return 0;
}

int sse_functional_noalloc(sse_handle* h, const double* u, const double* dudt, int32_t which, double* out) {
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    if (!u || !out) return fail(h, SSE_ERR_INVALID_ARGUMENT, "u or out is NULL");
    if (which < SSE_MASS || which > SSE_ENTROPY) return fail(h, SSE_ERR_INVALID_ARGUMENT, "unknown functional");
    if (sse::lds_functionals(h->P, which + (dudt ? 3 : 0)) > 64 * 1024) return fail(h, SSE_ERR_UNSUPPORTED, "element too large for the functionals");
    const size_t ncomp = h->P.N_c, need = (size_t)h->P.N_e * ncomp + ncomp;
    if (need > (size_t)SSE_FUNCTIONAL_NOALLOC_MAX_DOUBLES)
        return fail(h, SSE_ERR_INVALID_ARGUMENT,
                    "sse_functional_noalloc: N_e * N_c + N_c = " + std::to_string(need) +
                        " exceeds its workspace of " + std::to_string(SSE_FUNCTIONAL_NOALLOC_MAX_DOUBLES) +
                        " doubles; use sse_functional");
    if (!h->noalloc_ws) return fail(h, SSE_ERR_HIP, "sse_functional_noalloc: " + h->noalloc_err);
    double* partial = h->noalloc_ws;                     // N_e x N_c, every entry written by the first kernel
    double* sums = h->noalloc_ws + (size_t)h->P.N_e * ncomp;  // N_c totals, written by the second
    const int kind = which + (dudt ? 3 : 0);
    // energy: the mass-matrix factors are formed in LDS on every call (no per-handle cache)
    hipError_t e = sse::launch_functionals(h->P, u, dudt, kind, partial, sums, h->stream, nullptr, 0);
    if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    const int nout = which == SSE_ENTROPY ? 1 : h->P.N_c;
    SSE_HIP_TRY(h, hipMemcpyAsync(out, sums, nout * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    return 0;
}

}  // extern "C"

namespace {

int vjp_check(sse_handle* h) {
    const DevParams& P = h->P;
    if (P.standard || P.N_c == 1) return fail(h, SSE_ERR_UNSUPPORTED, "the VJP covers the Euler flux-differencing residual");
    if (P.law.two_point_flux != SSE_TWO_POINT_EC)
        return fail(h, SSE_ERR_UNSUPPORTED, "the VJP covers the entropy-conservative two-point flux");
    // compile-time-shaped VJP on the line-structured tables of the forward kernel, where it
    // fits; the general kernels otherwise (or with SSE_VJP=generic)
    const char* env = getenv("SSE_VJP");
    const size_t lds = h->use_struct ? sse::vjp_tensor_config(h->TP.elem, h->TP.P) : 0;
    const bool tensor = lds > 0 && lds <= 160 * 1024 && !(env && !strcmp(env, "generic"));
    if (!tensor && sse::lds_vjp_v2(P) > 160 * 1024)
        return fail(h, SSE_ERR_UNSUPPORTED, "element too large for the VJP kernels");
    h->vjp_tensor = tensor;
    const size_t Ne = P.N_e;
    int rc;
    if (!h->vjp_rbar && ((rc = alloc(h, (size_t)P.N_q * P.N_c * Ne, &h->vjp_rbar)) ||
                         (rc = alloc(h, (size_t)P.N_f * P.N_c * Ne, &h->vjp_fbar)) || (rc = alloc(h, Ne, &h->vjp_lam)) ||
                         (rc = alloc(h, 1, &h->vjp_lamout))))
        return rc;
    if (h->vjp_tensor && !h->vjp_fext && (rc = alloc(h, (size_t)P.N_f * P.N_c * Ne, &h->vjp_fext))) return rc;
    return 0;
}

// Phase-1 outputs of the tensor kernels (projected volume and facet states and the
// exterior facet states): the handle's own buffers, or per-stage copies the RK
// adjoint keeps from its recomputation.
struct Phase1Bufs {
    double* uq;
    double* uf;
    double* ufx;
};

void with_phase1(const sse_handle* h, const Phase1Bufs& b, sse::TensorParams& tp, sse::StructTables& st) {
    tp = h->TP;
    st = h->ST;
    tp.base.uq = b.uq;
    tp.base.uf = b.uf;
    tp.uf_ext = b.ufx;
    st.uf_ext = b.ufx;
}

// One RK stage of the line-structured kernels with phase 1 written to b
int launch_struct_into(sse_handle* h, double* u, const RKStage& rk, const Phase1Bufs& b) {
    sse::TensorParams tp;
    sse::StructTables st;
    with_phase1(h, b, tp, st);
    hipError_t e = sse::launch_tensor(tp, u, nullptr, rk, h->stream, 1);
    if (e == hipSuccess) e = sse::launch_struct(tp, st, nullptr, rk, h->stream, h->struct_blocks);
    if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return 0;
}

// Tensor VJP: ubar (+)= scale (dR/du)' y, the element λ partials (+)= scale (dR/dλ)' y, with
// y = ybar, or y = a ybar + b x stored back into ybar when x is set (the RK adjoint's
// kbar update). With `stored`, phase 1 of u already sits in those buffers; otherwise it
// is run here.
int vjp_tensor_device(sse_handle* h, const double* u, double* ybar, double* ubar, double scale, int accumulate,
                      const Phase1Bufs* stored, const double* x = nullptr, double a = 1.0, double b = 0.0) {
    Phase1Bufs own{h->TP.base.uq, h->TP.base.uf, h->TP.uf_ext};
    if (!stored) {
        int rc = launch(h, u, nullptr, RKStage{}, 1);
        if (rc) return rc;
    }
    sse::TensorParams tp;
    sse::StructTables st;
    with_phase1(h, stored ? *stored : own, tp, st);
    hipError_t e = sse::launch_vjp_tensor(tp, st, u, ybar, ubar, h->vjp_rbar, h->vjp_fbar, h->vjp_fext, h->vjp_lam,
                                          scale, accumulate, x, a, b, x ? ybar : nullptr, h->stream);
    if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return 0;
}

// ubar = (dR/du)' ybar on device buffers; lam_out (device) gets or accumulates scale * (dR/dλ)' ybar
int vjp_device(sse_handle* h, const double* u, const double* ybar, double* ubar, double lam_scale, int accumulate) {
    hipError_t e = hipSuccess;
    if (h->vjp_tensor) {
        int rc = vjp_tensor_device(h, u, const_cast<double*>(ybar), ubar, 1.0, 0, nullptr);  // ybar only read
        if (rc) return rc;
        e = sse::launch_vjp_sum(h->vjp_lam, h->P.N_e, h->vjp_lamout, lam_scale, accumulate, h->stream);
    } else {
        int rc = launch(h, u, nullptr, RKStage{}, 1);  // u tilde: the handle's own phase 1
        if (rc) return rc;
        e = sse::launch_vjp(h->P, u, ybar, ubar, h->vjp_rbar, h->vjp_fbar, h->vjp_lam, h->vjp_lamout, lam_scale,
                            accumulate, h->stream);
    }
    if (e != hipSuccess) return fail(h, SSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return 0;
}

}  // namespace

extern "C" {

int sse_residual_vjp(sse_handle* h, const double* u, const double* ybar, double* ubar, double* lambda_bar,
                     int32_t where) {
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    if (!u || !ybar || !ubar || !lambda_bar) return fail(h, SSE_ERR_INVALID_ARGUMENT, "NULL argument");
    if (where != SSE_DEVICE && where != SSE_HOST) return fail(h, SSE_ERR_INVALID_ARGUMENT, "bad `where`");
    int rc;
    if ((rc = vjp_check(h))) return rc;
    const size_t n = state_size(h);
    const double *ud = u, *yd = ybar;
    double* od = ubar;
    if (where == SSE_HOST) {
        if (!h->stage_u && ((rc = alloc(h, n, &h->stage_u)) || (rc = alloc(h, n, &h->stage_dudt)))) return rc;
        if (!h->rk_tmp && (rc = alloc(h, n, &h->rk_tmp))) return rc;
        SSE_HIP_TRY(h, hipMemcpyAsync(h->stage_u, u, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        SSE_HIP_TRY(h, hipMemcpyAsync(h->stage_dudt, ybar, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        ud = h->stage_u;
        yd = h->stage_dudt;
        od = h->rk_tmp;
    }
    if ((rc = vjp_device(h, ud, yd, od, 1.0, 0))) return rc;
    if (where == SSE_HOST)
        SSE_HIP_TRY(h, hipMemcpyAsync(ubar, od, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SSE_HIP_TRY(h, hipMemcpyAsync(lambda_bar, h->vjp_lamout, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SSE_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return check_status(h);
}

int sse_rk_adjoint(sse_handle* h, const double* u0, double dt, int32_t nsteps, int32_t checkpoint_every,
                   int32_t objective, double* J, double* grad_u0, double* grad_lambda, int32_t where) {
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    if (!u0 || !J || !grad_u0 || !grad_lambda) return fail(h, SSE_ERR_INVALID_ARGUMENT, "NULL argument");
    if (nsteps < 0 || checkpoint_every <= 0) return fail(h, SSE_ERR_INVALID_ARGUMENT, "bad nsteps or checkpoint_every");
    if (objective != SSE_ENTROPY) return fail(h, SSE_ERR_UNSUPPORTED, "the adjoint objective is the entropy integral");
    if (where != SSE_DEVICE && where != SSE_HOST) return fail(h, SSE_ERR_INVALID_ARGUMENT, "bad `where`");
    int rc;
    if ((rc = vjp_check(h))) return rc;
    const size_t n = state_size(h);
    const int nseg = (nsteps + checkpoint_every - 1) / checkpoint_every;
    const int seglen = nsteps < checkpoint_every ? nsteps : checkpoint_every;
    // work buffers, freed at the end of the call
    std::vector<void*> mine;
    auto dalloc = [&](size_t count, double** p) -> int {
        void* q = nullptr;
        if (hipMalloc(&q, count * sizeof(double)) != hipSuccess) return fail(h, SSE_ERR_OUT_OF_MEMORY, "hipMalloc failed");
        mine.push_back(q);
        *p = static_cast<double*>(q);
        return 0;
    };
    const hipStream_t ms = h->stream;  // the caller's stream: forward and reverse sweeps
    hipEvent_t ev_fwd = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_rev[2] = {nullptr, nullptr};
    auto release = [&]() {
        h->stream = ms;
        if (h->adj_stream) (void)hipStreamSynchronize(h->adj_stream);
        (void)hipStreamSynchronize(ms);
        for (hipEvent_t e : {ev_fwd, ev_rec[0], ev_rec[1], ev_rev[0], ev_rev[1]})
            if (e) (void)hipEventDestroy(e);
        for (void* q : mine) (void)hipFree(q);
    };
    double *u = nullptr, *ktmp = nullptr, *ub = nullptr, *kb = nullptr, *vb = nullptr, *ckpt = nullptr,
           *stages = nullptr, *p1 = nullptr;
    if ((rc = dalloc(n, &u)) || (rc = dalloc(n, &ktmp)) || (rc = dalloc(n, &ub)) || (rc = dalloc(n, &kb)) ||
        (rc = dalloc(n, &vb)) || (rc = dalloc(n * (size_t)(nseg > 0 ? nseg : 1), &ckpt)) ||
        (rc = dalloc(n * 5 * (size_t)(seglen > 0 ? seglen : 1), &stages))) {
        release();
        return rc;
    }
    // tensor VJP: keep each recomputed stage's phase-1 outputs for its VJP (when they fit
    // in 32 GiB), so that the reverse sweep does not run phase 1 again
    const bool fast = h->vjp_tensor;
    const size_t p1q = (size_t)h->P.N_q * h->P.N_c * h->P.N_e, p1f = (size_t)h->P.N_f * h->P.N_c * h->P.N_e;
    const size_t p1_stage = p1q + 2 * p1f, n_stages = 5 * (size_t)(seglen > 0 ? seglen : 1);
    const bool keep_p1 = fast && p1_stage * n_stages * sizeof(double) <= ((size_t)32 << 30);
    if (keep_p1 && (rc = dalloc(p1_stage * n_stages, &p1))) {
        release();
        return rc;
    }
    // With a second set of segment buffers, segment s - 1 is recomputed on a second stream
    // while the reverse sweep runs through segment s (SSE_ADJ_OVERLAP=0: one stream). Each
    // buffer set is reused every other segment; events order a recomputation after the
    // reverse sweep that last read its buffers, and a reverse sweep after its recomputation.
    struct SegBufs {
        double* stages;
        double* p1;
    };
    SegBufs bufs[2] = {{stages, p1}, {nullptr, nullptr}};
    double *u2 = nullptr, *ktmp2 = nullptr;
    bool overlap = keep_p1 && nseg > 1;
    if (const char* env = getenv("SSE_ADJ_OVERLAP")) overlap = overlap && env[0] != '0';
    if (overlap) {  // optional: without the memory (or a stream), one stream does both
        const size_t want[4] = {n * n_stages, p1_stage * n_stages, n, n};
        double** into[4] = {&bufs[1].stages, &bufs[1].p1, &u2, &ktmp2};
        for (int i = 0; i < 4 && overlap; ++i) {
            void* q = nullptr;
            if (hipMalloc(&q, want[i] * sizeof(double)) != hipSuccess) {
                (void)hipGetLastError();  // clear the allocation failure
                overlap = false;
                break;
            }
            mine.push_back(q);
            *into[i] = static_cast<double*>(q);
        }
        if (overlap && !h->adj_stream && hipStreamCreateWithFlags(&h->adj_stream, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            h->adj_stream = nullptr;
            overlap = false;
        }
        for (hipEvent_t* e : {&ev_fwd, &ev_rec[0], &ev_rec[1], &ev_rev[0], &ev_rev[1]})
            if (overlap && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                overlap = false;
            }
    }
    auto buf_of = [&](int seg) -> const SegBufs& { return bufs[overlap ? (seg & 1) : 0]; };
    auto p1_of = [&](const SegBufs& sb, int idx) {
        double* b = sb.p1 + p1_stage * (size_t)idx;
        return Phase1Bufs{b, b + p1q, b + p1q + p1f};
    };
    auto hip_step = [&](hipError_t e, const char* what) {
        return e == hipSuccess ? 0 : fail(h, SSE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
#define ADJ_TRY(expr)      \
    do {                    \
        if ((rc = (expr))) { \
            release();      \
            return rc;      \
        }                   \
    } while (0)
    ADJ_TRY(hip_step(hipMemcpyAsync(u, u0, n * sizeof(double),
                                    where == SSE_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ms),
                     "copy u0"));
    // one stage of a segment that keeps what its VJP reads in sb: the stage state and, on
    // the fast path, the phase-1 outputs; us and kt are the RK state and register, and the
    // launches go to h->stream
    auto stage_keep = [&](const SegBufs& sb, double* us, double* kt, int step, int first, int s) -> int {
        const int idx = 5 * (step - first) + s;
        double* keep = sb.stages + n * (size_t)idx;
        if (keep_p1) {  // the stage kernel also stores the stage state
            RKStage rk{1, CK_A[s], CK_B[s], dt, kt, us, nullptr, keep};
            return launch_struct_into(h, us, rk, p1_of(sb, idx));
        }
        int r = hip_step(hipMemcpyAsync(keep, us, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream),
                         "stage state");
        if (r) return r;
        RKStage rk{1, CK_A[s], CK_B[s], dt, kt, us, nullptr};
        return launch(h, us, nullptr, rk, 3);
    };
    // forward sweep with checkpoints at the start of each segment; the last segments (one
    // per buffer set) keep their stages as they go, so the reverse sweep does not recompute them
    const int kept_first = overlap ? std::max(nseg - 2, 0) : nseg - 1;
    for (int step = 0; step < nsteps; ++step) {
        if (step % checkpoint_every == 0)
            ADJ_TRY(hip_step(hipMemcpyAsync(ckpt + n * (size_t)(step / checkpoint_every), u, n * sizeof(double),
                                            hipMemcpyDeviceToDevice, ms), "checkpoint"));
        for (int s = 0; s < 5; ++s) {
            const int seg = step / checkpoint_every;
            if (seg >= kept_first) {
                ADJ_TRY(stage_keep(buf_of(seg), u, ktmp, step, seg * checkpoint_every, s));
                continue;
            }
            RKStage rk{1, CK_A[s], CK_B[s], dt, ktmp, u, nullptr};
            ADJ_TRY(launch(h, u, nullptr, rk, 3));
        }
    }
    if (overlap) ADJ_TRY(hip_step(hipEventRecord(ev_fwd, ms), "event"));  // the checkpoints are written
    // a segment's stages again from its checkpoint, on stream st with its own RK state
    auto recompute = [&](int seg, hipStream_t st, double* us, double* kt) -> int {
        const int first = seg * checkpoint_every, last = std::min(nsteps, first + checkpoint_every);
        h->stream = st;
        int r = hip_step(hipMemcpyAsync(us, ckpt + n * (size_t)seg, n * sizeof(double), hipMemcpyDeviceToDevice, st),
                         "restore");
        for (int step = first; step < last && !r; ++step)
            for (int s = 0; s < 5 && !r; ++s) r = stage_keep(buf_of(seg), us, kt, step, first, s);
        h->stream = ms;
        return r;
    };
    // objective and its gradient at the final state
    ADJ_TRY(hip_step(sse::launch_functionals(h->P, u, nullptr, SSE_ENTROPY, h->fpartial ? h->fpartial : ktmp,
                                             vb, ms), "entropy"));
    ADJ_TRY(hip_step(hipMemcpyAsync(J, vb, sizeof(double), hipMemcpyDeviceToHost, ms), "copy J"));
    ADJ_TRY(hip_step(hipStreamSynchronize(ms), "sync"));
    ADJ_TRY(hip_step(sse::launch_entropy_grad(h->P, u, ub, ms), "entropy gradient"));
    ADJ_TRY(hip_step(hipMemsetAsync(h->vjp_lamout, 0, sizeof(double), ms), "zero"));
    if (fast) ADJ_TRY(hip_step(hipMemsetAsync(h->vjp_lam, 0, h->P.N_e * sizeof(double), ms), "zero"));
    // reverse sweep, one segment at a time
    bool rev_recorded[2] = {false, false};
    if (overlap) ADJ_TRY(hip_step(hipStreamWaitEvent(h->adj_stream, ev_fwd, 0), "wait"));
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int first = seg * checkpoint_every;
        const int last = std::min(nsteps, first + checkpoint_every);
        if (overlap) {
            if (seg - 1 >= 0 && seg - 1 < kept_first) {  // the previous segment, beside this one's reverse sweep
                const int b = (seg - 1) & 1;
                if (rev_recorded[b]) ADJ_TRY(hip_step(hipStreamWaitEvent(h->adj_stream, ev_rev[b], 0), "wait"));
                ADJ_TRY(recompute(seg - 1, h->adj_stream, u2, ktmp2));
                ADJ_TRY(hip_step(hipEventRecord(ev_rec[b], h->adj_stream), "event"));
            }
            if (seg < kept_first) ADJ_TRY(hip_step(hipStreamWaitEvent(ms, ev_rec[seg & 1], 0), "wait"));
        } else if (seg < kept_first) {  // the kept segments' stages are still there from the forward sweep
            ADJ_TRY(recompute(seg, ms, u, ktmp));
        }
        const SegBufs& sb = buf_of(seg);
        // adjoint of the 2N stages: kbar += B_s ubar; ubar += dt R'(u_s)' kbar; kbar *= A_s
        for (int step = last - 1; step >= first; --step) {
            if (fast) {
                // per stage: kbar = A_{s+1} kbar + B_s ubar (A_5 = 0; done by the VJP's first
                // kernel), then ubar += dt R'(u_s)' kbar (written by its last kernel)
                for (int s = 4; s >= 0; --s) {
                    const int idx = 5 * (step - first) + s;
                    const double* us = sb.stages + n * (size_t)idx;
                    const Phase1Bufs b = keep_p1 ? p1_of(sb, idx) : Phase1Bufs{nullptr, nullptr, nullptr};
                    ADJ_TRY(vjp_tensor_device(h, us, kb, ub, dt, 1, keep_p1 ? &b : nullptr, ub,
                                              s == 4 ? 0.0 : CK_A[s + 1], CK_B[s]));
                }
                continue;
            }
            ADJ_TRY(hip_step(hipMemsetAsync(kb, 0, n * sizeof(double), ms), "zero"));
            for (int s = 4; s >= 0; --s) {
                const double* us = sb.stages + n * (size_t)(5 * (step - first) + s);
                ADJ_TRY(hip_step(sse::launch_axpby(n, CK_B[s], ub, 1.0, kb, ms), "axpby"));
                ADJ_TRY(vjp_device(h, us, kb, vb, dt, 1));
                ADJ_TRY(hip_step(sse::launch_axpby(n, dt, vb, 1.0, ub, ms), "axpby"));
                if (s > 0) ADJ_TRY(hip_step(sse::launch_axpby(n, 0.0, kb, CK_A[s], kb, ms), "scale"));
            }
        }
        if (overlap) {  // this segment's buffers are free once its reverse sweep is done
            ADJ_TRY(hip_step(hipEventRecord(ev_rev[seg & 1], ms), "event"));
            rev_recorded[seg & 1] = true;
        }
    }
    if (fast) ADJ_TRY(hip_step(sse::launch_vjp_sum(h->vjp_lam, h->P.N_e, h->vjp_lamout, 1.0, 0, h->stream), "λ sum"));
    ADJ_TRY(hip_step(hipMemcpyAsync(grad_u0, ub, n * sizeof(double),
                                    where == SSE_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, h->stream),
                     "copy gradient"));
    ADJ_TRY(hip_step(hipMemcpyAsync(grad_lambda, h->vjp_lamout, sizeof(double), hipMemcpyDeviceToHost, h->stream),
                     "copy lambda gradient"));
    ADJ_TRY(hip_step(hipStreamSynchronize(h->stream), "sync"));
#undef ADJ_TRY
    release();
    return check_status(h);
}

int sse_set_stream(sse_handle* h, void* stream) {
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    h->stream = stream ? static_cast<hipStream_t>(stream) : h->own_stream;
    return 0;
}

int sse_set_diffusion(sse_handle* h, double b) {
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    if (!h->P.second_order) return fail(h, SSE_ERR_UNSUPPORTED, "sse_set_diffusion: not an advection-diffusion law");
    h->P.b = b;
    h->TP.base.b = b;
    if (h->rk_exec) (void)hipGraphExecDestroy(h->rk_exec);  // it holds the old b in its kernel arguments
    if (h->rk_graph) (void)hipGraphDestroy(h->rk_graph);
    h->rk_exec = nullptr;
    h->rk_graph = nullptr;
    return 0;
}

int sse_synchronize(sse_handle* h) {
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    SSE_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return check_status(h);
}

int sse_residual(sse_handle* h, const double* u, double* dudt, double t, int32_t where) {
    (void)t;  // the residual is autonomous (no source terms on this path)
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    if (!u || !dudt) return fail(h, SSE_ERR_INVALID_ARGUMENT, "u or dudt is NULL");
    RKStage rk{};
    if (where == SSE_DEVICE) return launch(h, u, dudt, rk, 3);
    if (where != SSE_HOST) return fail(h, SSE_ERR_INVALID_ARGUMENT, "where must be SSE_DEVICE or SSE_HOST");
    const size_t n = state_size(h);
    int rc;
    if (!h->stage_u && ((rc = alloc(h, n, &h->stage_u)) || (rc = alloc(h, n, &h->stage_dudt)))) return rc;
    SSE_HIP_TRY(h, hipMemcpyAsync(h->stage_u, u, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if ((rc = launch(h, h->stage_u, h->stage_dudt, rk, 3))) return rc;
    SSE_HIP_TRY(h, hipMemcpyAsync(dudt, h->stage_dudt, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SSE_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return check_status(h);
}

int sse_rk_steps(sse_handle* h, double* u, double t0, double dt, int32_t nsteps, int32_t where) {
    (void)t0;
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    if (!u || nsteps < 0) return fail(h, SSE_ERR_INVALID_ARGUMENT, "u is NULL or nsteps < 0");
    const size_t n = state_size(h);
    int rc;
    if (!h->rk_tmp && (rc = alloc(h, n, &h->rk_tmp))) return rc;
    double* ud = u;
    if (where == SSE_HOST) {
        if (!h->stage_u && ((rc = alloc(h, n, &h->stage_u)) || (rc = alloc(h, n, &h->stage_dudt)))) return rc;
        SSE_HIP_TRY(h, hipMemcpyAsync(h->stage_u, u, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        ud = h->stage_u;
    } else if (where != SSE_DEVICE) {
        return fail(h, SSE_ERR_INVALID_ARGUMENT, "where must be SSE_DEVICE or SSE_HOST");
    }
    auto one_step = [&]() -> int {
        for (int s = 0; s < 5; ++s) {
            RKStage rk{1, CK_A[s], CK_B[s], dt, h->rk_tmp, ud};
            int r = launch(h, ud, nullptr, rk, 3);
            if (r) return r;
        }
        return 0;
    };
    // The first step runs eagerly (it also sets kernel attributes); the rest replay one
    // step captured as a HIP graph (ten kernel launches), recaptured when the state,
    // the step or the stream differ from the captured ones. SSE_RK_GRAPH=0 launches
    // every step eagerly.
    int step = 0;
    if (nsteps > 0) {
        if ((rc = one_step())) return rc;
        step = 1;
    }
    const char* genv = getenv("SSE_RK_GRAPH");
    const bool graph = !(genv && genv[0] == '0') && h->members == 1;
    if (graph && step < nsteps) {
        if (!h->rk_exec || h->rk_graph_u != ud || h->rk_graph_dt != dt || h->rk_graph_stream != h->stream) {
            if (h->rk_exec) (void)hipGraphExecDestroy(h->rk_exec);
            if (h->rk_graph) (void)hipGraphDestroy(h->rk_graph);
            h->rk_exec = nullptr;
            h->rk_graph = nullptr;
            SSE_HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            const int r = one_step();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(h->stream, &g);
            if (r) {
                if (g) (void)hipGraphDestroy(g);
                return r;
            }
            SSE_HIP_TRY(h, e);
            h->rk_graph = g;
            SSE_HIP_TRY(h, hipGraphInstantiate(&h->rk_exec, g, nullptr, nullptr, 0));
            h->rk_graph_u = ud;
            h->rk_graph_dt = dt;
            h->rk_graph_stream = h->stream;
        }
        for (; step < nsteps; ++step) SSE_HIP_TRY(h, hipGraphLaunch(h->rk_exec, h->stream));
    }
    for (; step < nsteps; ++step)
        if ((rc = one_step())) return rc;
    if (where == SSE_HOST) {
        SSE_HIP_TRY(h, hipMemcpyAsync(u, h->stage_u, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        SSE_HIP_TRY(h, hipStreamSynchronize(h->stream));
        return check_status(h);
    }
    return 0;
}

}  // extern "C"

namespace {

// out = base + sum_q c[q] K[idx[q]] over n entries (a Runge-Kutta stage combination;
// out may alias base). The nonzero coefficients only, at most SSE_ERK_MAX_STAGES.
struct ErkTerms {
    int count;
    int idx[SSE_ERK_MAX_STAGES];
    double c[SSE_ERK_MAX_STAGES];
};

__global__ void __launch_bounds__(256) k_erk_combine(size_t n, const double* base, const double* __restrict__ K,
                                                     ErkTerms terms, double* out) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int q = 0; q < terms.count; ++q) acc += terms.c[q] * K[(size_t)terms.idx[q] * n + t];
        out[t] = base[t] + acc;
    }
}

int erk_combine(sse_handle* h, size_t n, const double* base, const double* K, const ErkTerms& terms, double* out) {
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_erk_combine, dim3(blocks), dim3(256), 0, h->stream, n, base, K, terms, out);
    SSE_HIP_TRY(h, hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int sse_erk_steps(sse_handle* h, double* u, double t0, double dt, int32_t nsteps, int32_t stages, const double* A,
                  const double* b, int32_t where) {
    (void)t0;  // the residual is autonomous
    if (!h) return SSE_ERR_INVALID_ARGUMENT;
    if (!u || !A || !b || nsteps < 0) return fail(h, SSE_ERR_INVALID_ARGUMENT, "u, A or b is NULL or nsteps < 0");
    if (stages < 1 || stages > SSE_ERK_MAX_STAGES)
        return fail(h, SSE_ERR_INVALID_ARGUMENT, "stages must be 1.." + std::to_string(SSE_ERK_MAX_STAGES));
    if (where != SSE_DEVICE && where != SSE_HOST)
        return fail(h, SSE_ERR_INVALID_ARGUMENT, "where must be SSE_DEVICE or SSE_HOST");
    if (h->members != 1) return fail(h, SSE_ERR_UNSUPPORTED, "an ensemble handle advances with sse_batch_rk_steps");
    for (int i = 0; i < stages; ++i)
        for (int j = i; j < stages; ++j)
            if (A[i * stages + j] != 0.0)
                return fail(h, SSE_ERR_INVALID_ARGUMENT, "A must be strictly lower triangular (an explicit method)");
    const size_t n = state_size(h);
    int rc;
    if (h->erk_stages < stages) {  // stage derivatives and the stage state (kept for later calls)
        if ((rc = alloc(h, (size_t)(stages + 1) * n, &h->erk_k))) return rc;
        h->erk_stages = stages;
    }
    double* K = h->erk_k;
    double* y = h->erk_k + (size_t)h->erk_stages * n;
    double* ud = u;
    if (where == SSE_HOST) {
        if (!h->stage_u && ((rc = alloc(h, n, &h->stage_u)) || (rc = alloc(h, n, &h->stage_dudt)))) return rc;
        SSE_HIP_TRY(h, hipMemcpyAsync(h->stage_u, u, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        ud = h->stage_u;
    }
    ErkTerms fin{};
    for (int i = 0; i < stages; ++i)
        if (b[i] != 0.0) {
            fin.idx[fin.count] = i;
            fin.c[fin.count++] = dt * b[i];
        }
    for (int step = 0; step < nsteps; ++step) {
        for (int i = 0; i < stages; ++i) {
            ErkTerms st{};
            for (int j = 0; j < i; ++j)
                if (A[i * stages + j] != 0.0) {
                    st.idx[st.count] = j;
                    st.c[st.count++] = dt * A[i * stages + j];
                }
            const double* yi = ud;
            if (st.count) {
                if ((rc = erk_combine(h, n, ud, K, st, y))) return rc;
                yi = y;
            }
            if ((rc = launch(h, yi, K + (size_t)i * n, RKStage{}, 3))) return rc;
        }
        if ((rc = erk_combine(h, n, ud, K, fin, ud))) return rc;
    }
    if (where == SSE_HOST) {
        SSE_HIP_TRY(h, hipMemcpyAsync(u, h->stage_u, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        SSE_HIP_TRY(h, hipStreamSynchronize(h->stream));
        return check_status(h);
    }
    return 0;
}

int sse_profile_residual(sse_handle* h, const double* u, double* dudt, int32_t nrep, double* ms_out) {
    if (!h || !u || !dudt || !ms_out || nrep <= 0) return fail(h, SSE_ERR_INVALID_ARGUMENT, "bad arguments");
    // the nrep residuals back to back, as a solve runs them, with an event after each
    // launch; one wait at the end
    std::vector<hipEvent_t> ev(2 * (size_t)nrep + 1, nullptr);
    int rc = 0;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            rc = fail(h, SSE_ERR_HIP, "hipEventCreate failed");
            break;
        }
    RKStage rk{};
    if (!rc && hipEventRecord(ev[0], h->stream) != hipSuccess) rc = fail(h, SSE_ERR_HIP, "hipEventRecord failed");
    for (int r = 0; r < nrep && !rc; ++r) {
        if ((rc = launch(h, u, dudt, rk, 1))) break;
        if (hipEventRecord(ev[2 * r + 1], h->stream) != hipSuccess) rc = fail(h, SSE_ERR_HIP, "hipEventRecord failed");
        if (rc || (rc = launch(h, u, dudt, rk, 2))) break;
        if (hipEventRecord(ev[2 * r + 2], h->stream) != hipSuccess) rc = fail(h, SSE_ERR_HIP, "hipEventRecord failed");
    }
    double acc[SSE_NUM_KERNELS] = {0.0, 0.0};
    if (!rc && hipEventSynchronize(ev[2 * (size_t)nrep]) != hipSuccess) rc = fail(h, SSE_ERR_HIP, "hipEventSynchronize failed");
    for (int r = 0; r < nrep && !rc; ++r) {
        float a = 0.f, b = 0.f;
        if (hipEventElapsedTime(&a, ev[2 * r], ev[2 * r + 1]) != hipSuccess ||
            hipEventElapsedTime(&b, ev[2 * r + 1], ev[2 * r + 2]) != hipSuccess)
            rc = fail(h, SSE_ERR_HIP, "hipEventElapsedTime failed");
        acc[0] += a;
        acc[1] += b;
    }
    (void)hipStreamSynchronize(h->stream);
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc) return rc;
    for (int i = 0; i < SSE_NUM_KERNELS; ++i) ms_out[i] = acc[i] / nrep;
    return 0;
}

void sse_destroy(sse_handle* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->rk_exec) (void)hipGraphExecDestroy(h->rk_exec);
    if (h->rk_graph) (void)hipGraphDestroy(h->rk_graph);
    for (void* p : h->allocations) (void)hipFree(p);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    if (h->adj_stream) (void)hipStreamDestroy(h->adj_stream);
    delete h;
}

}  // extern "C"

// ------------------------------------------------------------------ batched problems
struct sse_batch {
    std::vector<sse_handle*> groups;
    std::vector<std::vector<int>> members;          // problem indices of each group, in element order
    std::vector<std::vector<long long>> offsets;    // element offset of each member inside its group
    std::vector<size_t> elems;                      // N_e of each problem
    std::vector<size_t> per_elem;                   // N_p * N_c of each problem
    std::vector<double*> state, tmp, dt_e;          // per group, device
    // ensemble groups (members share one mesh and differ only in b): lanes per
    // launch (0 for a group of concatenated meshes) and each member's b
    std::vector<int> chunk;
    std::vector<double*> b_dev;
    // tangent pairs of ensemble groups, allocated on first use: (u, v) per member
    std::vector<double*> t_state, t_tmp, t_dt, t_b;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // the captured RK step runs each group's five stages on a stream of its own (a branch
    // of the graph, forked from and joined to `stream`), so that groups overlap
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    std::string err;
};

namespace {

thread_local std::string g_batch_create_error;

// device memory the intermediates of one ensemble launch may take
constexpr size_t ENSEMBLE_BUDGET = (size_t)16 << 30;

int bfail(sse_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else g_batch_create_error = msg;
    return code;
}

bool same_array(const void* a, const void* b, size_t bytes) {
    if (a == b) return true;
    if (!a || !b) return false;
    return std::memcmp(a, b, bytes) == 0;
}

// true when problems a and b can share one handle (same reference element and
// operators, law, form and fluxes; geometry may differ). With any_b the
// diffusion coefficients may differ too.
bool same_discretisation(const sse_desc* a, const sse_desc* b, bool any_b = false) {
    const int32_t* ia = &a->abi_version;
    const int32_t* ib = &b->abi_version;
    // integer header fields up to viscous_flux, except N_e
    const int n_int = (int)(offsetof(sse_desc, gamma) / sizeof(int32_t));
    for (int t = 0; t < n_int; ++t) {
        if ((const char*)(ia + t) == (const char*)&a->N_e) continue;
        if (ia[t] != ib[t]) return false;
    }
    if (a->gamma != b->gamma || a->half_lambda != b->half_lambda || (!any_b && a->b != b->b)) return false;
    for (int m = 0; m < 3; ++m)
        if (a->a[m] != b->a[m]) return false;
    if (a->V_p != b->V_p) return false;
    for (int m = 0; m < 3; ++m)
        if (a->V_M[m] != b->V_M[m]) return false;
    const size_t Nq = a->N_q, Np = a->N_p, Nf = a->N_f, d = a->d, q1 = a->V_p + 1;
    const size_t dd = sizeof(double);
    return same_array(a->V, b->V, Nq * Np * dd) && same_array(a->R, b->R, Nf * Nq * dd) &&
           same_array(a->R_selection, b->R_selection, Nf * sizeof(int64_t)) && same_array(a->W, b->W, Nq * dd) &&
           same_array(a->B, b->B, Nf * dd) && same_array(a->S, b->S, Nq * Nq * d * dd) &&
           same_array(a->C, b->C, Nq * Nf * dd) && same_array(a->Minv, b->Minv, Np * Np * dd) &&
           same_array(a->D, b->D, Nq * Nq * d * dd) &&
           same_array(a->V_A, b->V_A, (size_t)a->V_M[0] * q1 * dd) &&
           same_array(a->V_B, b->V_B, (size_t)a->V_M[1] * q1 * q1 * dd) &&
           same_array(a->V_C, b->V_C, (size_t)a->V_M[2] * q1 * q1 * q1 * dd) &&
           same_array(a->face_normals, b->face_normals, d * a->N_fac * dd) &&
           ((a->nJq == nullptr) == (b->nJq == nullptr));
}

// true when a and b have the same mesh and geometric factors
bool same_mesh(const sse_desc* a, const sse_desc* b) {
    if (a->N_e != b->N_e) return false;
    const size_t Nq = a->N_q, Nf = a->N_f, d = a->d, Ne = a->N_e, dd = sizeof(double);
    return same_array(a->J_q, b->J_q, Nq * Ne * dd) && same_array(a->Lambda_q, b->Lambda_q, Nq * d * d * Ne * dd) &&
           same_array(a->nJf, b->nJf, d * Nf * Ne * dd) && same_array(a->J_f, b->J_f, Nf * Ne * dd) &&
           same_array(a->mapP, b->mapP, Nf * Ne * sizeof(int64_t)) &&
           same_array(a->nJq, b->nJq, d * a->N_fac * Nq * Ne * dd);
}

template <typename T>
void append(std::vector<T>& dst, const T* src, size_t n) {
    dst.insert(dst.end(), src, src + n);
}

// lanes per element one launch of an ensemble of R members handles (even, so
// that tangent pairs never straddle two launches)
int ensemble_chunk(const sse_desc* D, int R) {
    const size_t per_lane =
        sizeof(double) * (size_t)D->N_e * ((size_t)D->N_q + D->N_f + (size_t)D->d * (D->N_q + D->N_f));
    size_t cap = std::max<size_t>(2, ENSEMBLE_BUDGET / std::max<size_t>(per_lane, 1));
    if (const char* env = getenv("SSE_ENSEMBLE_CHUNK")) cap = std::max(2, atoi(env));  // smaller launches, for tests
    cap = std::min<size_t>(cap, 2 * (size_t)R);
    cap &= ~(size_t)1;
    return (int)std::max<size_t>(cap, 2);
}

// RK stage of members [c0, c0 + n) of ensemble group g (tangent: lanes of the pair buffers)
int ensemble_stage(sse_batch* b, size_t g, int s, bool tangent) {
    sse_handle* h = b->groups[g];
    const size_t n1 = (size_t)h->P.N_p * h->P.N_c * h->P.N_e, ne = h->P.N_e;
    const int R = (int)b->members[g].size() * (tangent ? 2 : 1);
    double* st = tangent ? b->t_state[g] : b->state[g];
    double* tm = tangent ? b->t_tmp[g] : b->tmp[g];
    double* de = tangent ? b->t_dt[g] : b->dt_e[g];
    const double* bb = tangent ? b->t_b[g] : b->b_dev[g];
    for (int c0 = 0; c0 < R; c0 += b->chunk[g]) {
        const int n = std::min(b->chunk[g], R - c0);
        RKStage rk{1, CK_A[s], CK_B[s], 0.0, tm + c0 * n1, st + c0 * n1, de + c0 * ne};
        int r = launch_ens(h, st + c0 * n1, nullptr, rk, 3, n, bb + c0, tangent ? 1 : 0);
        if (r) return bfail(b, r, h->err);
    }
    return 0;
}

int batch_alloc(double** p, size_t n) { return hipMalloc(p, std::max<size_t>(n, 1) * sizeof(double)) == hipSuccess; }

}  // namespace

extern "C" {

int sse_batch_create(const sse_desc* const* descs, int32_t n, sse_batch** out) {
    if (!out) return bfail(nullptr, SSE_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!descs || n <= 0) return bfail(nullptr, SSE_ERR_INVALID_ARGUMENT, "need at least one descriptor");
    for (int i = 0; i < n; ++i) {
        std::string msg;
        int code = validate(descs[i], msg);
        if (code) return bfail(nullptr, code, "problem " + std::to_string(i) + ": " + msg);
    }
    sse_batch* b = new sse_batch();
    auto cleanup = [&](int code, const std::string& msg) {
        g_batch_create_error = msg;
        sse_batch_destroy(b);
        return code;
    };
    b->elems.resize(n);
    b->per_elem.resize(n);
    for (int i = 0; i < n; ++i) {
        b->elems[i] = descs[i]->N_e;
        b->per_elem[i] = (size_t)descs[i]->N_p * descs[i]->N_c;
    }
    if (hipStreamCreate(&b->stream) != hipSuccess) return cleanup(SSE_ERR_HIP, "hipStreamCreate failed");
    // ensembles first: StandardForm scalar problems on one mesh that differ at most in b
    std::vector<std::vector<int>> ens;
    for (int i = 0; i < n; ++i) {
        const sse_desc* D = descs[i];
        if (D->form != SSE_STANDARD || D->N_c != 1) continue;
        bool placed = false;
        for (auto& e : ens)
            if (same_discretisation(descs[e[0]], D, true) && same_mesh(descs[e[0]], D)) {
                e.push_back(i);
                placed = true;
                break;
            }
        if (!placed) ens.push_back({i});
    }
    std::vector<char> grouped(n, 0);
    for (auto& e : ens) {
        if (e.size() < 2) continue;
        const sse_desc* R = descs[e[0]];
        const int R_n = (int)e.size();
        const int chunk = ensemble_chunk(R, R_n);
        sse_handle* h = nullptr;
        int code = create_handle(R, chunk, &h);
        if (code == SSE_ERR_UNSUPPORTED) continue;  // not a compiled shape: concatenate instead
        if (code) return cleanup(code, "ensemble: " + g_create_error);
        b->groups.push_back(h);
        sse_set_stream(h, b->stream);
        b->members.push_back(e);
        b->offsets.emplace_back();
        for (int t = 0; t < R_n; ++t) b->offsets.back().push_back((long long)t * R->N_e);
        b->chunk.push_back(chunk);
        const size_t n1 = (size_t)R->N_p * R->N_c * R->N_e;
        double *st = nullptr, *tm = nullptr, *de = nullptr, *bd = nullptr;
        const bool ok = batch_alloc(&st, n1 * R_n) && batch_alloc(&tm, n1 * R_n) &&
                        batch_alloc(&de, (size_t)R->N_e * R_n) && batch_alloc(&bd, R_n);
        b->state.push_back(st);
        b->tmp.push_back(tm);
        b->dt_e.push_back(de);
        b->b_dev.push_back(bd);
        b->t_state.push_back(nullptr);
        b->t_tmp.push_back(nullptr);
        b->t_dt.push_back(nullptr);
        b->t_b.push_back(nullptr);
        if (!ok) return cleanup(SSE_ERR_OUT_OF_MEMORY, "hipMalloc failed for the ensemble state");
        std::vector<double> bv;
        for (int i : e) bv.push_back(descs[i]->b);
        if (hipMemcpy(bd, bv.data(), bv.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(SSE_ERR_HIP, "hipMemcpy failed");
        for (int i : e) grouped[i] = 1;
    }
    // the rest: group problems by discretisation, in order of first appearance,
    // and concatenate their meshes
    std::vector<int> rep;  // representative problem of each group
    const size_t first = b->groups.size();
    for (int i = 0; i < n; ++i) {
        if (grouped[i]) continue;
        int g = -1;
        for (size_t t = 0; t < rep.size(); ++t)
            if (same_discretisation(descs[rep[t]], descs[i])) { g = (int)t; break; }
        if (g < 0) {
            rep.push_back(i);
            b->members.emplace_back();
            b->offsets.emplace_back();
            g = (int)rep.size() - 1;
        }
        b->members[first + g].push_back(i);
    }
    for (size_t g = 0; g < rep.size(); ++g) {
        const size_t G = first + g;
        const sse_desc* R = descs[rep[g]];
        sse_desc C = *R;
        const size_t Nq = R->N_q, Nf = R->N_f, d = R->d, Nfac = R->N_fac;
        std::vector<double> Jq, Lam, nJf, Jf, nJq;
        std::vector<int64_t> mapP;
        long long off = 0;
        for (int i : b->members[G]) {
            const sse_desc* D = descs[i];
            const size_t Ne = D->N_e;
            b->offsets[G].push_back(off);
            append(Jq, D->J_q, Nq * Ne);
            append(Lam, D->Lambda_q, Nq * d * d * Ne);
            append(nJf, D->nJf, d * Nf * Ne);
            append(Jf, D->J_f, Nf * Ne);
            if (D->nJq) append(nJq, D->nJq, d * Nfac * Nq * Ne);
            for (size_t t = 0; t < Nf * Ne; ++t) mapP.push_back(D->mapP[t] + (int64_t)Nf * off);
            off += (long long)Ne;
        }
        if (off > 2147483647LL) return cleanup(SSE_ERR_UNSUPPORTED, "group too large");
        C.N_e = (int32_t)off;
        C.J_q = Jq.data();
        C.Lambda_q = Lam.data();
        C.nJf = nJf.data();
        C.J_f = Jf.data();
        C.nJq = R->nJq ? nJq.data() : nullptr;
        C.mapP = mapP.data();
        sse_handle* h = nullptr;
        int code = sse_create(&C, &h);
        if (code) return cleanup(code, "group " + std::to_string(G) + ": " + g_create_error);
        b->groups.push_back(h);
        sse_set_stream(h, b->stream);
        b->chunk.push_back(0);
        const size_t nstate = (size_t)C.N_p * C.N_c * C.N_e;
        double *st = nullptr, *tm = nullptr, *de = nullptr;
        const bool ok = batch_alloc(&st, nstate) && batch_alloc(&tm, nstate) && batch_alloc(&de, (size_t)C.N_e);
        b->state.push_back(st);
        b->tmp.push_back(tm);
        b->dt_e.push_back(de);
        b->b_dev.push_back(nullptr);
        b->t_state.push_back(nullptr);
        b->t_tmp.push_back(nullptr);
        b->t_dt.push_back(nullptr);
        b->t_b.push_back(nullptr);
        if (!ok) return cleanup(SSE_ERR_OUT_OF_MEMORY, "hipMalloc failed for the group state");
    }
    *out = b;
    return 0;
}

int sse_batch_rk_steps(sse_batch* b, double* const* u, const double* dt, int32_t nsteps, int32_t where,
                       int32_t use_graph) {
    if (!b) return SSE_ERR_INVALID_ARGUMENT;
    if (!u || !dt || nsteps < 0) return bfail(b, SSE_ERR_INVALID_ARGUMENT, "u or dt is NULL, or nsteps < 0");
    if (where != SSE_DEVICE && where != SSE_HOST) return bfail(b, SSE_ERR_INVALID_ARGUMENT, "bad `where`");
    const hipMemcpyKind in = where == SSE_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind outk = where == SSE_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return 0;
        return bfail(b, SSE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    int rc;
    // gather states and per-element time steps into the group buffers
    for (size_t g = 0; g < b->groups.size(); ++g) {
        std::vector<double> de;
        for (size_t t = 0; t < b->members[g].size(); ++t) {
            const int i = b->members[g][t];
            if (!u[i]) return bfail(b, SSE_ERR_INVALID_ARGUMENT, "u[" + std::to_string(i) + "] is NULL");
            const size_t n_i = b->per_elem[i] * b->elems[i];
            if ((rc = hip_ok(hipMemcpyAsync(b->state[g] + b->per_elem[i] * b->offsets[g][t], u[i],
                                            n_i * sizeof(double), in, b->stream), "state gather")))
                return rc;
            de.insert(de.end(), b->elems[i], dt[i]);
        }
        if ((rc = hip_ok(hipMemcpyAsync(b->dt_e[g], de.data(), de.size() * sizeof(double), hipMemcpyHostToDevice,
                                        b->stream), "dt upload")))
            return rc;
        // the host vector must outlive the copy
        if ((rc = hip_ok(hipStreamSynchronize(b->stream), "sync"))) return rc;
    }
    auto one_step = [&]() -> int {
        for (int s = 0; s < 5; ++s)
            for (size_t g = 0; g < b->groups.size(); ++g) {
                if (b->chunk[g]) {
                    int r = ensemble_stage(b, g, s, false);
                    if (r) return r;
                    continue;
                }
                RKStage rk{1, CK_A[s], CK_B[s], 0.0, b->tmp[g], b->state[g], b->dt_e[g]};
                int r = launch(b->groups[g], b->state[g], nullptr, rk, 3);
                if (r) return bfail(b, r, b->groups[g]->err);
            }
        return 0;
    };
    int step = 0;
    if (nsteps > 0) {  // the first step runs eagerly (it also sets kernel attributes)
        if ((rc = one_step())) return rc;
        step = 1;
    }
    if (use_graph && step < nsteps) {
        if (!b->exec) {
            // one branch per group (SSE_BATCH_BRANCHES=0: all groups in sequence on one stream)
            const char* env = getenv("SSE_BATCH_BRANCHES");
            const bool branches = !(env && env[0] == '0') && b->groups.size() > 1;
            if (branches && b->branch.empty()) {
                b->branch.assign(b->groups.size(), nullptr);
                b->join.assign(b->groups.size(), nullptr);
                if ((rc = hip_ok(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "event create"))) return rc;
                for (size_t g = 0; g < b->groups.size(); ++g)
                    if ((rc = hip_ok(hipStreamCreate(&b->branch[g]), "stream create")) ||
                        (rc = hip_ok(hipEventCreateWithFlags(&b->join[g], hipEventDisableTiming), "event create")))
                        return rc;
            }
            if ((rc = hip_ok(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture")))
                return rc;
            int r = 0;
            hipError_t e = hipSuccess;
            if (branches) {
                e = hipEventRecord(b->fork, b->stream);
                for (size_t g = 0; g < b->groups.size() && e == hipSuccess; ++g) {
                    e = hipStreamWaitEvent(b->branch[g], b->fork, 0);
                    sse_set_stream(b->groups[g], b->branch[g]);
                }
                for (size_t g = 0; g < b->groups.size() && e == hipSuccess && !r; ++g)
                    for (int s = 0; s < 5 && !r; ++s) {
                        if (b->chunk[g]) {
                            r = ensemble_stage(b, g, s, false);
                            continue;
                        }
                        RKStage rk{1, CK_A[s], CK_B[s], 0.0, b->tmp[g], b->state[g], b->dt_e[g]};
                        r = launch(b->groups[g], b->state[g], nullptr, rk, 3);
                        if (r) r = bfail(b, r, b->groups[g]->err);
                    }
                for (size_t g = 0; g < b->groups.size(); ++g) {
                    if (e == hipSuccess) e = hipEventRecord(b->join[g], b->branch[g]);
                    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join[g], 0);
                    sse_set_stream(b->groups[g], b->stream);
                }
            } else {
                r = one_step();
            }
            hipGraph_t graph = nullptr;
            const hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
            if (r) return r;
            if ((rc = hip_ok(e, "branch capture")) || (rc = hip_ok(e2, "end capture"))) return rc;
            b->graph = graph;
            if ((rc = hip_ok(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0), "graph instantiate")))
                return rc;
        }
        for (; step < nsteps; ++step)
            if ((rc = hip_ok(hipGraphLaunch(b->exec, b->stream), "graph launch"))) return rc;
    }
    for (; step < nsteps; ++step)
        if ((rc = one_step())) return rc;
    // scatter back
    for (size_t g = 0; g < b->groups.size(); ++g)
        for (size_t t = 0; t < b->members[g].size(); ++t) {
            const int i = b->members[g][t];
            const size_t n_i = b->per_elem[i] * b->elems[i];
            if ((rc = hip_ok(hipMemcpyAsync(u[i], b->state[g] + b->per_elem[i] * b->offsets[g][t],
                                            n_i * sizeof(double), outk, b->stream), "state scatter")))
                return rc;
        }
    if ((rc = hip_ok(hipStreamSynchronize(b->stream), "sync"))) return rc;
    for (sse_handle* h : b->groups)
        if ((rc = check_status(h))) return bfail(b, rc, h->err);
    return 0;
}

int sse_batch_rk_tangent(sse_batch* b, double* const* u, double* const* v, const double* dt, int32_t nsteps,
                         int32_t where) {
    if (!b) return SSE_ERR_INVALID_ARGUMENT;
    if (!u || !v || !dt || nsteps < 0) return bfail(b, SSE_ERR_INVALID_ARGUMENT, "u, v or dt is NULL, or nsteps < 0");
    if (where != SSE_DEVICE && where != SSE_HOST) return bfail(b, SSE_ERR_INVALID_ARGUMENT, "bad `where`");
    for (size_t g = 0; g < b->groups.size(); ++g) {
        if (!b->chunk[g] || !b->groups[g]->P.second_order)
            return bfail(b, SSE_ERR_UNSUPPORTED,
                         "tangents with respect to b need every problem in an ensemble of advection-diffusion "
                         "problems on one mesh");
    }
    const hipMemcpyKind in = where == SSE_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind outk = where == SSE_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return 0;
        return bfail(b, SSE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    int rc;
    for (size_t g = 0; g < b->groups.size(); ++g) {
        const sse_handle* h = b->groups[g];
        const size_t n1 = (size_t)h->P.N_p * h->P.N_c * h->P.N_e, ne = h->P.N_e;
        const size_t R2 = 2 * b->members[g].size();
        if (!b->t_state[g]) {
            if (!batch_alloc(&b->t_state[g], n1 * R2) || !batch_alloc(&b->t_tmp[g], n1 * R2) ||
                !batch_alloc(&b->t_dt[g], ne * R2) || !batch_alloc(&b->t_b[g], R2))
                return bfail(b, SSE_ERR_OUT_OF_MEMORY, "hipMalloc failed for the tangent state");
        }
        std::vector<double> de, bv;
        for (size_t t = 0; t < b->members[g].size(); ++t) {
            const int i = b->members[g][t];
            if (!u[i] || !v[i]) return bfail(b, SSE_ERR_INVALID_ARGUMENT, "u[i] or v[i] is NULL");
            if ((rc = hip_ok(hipMemcpyAsync(b->t_state[g] + 2 * t * n1, u[i], n1 * sizeof(double), in, b->stream),
                             "state gather")) ||
                (rc = hip_ok(hipMemcpyAsync(b->t_state[g] + (2 * t + 1) * n1, v[i], n1 * sizeof(double), in,
                                            b->stream), "tangent gather")))
                return rc;
            de.insert(de.end(), 2 * ne, dt[i]);
        }
        std::vector<double> bh(b->members[g].size());
        if ((rc = hip_ok(hipMemcpy(bh.data(), b->b_dev[g], bh.size() * sizeof(double), hipMemcpyDeviceToHost),
                         "b download")))
            return rc;
        for (double x : bh) {
            bv.push_back(x);
            bv.push_back(x);
        }
        if ((rc = hip_ok(hipMemcpyAsync(b->t_dt[g], de.data(), de.size() * sizeof(double), hipMemcpyHostToDevice,
                                        b->stream), "dt upload")) ||
            (rc = hip_ok(hipMemcpyAsync(b->t_b[g], bv.data(), bv.size() * sizeof(double), hipMemcpyHostToDevice,
                                        b->stream), "b upload")) ||
            (rc = hip_ok(hipStreamSynchronize(b->stream), "sync")))
            return rc;
    }
    for (int step = 0; step < nsteps; ++step)
        for (int s = 0; s < 5; ++s)
            for (size_t g = 0; g < b->groups.size(); ++g)
                if ((rc = ensemble_stage(b, g, s, true))) return rc;
    for (size_t g = 0; g < b->groups.size(); ++g) {
        const sse_handle* h = b->groups[g];
        const size_t n1 = (size_t)h->P.N_p * h->P.N_c * h->P.N_e;
        for (size_t t = 0; t < b->members[g].size(); ++t) {
            const int i = b->members[g][t];
            if ((rc = hip_ok(hipMemcpyAsync(u[i], b->t_state[g] + 2 * t * n1, n1 * sizeof(double), outk, b->stream),
                             "state scatter")) ||
                (rc = hip_ok(hipMemcpyAsync(v[i], b->t_state[g] + (2 * t + 1) * n1, n1 * sizeof(double), outk,
                                            b->stream), "tangent scatter")))
                return rc;
        }
    }
    if ((rc = hip_ok(hipStreamSynchronize(b->stream), "sync"))) return rc;
    for (sse_handle* h : b->groups)
        if ((rc = check_status(h))) return bfail(b, rc, h->err);
    return 0;
}

int sse_batch_describe(const sse_batch* b, char* buf, int32_t len) {
    if (!b || !buf || len <= 0) return SSE_ERR_INVALID_ARGUMENT;
    std::string out = "batch groups=" + std::to_string(b->groups.size());
    for (size_t g = 0; g < b->groups.size(); ++g) {
        char one[256];
        sse_describe(b->groups[g], one, sizeof(one));
        out += " | " + std::to_string(b->members[g].size()) + (b->chunk[g] ? " problems on one mesh" : " problems") +
               ", N_e=" + std::to_string(b->groups[g]->P.N_e) + ": " + one;
    }
    snprintf(buf, len, "%s", out.c_str());
    return 0;
}

const char* sse_batch_last_error(const sse_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

void sse_batch_destroy(sse_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (sse_handle* h : b->groups) sse_destroy(h);
    for (double* p : b->state) (void)hipFree(p);
    for (double* p : b->tmp) (void)hipFree(p);
    for (double* p : b->dt_e) (void)hipFree(p);
    for (auto* v : {&b->b_dev, &b->t_state, &b->t_tmp, &b->t_dt, &b->t_b})
        for (double* p : *v)
            if (p) (void)hipFree(p);
    for (hipStream_t st : b->branch)
        if (st) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : b->join)
        if (ev) (void)hipEventDestroy(ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
