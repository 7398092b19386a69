// sse_kernels.h -- device parameter blocks shared by the host API and kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "sse_physics.h"

namespace sse {

// Everything a residual kernel reads, as device pointers. Operator tables are
// compressed-row lists built once on the host from the dense reference
// operators (nonzero entries only).
struct DevParams {
    int d, N_p, N_q, N_f, N_fac, npf, N_c, N_e;
    int has_V;       // modal (V != identity)
    int R_sel;       // diagonal-E selection extrapolation
    int project;     // entropy projection performed in nodal_values
    int C_kind;      // 0 none, 1/2 facet correction present
    int mass_diag;   // DiagonalSolver (W J)^{-1}
    int standard;    // StandardForm (1) or FluxDifferencingForm (0)
    int skew;        // StandardForm: skew-symmetric (1) or standard (0) mapping form
    int second_order;  // BR1 auxiliary-variable path (linear advection-diffusion)
    double b;          // diffusion coefficient
    LawParams law;
    const double* V;  // N_q x N_p column-major
    const double* Vt; // V' (N_p x N_q column-major): coalesced reads for threads over modes
    // R rows: facet node j -> (volume node, value)
    const int* R_ptr; const int* R_idx; const double* R_val;
    // R^T rows: volume node i -> (facet node, value)
    const int* Rt_ptr; const int* Rt_idx; const double* Rt_val;
    // union sparsity of S_1..S_d by rows: volume node i -> (node j != i, S_m[i,j] for m < d)
    const int* S_ptr; const int* S_idx; const double* S_val;
    // C rows: volume node i -> (facet node j, C_ij); C columns: facet node j -> (volume node i, C_ij)
    const int* Cr_ptr; const int* Cr_idx; const double* Cr_val;
    const int* Cc_ptr; const int* Cc_idx; const double* Cc_val;
    // flux work items of the general time-derivative kernel: volume pairs i < j
    // (item_val = S_m[i,j], m < d) then facet-correction pairs (item_val[0] = C_ij);
    // item_ij = i | j << 16
    int n_items, n_vol_items;
    const int* item_ij; const double* item_val;
    // dense R (N_f x N_q column-major) and the switch of the matrix-core nodal values
    // (Euler, entropy projection, modal V: k_nodal_values_mfma)
    const double* Rd;
    int mfma;
    // deterministic mode (det = 1): the items of each node, volume nodes then facet nodes,
    // own_item[w] = c << 1 | (1 if the node is the second node of volume pair c)
    int det;
    const int* own_ptr; const int* own_item;
    // StandardForm: union sparsity of D_1..D_d by rows (node i -> (node j, D_m[i,j] for m < d))
    // and of D_1^T..D_d^T (node i -> (node j, D_m[j,i] for m < d))
    const int* D_ptr; const int* D_idx; const double* D_val;
    const int* Dt_ptr; const int* Dt_idx; const double* Dt_val;
    const double* W;   // N_q
    const double* B;   // N_f
    const double* fn;  // d x N_fac reference face normals
    // geometry (reference layouts)
    const double* Jq;   // N_q x N_e
    const double* Lam;  // N_q x d x d x N_e
    const double* nJf;  // d x N_f x N_e
    const double* Jf;   // N_f x N_e
    const int* mapP;    // N_f x N_e, 0-based linear index into (N_f, N_e)
    // phase-1 outputs: u_q [k][e][i], u_f [k][e][j]
    double* uq;
    double* uf;
    // auxiliary variable of second-order laws: q_q [k][m][e][i], q_f [k][m][e][j]
    double* qq;
    double* qf;
    int* status;
};

// Tables of the compile-time-shaped tensor-product kernels (Tri/Tet,
// ModalTensor, Euler). Built once on the host from the reference operators.
struct TensorParams {
    DevParams base;
    int elem;        // 2 = Tri, 3 = Tet
    int P;           // polynomial degree
    int variant;     // compiled launch-shape variant
    int n_pairs;     // volume pairs (union sparsity of S, i < j) followed by facet pairs (nonzeros of C)
    int n_vol_pairs;
    int n_pass, pass_size, wv, wf, wrt, wr;
    const int* pk;       // n_pairs: i | j << 12 | kind << 24 (kind 0 volume pair, 1 facet pair)
    const double* cf;    // d x n_pairs: S_m[i,j] (volume) or C_ij in row 0 (facet)
    const int* vt;       // n_pass x wv x N_q: slot * 2 + (1 if subtracted), -1 = none
    const int* ft;       // n_pass x wf x N_f
    const int* rt_idx;   // wrt x N_q: R' rows (facet node), -1 = none
    const double* rt_val;
    const int* r_idx;    // wr x N_f: R rows (volume node), -1 = none
    const double* r_val;
    const double* VA;    // warped-product factors (column-major, as in the reference)
    const double* VB;
    const double* VC;
    const double* D1;    // StandardForm scalar kernels: 1D derivative matrix of each collapsed
                         // direction, d x N1 x N1 row-major (D_m = sum_l diag(c_lm) D1_l along lines)
    const double* Mw;    // weight-adjusted inverse mass matrix of every element, V' diag(W / J_k) V
                         // (N_p x N_p column-major per element), or null: the kernels then apply
                         // V, W / J_k and V' by sum factorisation
    int xcd_remap;       // 1: spread contiguous element ranges over the 8 XCDs
    double* uf_ext;      // when set, nodal_values also writes each facet state into the
                         // neighbour's exterior slot [k][e][j] (mapP is an involution)
};

// Low-storage RK stage fused into the time-derivative epilogue.
struct RKStage {
    int active;     // 0: write dudt
    double A, B, dt;
    double* tmp;    // low-storage register, (N_p, N_c, N_e)
    double* u;      // state updated in place
    const double* dt_e;  // per-element time step (batched independent problems), or null for dt
    double* keep;        // line-structured kernel only: receives the state before the update, or null
    __device__ __forceinline__ double dt_of(size_t k) const { return dt_e ? dt_e[k] : dt; }
};

}  // namespace sse
