/*
 * sse_hip.h -- C-ABI of libsse_hip.so, the MI355X (gfx950) semi-discrete
 * residual of StableSpectralElements.jl (the reference), as a drop-in for the
 * reference's `semi_discrete_residual!` path.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns 0 on success or a nonzero sse_status; sse_last_error() then gives a
 * message. A handle owns one HIP stream and its device buffers; like the
 * reference's `Solver` (src/Solvers/Solvers.jl:259-272, which owns mutable
 * scratch) a handle is not reentrant: calls on one handle must not overlap.
 *
 * Array layouts are the reference's column-major (Julia) layouts, so a Julia
 * caller passes its arrays unchanged and a NumPy caller passes Fortran-ordered
 * arrays with the reference's shapes.
 */
#ifndef SSE_HIP_H
#define SSE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSE_ABI_VERSION 3

typedef enum {
    SSE_OK = 0,
    SSE_ERR_INVALID_ARGUMENT = 1,
    SSE_ERR_UNSUPPORTED = 2,
    SSE_ERR_HIP = 3,
    SSE_ERR_NONPHYSICAL_STATE = 4, /* rho <= 0 or p <= 0 met inside a kernel */
    SSE_ERR_OUT_OF_MEMORY = 5
} sse_status;

/* element shapes (StartUpDG Line/Tri/Tet/Hex) */
typedef enum { SSE_LINE = 1, SSE_TRI = 2, SSE_TET = 3, SSE_HEX = 5 } sse_element;

/* conservation laws (src/ConservationLaws/ in the reference) */
typedef enum {
    SSE_EULER = 1,                     /* EulerEquations{d}, N_c = d + 2 */
    SSE_LINEAR_ADVECTION = 2,          /* LinearAdvectionEquation{d}, N_c = 1 */
    SSE_LINEAR_ADVECTION_DIFFUSION = 3 /* LinearAdvectionDiffusionEquation{d} */
} sse_law;

/* residual forms (src/Solvers/Solvers.jl:90-115) */
typedef enum { SSE_FLUX_DIFFERENCING = 1, SSE_STANDARD = 2 } sse_form;

/* inviscid numerical fluxes (src/ConservationLaws/ConservationLaws.jl:51-62) */
typedef enum {
    SSE_LAX_FRIEDRICHS = 1,
    SSE_CENTRAL = 2,
    SSE_ENTROPY_CONSERVATIVE_NUMERICAL_FLUX = 3
} sse_inviscid_flux;

/* mapping forms of StandardForm (src/Solvers/Solvers.jl:75-76; operators.jl:1-37, 83-160) */
typedef enum { SSE_SKEW_SYMMETRIC_MAPPING = 1, SSE_STANDARD_MAPPING = 2 } sse_mapping_form;

/* viscous numerical fluxes (src/ConservationLaws/ConservationLaws.jl:64-67) */
typedef enum { SSE_NO_VISCOUS_FLUX = 0, SSE_BR1 = 1 } sse_viscous_flux;

/* two-point fluxes (src/ConservationLaws/ConservationLaws.jl:69-73) */
typedef enum { SSE_TWO_POINT_EC = 1, SSE_TWO_POINT_CONSERVATIVE = 2 } sse_two_point_flux;

/* mass-matrix solvers (src/Solvers/mass_matrix.jl:1-17) */
typedef enum { SSE_MASS_WEIGHT_ADJUSTED = 1, SSE_MASS_DIAGONAL = 2 } sse_mass_solver;

/* facet-correction operator kind (src/Solvers/operators.jl:162-221) */
typedef enum { SSE_C_NONE = 0, SSE_C_DENSE = 1, SSE_C_SPARSE = 2 } sse_c_kind;

/* where the state arrays of a call live */
typedef enum { SSE_DEVICE = 0, SSE_HOST = 1 } sse_memory;

/*
 * Everything the reference's `Solver{..., FluxDifferencingOperators, ...}`
 * holds for one discretisation (src/Solvers/operators.jl:39-81,
 * src/Solvers/mass_matrix.jl:11-17, 59-75, src/Solvers/Solvers.jl:311-333).
 * All pointers are host pointers; sse_create copies what it needs.
 */
typedef struct sse_desc {
    int32_t abi_version;  /* SSE_ABI_VERSION */
    int32_t element;      /* sse_element */
    int32_t d;            /* spatial dimension */
    int32_t N_p;          /* solution coefficients per element */
    int32_t N_q;          /* volume quadrature nodes */
    int32_t N_f;          /* facet quadrature nodes (all faces) */
    int32_t N_fac;        /* faces per element */
    int32_t N_c;          /* conservative variables */
    int32_t N_e;          /* elements */
    int32_t law;          /* sse_law */
    int32_t form;         /* sse_form */
    int32_t inviscid_flux;   /* sse_inviscid_flux */
    int32_t two_point_flux;  /* sse_two_point_flux */
    int32_t mass_solver;     /* sse_mass_solver */
    int32_t S_sparse;        /* 1: S_m stored as SparseMatrixCSC in the reference */
    int32_t C_kind;          /* sse_c_kind */
    int32_t mapping_form;    /* sse_mapping_form (StandardForm only) */
    int32_t viscous_flux;    /* sse_viscous_flux (second-order laws only) */
    double gamma;            /* Euler: ratio of specific heats */
    double half_lambda;      /* LaxFriedrichsNumericalFlux.halfλ */
    double a[3];             /* advection velocity */
    double b;                /* diffusion coefficient */

    /* reference-element operators, column-major */
    const double* V;            /* N_q x N_p Vandermonde; NULL = identity (nodal) */
    const double* R;            /* N_f x N_q extrapolation */
    const int64_t* R_selection; /* N_f, 1-based volume node of each facet node for
                                   diagonal-E operators (SelectionMap), else NULL */
    const double* W;            /* N_q volume weights */
    const double* B;            /* N_f facet weights */
    const double* S;            /* N_q x N_q x d, S[:, :, m] = 1/2 (W D_m - D_m' W) */
    const double* C;            /* N_q x N_f, R' B (NULL when C_kind == SSE_C_NONE) */
    const double* Minv;         /* N_p x N_p weight-adjusted M^{-1}; NULL = identity */
    const double* D;            /* N_q x N_q x d reference derivative operators D[:, :, m] = D_m
                                   (ReferenceApproximation.D; StandardForm only, else NULL) */

    /* Optional sum factorisation of V: the fields of the reference's
     * WarpedTensorProductMap2D/3D (src/MatrixFreeOperators/warped_product_2d.jl:2-23,
     * warped_product_3d.jl:2-35, built by src/SpatialDiscretizations/tensor_simplex.jl:84-140),
     * modes ordered by nested loops over (i, j[, k]) with i + j + k <= V_p.
     * V_p = 0 means "not given" (V is then applied densely). */
    const double* V_A;          /* M1 x (V_p+1) */
    const double* V_B;          /* M2 x (V_p+1) x (V_p+1) */
    const double* V_C;          /* M3 x (V_p+1)^3, 3D only */
    int32_t V_p;
    int32_t V_M[3];             /* M1, M2, M3: 1D volume quadrature sizes */

    /* geometric factors (GeometricFactors, SpatialDiscretizations.jl:283-289). For
     * StandardForm, Lambda_q, nJf and J_f are those after apply_reference_mapping
     * (SpatialDiscretizations.jl:398-411), as ReferenceOperators/PhysicalOperators hold them. */
    const double* J_q;       /* N_q x N_e */
    const double* Lambda_q;  /* N_q x d x d x N_e, J dxi_l/dx_m */
    const double* nJf;       /* d x N_f x N_e */
    const double* J_f;       /* N_f x N_e */
    const double* nJq;       /* d x N_fac x N_q x N_e; NULL = derived from Lambda_q */
    const double* face_normals; /* d x N_fac reference normals nrstJ of each face */
    const int64_t* mapP;     /* N_f x N_e, 1-based linear index into (N_f, N_e) */

    /* 1: every sum of the line-structured kernels (ModalTensor Tri/Tet Euler: the time
     * derivative and the VJP) in a fixed order, so that repeated calls are bitwise equal; the
     * default (0) accumulates the partner shares and the facet-correction sums with LDS
     * atomics (faster, equal to the last bits only). No counterpart in the reference,
     * whose threaded loops have a fixed order. */
    int32_t deterministic;
} sse_desc;

typedef struct sse_handle sse_handle;

/* Kernel slots reported by sse_profile_residual. */
#define SSE_KERNEL_NODAL_VALUES 0   /* entropy projection + facet extrapolation */
#define SSE_KERNEL_TIME_DERIVATIVE 1 /* interface flux, Hadamard sum, facet correction,
                                        lift, V' and mass solve */
#define SSE_NUM_KERNELS 2

/* Replaces Solver construction (src/Solvers/Solvers.jl:311-333) and
 * WeightAdjustedSolver (src/Solvers/mass_matrix.jl:11-17, 59-75): validates the
 * descriptor, copies operators and geometry to the device, builds the kernels'
 * sparse tables. On failure *out is NULL and sse_last_error(NULL) explains. */
int sse_create(const sse_desc* desc, sse_handle** out);

/* Replaces semi_discrete_residual!(dudt, u, solver, t)
 * (src/Solvers/Solvers.jl:474-514). u and dudt are (N_p, N_c, N_e) column-major;
 * `where` says whether they are device or host pointers. dudt is overwritten.
 * The call is asynchronous on the handle's stream for SSE_DEVICE. */
int sse_residual(sse_handle* h, const double* u, double* dudt, double t, int32_t where);

/* Advances u (N_p, N_c, N_e) by nsteps steps of the 5-stage, 4th-order
 * low-storage Runge-Kutta method of Carpenter & Kennedy (1994), the
 * `CarpenterKennedy2N54()` integrator the reference hands its ODEProblem to
 * (test/test_driver.jl:77-83), with the stage update fused into the residual. */
int sse_rk_steps(sse_handle* h, double* u, double t0, double dt, int32_t nsteps, int32_t where);

/* Largest stage count sse_erk_steps accepts. */
#define SSE_ERK_MAX_STAGES 16

/* Advances u (N_p, N_c, N_e) by nsteps steps of size dt of the explicit Runge-Kutta
 * method with the Butcher tableau (A: stages x stages, row-major, strictly lower
 * triangular; b: stages weights), as OrdinaryDiffEq runs an explicit method with
 * adaptive=false: k_i = R(u + dt sum_{j<i} A[i][j] k_j), then u += dt sum_i b_i k_i.
 * With the Dormand-Prince 8(5,3) tableau this replaces solve(ode, DP8(),
 * adaptive=false, dt=dt) of the reference's 3D Euler test (test/euler_3d.jl:45-53).
 * The stage states are combined on the device (one kernel per stage); the library
 * keeps stages + 1 state-sized buffers on the handle. A and b are host arrays. */
int sse_erk_steps(sse_handle* h, double* u, double t0, double dt, int32_t nsteps, int32_t stages,
                  const double* A, const double* b, int32_t where);

/* Runs the residual nrep times back to back on device buffers and reports the
 * mean duration of each kernel slot in milliseconds (ms_out has SSE_NUM_KERNELS
 * entries), timed with HIP events on the handle's stream. No counterpart in the
 * reference (which times with TimerOutputs, src/Solvers/Solvers.jl:474). */
int sse_profile_residual(sse_handle* h, const double* u_dev, double* dudt_dev, int32_t nrep,
                         double* ms_out);

/* Writes a one-line description of the kernels a handle runs (kernel family,
 * launch variant, pair and table sizes) into buf (NUL-terminated, at most len
 * bytes). No counterpart in the reference; for tests and benchmarks. */
int sse_describe(const sse_handle* h, char* buf, int32_t len);

/* Vector-Jacobian product of the residual (the pull-back a ChainRulesCore.rrule of
 * semi_discrete_residual! returns; no counterpart in the reference, which has no
 * derivatives): ubar = (dR/du)' ybar and *lambda_bar = (dR/dλ)' ybar, λ the
 * LaxFriedrichsNumericalFlux coefficient (halfλ = λ / 2). u, ybar, ubar are
 * (N_p, N_c, N_e), device or host per `where`; lambda_bar is a host scalar.
 * Implemented for the Euler equations in FluxDifferencingForm with the
 * entropy-conservative two-point flux (any element, modal or nodal); returns
 * SSE_ERR_UNSUPPORTED otherwise. Blocks until done. */
int sse_residual_vjp(sse_handle* h, const double* u, const double* ybar, double* ubar, double* lambda_bar,
                     int32_t where);

/* Discrete adjoint of nsteps CarpenterKennedy2N54 steps of size dt from u0 for the
 * objective J = entropy integral of the final state (SSE_ENTROPY; the only
 * objective so far): writes J, grad_u0 = dJ/du0 (N_p, N_c, N_e) and *grad_lambda =
 * dJ/dλ. The forward states are checkpointed every `checkpoint_every` steps and
 * each segment is recomputed once, keeping its stage states; with the tensor VJP
 * kernels, the recomputation runs on a second stream of the library's beside the
 * reverse sweep, and the forward sweep keeps its last two segments (environment
 * SSE_ADJ_OVERLAP=0: one stream, the last segment kept). No counterpart in the
 * reference. u0 and grad_u0 are device or host per `where`. Blocks. */
int sse_rk_adjoint(sse_handle* h, const double* u0, double dt, int32_t nsteps, int32_t checkpoint_every,
                   int32_t objective, double* J, double* grad_u0, double* grad_lambda, int32_t where);

/* Functionals of src/Analysis/conservation.jl:113-189 */
typedef enum {
    SSE_MASS = 0,    /* sum_k 1' WJ_k V u_k, N_c values (PrimaryConservationAnalysis) */
    SSE_ENERGY = 1,  /* 1/2 sum_k u_k' M_k u_k, N_c values (EnergyConservationAnalysis) */
    SSE_ENTROPY = 2  /* sum_k sum_i WJ_k,i S((V u_k)_i), 1 value (EntropyConservationAnalysis) */
} sse_functional_kind;

/* Replaces evaluate_conservation (dudt == NULL) and evaluate_conservation_residual
 * (dudt != NULL: the functional's rate along dudt) of src/Analysis/conservation.jl:113-189,
 * with M_k the handle's mass matrix (src/Solvers/mass_matrix.jl:140-153). u and dudt
 * are (N_p, N_c, N_e), device or host memory per `where`; out (host) receives N_c
 * values for SSE_MASS and SSE_ENERGY, 1 for SSE_ENTROPY. Sums run in a fixed order.
 * Blocks until the values are in out. */
int sse_functional(sse_handle* h, const double* u, const double* dudt, int32_t which, int32_t where, double* out);

/* Replaces analyze(ErrorAnalysis(results_path, law, spatial_discretization), sol,
 * exact_solution, t; normalize) with the default error quadrature, the volume quadrature
 * (src/Analysis/error.jl:13-57, 58-91): out[e] = sqrt(sum_k sum_i W_i J_k,i
 * (u_exact[i, e, k] - (V u_k)_{i,e})^2), the sum divided by the total volume sum W J
 * inside the square root when normalize != 0. u is (N_p, N_c, N_e); u_exact is
 * (N_q, N_c, N_e), the exact solution evaluated at mesh.xyzq at time t; both at `where`.
 * out (host) receives N_c values. Blocks until they are there. */
int sse_l2_error(sse_handle* h, const double* u, const double* u_exact, int32_t normalize, int32_t where,
                 double* out);

/* Largest N_e * N_c + N_c that sse_functional_noalloc accepts (doubles of its workspace). */
#define SSE_FUNCTIONAL_NOALLOC_MAX_DOUBLES (1 << 20)

/* sse_functional for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the
 * reference. Same functionals and the same kernels as sse_functional, but u, dudt
 * (NULL for the functional itself) and out are DEVICE pointers, out receives N_c values
 * (SSE_MASS, SSE_ENERGY) or 1 (SSE_ENTROPY), and the call is asynchronous on the
 * handle's stream: it allocates nothing, copies nothing to the host and does not wait.
 * The per-element partial sums live in one fixed workspace per GPU inside the library
 * (SSE_FUNCTIONAL_NOALLOC_MAX_DOUBLES doubles, held until the process ends); a handle
 * with N_e * N_c + N_c above that is rejected with SSE_ERR_INVALID_ARGUMENT. Every
 * call on one GPU shares that workspace, so calls that may overlap on one GPU (other
 * handles, other streams) race: order them on one stream. The energy functionals form
 * each element's mass-matrix factor in LDS on every call (sse_functional keeps them). */
int sse_functional_noalloc(sse_handle* h, const double* u, const double* dudt, int32_t which, double* out);

/* Uses the given hipStream_t (0 = the handle's own stream) for later calls. */
int sse_set_stream(sse_handle* h, void* hip_stream);

/* Sets the diffusion coefficient b of an advection-diffusion handle for later calls
 * (SSE_ERR_UNSUPPORTED for other laws). No counterpart in the reference, whose laws are
 * values (a new Solver per coefficient, Solvers.jl:429-452); it is for callers that sweep
 * b on one mesh (the adjoint in b), so that the operators and geometric factors are not
 * copied to the device again for every coefficient. A captured RK step is re-captured. */
int sse_set_diffusion(sse_handle* h, double b);

/* Blocks until work queued on the handle's stream has finished. */
int sse_synchronize(sse_handle* h);

/* Message of the last failure on h (or of the last failed sse_create when h is NULL). */
const char* sse_last_error(const sse_handle* h);

/* Frees the handle's device memory and stream. */
void sse_destroy(sse_handle* h);

/* ABI version the library was built with. */
int32_t sse_abi_version(void);

/*
 * Batched independent problems. No counterpart in the reference, which runs one
 * problem per Solver; a caller would otherwise loop over per-problem solves.
 * Problems whose descriptors share the element type, reference operators, law,
 * form and fluxes are concatenated into one group (one kernel launch per phase
 * for all of them); a batch holds one group per distinct discretisation.
 * Geometry and mapP are per problem, exactly as for sse_create.
 *
 * Ensembles: StandardForm problems of a scalar law (linear advection or
 * advection-diffusion, ModalTensor Tri/Tet) with the same discretisation and
 * the same mesh (equal geometric factors and mapP, compared by content) that
 * differ at most in the diffusion coefficient b -- a parameter sweep -- form one
 * group that keeps one copy of the geometry; each launch advances up to a
 * bounded number of members side by side (16 GiB of intermediates at most).
 */
typedef struct sse_batch sse_batch;

/* descs[i] describes problem i (i < n). On failure *out is NULL and
 * sse_batch_last_error(NULL) explains. */
int sse_batch_create(const sse_desc* const* descs, int32_t n, sse_batch** out);

/* Advances every problem i by nsteps CarpenterKennedy2N54 steps of its own
 * time step dt[i]. u[i] points to problem i's (N_p, N_c, N_e_i) state, device or
 * host memory per `where`; it is read at the start and written at the end. With
 * use_graph != 0 the RK step of all groups is captured once in a HIP graph and
 * replayed; otherwise the kernels are launched step by step. Blocks until done. */
int sse_batch_rk_steps(sse_batch* b, double* const* u, const double* dt, int32_t nsteps, int32_t where,
                       int32_t use_graph);

/* Discrete tangent with respect to the diffusion coefficient, for batches
 * whose groups are all ensembles of advection-diffusion problems: advances each
 * problem's state u[i] and v[i] = d u[i] / d b_i together by nsteps
 * CarpenterKennedy2N54 steps of dt[i] (v[i] usually starts at zero). The
 * residual is linear in u and in b, so the pair obeys u' = R(u; b),
 * v' = R(v; b) + dR/db(u) with dR/db(u) the viscous terms of u at unit b; the
 * same RK stages applied to both give the exact derivative of the discrete
 * solution, and the gradient of any functional J(u_N) with respect to b_i is
 * <dJ/du, v_N> (for the energy 1/2 u' M u it is the energy rate of u along v,
 * sse_functional with SSE_ENERGY and dudt = v). For one parameter this forward sweep is
 * the cheapest exact gradient: two residuals per stage, no checkpoints. No
 * counterpart in the reference (ForwardDiff through solve would compute the
 * same derivative). Blocks until done. */
int sse_batch_rk_tangent(sse_batch* b, double* const* u, double* const* v, const double* dt, int32_t nsteps,
                         int32_t where);

/* One-line description of the groups (count, elements, kernel family). */
int sse_batch_describe(const sse_batch* b, char* buf, int32_t len);

/* Message of the last failure on b (or of the last failed sse_batch_create when b is NULL). */
const char* sse_batch_last_error(const sse_batch* b);

/* Frees the batch, its groups and its HIP graph. */
void sse_batch_destroy(sse_batch* b);

/* ------------------------------------------------------------------ set-up on the device
 * The steps that build a problem's arrays before the residual path: geometric factors of
 * curved elements, the reference mapping of StandardForm, and the L2 projection of the
 * initial data. One workgroup per element; the calls block until done and report errors
 * through sse_last_error(NULL). `where` says where the element arrays live; the reference
 * operators in sse_geometry_ops are host pointers. All arrays are column-major. */

/* metric types (src/SpatialDiscretizations/SpatialDiscretizations.jl:154-156) */
typedef enum { SSE_METRICS_EXACT = 1, SSE_METRICS_CURL = 2 } sse_metrics;

/* The RefElemData fields GeometricFactors reads (StartUpDG's RefElemData: Drst, Vq, Vf,
 * nrstJ), the Jacobian projector of project_jacobian!, and for conservative-curl metrics on
 * tetrahedra the degree-(N+1) nodal operators of mesh.jl:416-433. */
typedef struct sse_geometry_ops {
    int32_t d;           /* 1, 2 or 3 */
    int32_t metrics;     /* sse_metrics */
    int32_t N_p;         /* mapping nodes per element (degree-N nodal set) */
    int32_t N_q, N_f, N_fac;
    const double* D;     /* N_p x N_p x d: Dr, Ds, Dt at the mapping nodes */
    const double* Vq;    /* N_q x N_p: mapping nodes -> volume quadrature nodes */
    const double* Vf;    /* N_f x N_p: mapping nodes -> facet quadrature nodes */
    const double* nrstJ; /* N_f x d: scaled reference normals at the facet nodes */
    const double* P_J;   /* N_q x N_q: V (V' W V)^-1 V' W applied to J_q, or NULL */
    int32_t N_p1;        /* degree-(N+1) nodes (curl metrics on tetrahedra, else 0) */
    const double* I1;    /* N_p1 x N_p: degree-N nodal -> degree-(N+1) nodal */
    const double* D1;    /* N_p1 x N_p1 x 3: derivatives at the degree-(N+1) nodes */
    const double* Vq1;   /* N_q x N_p1: Vq times the degree-(N+1) -> degree-N interpolation */
    const double* Vf1;   /* N_f x N_p1 */
} sse_geometry_ops;

/* Replaces GeometricFactors(mesh, reference_element, metric_type)
 * (src/SpatialDiscretizations/mesh.jl:229-282 exact, :284-339 curl on triangles, :410-506
 * curl on tetrahedra) followed by project_jacobian! when P_J is given
 * (SpatialDiscretizations.jl:310-317). x, y, z: N_p x N_e mapping-node coordinates (y, z
 * NULL below d = 2, 3). Outputs: J_q (N_q, N_e), Lambda_q (N_q, d, d, N_e), J_f (N_f, N_e),
 * nJf (d, N_f, N_e), nJq (d, N_fac, N_q, N_e). */
int sse_geometric_factors(const sse_geometry_ops* ops, int32_t N_e, const double* x, const double* y,
                          const double* z, double* J_q, double* Lambda_q, double* J_f, double* nJf,
                          double* nJq, int32_t where);

/* Replaces apply_reference_mapping (SpatialDiscretizations.jl:395-411):
 * out[i, m, n, k] = sum_l Lambda_ref[i, m, l] Lambda_q[i, l, n, k] / J_ref[i].
 * Lambda_ref (N_q, d, d) and J_ref (N_q) are host pointers. */
int sse_apply_reference_mapping(int32_t N_q, int32_t d, int32_t N_e, const double* Lambda_ref,
                                const double* J_ref, const double* Lambda_q, double* out, int32_t where);

/* Replaces project_function with a modal V (src/Solvers/Solvers.jl:396-417): per element,
 * u = (V' W J V)^-1 V' W J u_q by a Cholesky factorisation. V (N_q, N_p) and W (N_q) are host
 * pointers; J_q (N_q, N_e), u_q (N_q, N_c, N_e) and u (N_p, N_c, N_e) live at `where`.
 * SSE_ERR_INVALID_ARGUMENT names the first element whose matrix is not positive definite
 * (the reference's cholesky throws there). */
int sse_project_l2(int32_t N_q, int32_t N_p, int32_t N_c, int32_t N_e, const double* V, const double* W,
                   const double* J_q, const double* u_q, double* u, int32_t where);

/* The device's entropy-conservative two-point flux of the Euler equations, contracted with
 * a vector: c[t] = sum_n a[t, n] F(uL[t], uR[t])[:, n], F the Ranocha flux of
 * compute_two_point_flux (src/ConservationLaws/euler_navierstokes.jl:171-195), evaluated
 * by the same device code as the residual kernels (hoisted logarithms, one refined
 * reciprocal per pair). uL, uR, c are n x (d + 2) and a is n x d, row-major (one pair per
 * row), at `where`. No counterpart in the reference: a test aid that checks the flux the
 * kernels evaluate against the oracle's, pair by pair. Blocks until done. */
int sse_ec_flux_batch(int32_t d, double gamma, int32_t n, const double* uL, const double* uR, const double* a,
                      double* c, int32_t where);

#ifdef __cplusplus
}
#endif

#endif /* SSE_HIP_H */
