/*
 * mumag_hip.h -- C ABI of libmumag_hip.so, the MI355X (gfx950) implementation of the
 * mammos-mumag per-field-step hot path.
 *
 * Conventions
 *  - Every entry point returns int: 0 = ok, < 0 = error (see MM_E_*), > 0 = warning
 *    (MM_W_MAXITER: an iteration cap was reached; the result is still returned, as the
 *    reference does, since it discards jax CG's info: scripts/solvers.py:222-231).
 *  - Vectors are caller-owned DEVICE pointers (e.g. torch.Tensor.data_ptr() on cuda:0)
 *    unless the name ends in _host.  Nodal 3-vectors are interleaved f64[3N]
 *    ([3i+a] = component a of node i), scalars f64[N], in mesh node order.
 *  - The context owns the operators and all workspace; no entry point allocates after
 *    mm_ctx_create, except the set-up calls of the accelerated forms (mm_multigrid,
 *    mm_mlp_create, and mm_warm_start the first time it enables a network).  One HIP stream per context; calls return after the stream is
 *    synchronised whenever a scalar is returned to the host.  A context is not
 *    thread-safe.
 *  - mm_last_error() returns a message for the last failure on the calling thread.
 */
#ifndef MUMAG_HIP_H
#define MUMAG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM_OK 0
#define MM_E_ARG (-1)      /* invalid argument (null pointer, bad size, bad option) */
#define MM_E_HIP (-2)      /* HIP runtime failure */
#define MM_E_MESH (-3)     /* degenerate mesh (zero-volume element, no magnetic element) */
#define MM_E_STATE (-4)    /* call not valid in this context state */
#define MM_W_MAXITER 1     /* iteration cap reached */

#define MM_PREC_FP64 0     /* every operator and vector in FP64 */
#define MM_PREC_MIXED 1    /* FP32 inner PCG and truncated CG, FP64 dots, energies, outer PCG residual */

typedef struct mm_ctx mm_ctx;

/* Mesh and materials.  Replaces the escript set-up of scripts/mapping.py:171-236
 * (Materials: scripts/materials.py:12-41; operators: scripts/matrix.py:9-74,
 * scripts/trafo.py:4-45; meas: scripts/escript_tools.py:13-17). */
typedef struct {
    int32_t n_nodes;          /* N */
    int32_t n_elems;          /* NE, all Tet4 elements (magnet and air) */
    const double *coords;     /* host f64[N*3], nm */
    const int32_t *tets;      /* host i32[NE*4], 0-based node indices */
    const int32_t *elem_mat;  /* host i32[NE], index into the material table */
    int32_t n_mats;
    const double *mat_Js;     /* host f64[n_mats], T   (Js <= 0: non-magnetic) */
    const double *mat_A;      /* host f64[n_mats], J/m */
    const double *mat_K1;     /* host f64[n_mats], J/m^3 */
    const double *mat_u;      /* host f64[n_mats*3], easy axis (unit) */
    const uint8_t *dirichlet; /* host u8[N]: 1 = u fixed to 0 (identity row/col in K) */
    const uint8_t *shell;     /* host u8[NE] or NULL: 1 = Imhoff shell tensor element */
    double shell_R;           /* inner radius of the shell transform (if shell != NULL) */
    double shell_Rinf;        /* outer radius */
    double length_scale;      /* metres per coordinate unit; the reference fixes 1e-9 */
    /* Optional structured Kuhn-grid description (grid_n all zero = unstructured).  When
     * set, the mesh must be the one the structured generator writes: nodes numbered
     * i + (nx+1) (j + (ny+1) k) at (grid_x[i], grid_y[j], grid_z[k]); elements voxel-major
     * (voxel v = vx + nx (vy + ny vz)) with the 6 Kuhn tetrahedra of each voxel in the
     * generator's order; the magnetic elements exactly the voxels of the cube
     * [grid_mag0, grid_mag0 + grid_magn)^3, of uniform edge; an air box (Dirichlet on the
     * six faces).  mm_ctx_create validates all of this and fails with MM_E_MESH if a
     * structured context was requested for a mesh that does not match. */
    int32_t grid_n[3];        /* voxels per axis of the whole grid */
    const double *grid_x;     /* host f64[nx+1], f64[ny+1], f64[nz+1] node coordinates */
    const double *grid_y;
    const double *grid_z;
    int32_t grid_mag0[3];     /* first voxel of the magnet block per axis */
    int32_t grid_magn;        /* magnet voxels per axis */
} mm_mesh_desc;

typedef struct {
    int32_t precision;        /* MM_PREC_FP64 or MM_PREC_MIXED */
    int32_t device;           /* HIP device ordinal */
    double drop_tol;          /* drop |K_ij| < drop_tol * max(|K_ii|,|K_jj|); 0 keeps all */
    int32_t pcg_batch;        /* PCG iterations launched between host checks (0 = default) */
    int32_t use_structured;   /* 1: use the Kuhn-grid stencil kernels (needs grid_n) */
    double inner_tol;         /* MM_PREC_MIXED: FP32 residual reduction between two FP64
                                 reliable updates (0 = default 0.1) */
    int32_t eval_precision;   /* MM_PREC_MIXED: arithmetic of the energy evaluation's operator
                                 applications (C m, D m, G u of total_eg, and mm_apply_C /
                                 mm_divergence / mm_gradient): MM_PREC_FP64 (0, default), so
                                 that energies, their line-search differences and the PCG
                                 right-hand side carry no FP32 rounding, or MM_PREC_MIXED
                                 (FP32).  The PCG inner iterations and the truncated CG run in
                                 FP32 either way. */
} mm_options;

typedef struct {
    int64_t n_nodes, n_elems, n_mag_elems, n_mag_nodes;
    int64_t k_nnz, k_sell_entries;   /* stored K entries, SELL-64 padded entries */
    double volume;                   /* V, nm^3 (scripts/materials.py:36-37) */
    int32_t structured;              /* 1 if the stencil kernels are active */
    int32_t precision;
    int64_t device_bytes;            /* device memory held by the context */
} mm_info;

/* Minimiser parameters: scripts/tools.py:64-120 (read_params) and loop.py:58-65. */
typedef struct {
    double tol_fun;           /* NCG tolerance (minimizers.py:164-166) */
    double tol_u;             /* PCG tolerance of the potential solve */
    int32_t precond_iter;     /* truncated-CG iterations of M_inv (loop.py:40-55) */
    int32_t hmag_on;          /* 0: no magnetostatics (energies.py:89-94) */
    int32_t iter_max;         /* NCG cap; the reference hard-codes 100000 (minimizers.py:142) */
} mm_min_params;

typedef struct {
    int64_t ncg_iters;        /* NCG iterations (the reference's cg_iter, loop.py:64, 119) */
    int64_t fcalls;           /* total_eg evaluations (the reference's stats) */
    int64_t pcg_iters;        /* Poisson PCG iterations */
    int64_t pcg_calls;
    int64_t tcg_iters;        /* truncated-CG iterations inside M_inv */
    int64_t c_applies;        /* exchange/anisotropy operator applications */
    int64_t backtracks;       /* line-search halvings (minimizers.py:121-127) */
    int64_t restarts_g;       /* direction reset to -g before the line search (minimizers.py:171-173) */
    int64_t restarts_z;       /* direction reset to -z after the HS update (minimizers.py:182-184) */
} mm_stats;

const char *mm_last_error(void);

/* Build the operators on the device.  Replaces escript2arrays (scripts/mapping.py:171-236). */
int mm_ctx_create(const mm_mesh_desc *mesh, const mm_options *opts, mm_ctx **out);
int mm_ctx_destroy(mm_ctx *ctx);
int mm_ctx_info(mm_ctx *ctx, mm_info *out);
/* Copies meas (f64[N], T nm^3) and the exchange diagonal D (f64[3N]) to host arrays. */
int mm_get_meas_host(mm_ctx *ctx, double *meas_host);
int mm_get_diag_host(mm_ctx *ctx, double *dex_host, double *diagK_host);

/* Zeeman energy and gradient: scripts/energies.py:7-21 (external_g / external_eg). */
int mm_external_eg(mm_ctx *ctx, const double *m, double hext, const double hdir[3],
                   double *E_host, double *g);
/* Exchange+anisotropy energy and gradient g = C m, E = m.g/2: energies.py:22-33. */
int mm_exani_eg(mm_ctx *ctx, const double *m, double *E_host, double *g);
/* Magnetostatic energy/gradient/potential: energies.py:36-59 (hmag_g / hmag_eg).
 * u_inout: warm start in, potential out.  pcg_iters_host may be NULL. */
int mm_hmag_eg(mm_ctx *ctx, const double *m, double *u_inout, double tol,
               double *E_host, double *g, int32_t *pcg_iters_host);
/* total_eg: energies.py:67-99.  Writes E, the projected gradient P(m,g) = m x (g x m)
 * (energies.py:61-65) and the full gradient g_full; u_inout is the warm start/result. */
int mm_total_eg(mm_ctx *ctx, const double *m, double hext, const double hdir[3],
                double tol_u, int32_t hmag_on, double *u_inout, double *E_host,
                double *g_proj, double *g_full, int32_t *pcg_iters_host);

/* Jacobi PCG on K with jax.scipy.sparse.linalg.cg semantics (solvers.py:222-231):
 * stop when r.r <= max(tol^2 b.b, atol^2) or after maxiter iterations. */
int mm_pcg(mm_ctx *ctx, const double *b, double *x_inout, double tol, double atol,
           int32_t maxiter, int32_t *iters_host);

/* M_inv: truncated diagonally preconditioned CG (solvers.py:7-131) on
 * H v = m_ref x ((C v) x m_ref) - (m_ref . g_full)_i v_i (loop.py:40-55), called with
 * -g.  Writes z (the NCG then uses d = -z). */
int mm_precond(mm_ctx *ctx, const double *m_ref, const double *g_full, const double *g,
               int32_t max_iter, double *z, int32_t *iters_host);

/* Retraction m_out = normalize(x + alpha d) per node, zero vectors kept
 * (jax_tools.py:4-18 update_m).  m_out may alias x. */
int mm_update_m(mm_ctx *ctx, const double *x, double alpha, const double *d, double *m_out);

/* Reductions over f64[n] device vectors. */
int mm_dot(mm_ctx *ctx, const double *a, const double *b, int64_t n, double *out_host);
int mm_norm_inf(mm_ctx *ctx, const double *a, int64_t n, double *out_host);

/* mm_dot without a context, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc); it has no
 * counterpart in the reference.  *out_host = a.b over n doubles on `device`, with the
 * same kernels and fixed summation order as mm_dot (bit-identical result).  It allocates
 * nothing: its partials and result live in a static device array of 16.3 KB per device,
 * enough for every n >= 0 (one partial per block, at most 2048 blocks).  Runs on the
 * device's null stream and returns after the result is on the host.  Calls that overlap
 * on one device share that array and race. */
int mm_dot_noalloc(int32_t device, const double *a, const double *b, int64_t n, double *out_host);

/* M.h = sum_i (h . m_i) meas_i / V in Tesla (loop.py:25-38 compute_mh). */
int mm_compute_mh(mm_ctx *ctx, const double *m, const double hdir[3], double *mh_host);

/* One field step: hestenes_stiefel_ncg (minimizers.py:139-190) with line_search
 * (92-137) on total_eg and M_inv, i.e. loop.py:58-65 minimize.  m_inout is the
 * iterate; u_inout and gF_inout are the reference's alt_args (potential and full
 * gradient) carried across calls; m_ref is the preconditioner's fixed magnetisation
 * (the module-level initial state that M_inv reads, loop.py:40-55 with 153-158).
 * Writes the final energy. */
int mm_minimize(mm_ctx *ctx, double *m_inout, double hext, const double hdir[3],
                const mm_min_params *params, const double *m_ref, double *u_inout,
                double *gF_inout, double *E_host, mm_stats *stats_inout);

/* Low-level operator applications (device vectors), for tests and callers that drive
 * their own iteration:  y = C v (3N);  y = K x (N);  b = Dx mx + Dy my + Dz mz (N);
 * g = interleave(Gx u, Gy u, Gz u) (3N). */
int mm_apply_C(mm_ctx *ctx, const double *v, double *y);
int mm_apply_K(mm_ctx *ctx, const double *x, double *y);
int mm_divergence(mm_ctx *ctx, const double *m, double *b);
int mm_gradient(mm_ctx *ctx, const double *u, double *g);

/* Multigrid preconditioner for the FP32 inner solve of a structured MM_PREC_MIXED context
 * (no counterpart in the reference, whose potential solve is Jacobi-preconditioned,
 * solvers.py:222-231): a geometric V-cycle of up to max_levels levels (2..8; 0 restores
 * Jacobi) replaces z = r / diag(K) inside the reliable-update PCG.  Coarsening is
 * spacing-aware (the fine cells of a graded air box merge first) and stops at 16 nodes
 * per axis.  The stop test on the FP64 true residual is unchanged, so only the
 * iteration count changes. */
#define MM_MG_MAX_LEVELS 8
int mm_multigrid(mm_ctx *ctx, int32_t max_levels);
/* Levels of the V-cycle (0: none) and, if dims is not null, the node counts per axis of
 * each level, dims[3 l + a] for l < *levels (dims holds 3 * MM_MG_MAX_LEVELS). */
int mm_mg_levels(mm_ctx *ctx, int32_t *levels, int32_t *dims);
/* z = M r with the multigrid V-cycle (device f32[N]; r zero on the Dirichlet faces). */
int mm_mg_apply(mm_ctx *ctx, const float *r, float *z);

/* Algebraic multigrid for the FP32 inner solve of a general (unstructured) MM_PREC_MIXED
 * context (no counterpart in the reference, whose potential solve is Jacobi-preconditioned,
 * solvers.py:222-231): smoothed aggregation on K (strong couplings
 * |a_ij| > 0.02 sqrt(a_ii a_jj) between free nodes, three-pass aggregation, one
 * damped-Jacobi smoothing of the piecewise-constant prolongation, Galerkin coarse
 * operators) down to at most 500 unknowns, solved by their dense inverse; Chebyshev
 * smoothing of degree 2 on the finest level and 1 on the coarser ones; up to max_levels
 * levels (2..10; 0 restores Jacobi).  Set up on the
 * host once per context (a later call with the same bound re-enables it; another bound
 * returns MM_E_STATE).  MM_E_STATE on structured or MM_PREC_FP64 contexts and when K is
 * already at the coarsest size, MM_E_ARG for a bound outside 2..10.  The stop test on
 * the FP64 true residual is unchanged, so only the iteration count changes. */
#define MM_AMG_MAX_LEVELS 10
int mm_amg(mm_ctx *ctx, int32_t max_levels);
/* Levels of the AMG (0: none) and, if sizes is not null, the unknowns of each level
 * (sizes holds MM_AMG_MAX_LEVELS). */
int mm_amg_levels(mm_ctx *ctx, int32_t *levels, int64_t *sizes);
/* z = M r with the AMG V-cycle (device f32[N]; r zero on the Dirichlet nodes). */
int mm_amg_apply(mm_ctx *ctx, const float *r, float *z);

/* ---- Learned warm start (no counterpart in the reference: the learned part of the
 * chosen form, for callers that run many potential solves on one structured grid).
 *
 * A per-node MLP (MM_MLP_IN features of a node and its 1-ring -> 128 -> 128 -> 128 -> 1,
 * ReLU, bf16 MFMA with FP32 accumulation, FP32 master weights, Adam) predicts the
 * scalar potential u from m.  With mm_warm_start enabled, every potential solve of
 * mm_total_eg / mm_hmag_eg first moves u by the Galerkin step along the change of the
 * prediction (the first step of a flexible PCG whose first preconditioner application
 * is the network); the Jacobi PCG then ends at the same tolerance, so only the
 * iteration count changes.  Structured (Kuhn grid) contexts only.
 *
 * Features (bf16, node-major [N][MM_MLP_IN]): m of the node and its 6 axis
 * neighbours, coordinates / half box edge, mean magnetisation, magnetic fraction and
 * mean Js of the 8 adjacent voxels.  Parameters: flat FP32 [MM_MLP_PARAMS], W_l
 * row-major [in][out] then b_l, layers 1..4. */
#define MM_MLP_IN 32
#define MM_MLP_PARAMS 37377
typedef struct mm_mlp mm_mlp;
/* A network on ctx's device and stream, He-uniform initialised from seed; training
 * batches of up to max_batch nodes.  Destroy it before its context. */
int mm_mlp_create(mm_ctx *ctx, int64_t max_batch, uint64_t seed, mm_mlp **out);
int mm_mlp_destroy(mm_mlp *net);
int mm_mlp_set_params(mm_mlp *net, const float *params_host);   /* also resets Adam */
int mm_mlp_get_params(mm_mlp *net, float *params_host_out);
/* Features of every node for magnetisation m (device f64[3N]) into X (device u16[N*32]). */
int mm_mlp_features(mm_ctx *ctx, const double *m, uint16_t *X);
/* y[i] = network(X[i]) for n nodes (device pointers). */
int mm_mlp_forward(mm_mlp *net, int64_t n, const uint16_t *X, float *y);
/* Gradients of the loss mean((y - target)^2) over a batch of n <= max_batch nodes
 * (flat, host) and the loss; no update. */
int mm_mlp_grads(mm_mlp *net, int64_t n, const uint16_t *X, const float *target, float *grad_host,
                 double *loss_host);
/* One Adam step (lr, beta = 0.9 / 0.999) on a batch; returns when the step is done;
 * loss_host (the batch loss before the step) may be NULL. */
int mm_mlp_train_step(mm_mlp *net, int64_t n, const uint16_t *X, const float *target, float lr,
                      double *loss_host);
/* Enable (net != NULL) or disable the learned warm start; u = scale * network output. */
int mm_warm_start(mm_ctx *ctx, mm_mlp *net, double scale);
/* Warm-started solves since mm_warm_start and the last Galerkin coefficient. */
int mm_warm_start_stats(mm_ctx *ctx, int64_t *calls_host, double *last_gamma_host);
/* When the network is applied: MM_WS_EVERY (default) before every potential solve;
 * MM_WS_MEASURED per solve in the form, warm-started or plain, with the lower mean wall
 * time (warm start plus PCG, timed on the host) over the solves that used it; the first
 * 3 solves of each form alternate, then the other form is re-measured on solve 15 and at
 * an interval that doubles (16 .. 256 solves) while it stays slower, back to 16 when it
 * wins.  The answer is the plain method's either way. */
#define MM_WS_EVERY 0
#define MM_WS_MEASURED 1
int mm_warm_start_policy(mm_ctx *ctx, int32_t policy);
/* Potential solves and warm-started ones since the policy was set, with the mean wall
 * seconds of plain and warm-started solves (0 for a form no solve has used). */
int mm_warm_start_report(mm_ctx *ctx, int64_t *solves, int64_t *warm_solves, double *mean_s_plain,
                         double *mean_s_warm);

/* Secant start of the line search's potential solves (no counterpart in the reference,
 * whose solves start from the previous solution, solvers.py:222-231): in mm_minimize, every
 * evaluation of a line search after its first starts from the previous solution moved by
 * the Galerkin step along the change of the solution since the line search's first
 * evaluation, gamma = d.(b - K u) / d.K d (d = u - u_first, on the device, no host sync).
 * The PCG ends at the unchanged tolerance, so the answer is the plain start's to within
 * it; only the iteration counts differ.  on = 0 (the default) restores the reference's
 * start.  Allocates N doubles on the first enable. */
int mm_secant_start(mm_ctx *ctx, int32_t on);

/* ---- Learned preconditioner (no counterpart in the reference, whose potential solve is
 * Jacobi-preconditioned, solvers.py:222-231): a second network of the same architecture
 * (mm_mlp_create) maps features of the residual r over a node's neighbourhood to the error
 * e = K^-1 r at the node, and serves as the preconditioner of a flexible CG.  Structured
 * (Kuhn grid) contexts only.
 *
 * Residual features (bf16, node-major [N][MM_MLP_IN]): w = r / diag(K) of the node and of
 * its neighbours one and two nodes away along each axis (13 values, 0 outside the grid),
 * over s = RMS(w); the six adjacent spacings over their mean; the face flag; the magnetic
 * fraction of the 8 adjacent voxels; the coordinates over the half box edge; 0 padding.
 * The network's output y gives z = scale s y (0 on the faces). */
/* Allocates the learned preconditioner's workspace (a set-up call). */
int mm_lp_setup(mm_ctx *ctx);
/* Residual features of r (device f64[N]) into X (device u16[N * MM_MLP_IN]). */
int mm_lp_features(mm_ctx *ctx, const double *r, uint16_t *X);
/* FP64 Jacobi CG on K from x_inout with the stop test of mm_pcg (r.r <= tol^2 b.b, the
 * recursive residual; jax CG semantics, solvers.py:222-231), in which the network (net,
 * output scale `scale`) contributes at the start and after every `every`-th iteration
 * (0: never, which is the Jacobi PCG): a Galerkin step along its predicted error d,
 * x += gamma d with gamma = d.r / d.K d (the K-norm of the error can only decrease), the
 * first step of a flexible CG whose preconditioner alternates between the network and
 * Jacobi; the CG then restarts from the corrected iterate.  The answer meets the Jacobi
 * PCG's tolerance; only the iteration count changes.  iters_host counts CG iterations,
 * lp_apps_host the network's steps.  Returns MM_W_MAXITER at the cap. */
int mm_lp_pcg(mm_ctx *ctx, mm_mlp *net, double scale, const double *b, double *x_inout, double tol,
              int32_t maxiter, int32_t every, int32_t *iters_host, int32_t *lp_apps_host);

/* HIP-event timing of the two fused PCG kernels since the last mm_timing_enable: the
 * K apply with the search-direction update (SELL SpMV or 7-point stencil, FP64 or the
 * FP32 inner solve) and the vector update, and with the multigrid preconditioner one
 * level-0 smoothing step of the iteration's V-cycle.  One iteration in 16 is timed (the
 * counts are of the sampled launches); only launches that did work (before
 * convergence) are counted. */
typedef struct {
    int64_t spmv_launches;
    double spmv_ms;           /* summed device time of those launches */
    int64_t spmv_bytes;       /* algorithmic bytes of one launch */
    int64_t update_launches;
    double update_ms;
    int64_t update_bytes;
    int64_t mg_smooth_launches;
    double mg_smooth_ms;
    int64_t mg_smooth_bytes;
    /* counts since mm_timing_enable: multigrid V-cycles that ran (-1: no coarse tail
     * kernel to count them), true FP64 residual passes and corrections x += dx of the
     * mixed-precision PCG */
    int64_t vcycles;
    int64_t ru_passes;
    int64_t corrections;
} mm_kernel_timing;
int mm_timing_enable(mm_ctx *ctx, int32_t on);
int mm_timing_get(mm_ctx *ctx, mm_kernel_timing *out);

#ifdef __cplusplus
}
#endif
#endif /* MUMAG_HIP_H */
