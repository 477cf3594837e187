/*
 * hz.h -- C ABI of the MI355X Helmholtz library (libhz.so).
 *
 * Matrix-free complex 7-point finite-volume Helmholtz operator with PML,
 * material (SDF-blended) coefficients, reflecting / open / soft / impedance
 * faces, the sinc-Neumann source columns, and restarted GMRES on the
 * device.  Plain pointers and sizes only; device buffers are owned by the
 * caller unless stated otherwise.
 *
 * Conventions
 *   - Unknowns are numbered Fortran-style: p = i + nx*(j + ny*k).
 *   - Complex data are interleaved (re, im).  Device vectors passed to the
 *     ABI are complex128 (16 bytes per value) unless a function says
 *     otherwise; k right-hand sides are stored column-major with a leading
 *     dimension ld >= N (column r starts at r*ld).
 *   - Every function returns an hz_status; on failure hz_last_error()
 *     returns a message (thread-local).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     One operator handle must not be used from two threads at once.
 *
 * Reference interfaces replaced (Rain-Kotsuzui/ultrasound_solver):
 *   hz_create          ConservativeHelmholtz.__init__ + assemble
 *                      (src/algorithm/physics/conservative_helmholtz.py:18-149);
 *                      no CSR is formed, coefficients stay per node.
 *   hz_apply           the SpMV `A_gpu @ vector` and the scaled/preconditioned
 *                      operators of src/algorithm/solvers/iterative_solver.py:37-56,
 *                      and the conj-transpose CSR of helmholtz_solver.py:613-620.
 *   hz_gmres           GpuIterativeSolver.solve (iterative_solver.py:30-104),
 *                      called from HelmholtzDirectSolver._solve_iterative /
 *                      _solve_gpu_adjoint (helmholtz_solver.py:589-632).
 *   hz_set_sources     WarpAssemblyEngine.source_columns
 *                      (warp_utils.py:95-99; sinc_aperture.py:36-85).
 *   hz_build_rhs       WarpAssemblyEngine.assemble_rhs (warp_utils.py:118-120)
 *                      and the basis-build RHS batch (phase_response_basis.py:106-107).
 *   hz_gather_sources  B^H lambda, the adjoint form of G^H c used by
 *                      PhaseResponseBasis.phase_vjp (phase_response_basis.py:208-253).
 *   hz_sample          trilinear rows of PhaseResponseBasis.target_response
 *                      (phase_response_basis.py:165-174).
 *   hz_rasterize_materials  the sphere SDF, smoothing and material blend of
 *                      WarpAssemblyEngine.assemble_system_gpu (warp_utils.py:44-79).
 *   hz_deposit_sources the aperture deposition of source_mass_matrix
 *                      (src/algorithm/physics/sinc_aperture.py:36-72).
 *   hz_z_transfer, hz_surface_*
 *                      HomogeneousSurfaceSolver (src/algorithm/solvers/
 *                      homogeneous_surface.py:41-259), the exact target
 *                      response of HelmholtzDirectSolver.target_response
 *                      (helmholtz_solver.py:179-208).
 */
#ifndef HZ_H
#define HZ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hz_operator hz_operator;

typedef enum hz_status {
    HZ_OK = 0,
    HZ_NOT_CONVERGED = 1,   /* stats are filled in */
    HZ_BAD_ARGUMENT = 2,
    HZ_OUT_OF_MEMORY = 3,
    HZ_DEVICE_ERROR = 4
} hz_status;

typedef enum hz_face_kind {
    HZ_FACE_REFLECTING = 0,
    HZ_FACE_OPEN = 1,
    HZ_FACE_SOFT = 2,
    HZ_FACE_IMPEDANCE = 3
} hz_face_kind;

typedef enum hz_precision {
    HZ_C128 = 0,            /* complex128 vectors, FP64 coefficients */
    HZ_C64 = 1              /* complex64 vectors, FP32 coefficients */
} hz_precision;

typedef enum hz_apply_mode {
    HZ_APPLY_A = 0,             /* y = A x                                   */
    HZ_APPLY_RAM = 1,           /* y = R A M x  (row scale R, Jacobi M)      */
    HZ_APPLY_AH = 2,            /* y = A^H x                                 */
    HZ_APPLY_RAM_H = 3,         /* y = R' A^H M' x (R', M' built from A^H)   */
    HZ_APPLY_RA = 4,            /* y = R A x (the Krylov apply of the
                                   multigrid-preconditioned mixed form)      */
    HZ_APPLY_RA_H = 5           /* y = R' A^H x                              */
} hz_apply_mode;

/* Grid, materials and faces.  All pointers are host memory read during
 * hz_create only.  Face order: -x, +x, -y, +y, -z, +z.                      */
typedef struct hz_grid_desc {
    int32_t nx, ny, nz;
    double dx;                  /* node spacing h                            */
    double omega;               /* angular frequency                         */
    const double *rho;          /* N densities                               */
    const double *kappa;        /* N compressibilities 1/(rho c^2)           */
    const double *weight[3];    /* per-axis half-volume weights, length n_a  */
    const double *stretch[3];   /* per-axis PML stretch, complex, length n_a */
    int32_t face_kind[6];       /* hz_face_kind per face                     */
    /* Per-face diagonal term i*omega*area*Y, complex, n_b*n_c values with the
     * lower remaining axis fastest; NULL for reflecting and soft faces.      */
    const double *face_term[6];
} hz_grid_desc;

typedef struct hz_gmres_opts {
    int32_t restart;            /* Arnoldi steps per cycle (reference: 30)   */
    int32_t maxiter;            /* inner-step limit (reference: 3000)        */
    double rtol;                /* scaled residual target (reference: 1e-6)  */
    double backward_error_tol;  /* componentwise limit (reference: 1e-4)     */
    int32_t precision;          /* HZ_C128 plain form, HZ_C64 mixed form     */
    int32_t adjoint;            /* 0: A x = b, 1: A^H x = b                  */
    int32_t chunk;              /* RHS solved together (mixed form)          */
    int32_t max_refinements;    /* FP64 refinement steps (mixed form)        */
    double inner_rtol;          /* inner c64 tolerance floor (mixed form)    */
    int32_t use_graph;          /* capture each cycle in a hipGraph          */
    int32_t preconditioner;     /* 0 Jacobi (reference), 1 multigrid (mixed) */
    int32_t arnoldi_block;      /* 1: one Arnoldi step at a time; 2: two
                                   basis vectors per Gram-Schmidt pass
                                   (mixed form only)                         */
    double relative_rtol;       /* 0: the reference's acceptance only; > 0:
                                   also ||b-Ax||/||b|| <= relative_rtol in
                                   FP64, which the mixed form's refinement
                                   then aims for as well                     */
} hz_gmres_opts;

typedef struct hz_gmres_stats {
    int32_t info;               /* 0 converged, else the inner-step count    */
    int32_t iterations;         /* inner (Arnoldi) steps taken               */
    int32_t cycles;             /* restart cycles                            */
    int32_t refinements;        /* FP64 refinement steps (mixed form)        */
    double scaled_residual;     /* ||R(Ax-b)|| / ||R b||  (FP64)             */
    double relative_residual;   /* ||Ax-b|| / ||b||       (FP64)             */
    double backward_error;      /* max_i |Ax-b|_i / (|A||x|+|b|)_i           */
    int32_t finite;             /* all entries of x finite                   */
    int32_t converged;          /* reference acceptance test passed (and the
                                   relative residual, when relative_rtol > 0) */
} hz_gmres_stats;

/* One level of the shifted-Laplacian multigrid preconditioner: the same
 * operator rediscretised on a coarser node set.  Tables are normalised by
 * the fine spacing dx: weight = dual-cell width / dx, edge = edge length / dx
 * (length n_a - 1), face terms i*omega*(w_b w_c s_b s_c / dx)*Y.
 * rho and kappa NULL on every level: the library derives each level's
 * fields on the device, 1/rho and kappa of the level above (the operator's
 * own at level 0) by normalised full weighting, and the term of an open face
 * whose face_term is NULL from the derived fields (Y = sqrt(kappa / rho));
 * the host then supplies only the axis tables and the other face terms.    */
typedef struct hz_level_desc {
    int32_t nx, ny, nz;
    const double *rho;          /* n_l densities (host), or NULL             */
    const double *kappa;        /* n_l compressibilities (host), or NULL     */
    const double *weight[3];
    const double *stretch[3];   /* complex                                   */
    const double *edge[3];
    const double *face_term[6]; /* complex, NULL for reflecting/soft faces   */
} hz_level_desc;

hz_status hz_create(const hz_grid_desc *desc, hz_operator **out);
hz_status hz_destroy(hz_operator *op);

/* y[:, r] = op(x[:, r]) for r < nrhs; x and y are device buffers of the
 * precision's complex type (complex128 for HZ_C128, complex64 for HZ_C64). */
hz_status hz_apply(hz_operator *op, int32_t precision, int32_t mode,
                   const void *x, void *y, int32_t nrhs, int64_t ld, void *stream);

/* Solve op x = b for nrhs columns (device complex128, leading dimension ld).
 * x is overwritten (zero initial guess).  stats: host array of nrhs.       */
hz_status hz_gmres(hz_operator *op, const hz_gmres_opts *opts, const void *b,
                   void *x, int32_t nrhs, int64_t ld, hz_gmres_stats *stats, void *stream);

/* Source columns B (N x M, real): nnz entries (node, column, value), host. */
/* Several independent scenes in one call: scene s solves nrhs[s] columns
 * b[s] (leading dimension ld[s] >= N_s) into x[s] with the same options and
 * writes stats[s][0 .. nrhs[s]).  Scenes run concurrently on up to `workers`
 * internal streams (<= 0: 4; at most 16), longest first by N x nrhs, each
 * as hz_gmres would run it (graph replay included), and releases its
 * cached workspace when done, so device memory is bounded by the scenes in
 * flight rather than the batch.  Work on the internal
 * streams starts after the work already queued on `stream`, and `stream` is
 * ordered after all of it on return.  Each operator may appear once.  No
 * single-call counterpart in the reference: its solver handles one scene per
 * process (src/algorithm/solvers/helmholtz_solver.py:589-604). */
hz_status hz_gmres_batch(int32_t nscenes, hz_operator *const *ops, const hz_gmres_opts *opts,
                         const void *const *b, void *const *x, const int32_t *nrhs, const int64_t *ld,
                         hz_gmres_stats *const *stats, int32_t workers, void *stream);

/* Several independent scenes in one lockstep solve (mixed form, HZ_C64):
 * arguments as hz_gmres_batch.  Every scene's chunk of columns iterates in
 * the same GMRES cycles, and each kernel of a cycle runs once for all the
 * scenes still iterating: one grid over every scene's workgroups, each
 * reading its scene's arguments from a per-scene descriptor table (up to 32
 * scenes per launch; more take several).  With opts->use_graph each cycle
 * body of the whole batch is one graph launch.  Columns are taken in rounds
 * of opts->chunk per scene (fewer when the scenes' workspaces together would
 * not fit in 90 % of the free device memory); the workspaces of all scenes
 * are held together during the call and released on return.  Every scene runs with the slab
 * tile shapes of the largest scene that has the level (so that the scenes'
 * launches are one kernel); each scene's own shapes are restored on return.  Per-scene results are those of
 * hz_gmres on the scene alone.  No counterpart in the reference: its solver
 * handles one scene per process (src/algorithm/solvers/helmholtz_solver.py:
 * 589-604). */
hz_status hz_gmres_fused(int32_t nscenes, hz_operator *const *ops, const hz_gmres_opts *opts,
                         const void *const *b, void *const *x, const int32_t *nrhs, const int64_t *ld,
                         hz_gmres_stats *const *stats, void *stream);

/* Launch counters of the solver since the last reset, process-wide:
 * counts[0] graph launches, counts[1] fused multi-scene kernel launches,
 * counts[2] single-scene kernel launches (launches recorded into a graph
 * count once per graph launch, as counts[0]).  reset != 0 zeroes them after
 * reading.  Instrumentation of hz_gmres and hz_gmres_fused; no counterpart
 * in the reference. */
hz_status hz_launch_counts(int64_t *counts, int32_t reset);

hz_status hz_set_sources(hz_operator *op, int32_t num_columns, int64_t nnz,
                         const int64_t *node, const int32_t *column, const double *value);

/* b[:, r] = B w[:, r] for r < nrhs; w is device complex128 (M x nrhs,
 * leading dimension ld_w), b device complex128 (N x nrhs, ld).  With
 * w = a*exp(i*phi) this is b(phi); with w = diag(a) columns it is the
 * basis-build batch.                                                      */
hz_status hz_build_rhs(hz_operator *op, const void *w, int64_t ld_w, void *b,
                       int32_t nrhs, int64_t ld, void *stream);

/* out[c, r] = sum_p B[p, c] lambda[p, r], i.e. B^T lambda = B^H lambda for
 * the real source columns; out is device complex128 (M x nrhs, ld_out).   */
hz_status hz_gather_sources(hz_operator *op, const void *lambda, int64_t ld,
                            void *out, int64_t ld_out, int32_t nrhs, void *stream);

/* out[q, r] = sum_c w[q, c] u[rows[q, c], r] for 8 trilinear corners;
 * rows (int64) and weights (double) are device arrays of npoints*8.        */
hz_status hz_sample(hz_operator *op, const int64_t *rows, const double *weights,
                    int32_t npoints, const void *u, int64_t ld, void *out,
                    int64_t ld_out, int32_t nrhs, void *stream);

/* Install a multigrid preconditioner (levels[0] is the operator's own grid;
 * level l+1 keeps the even nodes of level l plus its last node on each axis).
 * The mass term of every level is shifted by omega^2 -> omega^2 (1 + i*shift);
 * the smoother is damped Jacobi (weight) with `pre` sweeps, `post` sweeps on
 * the finest level and `post_coarse` on the levels below it, and
 * `coarse_sweeps` on the coarsest level; `cycle` coarse-grid corrections per
 * level below the finest (1: V-cycle, 2: W-cycle).  No counterpart in the reference,
 * which only offers Jacobi (src/algorithm/solvers/iterative_solver.py:46-56). */
hz_status hz_set_multigrid(hz_operator *op, int32_t nlevels, const hz_level_desc *levels, double shift,
                           int32_t pre, int32_t post, int32_t post_coarse, int32_t coarse_sweeps, double weight,
                           int32_t cycle);

/* Autotuner.  No counterpart in the reference (its CuPy GMRES has no
 * tunables).  Times every slab tile shape of the complex64 stencil on this
 * operator's grid for the fine apply and, with `preconditioner` =
 * multigrid, for each level's smoother, residual and first-sweep residual,
 * and keeps the fastest for each kernel.  Then picks the right-hand-side
 * chunk width of the mixed form: the candidate in {4, 8, 16, 32, 64} (at
 * most max_chunk and nrhs, workspace within memory_budget bytes, or half
 * of the free device memory when memory_budget <= 0) with the least time
 * per column of one Arnoldi step at j = restart / 2, preconditioner
 * included; of the widths within 2 % of that least time, the widest.  Writes the
 * width to *chunk_out and, when shapes_out is not null, the shape indices
 * (apply, then smoother / residual / first-sweep residual of each level) to
 * shapes_out[0 .. shapes_len).  The workspace used for timing is released. */
hz_status hz_autotune(hz_operator *op, int32_t nrhs, int32_t max_chunk, int32_t restart,
                      int32_t preconditioner, int64_t memory_budget, int32_t *chunk_out,
                      int32_t *shapes_out, int32_t shapes_len, void *stream);

/* Online response-basis path (phase_response_basis.py:177-253).  G is the
 * phase response basis on the device: nrows rows (nodes, Fortran order) of
 * m complex128 emitter entries, row stride ldg >= m, m <= 2048.  All
 * vectors are device pointers.  Partial sums live in a per-device workspace
 * kept by the library: calls that overlap on one device race on it.
 *   hz_basis_field:   u = G w                         (field, :177-206)
 *   hz_basis_adjoint: g = G^H c                       (phase_vjp, :208-253)
 *   hz_basis_field_match: one pass over G for the reference's field_match
 *     loss (phase_adjoint_optimizer.py:163-200): u = G w, loss
 *     0.5 sum (weight (|u| - target))^2 into *loss, cotangent
 *     c = weight^2 (|u| - target) u / (|u| + eps) and g = G^H c (m <= 512);
 *     u is written when not null.  The phase gradient is
 *     Re(conj(g) * i w) per emitter. */
hz_status hz_basis_field(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *w, void *u,
                         void *stream);
hz_status hz_basis_adjoint(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *c, void *g,
                           void *stream);
hz_status hz_basis_field_match(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *w,
                               const double *target, const double *weight, double eps, double *loss,
                               void *g, void *u, void *stream);

/* Fused loss evaluations on G, one pass each (phase_adjoint_optimizer.py:
 * 163-364; m <= 512): u = G w, the loss into *loss (device double), its
 * cotangent c per row and g = G^H c (m complex128); u is written when not
 * null.  The per-wave partials live in the per-device workspace (see above).
 *   HZ_LOSS_FIELD_MATCH  field_match (:163-177): target and weight per row.
 *   HZ_LOSS_ENERGY       the background term of focal_pressure (:244-258):
 *                        loss 0.5 scale sum |u|^2, c = scale u, with
 *                        scale = background_weight / N.
 *   HZ_LOSS_SOFTMAX      the sidelobe term of focal_contrast (:312-338): loss
 *                        scale (max + T (log sum exp((|u| - max)/T) - log count_a))
 *                        over rows with mask != 0, c = scale p u/(|u| + eps) with
 *                        p the softmax weights; scale = sidelobe weight,
 *                        T = temperature, count_a = number of masked rows.
 *   HZ_LOSS_RATIO        the focal contrast ratio mean_A |u|^2 / mean_B |u|^2
 *                        over rows of class mask == 1 (count_a rows) and
 *                        mask == 2 (count_b rows), c = 2u/(count_a mean_B) on A and
 *                        -2 mean_A u/(count_b mean_B^2) on B.  No reference
 *                        counterpart: the ratio form of a focal_contrast loss
 *                        as a 160^3 gradient workload states it.
 * The focal term -w mean_t |u_t| is not included: it needs only the target
 * rows (or exact samples) and is added by the caller
 * (src/algorithm/baselines/common.py:189-195). */
typedef enum hz_basis_loss_kind {
    HZ_LOSS_FIELD_MATCH = 0,
    HZ_LOSS_ENERGY = 1,
    HZ_LOSS_SOFTMAX = 2,
    HZ_LOSS_RATIO = 3
} hz_basis_loss_kind;

typedef struct hz_basis_loss_desc {
    int32_t kind;            /* hz_basis_loss_kind */
    const double *target;    /* FIELD_MATCH: target amplitude per row (device) */
    const double *weight;    /* FIELD_MATCH: weight per row (device) */
    const uint8_t *mask;     /* SOFTMAX: nonzero on background rows; RATIO: 1 / 2 classes (device) */
    double eps;              /* amplitude regulariser (1e-12 in the reference) */
    double scale;            /* ENERGY: background_weight / N; SOFTMAX: sidelobe weight */
    double temperature;      /* SOFTMAX */
    int64_t count_a;         /* SOFTMAX: masked rows; RATIO: class-1 rows */
    int64_t count_b;         /* RATIO: class-2 rows */
} hz_basis_loss_desc;

hz_status hz_basis_loss(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *w,
                        const hz_basis_loss_desc *desc, double *loss, void *g, void *u, void *stream);

/* hz_basis_loss without any allocation: the same kernels, launch shape and
 * results (bit for bit), with the per-wave partials in a static device array
 * of 8 MiB + 64 KiB that each GPU holds from the first call to process end.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving or optimiser loops,
 * harnesses that forbid hipMalloc).  Limits as hz_basis_loss (m <= 512;
 * arguments whose partials would exceed the array are rejected with
 * HZ_BAD_ARGUMENT).  Every call on one GPU shares that GPU's array: calls
 * that overlap on one device (two streams, two threads) race. */
hz_status hz_basis_loss_static(const void *G, int64_t ldg, int64_t nrows, int32_t m, const void *w,
                               const hz_basis_loss_desc *desc, double *loss, void *g, void *u, void *stream);

/* The device work of one phase-optimiser evaluation around the fused pass
 * over G (PhaseProblem.evaluate, src/algorithm/baselines/common.py:275-343),
 * one workgroup each, allocating nothing.  hz_phase_prologue: w = e^{i phi}
 * (m complex128) from m phases (FP64), and the exact focal samples
 * focal = R w of a row-major ntargets x m complex128 target response R.
 * hz_phase_epilogue: out[0] = loss[0] - target_weight mean_t |focal_t| (the
 * focal term; loss[0] is the grid loss of the fused pass) and, with
 * with_gradient, out[1 + j] = Re(conj(g_j + (R^H c)_j) i w_j) with
 * c_t = -target_weight/ntargets focal_t/(|focal_t| + eps) (common.py:189-195)
 * and g = G^H c of the grid loss.  m <= 512, ntargets <= 256 (0: no focal
 * term).  Device pointers; stream-ordered. */
hz_status hz_phase_prologue(const double *phases, int32_t m, const void *response, int32_t ntargets, void *w,
                            void *focal, void *stream);
hz_status hz_phase_epilogue(const void *w, int32_t m, const void *g, const void *response, int32_t ntargets,
                            const void *focal, double target_weight, double eps, const double *loss,
                            int32_t with_gradient, double *out, void *stream);

/* Material gradient, replacing ConservativeHelmholtz.material_vjp
 * (src/algorithm/physics/conservative_helmholtz.py:151-175): for
 * dL = -Re(lambda^H dA u) writes dL/drho and dL/dc per node (FP64, N each)
 * with lambda zeroed on identity rows.  u and lambda hold nrhs complex128
 * columns (leading dimension ld >= N); the gradients of the nrhs pairs are
 * summed (nrhs = 1 is the reference call). */
hz_status hz_material_vjp(hz_operator *op, const void *u, const void *lambda, int32_t nrhs, int64_t ld,
                          double *g_rho, double *g_c, void *stream);

/* The multigrid transfers between level `level` (0 = the operator's grid)
 * and level + 1 of the installed hierarchy, on nrhs complex64 vectors stored
 * one after another (device pointers, fine vectors 16-byte aligned):
 *   direction 0: out = P^T in   (restriction, fine -> coarse),
 *   direction 1: out += P in    (prolongation, coarse -> fine),
 * with P the trilinear prolongation of the even-node coarsening; identity
 * rows (pressure-release faces) take no part, as in the V-cycle.  No
 * counterpart in the reference: it exposes the preconditioner's transfers
 * to tests and to callers composing their own cycles. */
hz_status hz_mg_transfer(hz_operator *op, int32_t level, int32_t direction, const void *in, void *out, int32_t nrhs,
                         void *stream);

/* Copies level `level` (0 = the operator's grid) of the installed multigrid
 * hierarchy into caller device buffers: rho and kappa (n_l floats each,
 * Fortran order) and, for each face f that has a term, face[f] (complex64,
 * n_b x n_c with the lower-numbered in-plane axis fastest; face or face[f]
 * NULL skips it).  *face_mask receives bit f for each face that has a term.
 * No counterpart in the reference (it has no multigrid): for tests and
 * inspection of the hierarchy, such as comparing a device-derived level
 * with a host-built one. */
hz_status hz_mg_level(hz_operator *op, int32_t level, float *rho, float *kappa, void *const *face,
                      int32_t *face_mask, void *stream);

/* Free the cached solver workspace (Krylov basis, the multigrid form's
 * preconditioned basis, multigrid level vectors, captured graphs); the next
 * solve allocates what it needs.  Waits for the
 * device first.  No counterpart in the reference (CuPy's pool frees on its
 * own schedule). */
hz_status hz_release_workspace(hz_operator *op);

/* Pin the slab tile shapes (the layout hz_autotune reports: the apply, then
 * smoother / residual / first-sweep residual per multigrid level; len must
 * be 1 + 3 * levels).  Every shape computes the same arithmetic per node;
 * only the speed differs.  No counterpart in the reference. */
hz_status hz_set_tile_shapes(hz_operator *op, const int32_t *shapes, int32_t len);

/* Device bytes held by the operator (coefficients plus cached workspace). */
int64_t hz_device_bytes(const hz_operator *op);

/* Scene rasterisation: materials of a union of obstacle primitives in a
 * medium (warp_utils.py:11-23 blend, :52-77 sphere SDF and smoothing).  Per
 * node (i dx, j dx, k dx): sdf = min over the primitives, h = smoothed
 * Heaviside of sdf over `width` (m; 0 for a sharp step), and the blend of
 * 1/rho and 1/(rho c^2) linear in h between the obstacle material and the
 * medium (constants, or per-node device fields when rho_field / c_field
 * are given).  With no primitive the medium is copied.  Outputs rho, c and
 * sdf are device FP64 arrays of nx*ny*nz (Fortran order); each may be NULL.
 * The reference rasterises one sphere (or a precomputed mesh SDF); the
 * union, boxes and ellipsoids serve the benchmark scenes. */
typedef struct hz_primitive {
    int32_t kind;               /* 0 sphere, 1 box, 2 first-order ellipsoid  */
    double center[3];           /* m                                         */
    double size[3];             /* radius in size[0] / half extents / semi-axes (m) */
} hz_primitive;

hz_status hz_rasterize_materials(int32_t nx, int32_t ny, int32_t nz, double dx, int32_t nprim,
                                 const hz_primitive *prims, double width, double rho_obstacle, double c_obstacle,
                                 double rho_medium, double c_medium, const double *rho_field, const double *c_field,
                                 double *rho, double *c, double *sdf, void *stream);

/* In place: x (device FP64, nx*ny*nz, Fortran order) correlated along x,
 * then y, then z with the 2*radius+1 symmetric weights (host array, centre
 * weights[radius]), periodic boundaries: the Gaussian filter that turns white
 * noise into the random medium of the benchmark scenes.  No counterpart in
 * the reference, which takes media as given arrays. */
hz_status hz_gaussian_filter(int32_t nx, int32_t ny, int32_t nz, int32_t radius, const double *weights, double *x,
                             void *stream);

/* Bilinear deposition of sinc apertures on the z = 0 plane
 * (sinc_aperture.py:36-72, the masses of WarpAssemblyEngine.source_columns):
 * emitter e has centre (centers[2e], centers[2e+1]) and the q quadrature
 * points centre + (offsets[2p], offsets[2p+1]) with weights[p].  mass
 * (m x capacity, row-major) receives, for every node of the emitter's box
 * boxes[4e..4e+3] = (i0, j0, width, height) in the order t = di + width*dj,
 * the mass of node (i0 + di, j0 + dj): one sequential FP64 sum in the
 * reference's order (corners (0,0), (0,1), (1,0), (1,1), quadrature index
 * ascending), bit-identical to it; 0 where no point reaches the node.  All
 * arrays are host memory; the call is synchronous. */
hz_status hz_deposit_sources(int32_t m, const double *centers, int32_t q, const double *offsets,
                             const double *weights, double dx, const int32_t *boxes, int32_t capacity,
                             double *mass);

/* ---- Homogeneous surface condensation -------------------------------------
 * The exact strategy of HomogeneousSurfaceSolver
 * (src/algorithm/solvers/homogeneous_surface.py) for a homogeneous medium
 * over a reflecting -z baffle: modes of the x and y axis operators decouple
 * into tridiagonal z systems, solved by transfer ratios H[j] = u[j+1]/u[0]
 * and the surface Schur complement.  The eigendecompositions of the two
 * axis operators are computed by the caller (host LAPACK, as the reference
 * does, :20-38) and passed in.  Planes are nx x ny complex128, x fastest. */
typedef struct hz_surface hz_surface;

typedef struct hz_surface_desc {
    int32_t nx, ny, nz;
    double dx;
    const double *vx, *ix;      /* complex nx x nx, column-major (host)      */
    const double *vy, *iy;      /* complex ny x ny: eigenvectors, inverse    */
    const double *shift;        /* complex nx x ny, x fastest:
                                   lambda_x + lambda_y + (omega / c)^2       */
    const double *z_diagonal;   /* complex nz: diagonal of the z operator    */
    const double *z_upper;      /* complex nz-1: z[j, j+1]                   */
    const double *z_lower;      /* complex nz-1: z[j+1, j]                   */
} hz_surface_desc;

/* Transfer ratios and Schur complement of `modes` independent z systems
 * (device complex128): h is (nz-1) x modes, mode fastest; schur is modes.
 * Replaces the CUDA kernel z_transfer (homogeneous_surface.py:41-82) and
 * the Schur line (:164). */
hz_status hz_z_transfer(const void *shift, const void *diagonal, const void *upper, const void *lower,
                        int32_t nz, int64_t modes, void *h, void *schur, void *stream);

/* HomogeneousSurfaceSolver.__init__ after the eigendecompositions
 * (homogeneous_surface.py:140-164): uploads the tables and runs
 * hz_z_transfer.  nx, ny >= 2, nz >= 2. */
hz_status hz_surface_create(const hz_surface_desc *desc, hz_surface **out);
hz_status hz_surface_destroy(hz_surface *s);

/* Copy H ((nz-1) x nx*ny) and the Schur complement (nx*ny) into device buffers. */
hz_status hz_surface_tables(const hz_surface *s, void *h, void *schur, void *stream);

/* Source masses (nx*ny x M, real) as nnz (plane node, column, value)
 * triplets on the host: the masses argument of flux_target_response and
 * solve_flux (homogeneous_surface.py:175-187). */
hz_status hz_surface_set_sources(hz_surface *s, int32_t num_columns, int64_t nnz, const int64_t *node,
                                 const int32_t *column, const double *value);

/* y[p] = ix x[p] iy^T (inverse = 0) or vx x[p] vy^T (inverse = 1) for
 * `count` device planes: transform / inverse_transform (:166-173). */
hz_status hz_surface_transform(hz_surface *s, int32_t inverse, const void *x, void *y, int32_t count,
                               void *stream);

/* solve_flux (:207-229) for the emitter weights w = a*exp(i*phi) (device
 * complex128, M): surface load -4 pi/dx^3 masses w into rhs, surface modes
 * (also kept for hz_surface_planes / hz_surface_sample) into modes and the
 * surface plane into `surface`; each output may be NULL. */
hz_status hz_surface_solve_flux(hz_surface *s, const void *w, void *rhs, void *modes, void *surface, void *stream);

/* Planes z[0..count) of the last solved field into out (count planes;
 * z is a host array): plane (:231-236). */
hz_status hz_surface_planes(hz_surface *s, const int32_t *z, int32_t count, void *out, void *stream);

/* Trilinear samples of the last solved field at npoints points given by
 * their low grid corner (host int32, 3 per point) and fractions (host
 * double, 3 per point); out is device complex128 (npoints): sample (:250-259). */
hz_status hz_surface_sample(hz_surface *s, const int32_t *low, const double *frac, int32_t npoints, void *out,
                            void *stream);

/* out[p, m] (device complex128, npoints x M, row-major): the response of
 * point p to emitter m with ordinary transposes, so that p = out w
 * (flux_target_response, :189-205). */
hz_status hz_surface_target_response(hz_surface *s, const int32_t *low, const double *frac, int32_t npoints,
                                     void *out, void *stream);

const char *hz_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HZ_H */
