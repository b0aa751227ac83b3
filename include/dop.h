/*
 * dop.h -- C ABI of libdop.so, the MI355X (gfx950) operator-application library.
 *
 * It replaces the application step of DiffEqOperators.jl's lazy finite-difference
 * operators: the host (Julia via `ccall`, Python via ctypes) keeps building the
 * operators and boundary conditions exactly as before and hands their FP64 tables to
 * dop_plan_create(); dop_apply() then evaluates
 *
 *      du  = sum_t  L_t * Q * u          (GhostDerivativeOperator / Combination)
 *      du  = sum_t  L_t * M              (DerivativeOperator on a padded array)
 *
 * on device memory in one fused pass.  Paths in citations are relative to the
 * DiffEqOperators.jl v4.45.0 source tree.
 *
 * Conventions
 *   - Arrays are column-major (Julia order): size[0] is the fastest axis.  An operator's
 *     `axis` is Julia's type parameter N minus one.
 *   - u / du are device pointers to contiguous arrays of the plan's storage type.  The
 *     caller owns them and the stream.  du must not alias u.
 *   - Every function returns DOP_OK (0) or a negative DOP_E* code; dop_last_error()
 *     returns a thread-local message for the last failure.  No exception crosses the ABI.
 *   - dop_apply only reads the plan: concurrent dop_apply calls on one plan from several
 *     threads/streams are safe.  dop_plan_update and dop_plan_autotune change the plan and
 *     must not run while another thread uses it.
 */
#ifndef DOP_H
#define DOP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DOP_ABI_VERSION 2

/* error codes */
#define DOP_OK 0
#define DOP_EINVAL (-1)       /* bad shape / argument: the reference's AssertionError,
                                 derivative_operator_functions.jl:33-40 */
#define DOP_EDIM (-2)         /* length mismatch: DimensionMismatch, :182-184 */
#define DOP_EHIP (-3)         /* HIP runtime error */
#define DOP_ENOMEM (-4)       /* device allocation failed */
#define DOP_EUNSUPPORTED (-5) /* valid in the reference, not provided by this library */
#define DOP_EBUDGET (-6)      /* request exceeds the device-memory budget */
#define DOP_ESINGULAR (-7)    /* zero pivot in A \\ u: LinearAlgebra.SingularException */

/* storage types of u / du */
#define DOP_F64 0
#define DOP_F32 1
#define DOP_BF16 2

/* accumulation */
#define DOP_ACC_AUTO 0        /* f64 for f64 storage; f32, or DOP_ACC_F32_WIDE when an operator
                                 has d >= 3, for f32 / bf16 storage */
#define DOP_ACC_F64 1
#define DOP_ACC_F32 2
#define DOP_ACC_F32_KAHAN 3   /* f32 with compensated (TwoProduct/TwoSum) accumulation */
#define DOP_ACC_F32_WIDE 4    /* f32 weights and values; their products, exact in f64, are
                                 summed in f64 (compensation for high derivative orders at a
                                 third of the ALU work of DOP_ACC_F32_KAHAN) */

/* input kinds */
#define DOP_INPUT_GHOST 0     /* u unpadded; ghosts from the per-axis affine BCs (Q*u) */
#define DOP_INPUT_PADDED 1    /* u is the padded array M of the plain mul! */

/* which reference kernels define the row arithmetic */
#define DOP_SEM_LINES 0       /* AbstractVector kernels, convolutions.jl:27-362 (N-D Q*u,
                                 plain padded mul!) */
#define DOP_SEM_BPV 1         /* BoundaryPaddedVector kernels, convolutions.jl:367-730
                                 (1-D Q*u with an atomic BC) */

/* coeff_func kind, for the non-uniform upwind interior branch (convolutions.jl:234-253) */
#define DOP_COEFF_VECTOR 0    /* Vector / Function coeff_func: wind from each coefficient */
#define DOP_COEFF_NUM_POS 1   /* coeff_func::Number >= 0 */
#define DOP_COEFF_NUM_NEG 2   /* coeff_func::Number <  0 */

/* One DerivativeOperator's fields (derivative_operator.jl:14-29).  Table layouts
 * (row-major, FP64):
 *   centered, uniform dx:      stencil_coefs[SL], low/high_boundary_coefs[BPC][BSL]
 *   centered, vector dx:       stencil_coefs[len-2BPC][SL], low/high[BPC][BSL]
 *   upwind, uniform dx:        stencil_coefs[SL], low[BPC][BSL], high[BPC+offside][BSL]
 *   upwind, vector dx:         stencil_coefs[2][len-2BPC][SL], low[2][BPC][BSL],
 *                              high[2][BPC+offside][BSL]  (first index: 0 upwind, 1 downwind)
 */
typedef struct dop_operator {
    int32_t axis;                     /* N - 1 */
    int32_t wind;                     /* Wind type parameter: 0 centered, 1 upwind */
    int32_t derivative_order;
    int32_t approximation_order;
    int32_t len;
    int32_t stencil_length;
    int32_t boundary_stencil_length;
    int32_t boundary_point_count;
    int32_t offside;
    int32_t uniform_grid;             /* dx isa Number */
    int32_t coeff_kind;               /* DOP_COEFF_* */
    int32_t reserved;
    const double* stencil_coefs;
    const double* low_boundary_coefs;
    const double* high_boundary_coefs;
    const double* coefficients;       /* [len] */
} dop_operator;

/* An affine BC along one axis (RobinBC and its Dirichlet/Neumann special cases,
 * bc_operators.jl:21-61, 173-183): ghost_l = a_l . u[1:n_l] + b_l,
 * ghost_r = a_r . u[end-n_r+1:end] + b_r, identical for every line of that axis
 * (multi_dim_bc_operators.jl:141-149). */
#define DOP_MAX_BC_TAPS 8
#define DOP_BC_AFFINE 0        /* ghosts a . u + b as above */
#define DOP_BC_PERIODIC 1      /* PeriodicBC on a line: ghost_l = u[end], ghost_r = u[1]
                                  (bc_operators.jl:192); taps ignored */
#define DOP_BC_PERIODIC_ND 2   /* PeriodicBC extended over an N-D array (MultiDimDirectionalBC /
                                  ComposedMultiDimBC of PeriodicBC): ghost planes are the first
                                  and last planes, ghost_l = u[1], ghost_r = u[end]
                                  (multi_dim_bc_operators.jl:221-252); taps ignored */
typedef struct dop_affine_bc {
    int32_t present;                  /* 0: axis has no BC (must carry no operator) */
    int32_t n_l, n_r;
    int32_t kind;                     /* DOP_BC_* */
    double a_l[DOP_MAX_BC_TAPS];
    double b_l;
    double a_r[DOP_MAX_BC_TAPS];
    double b_r;
} dop_affine_bc;

typedef struct dop_desc {
    int32_t ndim;                     /* 1..3 */
    int32_t storage;                  /* DOP_F64 / DOP_F32 / DOP_BF16 */
    int32_t accum;                    /* DOP_ACC_* */
    int32_t input;                    /* DOP_INPUT_* */
    int32_t semantics;                /* DOP_SEM_* */
    int32_t n_ops;
    int64_t size[3];                  /* unpadded size of du (and of u for GHOST) */
    /* PADDED input: padded[k] = 1 if M is size[k]+2 along axis k.  transverse_view = 1
     * reproduces the views 1:n of the composite mul! (derivative_operator_functions.jl:
     * 429-460, 780-800) instead of trimming padded transverse axes to 2:end-1 (:48-57). */
    int32_t padded[3];
    int32_t transverse_view;
    const dop_operator* ops;          /* [n_ops], applied as sum_t */
    dop_affine_bc bc[3];              /* GHOST input: per axis */
    /* GHOST input, optional per axis: one atomic BC per line along that axis (MultiDimBC{dim}
     * of an Array of atomic BCs, multi_dim_bc_operators.jl:93-95; line l of axis a is the
     * column-major index over the other axes), or NULL for the single BC bc[a] (whose
     * present flag must still be set, with kind DOP_BC_AFFINE; its taps are then unused).
     * An entry is DOP_BC_AFFINE (its own taps) or DOP_BC_PERIODIC: slice_rmul pads line l
     * with BC[l] * u[line] (:29-39), which for a PeriodicBC is the wrap u[end], u[1]
     * (bc_operators.jl:192). */
    const dop_affine_bc* line_bc[3];
} dop_desc;

typedef struct dop_plan dop_plan;

int dop_version(void);
const char* dop_last_error(void);

/* Builds the fused per-axis row programs from the operators' own tables and uploads them.
 * Replaces the operator/BC side of `L*Q` (ghost_derivative_operator.jl:7-13),
 * `L1 + L2` (composite_operators.jl:31-46) and `compose` (multi_dim_bc_operators.jl:
 * 178-192): the device-side form of what those lazy objects hold. */
int dop_plan_create(const dop_desc* desc, dop_plan** out);
void dop_plan_destroy(dop_plan* plan);

/* New values for an existing plan, in place: the BCs' taps and constants (also per line) and
 * the operators' tables and coefficients of `desc`, which must describe the same structure as
 * the descriptor the plan was created from (ndim, sizes, storage, accumulation, input kind,
 * semantics, the operators' axes, kinds, orders, lengths and grid kind, the BCs' kinds and tap
 * counts; DOP_EINVAL otherwise, the plan unchanged).  Replaces what a method-of-lines RHS does
 * when it rebuilds its BC (or calls update_coefficients!) at every call before `A*u`
 * (heat_equation.jl:38-43, KdV.jl:26-30, DiffEqOperators.jl:67-75): the plan's device memory
 * is reused, nothing is allocated.  When only the BC values change, the new values travel as a
 * kernel argument on `stream` (no host synchronization; capturable); changed tables are copied
 * on `stream` and the call waits for those copies.  Applies enqueued on `stream` before the
 * update see the old values, later ones the new; the caller orders applies on other streams.
 * Batches, RK4 integrators and servers created from the plan read its values on the device,
 * so their later launches see the update. */
int dop_plan_update(dop_plan* plan, const dop_desc* desc, void* stream);

/* Number of bytes of u and du the plan reads/writes per dop_apply (storage elements). */
int64_t dop_plan_input_elems(const dop_plan* plan);
int64_t dop_plan_output_elems(const dop_plan* plan);

/* du = sum_t L_t * Q * u  (accumulate = 0), or du += ... (accumulate = 1).
 * Replaces: mul!(out, A::GhostDerivativeOperator, u) (ghost_derivative_operator.jl:15-24),
 *           *(A::GhostDerivativeOperator, u) (:26-37),
 *           *(L::DiffEqOperatorCombination, x) (composite_operators.jl:64-65),
 *           mul!(du, L::DerivativeOperator, M; overwrite) (derivative_operator_functions.jl:
 *             18-25, convolutions.jl:17-22), *(L, M) (:150-163),
 *           composite mul! (derivative_operator_functions.jl:203, 466).
 * accumulate = 0 never reads du (the `!overwrite * x_temp[i]` strong zero,
 * convolutions.jl:49).  stream: hipStream_t (NULL = default stream). */
int dop_apply(const dop_plan* plan, void* du, const void* u, int accumulate, void* stream);

/* ---- batching (no counterpart in the reference: it applies one operator per call) ----
 * A batch binds n independent plans (any mix of ndim, sizes, operators, BCs; one storage
 * type) to their du / u device pointers once; dop_batch_apply then evaluates all of them
 * from one device work list: its items are sorted into groups of one kind (2-D tiles, 1-D
 * tiles, generic points), window width and accumulator, and each group is one launch of a
 * kernel specialised for it (a handful of launches for any number of problems). */
typedef struct dop_batch dop_batch;
int dop_batch_create(int n, const dop_plan* const* plans, void* const* du,
                     const void* const* u, dop_batch** out);
int dop_batch_apply(const dop_batch* batch, int accumulate, void* stream);
/* Times the candidate row-chunk heights of the batch's 2-D tiles (on its own buffers: du is
 * overwritten) and keeps the fastest; *choice receives the chunk height (may be NULL). */
int dop_batch_autotune(dop_batch* batch, void* stream, int* choice);
void dop_batch_destroy(dop_batch* batch);

/* ---- explicit time stepping (no counterpart in the reference; its RHS is called from
 * OrdinaryDiffEq, heat_equation.jl:30-33) ----
 * A batch of method-of-lines problems du/dt = scale_i * (sum_t L_t Q u) advanced with
 * classical RK4, every stage of the whole batch one grouped launch set (as dop_batch_apply; the
 * groups of a stage run on forked streams inside the graph), the four stages of
 * a step captured once in a hipGraph and replayed.  Storage f32 or f64.  Creation tunes
 * the tile shape (DOP_RK4_AUTOTUNE=0 skips it) by timing stage-0 launches, which read u and
 * write only the internal stage buffers. */
typedef struct dop_rk4 dop_rk4;
int dop_rk4_create(int n, const dop_plan* const* plans, const double* scale,
                   const double* dt, void* const* u, void* stream, dop_rk4** out);
int dop_rk4_step(dop_rk4* rk, int nsteps, void* stream);   /* advances every problem */
int dop_rk4_chunk(const dop_rk4* rk);   /* row-chunk height picked by the tile autotuner */
void dop_rk4_destroy(dop_rk4* rk);

/* ---- serving (no counterpart in the reference) ----
 * A continuous-batching method-of-lines service.  Each request is a Q*u plan with its
 * scale, step dt, number of RK4 steps, an initial field and a result buffer in pinned host
 * memory, and a release time (seconds after dop_server_run starts).  dop_server_run
 * releases requests at their times, admits released requests in FIFO order while their state
 * (u and three RK4 stage buffers, carved from one device arena of `hbm_budget_bytes` rounded
 * down to 256 bytes) fits and fewer than `max_active` are resident, and advances the resident
 * requests by RK4 steps: they are split into size classes by grid points (up to 256^2, up to
 * 1024^2, larger), each class on a stream of its own advancing all its requests one step per
 * iteration (each stage one grouped launch set over the class), the classes side by side on
 * the GPU, so a small request is not held to the pace of the largest resident grid.  A
 * finished state is copied back asynchronously.  Storage f32 or f64 (one type per server).
 * dop_server_submit may be called from any thread, also while dop_server_run runs: such a
 * request joins when its release time (dop_server_now gives the run's current time) has
 * passed.  dop_server_run returns once every submitted request has completed and the server
 * is not open (dop_server_set_open(srv, 1) keeps it waiting for submissions until it is set
 * to 0).  A request whose state alone exceeds the budget is refused at submission
 * (DOP_EBUDGET); on an error the run stops and the resident requests' memory is returned.
 * `stream` is unused (the classes run on the library's three side streams of the device, the
 * ones the second half of a 3-D apply also uses, synchronized before dop_server_run returns:
 * with the caller's stream they fill HIP's default four hardware queues, so no class waits
 * behind another; servers run at once share them and run in turn); dop_server_times / dop_server_stats describe the last run and are
 * read after it. */
typedef struct dop_server dop_server;
int dop_server_create(int64_t hbm_budget_bytes, int max_active, int chunk, void* stream,
                      dop_server** out);
int dop_server_submit(dop_server* srv, const dop_plan* plan, double scale, double dt, int nsteps,
                      const void* u0_host, void* result_host, double release_s, int64_t* id);
int dop_server_set_open(dop_server* srv, int open);
int dop_server_run(dop_server* srv);
/* seconds since the running dop_server_run started (-1 when it is not running) */
double dop_server_now(const dop_server* srv);
/* times (s, relative to the run start) of request `id`: release, admission, completion */
int dop_server_times(const dop_server* srv, int64_t id, double* release, double* admitted,
                     double* completed);
/* peak resident requests and peak arena bytes during the last run */
int dop_server_stats(const dop_server* srv, int* peak_active, int64_t* peak_bytes,
                     int64_t* iterations);
void dop_server_destroy(dop_server* srv);

/* ---- concretization (concretization.jl) ----
 * The map a plan applies, du = A * vec(src) + b (accumulate = 0; src = u for GHOST input, the
 * padded M for PADDED input), as a CSR matrix with FP64 values: n_rows =
 * dop_plan_output_elems, n_cols = dop_plan_input_elems, rows and columns in column-major
 * linear order, columns ascending within a row, exact zeros not stored.
 * Replaces, for the operator a plan was built from:
 *   Array / sparse / BandedMatrix of a DerivativeOperator on a line (concretization.jl:8-139,
 *     upwind :556-810) and its N-D Kron form on Mshape (:425-505)  -- PADDED plans, b = 0;
 *   Array / sparse of a GhostDerivativeOperator, (L*Q_L, L*Q_b) (:866-913), for 1-D and for
 *     an operator with a MultiDimBC along its own axis  -- GHOST plans;
 *   Array of a DiffEqOperatorCombination, the sum of its operators' matrices (:918-924).
 * GHOST plans of sums with BCs on several axes (which the reference cannot concretize: its
 * per-operator matrices differ in shape) export the matrix of the same fused apply.
 * Step 1: row_ptr (device, n_rows + 1 int64) receives the row pointers, *nnz (host) the
 *   number of entries, *kl / *ku (host, may be NULL) the bandwidths max(row - col) and
 *   max(col - row).  Synchronizes the stream.
 * Step 2: col (device int64[nnz]), val (device double[nnz]) and b (device double[n_rows], may
 *   be NULL) are filled on the stream. */
int dop_plan_csr_rows(const dop_plan* plan, int64_t* row_ptr, int64_t* nnz, int64_t* kl,
                      int64_t* ku, void* stream);
int dop_plan_csr_fill(const dop_plan* plan, const int64_t* row_ptr, int64_t* col, double* val,
                      double* b, void* stream);

/* Step 1 of the export without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Same results as dop_plan_csr_rows, but the scan's partial sums and the
 * bandwidth counters live in a static device array of the library (512 KiB + 256 bytes per
 * GPU, held until the process ends) instead of buffers allocated per call: at most 2^28 rows
 * (DOP_EBUDGET beyond).  Every call on a GPU shares that GPU's array, so calls that overlap on
 * one GPU (several host threads, several streams) race; serialize them.  Synchronizes the
 * stream, like dop_plan_csr_rows. */
int dop_plan_csr_rows_noalloc(const dop_plan* plan, int64_t* row_ptr, int64_t* nnz, int64_t* kl,
                              int64_t* ku, void* stream);

/* ---- direct solve: A \\ u = A_l \\ (u - A_b) (ghost_derivative_operator.jl:39-58) ----
 * dop_solver_create concretizes a square GHOST plan on the device and factors its band with
 * partial pivoting (LU, the LAPACK gbtf2 algorithm) once; DOP_ESINGULAR on a zero pivot (the
 * reference's sparse LU raises SingularException), DOP_EUNSUPPORTED when the band with its
 * fill-in exceeds 2^30 entries.  dop_solver_solve solves nrhs right-hand sides (columns of n
 * contiguous FP64 values each, device memory) into x, one block per right-hand side.
 * Arithmetic is FP64 whatever the plan's storage type.  The reference factors with a general
 * sparse LU (UMFPACK); both are backward stable, so solutions agree to about cond(A) * eps. */
typedef struct dop_solver dop_solver;
int dop_solver_create(const dop_plan* plan, void* stream, dop_solver** out);
int dop_solver_solve(const dop_solver* solver, void* x, const void* rhs, int nrhs, void* stream);
int dop_solver_band(const dop_solver* solver, int64_t* n, int64_t* kl, int64_t* ku);
void dop_solver_destroy(dop_solver* solver);

/* ---- nonlinear_diffusion! (derivative_operator.jl:34-66) ----
 * du = sum_{l=0..s} C(s, l) (D^{l+f} q) .* (D^{s-l} p) on padded vectors p, q (length n+2;
 * s = nterms - 1): dq[l] is the plan of CenteredDifference(l+f, approx, dx, n) on a padded
 * vector, or NULL for f + l = 0 (then the interior q[2:n+1]); dp[l] likewise for s - l.
 * All plans are 1-D PADDED plans of one length and storage (f64 or f32; the products and sums
 * are formed in f64 in the reference's order).  One launch. */
int dop_nonlinear_diffusion(const dop_plan* const* dq, const dop_plan* const* dp, int nterms,
                            void* du, const void* p, const void* q, void* stream);

/* ---- adaptive time stepping (no counterpart in the reference: its method-of-lines tests hand
 * `(u, p, t) -> A*bc*u` to OrdinaryDiffEq's Tsit5, heat_eqn.jl:4-60) ----
 * Integrates du/dt = scale * (sum_t L_t Q u) from t0 to t1 in place in u (the plan's storage,
 * f32 or f64; stage arithmetic and state in f64) with the 5(4) Runge-Kutta pair of Tsitouras
 * (2011), the error norm sqrt(mean((e / (abstol + reltol max(|u|, |unew|)))^2)) and the step
 * controller of Hairer, Norsett & Wanner's DOPRI5 (PI, beta 0.04, safety 0.9, ratio [0.2, 10]);
 * steps are clamped to end exactly at t1.  dt0 <= 0 picks the starting step size by Hairer's
 * rule.  Every attempted step (six stages, six fused applies, the error reduction and the
 * controller) is one hipGraph replay with no host round trip; dop_tsit5_run replays up to
 * max_attempts of them, reading the device state between batches of replays (1, 2, 4, ...
 * up to check_every, and no more than the current step size needs to reach the end time,
 * plus one), then commits the last accepted
 * step into u.  dop_tsit5_status: done = 1 at t1, -1 when the step size
 * collapsed.  dop_tsit5_set_end moves the end time (continuing from the current state). */
typedef struct dop_tsit5 dop_tsit5;
int dop_tsit5_create(const dop_plan* plan, double scale, void* u, double t0, double t1,
                     double dt0, double abstol, double reltol, void* stream, dop_tsit5** out);
int dop_tsit5_run(dop_tsit5* integ, int64_t max_attempts, int check_every, void* stream);
int dop_tsit5_set_end(dop_tsit5* integ, double t1, void* stream);
int dop_tsit5_status(const dop_tsit5* integ, double* t, double* dt, int64_t* naccept,
                     int64_t* nreject, int* done, void* stream);
void dop_tsit5_destroy(dop_tsit5* integ);

/* ---- tile autotuner (no counterpart in the reference) ----
 * Times the candidate tile shapes of the plan's fused kernel on scratch buffers and keeps
 * the fastest in the plan and in a process-wide cache keyed by device, storage,
 * accumulation, window width, grid kind and extents.  Returns the chosen candidate index via
 * *choice (may be NULL).  Plans created for a key already in the cache start with its
 * variant; a plan with a tiled path of at least 2^20 points that is still untuned tunes on its
 * first dop_apply with accumulate = 0 (timing the candidates on that call's own du and u, then
 * applying), unless the stream is being captured or DOP_AUTOTUNE=0 is set.  Every candidate
 * computes the same values bit for bit: the choice changes only the time. */
int dop_plan_autotune(dop_plan* plan, void* stream, int* choice);
/* The candidate the plan launches (the autotuner's choice), and whether it was tuned (1) or
 * is the default (0). */
int dop_plan_variant(const dop_plan* plan, int* tuned);

#ifdef __cplusplus
}
#endif
#endif
