/* C-ABI of the MI355X device solve for the octree FV curl-curl system
 *
 *     (Ce^T Mf^-1 Ce + i*omega*Me(sigma)) e = b
 *
 * solved by SSOR-preconditioned BiCGStab with divergence correction.  This is
 * the drop-in boundary of the forward / adjoint solve path: a host driver
 * built like ForwardSolverFV (src/forward/forward_solver_fv.cpp) forwards the
 * bodies of its inner_solve (:266-285), solve_system_ (:727-759),
 * solve_with_divcorr_ (:778-888) and divcorr_poisson_ (:623-722) here.  No C++
 * types or exceptions cross this boundary: plain pointers and sizes, int
 * status codes, and m3d_last_error() for the message.
 *
 * Complex vectors are interleaved (re, im) doubles.  Batched calls take
 * system-major arrays: system s, edge i at index 2*(s*n + i).
 *
 * Threading: one context per host thread; calls are blocking from the
 * caller's view and ordered on the context's stream.  Host buffers stay
 * caller-owned; the library owns device memory until destroy.
 */
#ifndef M3D_DEVICE_H
#define M3D_DEVICE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define M3D_OK 0
#define M3D_ERR_ARG 1     /* invalid argument */
#define M3D_ERR_HIP 2     /* HIP runtime error */
#define M3D_ERR_STATE 3   /* call order (e.g. no mass set) */
#define M3D_ERR_ALLOC 4   /* device allocation failed */

/* Solve modes. */
#define M3D_MODE_INNER 0        /* BiCGStab + dead-edge interpolation: inner_solve (:266-285) */
#define M3D_MODE_FORWARD 1      /* scattered-field solve_pol (:316-367): rhs_sec = b - A*E_pri,
                                   BiCGStab from 0, E = E_pri + E_sec, Dirichlet values
                                   re-imposed, post-hoc DivCorr */
#define M3D_MODE_SYSTEM 2       /* solve_system_ (:727-759): BiCGStab from x, zeroed if diverged */
#define M3D_MODE_INTERLEAVED 3  /* adjoint_solve / solve_rhs (:914-940) with
                                   solve_with_divcorr_ (:778-888); phi0 = G^T b / (i omega) */

/* Inner-solve precision. */
#define M3D_PREC_FP64 0  /* complex128 BiCGStab (the reference arithmetic) */
#define M3D_PREC_MIXED 1 /* complex64 inner BiCGStab inside FP64 iterative refinement */

typedef struct m3d_context m3d_context;
typedef struct m3d_operator m3d_operator;

/* Context per device.  Replaces nothing in the reference (it has no device);
 * created once per host thread by the driver. */
int m3d_context_create(int device, m3d_context** out);
int m3d_context_destroy(m3d_context* ctx);
/* Message of the last failing call on this host thread. */
const char* m3d_last_error(void);

/* Frequency-independent operator of one mesh after identity elimination of
 * the constrained (Dirichlet + dead) edges, in split form:
 *   A_ij (i != j) = offdiag (real, free rows x free columns only),
 *   A_ii          = diag_s[i] + i*omega*me[i]  (free rows),
 *   A_ii          = 1                          (constrained rows).
 * Replaces the per-frequency SparseMatC A_ of ForwardSolverFV
 * (forward_solver_fv.h:202; assembled by DiscreteOperators::assemble_system,
 * operators.cpp:556-837, and modified in place at forward_solver_fv.cpp:224-235). */
typedef struct {
    int n;                            /* edge DOFs */
    const int* rowptr;                /* n+1 */
    const int* colidx;                /* nnz, no diagonal, no constrained column */
    const double* values;             /* nnz */
    const double* diag_s;             /* n: stiffness diagonal */
    const unsigned char* constrained; /* n: identity rows */
    const unsigned char* bc;          /* n: Dirichlet rows (value re-imposed after FORWARD) */
    int n_hanging;                    /* dead coarse edges */
    const int* hanging;               /* 3*n_hanging: (coarse, child1, child2) */
} m3d_operator_desc;

int m3d_operator_create(m3d_context* ctx, const m3d_operator_desc* desc, m3d_operator** out);
int m3d_operator_destroy(m3d_operator* op);

/* Moves an operator to another context of the same device: waits for the
 * operator's current stream, then runs every later call on ctx's stream
 * (its workspaces, launch shapes and captured graphs stay; the old context
 * is freed if it was destroyed and nothing else depends on it).  No
 * counterpart in the reference, which has no device contexts: it is for a
 * scheduler that builds many models on one context and hands each to
 * whichever stream is free (the many-model batch that replaces the MPI
 * frequency parallelism of freq_parallel.cpp).  Calls on the operator must
 * not overlap the move. */
int m3d_operator_set_context(m3d_operator* op, m3d_context* ctx);

/* Edge mass Me(sigma) (operators.cpp:481-518), n values; call again after
 * every sigma update (ForwardSolverFV::update_sigma, :53-58). */
int m3d_operator_set_mass(m3d_operator* op, const double* me);

/* DivCorr data for the current sigma (compute_bc_flags_, :546-610):
 * node gradient G_node (2 entries per edge, operators.cpp:849-944), the edge
 * weights Me_sigma, the DivCorr Dirichlet edge / node sets, and the node
 * Laplacian L = G^T Me_sigma G after its Dirichlet rewrite with the Jacobi
 * diagonal taken before it. */
typedef struct {
    int nn;                           /* nodes */
    const int* g_col;                 /* 2n: node of each G entry */
    const double* g_val;              /* 2n */
    const double* me_sigma;           /* n */
    const unsigned char* bc_edge;     /* n: edges left uncorrected */
    const unsigned char* dead_edge;   /* n: excluded from div(sigma E) and phi0 */
    const unsigned char* bc_node;     /* nn: phi = 0 */
    const int* L_rowptr;              /* nn+1 */
    const int* L_colidx;
    const double* L_values;
    const double* L_diag;             /* nn */
} m3d_divcorr_desc;

int m3d_operator_set_divcorr(m3d_operator* op, const m3d_divcorr_desc* desc);

/* Cell-side data of compute_sensitivity (forward_solver_fv.cpp:989-1059):
 * per cell the list of its edges (caller numbering, CSR) and per edge the
 * weight A_dual * L / n_adj, 0 for edges the sensitivity skips (constrained
 * with the 1 m tolerance of :1017-1029, or without adjacent cells).  Set
 * again whenever sigma changes. */
typedef struct {
    int nc;                   /* cells */
    const int* cell_ptr;      /* [nc + 1] */
    const int* cell_edges;    /* [cell_ptr[nc]] */
    const double* edge_weight; /* [n] */
    const double* sigma;      /* [nc] */
} m3d_sensitivity_desc;
int m3d_operator_set_sensitivity(m3d_operator* op, const m3d_sensitivity_desc* desc);

/* d_grad[c] = sum over the nsys systems of
 *   Re(i omega_s sigma_c sum_{e in c} lambda_s[e] E_s[e] w_e)
 * (compute_sensitivity, :989-1059; lambda is not conjugated, :1008), the
 * gradient of the misfit in ln sigma_c when (E_s, lambda_s) are the forward
 * and adjoint fields of every frequency and polarisation.  d_E and d_lambda
 * are device complex128 [nsys][n] (caller order), d_grad device double [nc]
 * (overwritten).  Returns after the context stream drains. */
int m3d_sensitivity_device(m3d_operator* op, int nsys, const double* omega, const double* d_E,
                           const double* d_lambda, double* d_grad);

/* m3d_sensitivity_device that allocates no device memory.  No counterpart
 * in the reference: it is for callers that cannot allocate (tight serving or
 * inversion loops, harnesses that forbid hipMalloc).  The frequencies of a
 * call are staged in a static device array of the library, 4096 values on
 * each GPU, held until the process ends, instead of a buffer the operator
 * allocates and grows; nsys above 4096 is refused (M3D_ERR_ARG).  Every call
 * on one GPU stages through that GPU's one array, so calls that overlap in
 * time on one GPU (several host threads or contexts) race.  Same kernel and
 * results as m3d_sensitivity_device.  Returns after the context stream
 * drains. */
int m3d_sensitivity_device_noalloc(m3d_operator* op, int nsys, const double* omega, const double* d_E,
                                   const double* d_lambda, double* d_grad);

/* Perturbation right-hand side of the Gauss-Newton product J*v
 * (build_perturbation_rhs, forward_solver_fv.cpp:945-984) on the device:
 * for each system s
 *   rhs_s[e] = -sum_{cells c containing e} i omega_s dMe_ce dsigma_c E_bg,s[e],
 *   dMe_ce   = L_e^3 savg_e^2 / (n_adj(e) sigma_c^2)
 * with savg_e the harmonic mean of sigma over the edge's adjacent cells (the
 * reference's mass derivative; it equals the derivative of the edge mass
 * only in a uniform medium on a uniform mesh), the current sigma of the
 * sensitivity, and the reference's operation order, so the values equal the
 * host's bit for bit.  d_ebg, d_rhs: device complex128 [nsys][n] (caller
 * order); d_dsigma: device double [cells], sigma space.  Needs
 * m3d_operator_set_mass_structure and m3d_operator_set_sensitivity.  Returns
 * after the context stream drains. */
int m3d_perturbation_rhs_device(m3d_operator* op, int nsys, const double* omega, const double* d_ebg,
                                const double* d_dsigma, double* d_rhs);

/* Mirrors the solver fields of ForwardParamsFV (forward_solver_fv.h:31-69). */
typedef struct {
    double tol;              /* bicgstab_tol */
    int maxiter;             /* bicgstab_maxiter */
    int div_correction;      /* div_correction */
    int div_corr_iters;      /* div_corr_iters */
    double div_corr_tol;     /* div_corr_tol */
    int divcorr_outer_max;   /* divcorr_outer_max */
    int divcorr_iter_per_dc; /* divcorr_iter_per_dc */
    int precision;           /* M3D_PREC_FP64 or M3D_PREC_MIXED */
    double inner_tol;        /* MIXED: relative tolerance of each complex64 inner solve */
    int use_initial_guess;   /* FORWARD: x holds an E_sec guess (e.g. the learned e0).
                                1: used as given where ||b - A e0|| < ||b||;
                                2: scaled per system by gamma = (A e0, b) / (A e0, A e0), the
                                least-squares factor, and used where ||b - gamma A e0|| <
                                0.5 ||b||; 3: scaled as 2 and used where
                                ||b - gamma A e0|| < ||b||.  Elsewhere the zero start. */
    int interleaved_forward; /* FORWARD: 0 = the reference's full BiCGStab on E_sec
                                (forward_solver_fv.cpp:347-354); 1 = the interleaved
                                BiCGStab + DivCorr bursts of solve_with_divcorr_
                                (:778-888) on E_sec, with a full BiCGStab from the
                                current E_sec for any system that stagnates or
                                diverges.  The post-processing is the same. */
    double divcorr_pcg_tol;  /* PCG stop ||r|| < tol * ||div J|| of the DivCorr steps
                                between interleaved bursts (INTERLEAVED, and FORWARD
                                with interleaved_forward); 0 = the reference's 1e-10
                                (forward_solver_fv.cpp:700-702).  The FORWARD
                                post-hoc correction and m3d_divcorr always use 1e-10.
                                This PCG runs in FP64.  With the environment variable
                                M3D_DIVCORR_FP32=1, precision MIXED and a tolerance
                                >= 1e-5 it runs in FP32 instead (complex64 node
                                vectors, FP32 Laplacian, double reductions; E and
                                the stops stay FP64). */
} m3d_solve_params;

/* Per-system outcome; mirrors BiCGStabResult (bicgstab.h:23-28) plus what the
 * outer loops of solve_with_divcorr_ report. */
typedef struct {
    int iterations;       /* BiCGStab iterations (summed over bursts / refinement steps) */
    double residual;      /* recursive relative residual at exit */
    double true_residual; /* ||A x - b|| / ||b|| in FP64 at exit */
    int converged;
    int diverged;         /* solution zeroed because the residual exceeded 1 */
    int info;             /* 0 converged, 1 early (||s||), 2 breakdown, 3 max iterations */
    int outer;            /* outer iterations (interleaved bursts / refinement steps) */
    int guess_used;       /* FORWARD with use_initial_guess: 1 if the guess was kept */
    double guess_residual; /* FORWARD with use_initial_guess: ||b - A x0|| / ||b|| of the
                              (scaled) guess, whether kept or not; else 0 */
    int stagnated;        /* INTERLEAVED: stopped by the stagnation rule (a burst that
                             lowered the true residual by less than 5 %, accepted as is:
                             forward_solver_fv.cpp:866-873) */
} m3d_solve_result;

/* Solve nsys systems that share this operator, one omega each.
 *   rhs  [nsys][n] complex: b (FORWARD: rhs_total of build_rhs, :244-255)
 *   aux  [nsys][n] complex: E_pri for FORWARD, else NULL
 *   x    [nsys][n] complex: initial guess in (INNER / SYSTEM; FORWARD only
 *                           with use_initial_guess, as E_sec; INTERLEAVED
 *                           starts from 0), solution out
 * Host-memory variant. */
int m3d_solve_batch(m3d_operator* op, int nsys, int mode, const double* omega,
                    const double* rhs, const double* aux, double* x,
                    const m3d_solve_params* params, m3d_solve_result* results);

/* Same with device pointers (HBM-resident inputs and outputs), ordered on the
 * context's stream. */
int m3d_solve_batch_device(m3d_operator* op, int nsys, int mode, const double* omega,
                           const double* d_rhs, const double* d_aux, double* d_x,
                           const m3d_solve_params* params, m3d_solve_result* results);

/* divcorr_poisson_ (:623-722) on nsys fields in place; phi0 [nsys][nn]
 * complex or NULL. */
int m3d_divcorr(m3d_operator* op, int nsys, const double* omega, double* E, const double* phi0,
                const m3d_solve_params* params);

/* Building blocks, host memory, for parity tests of the kernels. */
/* y = A x (CSRMatrix::matvec, common.h:52-61). */
int m3d_spmv(m3d_operator* op, int nsys, const double* omega, const double* x, double* y);
/* z = M^-1 r with the multicolour symmetric Gauss-Seidel (ssor.h:53-79 in
 * the colour order returned by m3d_operator_permutation). */
int m3d_ssor_apply(m3d_operator* op, int nsys, const double* omega, const double* r, double* z);
/* Device row order: perm[k] = original edge at device row k (-1 = padding
 * row); colour_start[c] for c = 0..ncolours (device rows).  Either output
 * may be NULL; sizes via m3d_operator_info. */
int m3d_operator_permutation(m3d_operator* op, int* perm, int* colour_start);
/* n_padded, ncolours, stored off-diagonal entries (padding included) and
 * off-diagonal nonzeros. */
int m3d_operator_info(m3d_operator* op, int* n_padded, int* ncolours, long long* stored_entries,
                      long long* nnz);
/* Storage of the off-diagonal blocks (no reference counterpart: the
 * reference keeps CSR values and columns).  value_table > 0: besides the
 * separate value and column arrays the operator holds packed blocks, one
 * 32-bit word per entry with the column in the low column_bits bits and the
 * index of the entry's value in a table of value_table doubles, used at the
 * widths whose launch shape has packed = 1; value_table = 0: no packed blocks.
 * They are built at m3d_operator_create when the table fits (at most 4096
 * values, index and column within 32 bits); the environment variable
 * M3D_PACKED_SELL=0 switches them off. */
int m3d_operator_format(m3d_operator* op, int* value_table, int* column_bits);

/* Device memory of the operator (no reference counterpart; for callers that
 * keep several meshes resident): operator_bytes = the operator, DivCorr,
 * sigma-update and sensitivity arrays; workspace_bytes = the solve
 * workspaces of every batch width used so far, which stay allocated between
 * solves (about 20 GB per million edges once every width up to 32 systems
 * has run in FP64 with DivCorr).  Either output may be NULL. */
int m3d_operator_device_bytes(m3d_operator* op, long long* operator_bytes, long long* workspace_bytes);
/* Frees the solve workspaces and their captured graphs after the context
 * stream drains; the operator, its shapes and data stay, and the next solve
 * allocates what it needs again (with the same results). */
int m3d_operator_trim(m3d_operator* op);

/* Measurement (no reference counterpart): average duration in ms of one
 * launch of a kernel class over `reps` launches, timed with HIP events on the
 * context stream, and its algorithmic HBM bytes per launch: alg_bytes counts
 * every off-diagonal coupling as its value and its int32 column (w_v + 4
 * bytes), the diagonal pair and each vector once; stored_bytes (may be NULL)
 * the same with the bytes per coupling of the format the launch reads (one
 * 4-byte word in the packed blocks and the row tiles).
 * kind 0: batched SpMV y = A x; kind 1: one multicolour SSOR apply;
 * kind 2: one FP64 BiCGStab iteration (b = 1, tolerance 0, so every system
 * keeps running); kind 3: the same iteration in complex64; kind 4: one
 * backward-sweep launch (the mean over the backward launches of an SSOR
 * apply: every colour but the last, whose backward step is written by its
 * forward launch when it has no upper couplings). */
int m3d_bench_kernel(m3d_operator* op, int kind, int nsys, const double* omega, int reps,
                     double* avg_ms, double* alg_bytes, double* stored_bytes);

/* Edge-side data of the edge mass, fixed per mesh, for sigma updates on the
 * device.  build_edge_mass (operators.cpp:481-518) gives
 *   Me[e]       = mean(sigma over adj(e)) * (A_dual[e] * L[e]),
 *                 0 for massless edges (hanging coarse edges) and edges
 *                 without adjacent cells,
 * and the DivCorr weight (:956-973, :1034-1048) is
 *   Me_sigma[e] = (mean(sigma over adj(e)) * A_dual[e]) * L[e], 0 on dead
 *                 edges,
 * where the mean runs over the listed cells c with 0 <= c < nc, and is 1e-8
 * when there is none. */
typedef struct {
    int nc;                        /* cells */
    const int* adj_ptr;            /* [n + 1] cells adjacent to each edge (CSR) */
    const int* adj_cell;           /* [adj_ptr[n]] in the order the mean adds them */
    const double* dual_area;       /* [n] A_dual */
    const double* length;          /* [n] L */
    const unsigned char* massless; /* [n] edges whose Me is 0 (hanging coarse edges) */
    const unsigned char* dead;     /* [n] edges whose DivCorr weight is 0 */
} m3d_mass_structure;
int m3d_operator_set_mass_structure(m3d_operator* op, const m3d_mass_structure* desc);

/* Sigma update on the device (a line-search trial of the inversion,
 * inversion_fv.cpp:427-436, where only Me changes): from sigma [nc] it
 * recomputes Me (the operator diagonal, FP64 and the FP32 copy), the DivCorr
 * edge weights, the values of the DivCorr Laplacian L = G^T diag(Me_sigma) G
 * (operators.cpp:952-1019) with its Dirichlet rewrite and Jacobi diagonal
 * (forward_solver_fv.cpp:581-606), and the sigma of the sensitivity — the
 * same values, bit for bit, as recomputing them on the host and calling
 * m3d_operator_set_mass, m3d_operator_set_divcorr and
 * m3d_operator_set_sensitivity.  Needs m3d_operator_set_mass_structure and a
 * previous m3d_operator_set_divcorr (its sparsity pattern is kept).  Returns
 * M3D_ERR_STATE, with the device arrays undefined until the next host update,
 * if the new sigma would move an edge weight across the 1e-30 threshold that
 * decides the pattern of L.  sigma is host memory for
 * m3d_operator_update_sigma and device memory for
 * m3d_operator_update_sigma_device (ordered after earlier work on the
 * context stream; returns after the update is done). */
int m3d_operator_update_sigma(m3d_operator* op, const double* sigma);
int m3d_operator_update_sigma_device(m3d_operator* op, const double* d_sigma);

/* Launch shape of the row kernels at one batch width and precision (no
 * reference counterpart: the reference has one CPU thread per system).
 * Only the assignment of rows to GPU threads changes; the arithmetic of each
 * row does not, except that sweep_sub > 1 adds a row's couplings in sweep_sub
 * interleaved partial sums.  Reductions over rows are grouped by workgroup,
 * so a different shape can change the last bits of dot products, and with
 * them iteration counts where convergence is sensitive to rounding. */
typedef struct m3d_launch_shape {
    int sweep_sub; /* lanes per row in the SSOR sweeps: 1, 2, 4 or 8, with width * sub <= 32 */
    int sweep_cap; /* grid cap of one sweep launch, workgroups: a multiple of 8 in 8..65536 */
    int row_cap;   /* grid cap of the SpMV kernels that reduce: a multiple of 8 in 8..2048 */
    int sweep_chunk; /* 1: a sweep's workgroups grid-stride over its row blocks (sweep_cap
                        applies); 2, 4, 8 or 16: each workgroup takes that many consecutive
                        row blocks (sweep_cap is then unused) */
    int packed;      /* 1: gather couplings from the packed blocks (see
                        m3d_operator_format) when the operator has them; 0: from
                        the separate value and column blocks.  Same values, same
                        order of summation: results are identical either way */
    int wave;        /* widths 8-32 only (0 below): 1 = the width lanes of a row
                        load the couplings of up to width consecutive entries at
                        once, one entry per lane, and pass them to each other
                        through the LDS crossbar (a packed block's values from an
                        LDS copy of the value table), in the SpMV, the BiCGStab
                        products and the SSOR sweeps at sweep_sub 1; 0 = every
                        lane loads each coupling itself.  Same order of summation
                        either way */
    int tiled;       /* widths 8-32 only (0 below): 1 = the plain product (m3d_spmv,
                        the forward right-hand side b - A E_pri) runs over row
                        tiles of consecutive caller edges whose x rows are first
                        copied into LDS, every coupling then reading x from LDS;
                        built at m3d_operator_create with the packed blocks
                        (M3D_TILED_SPMV=0 switches them off).  0 = the product
                        gathers x from device memory.  Same order of summation
                        either way */
} m3d_launch_shape;

/* width is a batch width 1, 2, 4, 8, 16 or 32; precision M3D_PREC_FP64
 * (complex128 kernels) or M3D_PREC_MIXED (the complex64 inner kernels).  The
 * default shape of a new operator: sweep_sub 1, sweep_cap 4096, row_cap 2048,
 * sweep_chunk 1, packed 1, wave 1 and tiled 1 at widths 8-32 (0 below). */
int m3d_operator_get_shape(m3d_operator* op, int precision, int width, m3d_launch_shape* out);
int m3d_operator_set_shape(m3d_operator* op, int precision, int width, const m3d_launch_shape* shape);

/* Autotuner (the slice and block shape choice of the batched solver): for
 * every width whose bit is set in widths_mask (bit i = width 1 << i) and both
 * precisions, times one BiCGStab iteration (as m3d_bench_kernel kinds 2 and 3,
 * reps iterations per candidate, frequencies omega[0..width-1]) under candidate
 * shapes by coordinate descent (sweep_sub, sweep_cap, row_cap, sweep_chunk, packed) and
 * keeps a candidate only when it is at least 5 % faster than the incumbent.  best_ms, if not NULL, receives 2 * 6 values:
 * [precision * 6 + log2(width)] = ms per iteration of the kept shape, 0 where
 * not tuned.  Blocking; run it outside timed regions.  Choices stay with the
 * operator; read them with m3d_operator_get_shape to pin them elsewhere. */
int m3d_operator_autotune(m3d_operator* op, int widths_mask, const double* omega, int reps,
                          double* best_ms);

/* Stream the library launches on (hipStream_t), for callers that time or
 * order work around it. */
void* m3d_context_stream(m3d_context* ctx);

/* ---- Station side of the gradient, on the device ----------------------
 * The steps either side of the adjoint solve of InversionFV::compute_gradient_
 * (inversion_fv.cpp:250-280) on device-resident fields, so that the forward
 * and adjoint batches follow each other with no station data crossing the
 * bus: impedance extraction (extract_impedance_, forward_solver_fv.cpp:
 * 1490-1727), the weighted residual and misfit (inversion_fv.cpp:258-271,
 * :75-98) and the adjoint right-hand side (build_adjoint_rhs_from_residual,
 * forward_solver_fv.cpp:1804-2013).  A station set belongs to one mesh and
 * one list of stations; fields are complex128, system-major in caller edge
 * order, system 2f the first polarisation of frequency f and 2f + 1 the
 * second (ForwardSolverFV's E1_ / E2_). */
typedef struct m3d_stations m3d_stations;

/* Station table (the host producer's find_stations_ result, :1393-1485, and
 * the edge grouping of :1534-1562).  Per station: valid = 1 when it has a
 * station cell; the cell's edges in the cell-edge order of the staggered
 * grid, each with its role (bits 0-1 direction 0 x, 1 y, 2 z; bit 2 set for
 * an x / y edge above the cell centre or a z edge on the +x side; bit 3 set
 * for a z edge on the +y side); the cell sizes dx, dy, dz.  Per station edge
 * (sorted, unique): the contributions of the adjoint right-hand side in the
 * order the reference adds them (stations in order; in a station the E part
 * before the H part; for z edges the y part before the x part), 4 ints each:
 * station, kind (0: s_x we, 1: s_y we, 2: aH_y wh, 3: aH_x wh), axis * 2 +
 * (sign < 0), n, where wh = sign * i_axis / n, i_axis = 1 / (i omega mu0
 * h_axis) and axis 0 = z, 1 = x, 2 = y (:1928-2008); the E kinds carry their
 * real weight we in contrib_we.  max_nf bounds the frequencies of one call
 * (the device buffers are sized for it at creation). */
typedef struct {
    int n;                        /* edges of the mesh (length of one field) */
    int ns;                       /* stations */
    int max_nf;                   /* frequencies per call, at most */
    const int* valid;             /* [ns] */
    const int* edge_ptr;          /* [ns + 1] */
    const int* edge;              /* [edge_ptr[ns]] caller edge ids */
    const unsigned char* role;    /* [edge_ptr[ns]] */
    const double* cell_size;      /* [3 ns] */
    int nu;                       /* station edges */
    const int* uedge;             /* [nu] */
    const int* contrib_ptr;       /* [nu + 1] */
    const int* contrib;           /* [4 contrib_ptr[nu]] */
    const double* contrib_we;     /* [contrib_ptr[nu]] */
} m3d_station_desc;

/* Uploads the table; the set keeps its context alive until it is destroyed. */
int m3d_stations_create(m3d_context* ctx, const m3d_station_desc* desc, m3d_stations** out);
int m3d_stations_destroy(m3d_stations* st);
/* m3d_operator_set_context for a station set (no reference counterpart;
 * the same many-model scheduling). */
int m3d_stations_set_context(m3d_stations* st, m3d_context* ctx);

/* extract_impedance_ (:1490-1727) of nf frequencies: omega [nf] host,
 * station_sigma [ns] host (conductivity of each station cell, used by the
 * 1-D fallback when |det H| <= 1e-30), d_E device [2 nf][n] complex, d_Z
 * device [nf][ns][4] complex (Zxx, Zxy, Zyx, Zyy; 0 for stations without a
 * cell).  The set keeps each (frequency, station) H, det H and Z on the
 * device, as ForwardSolverFV's station_H_cache_ (forward_solver_fv.h:
 * 216-224), for the next m3d_adjoint_rhs_device.  Returns after the context
 * stream drains. */
int m3d_impedance_device(m3d_stations* st, int nf, const double* omega, const double* station_sigma,
                         const double* d_E, double* d_Z);

/* Weighted residual and misfit of the frequencies of the last
 * m3d_impedance_device call, its Z as the prediction: per datum
 *   wr = w * w * conj(obs - pred), w = 1 / err   (inversion_fv.cpp:258-271)
 * and d_phi[0] = sum over frequencies, stations and the four components of
 * 0.5 w w |obs - pred|^2 in that order (:75-98); data with err <= 0 or
 * weight <= 0 give wr = 0 and add nothing.  d_obs device [nf][ns][4]
 * complex, d_err and d_weight device [nf][ns][4] double (d_weight may be
 * NULL: all 1), d_wr device [nf][ns][4] complex, d_phi device double [1].
 * Returns after the context stream drains. */
int m3d_weighted_residual_device(m3d_stations* st, int nf, const double* d_obs, const double* d_err,
                                 const double* d_weight, double* d_wr, double* d_phi);

/* build_adjoint_rhs_from_residual (:1804-2013) for every frequency of the
 * last m3d_impedance_device call: d_wr device [nf][ns][4] complex in, d_rhs
 * device [2 nf][n] complex out (overwritten: the right-hand side of the
 * first polarisation in system 2f, of the second in 2f + 1, zero off the
 * station edges).  Returns after the context stream drains. */
int m3d_adjoint_rhs_device(m3d_stations* st, int nf, const double* d_wr, double* d_rhs);

#ifdef __cplusplus
}
#endif

#endif /* M3D_DEVICE_H */
