/* alps_ed.h -- C ABI of the MI355X exact-diagonalization engine (gfx950).
 *
 * This is the device boundary of the ALPS `sparsediag` hot path: Hamiltonian
 * construction over the fixed-Sz spin-1/2 basis, the sparse H.v product and
 * the ietl::lanczos recurrence.  In the reference the boundary does not
 * exist as an ABI: sparsediag instantiates the header-only
 *     ietl::lanczos<ublas::mapped_vector_of_mapped_vector<double>,
 *                   ietl::vectorspace<ublas::vector<double>>>
 * (applications/diag/sparsediag/sparsediag.h:26,96,104-113) and the work
 * happens in uBLAS prod / inner_prod / norm_2 (src/ietl/interface/ublas.h:27-63)
 * and in hamiltonian_matrix::build (src/alps/model/hamiltonian_matrix.hpp:423-450).
 * Each entry point below names the reference interface it replaces.  The C++
 * ietl vector space over these calls is include/alps_ed/ietl_gpu.hpp.
 *
 * Conventions
 *   - Every function returns ALPS_ED_OK (0) or a negative status; the text of
 *     the last error of the calling thread is alps_ed_last_error().
 *   - Handles are opaque; all sizes are element counts.
 *   - Vectors live in device memory (HBM) of the current HIP device and are
 *     stored as FP32 or FP64 (alps_ed_dtype).  Scalars crossing the ABI are
 *     FP64; dot products and norms accumulate in FP64 whatever the storage.
 *   - All work is issued on the stream set by alps_ed_set_stream (default:
 *     the null stream).  Functions that return host values synchronize it.
 *   - Results are deterministic: reductions use a fixed tree, never float
 *     atomics, so repeated runs give bitwise identical alpha/beta.
 */
#ifndef ALPS_ED_H
#define ALPS_ED_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALPS_ED_OK 0
#define ALPS_ED_ERR_ARGUMENT -1   /* invalid argument (std::invalid_argument in C++) */
#define ALPS_ED_ERR_DEVICE -2     /* HIP runtime error */
#define ALPS_ED_ERR_MEMORY -3     /* device allocation failed or over budget */
#define ALPS_ED_ERR_SOLVER -4     /* tridiagonal solve failed (tmatrix.h:144-145), Lanczos breakdown or
                                     non-finite coefficients, or a broken internal invariant */
#define ALPS_ED_ERR_STATE -5      /* call out of order (lanczos.h:362-363) */

typedef enum { ALPS_ED_F32 = 0, ALPS_ED_F64 = 1 } alps_ed_dtype;

typedef struct alps_ed_operator_s* alps_ed_operator;
typedef struct alps_ed_vector_s* alps_ed_vector;
typedef struct alps_ed_lanczos_s* alps_ed_lanczos;
typedef struct alps_ed_batch_s* alps_ed_batch;

/* Library ------------------------------------------------------------------ */

int alps_ed_version(void);                 /* 10000*major + 100*minor + patch */
const char* alps_ed_last_error(void);
int alps_ed_device_info(char* name, size_t name_len, uint64_t* hbm_bytes, int* n_cu);
/* Stream for all subsequent work of this thread (a hipStream_t; NULL = null stream). */
int alps_ed_set_stream(void* hip_stream);
int alps_ed_synchronize(void);

/* Model -------------------------------------------------------------------- */

/* A spin-1/2 model with conserved Sz on n_sites <= 32 sites, restricted to
 * the sector with n_up sites up (Sz_total = n_up - n_sites/2).  Bonds are
 * given in the graph's edge order (latticegraph.h:98-136) with per-bond
 * couplings; the local terms are those of the "spin" HAMILTONIAN
 * (lib/xml/models.xml.in:342-373):
 *   site s:  -h[s]*Sz + d[s]*Sz^2
 *   bond b:  jz[b]*Sz*Sz + jxy[b]*(S+S- + S-S+)/2 + k[b]*biquadratic
 * Null coupling arrays mean zero. */
typedef struct {
  int n_sites;
  int n_up;
  int n_bonds;
  const int* bond_site1;
  const int* bond_site2;
  const double* jz;
  const double* jxy;
  const double* k;
  const double* h;
  const double* d;
} alps_ed_model;

/* Operator (replaces hamiltonian_matrix<M>::build and its matrix_,
 * src/alps/model/hamiltonian_matrix.hpp:67-231,388-450; the operator is
 * matrix-free: nothing of size nnz is stored). ---------------------------- */

int alps_ed_operator_create(const alps_ed_model* model, alps_ed_operator* op);
/* The momentum (Bloch) sector of the model's fixed-Sz sector
 * (bloch_basis_states, src/alps/model/blochbasisstates.h:97-271, and the
 * matrix in it, hamiltonian_matrix.hpp:268-345) under n_translations
 * translations as graph_helper::translations(k) lists them
 * (hypercubic.h:425-462): site maps perm[t*n_sites + i] (image of site i;
 * t = 0 is the identity) with phases (phase_re[t], phase_im[t]).  The
 * sector's vectors are complex and stored as 2d reals with the real and
 * imaginary part of each element side by side (Re v_0, Im v_0, Re v_1, ...,
 * the layout of a std::complex<double> array): the operator's dimension is
 * 2d and every vector call
 * (matvec, Lanczos, Ritz vectors) works on that form.  Explicit states,
 * elements and spin sums are not available for it.  The model must be
 * invariant under the translations. */
int alps_ed_operator_create_momentum(const alps_ed_model* model, int n_translations, const int* perm,
                                     const double* phase_re, const double* phase_im, alps_ed_operator* op);
/* Releases the caller's reference to the operator: it is freed once no
 * Lanczos object created on it remains (each holds a reference until
 * alps_ed_lanczos_destroy). */
int alps_ed_operator_destroy(alps_ed_operator op);
/* No counterpart in the reference: the H.v schedule (waves per SIMD and
 * flip-loop chunk sizes) an operator uses with FP32 vectors.  Operators of
 * fixed-Sz sectors with at least 200000 states pick theirs at creation by
 * timing all ALPS_ED_HV_SCHEDULES of them; seconds[s] receives each one's
 * time per H.v (-1 when the operator was not timed; seconds may be null).
 * Schedules differ only in speed: every one gives the same y.
 * set_schedule forces one (for inspection and tests). */
#define ALPS_ED_HV_SCHEDULES 6
int alps_ed_operator_schedule(alps_ed_operator op, int* schedule, double* seconds);
int alps_ed_operator_set_schedule(alps_ed_operator op, int schedule);
/* hamiltonian_matrix::dimension (hamiltonian_matrix.hpp:92); the length of
 * the operator's vectors (twice the sector dimension for a momentum sector). */
int alps_ed_operator_dimension(alps_ed_operator op, uint64_t* dim);
/* Device bytes held by the operator's tables. */
int alps_ed_operator_bytes(alps_ed_operator op, uint64_t* bytes);
/* Basis states [first, first+count) as bit strings, bit (N-1-s) set iff site
 * s is up; the order is basis_states::build's (basisstates.h:176-278). */
int alps_ed_operator_states(alps_ed_operator op, uint64_t first, uint64_t count, uint32_t* states_out);
/* basis_states::index (basisstates.h:68-74): index of each state, or the
 * dimension when the state is not in the sector. */
int alps_ed_operator_index(alps_ed_operator op, const uint32_t* states, uint64_t count, uint64_t* index_out);
/* Explicit matrix rows [row0, row0+nrows): the elements H(i,j) that
 * hamiltonian_matrix::build accumulates (sites first, then bonds in edge
 * order, hamiltonian_matrix.hpp:388-397), columns ascending, the explicit
 * diagonal always present (:44-58).  Row r occupies cols/vals
 * [r*max_per_row, r*max_per_row + counts[r]).  Needs max_per_row >=
 * n_bonds + 1. */
int alps_ed_operator_elements(alps_ed_operator op, uint64_t row0, uint64_t nrows, int max_per_row,
                              uint32_t* cols, double* vals, int* counts);

/* Vectors (replace ietl::vectorspace<ublas::vector<double>>, new_vector,
 * vec_dimension, generate, clear, copy, dot, two_norm and the uBLAS vector
 * expressions; src/ietl/vectorspace.h:23-46,101-109,
 * src/ietl/interface/ublas.h:27-55, lanczos.h:383-402). ------------------- */

int alps_ed_vector_create(uint64_t n, alps_ed_dtype dtype, alps_ed_vector* v);
int alps_ed_vector_destroy(alps_ed_vector v);
int alps_ed_vector_size(alps_ed_vector v, uint64_t* n);
int alps_ed_vector_dtype(alps_ed_vector v, alps_ed_dtype* dtype);
/* Raw device pointer of the storage (float* or double*). */
int alps_ed_vector_data(alps_ed_vector v, void** device_ptr);
int alps_ed_vector_upload(alps_ed_vector v, const double* host, uint64_t n);
int alps_ed_vector_download(alps_ed_vector v, double* host, uint64_t n);
/* Counter-based start vector: v[i] = lo + (hi-lo)*u_i with u_i from the
 * splitmix64 output of seed + (i+1)*0x9e3779b97f4a7c15, uniform on [0,1). */
int alps_ed_vector_fill_random(alps_ed_vector v, uint64_t seed, double lo, double hi);
int alps_ed_vector_fill(alps_ed_vector v, double value);
int alps_ed_vector_copy(alps_ed_vector dst, alps_ed_vector src);
int alps_ed_vector_scale(alps_ed_vector v, double a);                       /* v *= a */
int alps_ed_vector_axpby(alps_ed_vector y, double a, alps_ed_vector x, double b); /* y = a*x + b*y */
int alps_ed_vector_dot(alps_ed_vector x, alps_ed_vector y, double* result);
int alps_ed_vector_nrm2(alps_ed_vector x, double* result);

/* ietl::mult(m, x, y) -> y = prod(m, x) (interface/ublas.h:57-63). */
int alps_ed_matvec(alps_ed_operator op, alps_ed_vector x, alps_ed_vector y);

/* A momentum-sector vector w ((re, im) pairs, of an operator from
 * alps_ed_operator_create_momentum) in the configuration basis of its
 * fixed-Sz sector: re + i im, FP64 vectors of C(N, n_up) elements.  State r
 * of a kept orbit gets phase_t * norm_a * w_a, with a its representative and
 * t the first translation taking r to it (the Bloch states of
 * src/alps/model/blochbasisstates.h:97-271); other states get 0.  The
 * expanded vector is an eigenvector of the sector's Hamiltonian when w is
 * one of the momentum sector's; eigenvector measurements
 * (DiagMatrix::perform_measurements in a momentum sector,
 * applications/diag/diag.h:191-314) run on it. */
int alps_ed_vector_bloch_expand(alps_ed_operator op, alps_ed_vector w, alps_ed_vector re, alps_ed_vector im);

/* Spin expectation sums of a vector v of the operator's sector, the parts
 * from which SparseDiagMatrix::calculate (sparsediag.h:151-159) evaluates
 * <v|O|v> for the measurements of DiagMatrix::perform_measurements
 * (applications/diag/diag.h:191-314).  With w_r = v_r^2 and Sz_i(r) = +-1/2:
 *   *norm2 = sum_r w_r;  sz[i] = sum_r w_r Sz_i(r);
 *   szsz[i*n_sites + j] = sum_r w_r Sz_i(r) Sz_j(r)   (i, j model sites).
 * Any output may be null.  Sums are FP64 and deterministic; divide by
 * *norm2 for expectation values of an unnormalised v. */
int alps_ed_vector_spin_moments(alps_ed_operator op, alps_ed_vector v, double* norm2, double* sz, double* szsz);
/* out[p] = <v| S+_{site_i[p]} S-_{site_j[p]} |v> = sum over the rows r with
 * site j up and site i down of v_r v_r', r' the row with both turned
 * (site_i[p] != site_j[p]).  <v|S-_i S+_j|v> is the pair (j, i). */
int alps_ed_vector_flip_pairs(alps_ed_operator op, alps_ed_vector v, int n_pairs, const int* site_i,
                              const int* site_j, double* out);

/* One ietl::lanczos::make_step on caller vectors (lanczos.h:392-406), fused:
 * w = H v_cur; alpha = <v_cur,w>; w -= alpha v_cur + beta_prev v_prev;
 * beta = |w|; w /= beta.  v_prev and v_cur are not modified.  Used by the
 * device vector space of include/alps_ed/ietl_gpu.hpp. */
int alps_ed_lanczos_step_vectors(alps_ed_operator op, alps_ed_vector v_prev, alps_ed_vector v_cur,
                                 alps_ed_vector w, double beta_prev, double* alpha, double* beta);

/* alps_ed_lanczos_step_vectors without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels, same
 * results bit for bit; the reduction partials and scalars live in a static
 * device array of 2 MiB + 8 KiB + 128 B per GPU, held until the process ends.
 * Operators whose H.v has more than 262144 blocks (segments, or blocks of 256
 * momentum-sector rows) are rejected with ALPS_ED_ERR_ARGUMENT.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (several
 * host threads or streams) race and must be serialized by the caller. */
int alps_ed_lanczos_step_vectors_noalloc(alps_ed_operator op, alps_ed_vector v_prev, alps_ed_vector v_cur,
                                         alps_ed_vector w, double beta_prev, double* alpha, double* beta);

/* Lanczos recurrence with the Krylov state resident on the device
 * (ietl::lanczos::make_first_step / make_step / generate_tmatrix,
 * src/ietl/lanczos.h:341-406).  Storage dtype applies to the three Lanczos
 * vectors; alpha, beta and the normalisations are FP64 on the device. ----- */

/* keep_basis != 0 keeps every Lanczos vector in HBM (max_steps+1 vectors,
 * allocated in chunks as the recurrence advances) for alps_ed_lanczos_ritz. */
int alps_ed_lanczos_create(alps_ed_operator op, alps_ed_dtype dtype, int max_steps, int keep_basis,
                           alps_ed_lanczos* lz);
/* Frees the Lanczos object and releases its reference to the operator. */
int alps_ed_lanczos_destroy(alps_ed_lanczos lz);
/* Reorthogonalization of the recurrence, set before alps_ed_lanczos_start.
 * mode 0: none, the reference's recurrence (make_step, lanczos.h:392-406,
 * reorthogonalizes nothing).  mode 1: partial reorthogonalization (Simon
 * 1984) for FP32 storage, which needs keep_basis: each step extends the
 * omega estimates of v_{j+1}.v_k in FP64 on the device, and when the largest
 * exceeds sqrt(eps) (eps the storage's unit roundoff) the new vector and the
 * next one are orthogonalized against the kept basis by one Gram-Schmidt
 * pass.  Steps stay free of host synchronization. */
int alps_ed_lanczos_set_reorthogonalization(alps_ed_lanczos lz, int mode);
/* Number of steps whose new vector was reorthogonalized (synchronizes). */
int alps_ed_lanczos_reorthogonalizations(alps_ed_lanczos lz, int* count);
/* make_first_step with the start vector given by fill_random(seed, -1, 1),
 * or by `start` when it is not NULL (must have the operator's dimension). */
int alps_ed_lanczos_start(alps_ed_lanczos lz, uint64_t seed, alps_ed_vector start);
/* nsteps more make_step calls, enqueued without host synchronization. */
int alps_ed_lanczos_steps(alps_ed_lanczos lz, int nsteps);
/* Copies alpha[0..n), beta[0..n) with n = steps done so far (synchronizes). */
int alps_ed_lanczos_coefficients(alps_ed_lanczos lz, double* alpha, double* beta, int max_n, int* n);
/* The current Lanczos vector v_n (normalized) and its predecessor, copied
 * into caller vectors of the same dimension. */
int alps_ed_lanczos_vectors(alps_ed_lanczos lz, alps_ed_vector v_prev, alps_ed_vector v_cur);

/* Snapshot and restore of a recurrence (lanczos::save / load,
 * src/ietl/lanczos.h:86-98, which archive the T matrix, the last two
 * Lanczos vectors and the step count), e.g. to free the device for other
 * work and continue later.  The snapshot is a host byte image: alpha, beta,
 * the normalizations, the reduction partials of the next step, the last two
 * Lanczos vectors (the whole kept basis with keep_basis) and the partial
 * reorthogonalization state.  alps_ed_lanczos_save synchronizes.
 * alps_ed_lanczos_load takes a fresh object (before lanczos_start) of an
 * operator of the same model, with the same dtype, max_steps, keep_basis and
 * reorthogonalization mode; steps after it give bit-identical coefficients to
 * an uninterrupted run.  Errors: ALPS_ED_ERR_ARGUMENT for a buffer that is
 * too short, not a snapshot, or of another operator or configuration;
 * ALPS_ED_ERR_STATE for a load into an object that has started. */
int alps_ed_lanczos_snapshot_size(alps_ed_lanczos lz, size_t* bytes);
int alps_ed_lanczos_save(alps_ed_lanczos lz, void* buf, size_t len);
int alps_ed_lanczos_load(alps_ed_lanczos lz, const void* buf, size_t len);

/* Ritz vectors from the kept basis (the basis transformation of
 * lanczos::eigenvectors, src/ietl/lanczos.h:242-281, without the replay):
 * outs[i] = sum_{j<m} coeffs[i*m + j] v_j for i < nvec, v_j the normalized
 * Lanczos vectors, m <= steps; outs are FP64 vectors of the dimension. */
int alps_ed_lanczos_ritz(alps_ed_lanczos lz, int nvec, int m, const double* coeffs, alps_ed_vector* outs);

/* T-matrix analysis (ietl::Tmatrix::compute, src/ietl/tmatrix.h:121-228):
 * distinct non-ghost eigenvalues of the n x n tridiagonal with diagonal
 * alpha and off-diagonal beta[0..n-2], their error bounds |beta[n-1]*z_n|
 * and multiplicities.  Host computation, no device work. */
int alps_ed_tmatrix_analyze(int n, const double* alpha, const double* beta, double* eigenvalues,
                            double* errors, int* multiplicities, int* n_out);

/* SparseDiagMatrix<double>::do_subspace (sparsediag.h:93-149) for one
 * sector: builds the operator, runs the Lanczos recurrence on the device
 * with the lanczos_iteration_nlowest stopping rule (iteration.h:107-134;
 * rtol = atol = 100*eps) or a fixed iteration count, and returns the lowest
 * `nev` distinct non-ghost eigenvalues.  max_iter <= 0 selects the
 * reference default min(10*D, 1000); D == 1 returns H(0,0).
 * With FP64 storage the eigenvalues are those of T, as in the reference.
 * With FP32 storage the basis is kept in HBM, near-copies of eigenvalues
 * (within 64 eps_f32 |lambda|) are merged keeping their smallest error bound,
 * the stopping tolerance is 1e-8 (the FP32 T-matrix bounds stall above
 * 100 eps), and each reported eigenvalue is the FP64-accumulated Rayleigh
 * quotient of its Ritz vector, whose error is about the square of the
 * residual.
 * FP32 storage is used only for sectors with D >= 1000 and nev <= 2; other
 * solves store FP64 whatever `dtype` asks: in small sectors the Krylov space
 * is a large part of the sector, and for more eigenvalues the copies of
 * converged levels that FP32 rounding creates early delay the higher ones
 * (about twice the FP64 iteration count for four), costing more than the
 * halved bytes save. */
typedef struct {
  int max_iter;
  int nev;
  int fixed_iterations;   /* fixed_lanczos_iteration semantics when non-zero */
  uint64_t seed;
  alps_ed_dtype dtype;    /* storage of the Lanczos vectors */
} alps_ed_solve_params;

typedef struct {
  uint64_t dim;
  int n_eigenvalues;
  double eigenvalues[64];
  int iterations;
  int error_code;         /* 1 when max_iter was reached (iteration.h:80); in a batch also -1 (can
                           * never fit the arena) and -2 (non-finite Lanczos coefficients, no
                           * eigenvalues), which a single solve reports as ALPS_ED_ERR_SOLVER.
                           * A zero beta ends the recurrence early with the exact T (error 0). */
  double seconds;         /* device wall time of the solve */
  alps_ed_dtype storage;  /* storage of the Lanczos vectors actually used: an FP32 request is served
                           * in FP64 for sectors of fewer than 1000 states (and, in a batch, for
                           * more than two requested levels) */
  int reorthogonalizations; /* Lanczos steps whose new vector was reorthogonalized (partial
                             * reorthogonalization, used with FP32 storage in single solves) */
} alps_ed_solve_result;

int alps_ed_sparsediag(const alps_ed_model* model, const alps_ed_solve_params* params,
                       alps_ed_solve_result* result);
/* do_subspace in a momentum sector (translations as for
 * alps_ed_operator_create_momentum); result->dim is the sector dimension. */
int alps_ed_sparsediag_momentum(const alps_ed_model* model, int n_translations, const int* perm,
                                const double* phase_re, const double* phase_im, const alps_ed_solve_params* params,
                                alps_ed_solve_result* result);

/* ---------------------------------------------------------------------------
 * Batched solves and the request loop.
 *
 * Many independent do_subspace solves (each its own Hamiltonian and sector,
 * as DiagMatrix::dostep produces one per sector, diag.h:86-138, and as the
 * reference's task farm runs one per task) advance together: one launch of
 * each batched kernel serves every problem in the batch, so small sectors
 * fill the GPU.  Requests are admitted first come first served while their
 * HBM footprint (tables, Lanczos vectors -- the whole basis for FP32 storage
 * -- scalars and refinement buffers) fits the batch's arena; requests
 * submitted while others run join at the next round.  Each problem's
 * arithmetic is independent of the rest of the batch: its eigenvalues are
 * bit-identical alone and in any mix.  The stopping rule, precision policy
 * and results are those of alps_ed_sparsediag; the start vector uses the
 * request's seed.  A request that can never fit is completed with
 * error_code -1 and no eigenvalues.
 * ------------------------------------------------------------------------- */
typedef struct alps_ed_batch_s* alps_ed_batch;

typedef struct {
  alps_ed_dtype dtype;    /* storage of the Lanczos vectors (FP32 as in alps_ed_sparsediag) */
  uint64_t hbm_bytes;     /* arena size; 0 = half of the free device memory */
  uint64_t host_bytes;    /* mapped host memory for alpha/beta mirrors; 0 = 256 MiB */
  int steps_per_round;    /* Lanczos steps between stopping-rule tests; 0 = 8 */
  int max_active;         /* problems admitted at once; 0 = no limit */
  int host_threads;       /* threads of the host analysis; 0 = 8 */
  int graphs;             /* 1: a round whose launches repeat the previous round's (same work
                           * lists, no new problem) replays them from a captured HIP graph */
} alps_ed_batch_config;

typedef struct {
  int n_active, n_pending;
  uint64_t hbm_used, hbm_bytes;
  uint64_t rounds, steps, solves_done;
  /* wall seconds spent in: admission (layouts, uploads), the Lanczos steps
   * (launch to completion), the stopping-rule tests, refinement */
  double seconds_admit, seconds_steps, seconds_decide, seconds_refine;
  /* rows processed by the Lanczos steps' H.v launches, by storage type
   * (sum over steps of the dimension of each problem stepped): the
   * algorithmic H.v bytes are 8 x hv_rows_f32 + 16 x hv_rows_f64 */
  uint64_t hv_rows_f32, hv_rows_f64;
  /* rounds whose steps ran from a HIP graph, and graphs captured */
  uint64_t graph_rounds, graph_captures;
} alps_ed_batch_status;

int alps_ed_batch_create(const alps_ed_batch_config* cfg, alps_ed_batch* out);
int alps_ed_batch_destroy(alps_ed_batch b);
/* Queues one solve; the model arrays are copied.  *ticket identifies it. */
int alps_ed_batch_submit(alps_ed_batch b, const alps_ed_model* model, const alps_ed_solve_params* params,
                         uint64_t* ticket);
/* Queues one solve in a momentum sector (translations as for
 * alps_ed_operator_create_momentum; the result's dim is the sector
 * dimension).  The same request as alps_ed_sparsediag_momentum, advanced
 * with the rest of the batch: DiagMatrix::dostep's momentum loop inside the
 * Sz loop (applications/diag/diag.h:85-138) as one batch.  Momentum sectors
 * of one Sz sector share the orbits, built once. */
int alps_ed_batch_submit_momentum(alps_ed_batch b, const alps_ed_model* model, int n_translations, const int* perm,
                                  const double* phase_re, const double* phase_im, const alps_ed_solve_params* params,
                                  uint64_t* ticket);
/* One round: admit, start, steps_per_round steps, stopping-rule tests,
 * refinement, retirement.  Returns the counts after the round. */
int alps_ed_batch_progress(alps_ed_batch b, int* n_active, int* n_pending);
/* Rounds until every submitted request is complete. */
int alps_ed_batch_run(alps_ed_batch b);
/* Tickets completed since the last call (at most max), oldest first. */
int alps_ed_batch_completed(alps_ed_batch b, uint64_t* tickets, int max, int* n);
/* Result of a ticket; *ready = 0 while it is not complete. */
int alps_ed_batch_result(alps_ed_batch b, uint64_t ticket, alps_ed_solve_result* result, int* ready);
int alps_ed_batch_stats(alps_ed_batch b, alps_ed_batch_status* stats);
/* HBM bytes a request would hold while admitted. */
int alps_ed_batch_footprint(const alps_ed_model* model, const alps_ed_solve_params* params, alps_ed_dtype dtype,
                            uint64_t* bytes);
/* The same for a momentum-sector request (alps_ed_batch_submit_momentum). */
int alps_ed_batch_footprint_momentum(const alps_ed_model* model, int n_translations, const int* perm,
                                     const double* phase_re, const double* phase_im,
                                     const alps_ed_solve_params* params, alps_ed_dtype dtype, uint64_t* bytes);

/* The sparsediag task for the spin-1/2 path: runs every task of a parameter
 * file (the grammar of src/alps/parameter/parameterlist.h:20-36; LATTICE,
 * L, W, H, MODEL="spin", local_S=1/2, J/Jz/Jxy/J'/J1/Jz#/Jxy#/K#/h#/D#,
 * CONSERVED_QUANTUMNUMBERS, Sz_total, TRANSLATION_SYMMETRY, TOTAL_MOMENTUM,
 * MAX_ITERATIONS, NUMBER_EIGENVALUES, MEASURE_AVERAGE/LOCAL/CORRELATIONS/
 * STRUCTURE_FACTOR[name], plus ED_PRECISION = "f64" | "f32" and ED_SEED) as
 * DiagMatrix::dostep does (applications/diag/diag.h:85-175): one Lanczos
 * solve per Sz sector and, on lattices with periodic directions unless
 * TRANSLATION_SYMMETRY is false, per momentum inside it (the Bloch sectors
 * of alps_ed_operator_create_momentum; TOTAL_MOMENTUM selects one).  The
 * sectors of a task, with or without momenta, advance together as one batch
 * (alps_ed_batch_*) when their footprint fits the free device memory, else
 * one after the other.  Then the ground-state energy and gap of
 * SparseDiagMatrix::write_xml_body (sparsediag.h:59-90), the eigenvector
 * measurements (diag.h:191-314; in momentum sectors on the eigenvectors
 * expanded by alps_ed_vector_bloch_expand, without local values, :212) and
 * the task's result XML.  The input may also be a task or parameter XML
 * document (<SIMULATION>/<PARAMETERS>, <PARAMETERLIST>).  Tasks that measure
 * solve their sectors one at a time.  The results are written to `out` as a
 * JSON document (at most
 * out_len bytes including the terminating zero; *needed receives the full
 * length). */
int alps_ed_run_parameter_file(const char* text, char* out, size_t out_len, size_t* needed);

/* The same parameter file resolved on the host only, without solving: per
 * task its parameters, the lattice's sites and bonds in edge order (the
 * graph of make_graph_from_lattice, src/alps/lattice/latticegraph.h:35-171),
 * the per-bond Jz, Jxy, K and per-site h, D after the spin model's defaults
 * (lib/xml/models.xml.in:342-373), the Sz sectors and momenta to solve
 * (applications/diag/diag.h:85-175) and the measurement requests with
 * their labels.  Needs no device. */
int alps_ed_describe_parameter_file(const char* text, char* out, size_t out_len, size_t* needed);

#ifdef __cplusplus
}
#endif

#endif /* ALPS_ED_H */
