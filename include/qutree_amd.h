/*
 * qutree_amd.h -- C ABI of the MI355X (gfx950) tree-tensor contraction engine.
 *
 * Conventions
 *  - Complex data is interleaved (re, im): complex128 as double[2], complex64
 *    as float[2].  Suffix _z = complex128, _c = complex64.
 *  - Tensors are column-major with the first index fastest; an order-f tensor
 *    seen on mode k is the 3-index array (before, n_k, after) with
 *    before = prod_{i<k} d_i and after = prod_{i>k} d_i
 *    (src/Core/TensorShape.cpp:3-40 of QuTree).  Matrices are column-major,
 *    (i, j) -> j * dim1 + i (include/Core/Matrix_Implementation.h:84-96).
 *  - Node-level entry points take DEVICE pointers.  The caller owns every
 *    buffer; nothing is allocated by a node-level call (grouped plans own
 *    their descriptor copies and split-K workspace, allocated at creation).
 *  - Every call returns a qtg_status; nothing calls exit().  qtg_last_error()
 *    returns a description of the most recent failure on the calling thread.
 *  - stream == NULL means the legacy default stream for node-level calls and
 *    the engine's own non-blocking stream for qtg_engine_* calls;
 *    qtg_synchronize(NULL) waits for the whole device.
 */
#ifndef QUTREE_AMD_H
#define QUTREE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  QTG_OK = 0,
  QTG_ESHAPE = 1,   /* dimensions inconsistent or unsupported */
  QTG_EDEVICE = 2,  /* HIP runtime error or no device */
  QTG_EARG = 3,     /* bad pointer / flag */
  QTG_ENOMEM = 4,   /* device allocation failed */
  QTG_ESTATE = 5    /* object not in the required state (e.g. not built) */
} qtg_status;

typedef void* qtg_stream;

/* ---------------------------------------------------------------- runtime */
int qtg_init(int device);
const char* qtg_last_error(void);
int qtg_device_alloc(void** ptr, size_t bytes);
int qtg_device_free(void* ptr);
/* dst = src for n doubles on the device by the library's own copy kernel
 * (n even, 16-byte aligned pointers): reads and writes 8 n bytes, which
 * also calibrates the HBM byte counters of the profiles. */
int qtg_copy_d(double* dst, const double* src, int64_t n, qtg_stream stream);
int qtg_memcpy_h2d(void* dst, const void* src, size_t bytes, qtg_stream stream);
int qtg_memcpy_d2h(void* dst, const void* src, size_t bytes, qtg_stream stream);
int qtg_memset(void* dst, int value, size_t bytes, qtg_stream stream);
int qtg_stream_create(qtg_stream* stream);
int qtg_stream_destroy(qtg_stream stream);
int qtg_synchronize(qtg_stream stream);
/* Elapsed milliseconds between two points recorded on `stream` around a
 * callback-free region: qtg_timer_start / qtg_timer_stop bracket the work. */
int qtg_timer_start(qtg_stream stream);
int qtg_timer_stop(qtg_stream stream, float* ms);

/* ------------------------------------------------------------- node level */
/* Mode-k product, replaces LA_lib.f matvec_ (src/Core/LA_lib.f:34-81) as
 * bound in include/Core/Tensor_Implementation.h:450-461 and called from
 * matrixTensor (Tensor_Implementation.h:554-577):
 *   C(b,k,c) (+)= sum_l M(k,l) B(b,l,c),  M a x a,  B and C (b, a, c).
 * add != 0 accumulates into C (the Fortran logical `add`). */
int qtg_matvec_z(double* C, const double* B, const double* M, int64_t a, int64_t b, int64_t c,
                 int add, qtg_stream stream);
/* Transposed product, replaces LA_lib.f tmatvec_ (LA_lib.f:104-151):
 *   C(b,k,c) (+)= sum_l M(l,k) B(b,l,c). */
int qtg_tmatvec_z(double* C, const double* B, const double* M, int64_t a, int64_t b, int64_t c,
                  int add, qtg_stream stream);
/* Adjoint product, replaces LA_lib.f ctmatvec_ (LA_lib.f:244-291) behind
 * tMatrixTensor (Tensor_Implementation.h:580-644):
 *   C(b,k,c) (+)= sum_l conj(M(l,k)) B(b,l,c). */
int qtg_ctmatvec_z(double* C, const double* B, const double* M, int64_t a, int64_t b, int64_t c,
                   int add, qtg_stream stream);
/* Hole contraction, replaces LA_lib.f rhomat_ (LA_lib.f:718-762) behind
 * contraction() (Tensor_Implementation.h:482-531):
 *   M(j,l) = sum_{n,m} conj(bra(n,j,m)) ket(n,l,m); always overwrites M. */
int qtg_rhomat_z(const double* bra, const double* ket, double* M, int64_t a, int64_t b, int64_t c,
                 qtg_stream stream);

/* Real products, replace LA_lib.f rmatvec_ (LA_lib.f:314-379, bound in
 * Tensor_Implementation.h:457 and called from matrixTensor for real M and B,
 * :569-572) and trmatvec_ (LA_lib.f:402-467, bound in Matrix_Implementation.h:
 * 459 and called from multAB for real square B, :481-483):
 *   rmatvec:  C(b,k,c) (+)= sum_l M(k,l) B(b,l,c),
 *   trmatvec: C(b,k,c) (+)= sum_l M(l,k) B(b,l,c),  all real FP64. */
int qtg_rmatvec_d(double* C, const double* B, const double* M, int64_t a, int64_t b, int64_t c, int add,
                  qtg_stream stream);
int qtg_trmatvec_d(double* C, const double* B, const double* M, int64_t a, int64_t b, int64_t c, int add,
                   qtg_stream stream);

/* Rectangular mode product with matrixTensorBLAS semantics
 * (src/Core/TensorBLAS.cpp:124-173,232-271): M is nC x nB, B (before,nB,after),
 * C (before,nC,after).  op: 0 = M, 1 = M^T, 2 = M^H (M then nB x nC). */
int qtg_mode_product_z(double* C, const double* M, const double* B, int64_t before, int64_t nB,
                       int64_t nC, int64_t after, int op, int add, qtg_stream stream);
int qtg_mode_product_c(float* C, const float* M, const float* B, int64_t before, int64_t nB,
                       int64_t nC, int64_t after, int op, int add, qtg_stream stream);
/* Hole contraction with contractionBLAS semantics (TensorBLAS.cpp:200-225,
 * 279-326): S (nA x nB) = (accumulate ? S : 0) + sum conj(A(b,j,a)) B(b,l,a). */
int qtg_hole_contraction_z(double* S, const double* A, const double* B, int64_t before, int64_t nA,
                           int64_t nB, int64_t after, int accumulate, qtg_stream stream);
int qtg_hole_contraction_c(float* S, const float* A, const float* B, int64_t before, int64_t nA,
                           int64_t nB, int64_t after, int accumulate, qtg_stream stream);

/* Grouped variants: `count` independent problems of uneven shape launched
 * together (no two may write the same output).  A plan copies the
 * descriptors to the device once; executing a plan allocates nothing.
 * Offsets are in complex elements from the base pointers passed at
 * execution.  No reference counterpart: these batch the per-term,
 * per-node calls of TreeFunctions
 * (include/TreeClasses/SparseMatrixTreeFunctions_Implementation.h:18-276). */
typedef struct {
  int64_t offC, offM, offB;       /* complex-element offsets */
  int64_t before, nB, nC, after;  /* view of B on the contracted mode */
  int32_t op, add;                /* op as in qtg_mode_product_z */
} qtg_mode_product_desc;
typedef struct {
  int64_t offS, offA, offB;
  int64_t before, nA, nB, after;
  int32_t accumulate, pad;
} qtg_hole_contraction_desc;
typedef void* qtg_plan;
int qtg_mode_product_plan_create(qtg_plan* plan, const qtg_mode_product_desc* descs, int64_t count,
                                 int complex64);
int qtg_hole_contraction_plan_create(qtg_plan* plan, const qtg_hole_contraction_desc* descs,
                                     int64_t count, int complex64);
/* Cbase/Sbase, Mbase/Abase, Bbase: double* for complex128 plans, float* for complex64. */
int qtg_plan_execute(qtg_plan plan, void* out_base, const void* first_base, const void* second_base,
                     qtg_stream stream);
int qtg_plan_destroy(qtg_plan plan);
/* Total algorithmic FLOPs (8 per complex multiply-add) and minimum HBM bytes
 * (each operand read once, each result written once) of a plan. */
int qtg_plan_counts(qtg_plan plan, double* flops, double* bytes);

/* ------------------------------------------------------ small matrices, FP64 */
/* Batches of `count` matrices stored back to back (device pointers,
 * complex128 column-major), one workgroup per matrix; 1 <= n <= 64.
 * qtg_heev_batch_z: Hermitian eigendecomposition A = U diag(lambda) U^H
 * (Matrix::cDiag, include/Core/Matrix_Implementation.h:284-308): lambda (n
 * doubles per matrix) ascending, the columns of U the eigenvectors, each
 * phased so that its first component is real and >= 0 (the reference only
 * flips columns whose first component has a negative real part; eigenvector
 * phases of its Eigen solver are otherwise the solver's).  U or lambda may be
 * null.  qtg_reg_inverse_batch_z: R = U diag(1 / (l + eps e^{-l/eps})) U^H
 * (SpectralDecompositionTree.cpp:30-48, Vector_Implementation.h:266-280).
 * qtg_inverse_batch_z: A^-1 by Gauss-Jordan with partial pivoting (Matrix::
 * cInv, Matrix_Implementation.h:311-328).  qtg_qr_batch_z: the thin Q (m x n,
 * m >= n) of a Householder QR (qr(Matrixcd), Matrix_Implementation.h:795-806;
 * LAPACK zgeqrf/zungqr sign convention, R with diagonal -sign(Re a_jj) |.|);
 * A and Q may be the same buffer. */
int qtg_heev_batch_z(const double* A, double* U, double* lambda, int64_t n, int64_t count, qtg_stream stream);
int qtg_reg_inverse_batch_z(const double* A, double* R, int64_t n, int64_t count, double eps, qtg_stream stream);
int qtg_inverse_batch_z(const double* A, double* Ainv, int64_t n, int64_t count, qtg_stream stream);
int qtg_qr_batch_z(const double* A, double* Q, int64_t m, int64_t n, int64_t count, qtg_stream stream);

/* No reference counterpart: for callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc).  The same grouped
 * hole contractions as qtg_hole_contraction_plan_create + qtg_plan_execute
 * (complex128; X from A_base, Y from B_base, offsets in complex elements),
 * but the descriptor tables are staged in a fixed device array of the library
 * (about 1.3 MB per GPU, held until the process ends) instead of a plan's
 * allocation: at most 4096 descriptors and 2^18 16 x 16 output tiles
 * (QTG_ESHAPE beyond).  Allocates no device memory; returns when the result
 * is written.  Every call on one GPU shares that array, so calls that overlap
 * on one GPU (other host threads) race. */
int qtg_hole_contraction_grouped_noalloc_z(const qtg_hole_contraction_desc* descs, int64_t count, double* S_base,
                                           const double* A_base, const double* B_base, qtg_stream stream);

/* ------------------------------------------------------------- tree engine */
/* A device-resident model of one tree and one sum-of-products Hamiltonian.
 * The tree is QuTree's compact text (Tree(istream), src/TreeShape/Node.cpp:
 * 101-165, followed by one "omega r0 wfr0 wfomega" line per mode; HO-DVR
 * leaves).  The wave function is the concatenation of the node tensors in
 * node-address order (post-order, last child first, Tree.cpp:92-115), complex128
 * interleaved.  precision: 0 = complex128 throughout, 1 = complex64 bulk
 * contractions with FP64 density inverse, projector and derivative output,
 * 2 = as 1 but with the slots of the large-coefficient terms (|c| >= 0.1
 * max |c|) and the densities in FP64 (grouped path only). */
typedef void* qtg_engine;
int qtg_engine_create(const char* tree_text, int precision, qtg_engine* engine);
int qtg_engine_destroy(qtg_engine engine);
/* SOP H = sum_l coeff[l] prod_f h_f (SumOfProductsOperator / MultiLeafOperator,
 * src/TreeOperators/MultiLeafOperator.cpp:11-64).  Term l owns factors
 * term_off[l] .. term_off[l+1]-1; factor f acts on leaf mode fmode[f] with kind
 * fkind[f]: 0 x, 1 x^2, 2 kinetic, 3 p (DVRBasis.cpp:8-58), 4 identity,
 * 5 explicit d x d complex matrix at pool + 2*fpool[f] (LeafMatrix.cpp:37-41),
 * 6 LeafMatrix a0 T + a1 x + a2 x^2 + a3 x^4 with (a0..a3) at pool + 2*fpool[f]. */
int qtg_engine_set_hamiltonian(qtg_engine engine, int64_t nterms, const double* coeff,
                               const int64_t* term_off, const int64_t* fmode, const int64_t* fkind,
                               const int64_t* fpool, const double* pool);
int64_t qtg_engine_psi_size(qtg_engine engine);  /* complex elements */
int64_t qtg_engine_num_nodes(qtg_engine engine);
/* info: parent, childIdx, bottom, depth, order, nchildren, mode (-1), offset,
 * dims[order], children[nchildren] */
int qtg_engine_node_info(qtg_engine engine, int64_t addr, int64_t* info, int64_t len);
/* TensorTree(mt19937(seed), tree, delta_lowest) (TensorTree_Implementation.h:39-76) */
int qtg_engine_fill_random(qtg_engine engine, uint64_t seed, int delta_lowest, double* psi_host);
int qtg_engine_set_psi(qtg_engine engine, const double* psi_host, qtg_stream stream);
int qtg_engine_set_psi_device(qtg_engine engine, const double* psi_device, qtg_stream stream);
/* HamiltonianRepresentation::build (include/TreeClasses/HamiltonianRepresentation.h:62-87):
 * density tree, regularised inverse, h-matrices and hole matrices of all terms. */
int qtg_engine_build(qtg_engine engine, qtg_stream stream);
/* LayerDerivative (src/TreeClasses/HamiltonianRepresentation.cpp:45-72) at every
 * node with the current representation. */
int qtg_engine_layer_derivative_all(qtg_engine engine, double phase_re, double phase_im,
                                    qtg_stream stream);
/* Derivative (HamiltonianRepresentation.cpp:74-81) = build + LayerDerivative
 * at every node: one H|Psi> sweep. */
int qtg_engine_derivative(qtg_engine engine, double phase_re, double phase_im, qtg_stream stream);
/* which: 0 = derivative, 1 = Apply() output (before projector), 2 = Psi */
int qtg_engine_get_dpsi(qtg_engine engine, double* host, int which, qtg_stream stream);
/* Expectation / S_top (HamiltonianRepresentation.cpp:45-51, CMFIntegrator.cpp:235-241)
 * of the last derivative evaluation (orthonormal lower nodes).  With several
 * states at the top node this is the state average trace(H) / trace(S); the
 * levels of each state come from qtg_cmf_eigenstates. */
int qtg_engine_energy(qtg_engine engine, double* energy, qtg_stream stream);
/* which: 0 rho[node], 1 rho^-1[node], 2 h-matrix[term][node], 3 hole[term][node] */
int qtg_engine_get_matrix(qtg_engine engine, int which, int64_t term, int64_t node, double* out,
                          qtg_stream stream);
/* FLOPs of one H|Psi> sweep of the reference algorithm (8 per complex
 * multiply-add; one represent, hole and Apply contraction per (term, active
 * node) as TreeFunctions and HamiltonianRepresentation::Apply loop) and item
 * counts (items[0] sandwich items, items[1] apply chunks, items[2] apply terms). */
int qtg_engine_counts(qtg_engine engine, double* sweep_flops, int64_t* items);
/* Gradient of the energy functional E = <Psi|H|Psi> / <Psi|Psi> with respect
 * to conj(A) at every node (complex128, grouped path), for the state set by
 * qtg_engine_set_psi with orthonormal non-top nodes:
 *   dE/dA_v* = (Apply(v) + M_v x_last A_v - E rho_v x_last A_v) / <A_top|A_top>,
 * Apply as in HamiltonianRepresentation.cpp:6-43, M_v the mean field of the
 * terms inactive at v (which the projector of LayerDerivative removes).  No
 * reference counterpart; rho^-1 (1 - P) of it is the reference's
 * imaginary-time LayerDerivative.  grad_host (engine layout, complex128
 * interleaved) and energy may be null.  One state at the top node
 * (QTG_ESHAPE otherwise). */
int qtg_engine_gradient(qtg_engine engine, double* grad_host, double* energy, qtg_stream stream);
/* The energy functional as a forward / backward pair (complex128 engine,
 * grouped path, one state at the top; no reference counterpart).
 * qtg_engine_forward: the sweep at phase 1 on the state set by
 * qtg_engine_set_psi, leaving E and 1/<Psi|Psi> on the device; energy (may be
 * null: then the call does not wait) receives E.  qtg_engine_backward:
 * dE/dA* at every node into grad_host (engine layout; may be null).  With
 * option "grad_recompute" 0 (default) backward uses the matrices of the last
 * forward, which must have run on the same state; with 1 it recomputes them
 * from the current state first, so nothing of the forward pass has to be
 * kept in between (other sweeps may run).  qtg_engine_gradient = forward +
 * backward.  qtg_engine_canonicalize: the QR gauge on the device: every
 * non-top node orthonormalised by Householder QR, deepest layer first, its R
 * pushed into the parent on the node's leg (the represented state is
 * unchanged); psi_host (may be null) receives the state.
 * qtg_engine_gradient_descent: `steps` steps A <- A - eta dE/dA* at every
 * node, each followed by the QR gauge, all on the device; energies (steps + 1
 * doubles, may be null) receives E before every step and after the last,
 * psi_host (may be null) the final state. */
int qtg_engine_forward(qtg_engine engine, double* energy, qtg_stream stream);
int qtg_engine_backward(qtg_engine engine, double* grad_host, qtg_stream stream);
int qtg_engine_canonicalize(qtg_engine engine, double* psi_host, qtg_stream stream);
int qtg_engine_gradient_descent(qtg_engine engine, int64_t steps, double eta, double* energies, double* psi_host,
                                qtg_stream stream);
/* FLOPs one sweep of this engine's plan executes (fewer than the reference
 * count where equal slots are shared and Apply summands merged). */
int qtg_engine_plan_flops(qtg_engine engine, double* executed);
/* Host-only variants (no device is touched): plan statistics, state size and
 * the random initial state of a tree. */
int qtg_host_plan_counts(const char* tree_text, int64_t nterms, const double* coeff,
                         const int64_t* term_off, const int64_t* fmode, const int64_t* fkind,
                         const int64_t* fpool, const double* pool, int compress, double* sweep_flops,
                         int64_t* items);
int64_t qtg_host_psi_size(const char* tree_text);
int qtg_host_fill_random(const char* tree_text, uint64_t seed, int delta_lowest, double* psi_host);
/* The random state drawn from mt19937(seed) after `discard` outputs, for a
 * generator that earlier states already drew from (the mctdh runner's
 * shared rng_, mctdh.cpp:164-166,272-275); returns the number of generator
 * outputs it drew, or a negative qtg_status. */
int64_t qtg_host_fill_random_stream(const char* tree_text, uint64_t seed, int64_t discard, int delta_lowest,
                                    double* psi_host);
/* Kernel timing: while on, sweeps run eagerly behind a short device-side
 * hold (so the host queues the whole sweep first) and HIP events time every
 * launch on its stream; the grouped kernels carry the events in their own
 * dispatch (hipExtLaunchKernelGGL), the others are bracketed.  read_timing
 * returns, per kernel group [0 sandwich (represent/holes/density),
 * 1 eigensolver, 2 apply, 3 finish, 4 precision conversion, 5 matrix
 * combination], the summed milliseconds, launch counts, algorithmic FLOPs
 * and (grouped kernels) algorithmic bytes since set_timing(1); the grouped
 * path adds 6 mode products (gm_grouped), 7 hole contractions (gc_grouped),
 * 8 batched solves, 9 fused hole layers (hole_fused_kernel).  Arrays hold
 * 10 entries; any
 * pointer may be null. */
int qtg_engine_set_timing(qtg_engine engine, int on);
/* waits for the engine's own stream */
int qtg_engine_synchronize(qtg_engine engine);
int qtg_engine_read_timing(qtg_engine engine, double* ms, int64_t* launches, double* flops, double* bytes);
/* "rho_eps": regularisation of the density inverse (default 1e-10).
 * "compress" (set before qtg_engine_set_hamiltonian; default 1): compute each
 * distinct h-matrix / hole matrix once (they depend only on the term's
 * factors inside / outside the node's subtree) and merge Apply summands that
 * share all slots but one; 0 keeps one work item per (term, node) as the
 * reference loops do.
 * "graph" (default 1): Derivative() on the engine's stream is captured once
 * into a HIP graph and replayed; 0 launches eagerly.
 * "path" (set before qtg_engine_set_hamiltonian; default 1): 1 runs every
 * node contraction as grouped MFMA tile GEMMs through a device workspace (any
 * node size); 0 keeps whole node tensors in LDS (nodes up to ~150 KB).
 * "split_waves", "split_min_k" (set before qtg_engine_set_hamiltonian;
 * defaults 2048, 16): a grouped hole-contraction launch with fewer 16 x 16
 * tiles than split_waves splits its reductions along K into chunks of at
 * least split_min_k until it holds about split_waves wavefronts.
 * "gc_long_k" (before qtg_engine_set_hamiltonian; default 512): hole
 * contractions longer than 2 gc_long_k always split along K.
 * "gm_split_k" (before qtg_engine_set_hamiltonian; default 32): mode products
 * longer than 2 gm_split_k split along K in chunks of at least gm_split_k.
 * "tuning" (before qtg_engine_set_hamiltonian; default 1): the four split
 * options not set explicitly come from the tuning table's nearest class
 * (qtg_engine_tuning); 0 keeps the defaults.
 * "grad_recompute" (default 0): see qtg_engine_backward.
 * "fuse_holes" (before qtg_engine_set_hamiltonian; default 1): complex128
 * sweeps of binary trees whose hole outputs are at most 16 x 16 run each
 * hole layer as one launch (ket tensors built in LDS beside the hole
 * contractions that read them); 0 keeps two launches per layer.
 * "heev_side" (default 1): 1 runs the grouped sweep's density eigensolver on
 * a second stream, overlapping the hole contractions; 0 runs it in line.
 * Plan options set after qtg_engine_set_hamiltonian return QTG_ESTATE;
 * rho_eps must be positive and reaches a captured sweep. */
int qtg_engine_set_option(qtg_engine engine, const char* key, double value);
/* The tuning-table line (qutree_amd/tuning/gfx950.txt, written by
 * qutree_amd/autotune.py) applied to this engine's plan, as text into buf
 * (cap bytes, NUL-terminated; empty when none matched). */
int qtg_engine_tuning(qtg_engine engine, char* buf, int64_t cap);

/* ----------------------------------------------------------- CMF integrator */
/* CMFIntegrator(H, tree, phase).Integrate(ivars) (src/Applications/
 * CMFIntegrator.cpp:25-249) on a complex128 engine: constant-mean-field
 * steps with the reference's step-size control, per-node Bulirsch-Stoer
 * integration (include/Util/BS_integrator.h:31-256) of LayerDerivative with
 * the matrices of the last build, CMF error and orthogonal() after accepted
 * steps.  psi_host (complex128 interleaved, engine layout) is the initial
 * state on entry and the final state on return.  At t0 and after every
 * accepted step that reaches the next output time: times, energies
 * (h(0,0) / S(0,0) of Expectation and S_top: the first state's) and
 * autocorr (<Psi_0(t0)|Psi_0(t)>, interleaved).
 * stats (5 entries): steps, rejected steps, builds, node derivative
 * evaluations, lockstep ticks.  Returns the number of outputs, or a negative
 * qtg_status. */
typedef void* qtg_cmf;
int qtg_cmf_create(qtg_engine engine, qtg_cmf* cmf);
/* Psi.write(job.ofname, job.append_) of CMFIntegrator::Integrate
 * (CMFIntegrator.cpp:46,153): while set, qtg_cmf_integrate writes the state
 * as a TensorTree (the format of qtg_tensor_tree_write) at the start and at
 * every output time.  append != 0 appends (creating the file); append == 0
 * opens an existing file in place and overwrites it from its start without
 * truncating, and writes nothing when the file is missing
 * (TensorTree_Implementation.h:93-100, ios::in).  filename NULL: off. */
int qtg_cmf_set_output_file(qtg_cmf cmf, const char* filename, int append);
/* "graph" (default 0): 1 replays each lockstep tick (pre programs, masked
 * Apply and finish, post programs) from a HIP graph captured once per state
 * buffer and phase, uploading only the tick's programs; 0 launches every
 * tick eagerly (measured faster on MI355X).  "fuse" (default 1): the vector
 * op sequences a tick emits most (the Bulirsch-Stoer start, the modified
 * midpoint steps, the final midpoint with rational extrapolation) run as one
 * pass over the node each; 0 runs one pass per op.  "lookahead" (default
 * 1): the tick after a tick that computes node errors is uploaded and
 * launched before those errors reach the host, with both outcomes of each
 * waiting node's accept test staged (the device picks one by the same
 * comparison and sets the node's Apply mask); 0 waits for the errors before
 * staging the next tick.  "device_bs" (default 1): every node's
 * Bulirsch-Stoer state machine runs on the device, one launch per tick
 * boundary (post ops, accept test, next pre op, Apply mask), and the host
 * reads a small per-tick record only where the ticks known to have a
 * derivative run out; 0 runs the state machines on the host, which the
 * three options above then shape.  Results are the same either way for all
 * four options. */
int qtg_cmf_set_option(qtg_cmf cmf, const char* key, double value);
int qtg_cmf_destroy(qtg_cmf cmf);
int64_t qtg_cmf_integrate(qtg_cmf cmf, double* psi_host, double t0, double tend, double dt, double out,
                          double acc_cmf, double acc_bs, double phase_re, double phase_im,
                          int64_t max_steps, int64_t max_out, double* times, double* energies,
                          double* autocorr, int64_t* stats);

/* Imaginary-time eigenstates, Eigenstates(IntegratorVariables&)
 * (src/Applications/Eigenstates.cpp:60-103) on a CMF integrator: Eigenstate()
 * (build, diagonalise the top node's <Psi_i|H|Psi_j>, rotate the top node into
 * its eigenbasis; Eigenstates.cpp:18-31), then `iterations` times: CMF with
 * phase -i over [0, out - 1e-5] (the step size carries over between
 * iterations, as job.dt does), propagatorEnergies (-log(ev(S_top)) / (2 out),
 * ascending ev; Eigenstates.cpp:8-16), orthonormal() (Gram-Schmidt at every
 * node), Eigenstate(), and converged(i) = sum_j |<Psi_i|Psi_prev_j>|^2 (the s(i)
 * of Status, Eigenstates.cpp:33-45).  Eigenvector columns are phased so that
 * their first component is real and non-negative.  psi_host: initial state in,
 * final state out.  energies: (iterations + 1) x nstates ascending eigenvalues;
 * propagator, converged: iterations x nstates.  stats (4 entries): CMF steps,
 * rejected steps, CMF builds, node derivative evaluations; dt_end: the step
 * size on return.  nstates <= 64.  Returns nstates, or a negative qtg_status. */
int64_t qtg_cmf_eigenstates(qtg_cmf cmf, double* psi_host, double dt, double out, int64_t iterations, double acc_cmf,
                            double acc_bs, int64_t max_steps, double* energies, double* propagator,
                            double* converged, int64_t* stats, double* dt_end);

/* ------------------------------------------------------------ SCF sweeps */
/* Krylov-sweep SCF ground-state search, scf(SCF_parameters&)
 * (Applications/SCF.cpp:270-348).  The engine's tree and Hamiltonian
 * (complex128 engine, Hamiltonian set) are used with the correction
 * matrices of HamiltonianRepresentation(H, tree, tree, false)
 * (HamiltonianRepresentation.h:19-36,131-177, .cpp:119-218).
 * qtg_scf_create plans a sweep with Krylov spaces of n_krylov (1..64)
 * vectors, capped per node at the node tensor's size (SCF.cpp:301-302).
 * qtg_scf_run performs n_iter iterations (SCF_parameters::nIter), the first
 * n_itp of them imaginary-time updates with step beta (nITP, beta), the rest
 * replacing each node by its lowest Krylov eigenvector; psi (engine layout,
 * non-top nodes orthonormal as TensorTree(..., delta_lowest) leaves them) is
 * the start on entry and the final state on return.  energies receives, per
 * iteration and microstep, the lowest Ritz value in hartree (what the
 * reference prints, before its unit conversion), at most cap values.
 * Each Krylov ground-state eigenvector is phased with its first component
 * real and non-negative (the reference leaves this sign to its
 * eigensolver; energies do not depend on it).  Every node toward a
 * neighbour must have at least as many rows as columns (QTG_ESHAPE). */
typedef void* qtg_scf;
int qtg_scf_create(qtg_engine engine, int64_t n_krylov, qtg_scf* scf);
int qtg_scf_destroy(qtg_scf scf);
/* microsteps per iteration (2 nodes - 1) */
int64_t qtg_scf_sweep_length(qtg_scf scf);
int qtg_scf_run(qtg_scf scf, double* psi_host, int64_t n_iter, int64_t n_itp, double beta, double* energies,
                int64_t cap);
/* "graph" (1: replay iterations from a captured HIP graph, the default; 0:
 * launch eagerly); "launches" (read only via qtg_scf_info) */
int qtg_scf_set_option(qtg_scf scf, const char* key, double value);
/* info[0] microsteps per iteration, info[1] launches per iteration,
 * info[2] device bytes */
int qtg_scf_info(qtg_scf scf, int64_t* info);
/* Host-only planning of the same sweep (no device needed): checks that every
 * launch of the plan stays inside its buffers, that no two outputs of one
 * launch overlap and that no launch writes what it reads.  info (6 entries):
 * microsteps per iteration, launches per iteration, work and pool elements,
 * mode-product and hole-contraction tasks. */
int qtg_host_scf_plan(const char* tree_text, int64_t nterms, const double* coeff, const int64_t* term_off,
                      const int64_t* fmode, const int64_t* fkind, const int64_t* fpool, const double* pool,
                      int64_t n_krylov, int64_t* info);

/* ------------------------------------------- SOP application fitted into the tree */
/* applyOperatorIteration(Psi, rho, sop, tree) (Applications/ApplyFirstOrder.cpp:
 * 43-147): H|Psi> fitted into a tree of Psi's shape by alternating sweeps,
 * with fidelity |<fit|H Psi>|^2 (TreeApplyOperator.cpp:22-35).  The operator
 * is the engine's Hamiltonian (complex128 engine) with its coefficients NOT
 * applied, as the reference sums the bare products sop(i): fold a
 * coefficient into one of the product's leaf matrices.  The fit starts from
 * TensorTree(mt19937(seed), tree, false) (the reference seeds from the
 * clock); the loop stops when 1 - f < 1e-6, when f rose by less than 1e-10,
 * or after max_iter iterations (the reference's 100).  psi_host holds the
 * state on entry and the fit on return; fidelities (cap) the fidelity after
 * each iteration, *iterations their number.  A one-product operator is
 * applied to the bottom tensors directly (applyReference) with fidelity 1.
 * Supported for operators whose tail-cut sparse tree reaches the top node,
 * with one state there (QTG_ESHAPE otherwise). */
typedef void* qtg_apply_fit;
int qtg_apply_fit_create(qtg_engine engine, qtg_apply_fit* fit);
int qtg_apply_fit_destroy(qtg_apply_fit fit);
int qtg_apply_fit_run(qtg_apply_fit fit, double* psi_host, uint64_t seed, int64_t max_iter, double* fidelities,
                      int64_t cap, int64_t* iterations);
/* "graph" (1: replay iterations from a captured HIP graph, the default) */
int qtg_apply_fit_set_option(qtg_apply_fit fit, const char* key, double value);
/* info[0] launches of the first iteration's h-matrices, info[1] launches
 * per iteration, info[2] device bytes */
int qtg_apply_fit_info(qtg_apply_fit fit, int64_t* info);
/* Host-only planning with the checks of qtg_host_scf_plan; info (4
 * entries): launches per iteration, work and pool elements, GM tasks. */
int qtg_host_apply_fit_plan(const char* tree_text, int64_t nterms, const double* coeff, const int64_t* term_off,
                            const int64_t* fmode, const int64_t* fkind, const int64_t* fpool, const double* pool,
                            int64_t* info);

/* ---------------------------------------------------------------- formats */
/* QuTree's binary files, byte for byte (host only; complex<double>,
 * little-endian).  TensorTree<cd>::write/read (TensorTree_Implementation.h:
 * 79-132, Tensor_Implementation.h:204-258, TensorShape.cpp:68-111): "TTre",
 * int32 node count, then per node in address order "TENS", "TDIM", int32
 * order, int32 dims[order], int32 16, the coefficients (first index fastest).
 * qtg_tensor_tree_write: psi in engine layout (node-address order, as
 * qtg_engine_psi_size) of the tree in tree_text into buf; returns the byte
 * count (buf may be null to ask for it; a smaller cap is QTG_EARG).
 * qtg_tensor_tree_read: one TensorTree from buf, every node's order and
 * dimensions checked against the tree (QTG_ESHAPE names the first mismatch);
 * returns the bytes consumed, so that files the reference appended to
 * (TensorTree::write(file, append)) are read tree by tree.  psi is undefined
 * after an error.  Both return a negative qtg_status on error. */
int64_t qtg_tensor_tree_write(const char* tree_text, const double* psi, unsigned char* buf, int64_t cap);
int64_t qtg_tensor_tree_read(const char* tree_text, const unsigned char* buf, int64_t len, double* psi);
/* MatrixTree<cd> ("MTre", MatrixTree.cpp:52-90) and SparseMatrixTree<cd>
 * ("SMTr", SparseMatrixTree.cpp:27-70; sparse != 0): int32 count, then per
 * matrix Matrix::write (Matrix_Implementation.h:365-378): "MATR", uint64
 * dim1, uint64 dim2, int32 16, the coefficients (column-major).  dims holds
 * (dim1, dim2) per matrix, mats the matrices back to back.  write: returns
 * the byte count (buf may be null).  read: *count = number of matrices; at
 * most max_count matrices and mats_cap complex coefficients are stored;
 * returns the bytes consumed.  Negative qtg_status on error. */
int64_t qtg_matrix_tree_write(int sparse, int64_t count, const int64_t* dims, const double* mats, unsigned char* buf,
                              int64_t cap);
int64_t qtg_matrix_tree_read(int sparse, const unsigned char* buf, int64_t len, int64_t max_count, int64_t* dims,
                             double* mats, int64_t mats_cap, int64_t* count);

/* ------------------------------------------------------------------- TTNO */
/* Tree tensor network operator on an engine's tree (TensorTreeOperator,
 * include/TreeOperators/TensorOperators/TensorTreeOperator.h:23-58; complex128
 * engines).  bond[a] = bond dimension of node a (1 at the top); B = the
 * operator tensors in node-address order, complex128 interleaved: bottom
 * (dim^2, D_a) with the leaf matrices as columns (toMatrix(B, l),
 * TensorTreeOperator.cpp:40-47), upper (D_child0, ..., D_childk, D_a).
 * qtg_ttno_apply: TTOrepresentation::calculate (TTOrepresentation.h:81-103)
 * with <bra| . |ket>, TTOcontraction::calculate (TTOcontraction.h:58-112),
 * then apply() (TTOcontraction_Implementation.h:7-30) at every node to the
 * ket's node tensors into out (engine layout); expect (2 doubles) receives
 * <bra|H|ket>.  bra, ket, out are host arrays; null out/expect skip them. */
typedef void* qtg_ttno;
int qtg_ttno_create(qtg_engine engine, const int64_t* bond, const double* B, qtg_ttno* ttno);
int qtg_ttno_destroy(qtg_ttno ttno);
int qtg_ttno_apply(qtg_ttno ttno, const double* bra, const double* ket, double* out, double* expect,
                   qtg_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* QUTREE_AMD_H */
