/*
 * qsv.h — C ABI of libqsv, the MI355X state-vector engine behind QClojure's
 * backend protocol.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.
 * Every call returns QSV_OK (0) or a negative QSV_E* code; the message of the
 * last failure on the calling thread is returned by qsv_last_error().
 *
 * Conventions shared with the reference (lsolbach/qclojure):
 *   - Qubit q is the MOST significant index bit: basis index
 *     i = sum_q b_q * 2^(n-1-q)            (src/org/soulspace/qclojure/domain/state.clj:114-136)
 *   - Amplitudes cross the ABI as interleaved FP64 pairs (re, im), the
 *     layout of the reference's fastmath Vec2 state vectors
 *     (src/org/soulspace/qclojure/domain/state.clj:9-35).
 *   - Operation kinds and operand order follow the dispatcher
 *     apply-gate-to-state (src/org/soulspace/qclojure/domain/circuit.clj:933-1072).
 *
 * Each entry point cites the reference function it replaces.
 */
#ifndef QSV_H
#define QSV_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define QSV_OK              0
#define QSV_E_INVALID      -1   /* bad argument (qubit out of range, null pointer ...) */
#define QSV_E_UNSUPPORTED  -2   /* operation the reference also rejects              */
#define QSV_E_NOMEM        -3   /* device allocation / HBM budget exhausted            */
#define QSV_E_DEVICE       -4   /* HIP runtime error                                   */
#define QSV_E_STATE        -5   /* wrong object state (not initialised, job not done)  */
#define QSV_E_NOTFOUND     -6   /* unknown job id                                       */

/* ---- precision ----------------------------------------------------------- */
#define QSV_C128 0   /* complex128: FP64 re/im, 16 B per amplitude */
#define QSV_C64  1   /* complex64 : FP32 re/im,  8 B per amplitude; FP64 reductions */

/* ---- operation kinds (src/org/soulspace/qclojure/domain/circuit.clj:964-1072) ----
 * Operand slots q[] per kind:
 *   1-qubit gates ..................... q[0]=target
 *   CNOT/CZ/CY/CRX/CRY/CRZ/RYDBERG_* .. q[0]=control, q[1]=target
 *   SWAP/ISWAP ........................ q[0]=qubit1,  q[1]=qubit2
 *   TOFFOLI ........................... q[0]=control1, q[1]=control2, q[2]=target
 *   FREDKIN ........................... q[0]=control, q[1]=target1, q[2]=target2
 *   RYDBERG_BLOCKADE, MEASURE ......... mask = set of qubits (bit q = qubit q)
 *   U ................................. q[0]=target, m[] = row-major 2x2 (re,im)
 *   CU ................................ q[0]=control, q[1]=target, m[] as U
 */
enum qsv_op_kind {
  QSV_OP_X = 1, QSV_OP_Y, QSV_OP_Z, QSV_OP_H,
  QSV_OP_S, QSV_OP_SDG, QSV_OP_T, QSV_OP_TDG,
  QSV_OP_RX, QSV_OP_RY, QSV_OP_RZ, QSV_OP_PHASE,
  QSV_OP_CNOT, QSV_OP_CZ, QSV_OP_CY, QSV_OP_CRX, QSV_OP_CRY, QSV_OP_CRZ,
  QSV_OP_SWAP, QSV_OP_ISWAP, QSV_OP_TOFFOLI, QSV_OP_FREDKIN,
  QSV_OP_RYDBERG_CZ, QSV_OP_RYDBERG_CPHASE, QSV_OP_RYDBERG_BLOCKADE,
  QSV_OP_GLOBAL_RX, QSV_OP_GLOBAL_RY, QSV_OP_GLOBAL_RZ, QSV_OP_GLOBAL_H,
  QSV_OP_GLOBAL_X, QSV_OP_GLOBAL_Y, QSV_OP_GLOBAL_Z,
  QSV_OP_MEASURE,
  QSV_OP_U, QSV_OP_CU,
  QSV_OP_KIND_END
};

/* One decoded circuit operation: a fixed-size POD (96 bytes). */
typedef struct qsv_op {
  int32_t  kind;     /* enum qsv_op_kind                           */
  int32_t  q[3];     /* qubit operands, reference (MSB) numbering  */
  uint64_t mask;     /* qubit set for BLOCKADE / MEASURE           */
  double   angle;    /* rotation / phase angle in radians          */
  double   m[8];     /* explicit 2x2 matrix for U / CU             */
} qsv_op;

/* apply flags */
#define QSV_FLAG_MSB_SWAP  0x1u  /* SWAP/ISWAP on MSB bit positions. Default is the
                                    reference's LSB positions (gate.clj:768-778,823-833). */
#define QSV_FLAG_FUSE      0x2u  /* batch runs of gates into LDS-tiled passes */
#define QSV_FLAG_FP64_PASSES 0x4u /* complex64 tiled passes compute in FP64. By default
                                    they compute in FP32 and remove the resulting norm
                                    drift pass by pass (FP64 per-tile norm sums). */
#define QSV_FLAG_VALU_PASSES 0x8u /* FP32 complex64 passes keep every gate on the vector
                                    units. By default runs of register-bit gates in a
                                    stage become dense 16x16 blocks on FP32 MFMA. */

/* ---- library / device --------------------------------------------------- */
typedef struct qsv_device_info {
  char     name[64];
  char     arch[32];        /* e.g. "gfx950"                         */
  int32_t  compute_units;
  int32_t  wavefront_size;
  uint64_t hbm_bytes;       /* total device memory                    */
  uint64_t hbm_free_bytes;
} qsv_device_info;

/* Replaces the implicit JVM heap of LocalQuantumSimulator
 * (src/org/soulspace/qclojure/adapter/backend/ideal_simulator.clj:181-195). */
int         qsv_init(int device);
int         qsv_shutdown(void);
int         qsv_device_query(qsv_device_info* out);
const char* qsv_last_error(void);
const char* qsv_version(void);

/* ---- state vectors ------------------------------------------------------ */
typedef struct qsv_state qsv_state;

/* zero-state / computational-basis-state (state.clj:251-284, 484-519). */
int qsv_state_create(int num_qubits, int precision, qsv_state** out);
int qsv_state_destroy(qsv_state* s);
int qsv_state_num_qubits(const qsv_state* s);
int qsv_state_precision(const qsv_state* s);
int qsv_state_init_basis(qsv_state* s, uint64_t index);
/* :initial-state option (ideal_simulator.clj:85-87): interleaved FP64 (re,im). */
int qsv_state_upload(qsv_state* s, const double* amps, uint64_t offset, uint64_t count);
/* :final-state materialisation (circuit.clj:1783): interleaved FP64 (re,im). */
int qsv_state_download(const qsv_state* s, double* amps, uint64_t offset, uint64_t count);
int qsv_state_copy(qsv_state* dst, const qsv_state* src);
int qsv_state_sync(qsv_state* s);

/* ---- gate application --------------------------------------------------- */
/* reduce apply-operation-to-state over the ops (circuit.clj:1782, 1094-1111);
 * seed drives QSV_OP_MEASURE collapses (state.clj:915-1014). */
int qsv_apply_program(qsv_state* s, const qsv_op* ops, uint64_t n_ops,
                      uint32_t flags, uint64_t seed);

/* ---- result extraction -------------------------------------------------- */
/* cla/norm2 squared, FP64 accumulation (complex_linear_algebra.clj:435-446). */
int qsv_norm2(const qsv_state* s, double* out);
/* normalize-state (state.clj:524-551). */
int qsv_normalize(qsv_state* s);
/* measurement-probabilities / probabilities (state.clj:658-682, 1071-1092), FP64. */
int qsv_probabilities(const qsv_state* s, uint64_t offset, uint64_t count, double* out);
/* measure-state repeated `shots` times (result.clj:224, state.clj:869-913):
 * outcome k uses uniform u_k = qsv_uniform(seed, k). */
int qsv_sample(const qsv_state* s, uint64_t shots, uint64_t seed, uint64_t* outcomes);
/* qsv_sample without any device allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc). The sampler's
 * chunk prefix and the device copy of the outcomes live in a fixed per-GPU
 * scratch array of 12 MiB, so the state may hold at most
 * QSV_NOALLOC_MAX_QUBITS qubits and at most QSV_NOALLOC_MAX_SHOTS shots are
 * drawn; larger requests return QSV_E_INVALID. Outcomes equal qsv_sample's
 * for the same state, shots and seed. Calls that overlap on one GPU share
 * the scratch and race: the caller serialises them. */
#define QSV_NOALLOC_MAX_QUBITS 30
#define QSV_NOALLOC_MAX_SHOTS  (1ull << 19)
int qsv_sample_noalloc(const qsv_state* s, uint64_t shots, uint64_t seed, uint64_t* outcomes);
/* measure-specific-qubits (state.clj:915-1014): collapse `mask` using uniform u.
 * outcome bit j = value of the j-th measured qubit in ascending qubit order;
 * probs (optional, 2^popcount(mask) doubles) receives every outcome probability. */
int qsv_measure_qubits(qsv_state* s, uint64_t mask, double u,
                       uint64_t* outcome, double* probs);
/* pauli-string-expectation per term (hamiltonian.clj:82-94, observables.clj:216-251):
 * `strings` holds n_terms strings of num_qubits chars from "IXYZ", packed back to back;
 * out[k] = Re <psi|P_k|psi> in FP64. The caller sums c_k*out[k] in term order. */
int qsv_expect_pauli(const qsv_state* s, const char* strings, uint64_t n_terms, double* out);
/* expectation-value of a 1-qubit observable embedded as I..O..I
 * (result.clj:254-288, observables.clj:234-251): out = {re, im} of <psi|O_q|psi>. */
int qsv_expect_1q(const qsv_state* s, int qubit, const double m[8], double out[2]);
/* inner product <a|b> with FP64 accumulation (state.clj:1164-1185 takes |.|). */
int qsv_inner(const qsv_state* a, const qsv_state* b, double out[2]);
/* One pass over two states of the same size and any precisions (precision
 * studies): out = {Re <a|b>, Im <a|b>, sum |a_i|^2, sum |b_i|^2, max |a_i - b_i|}.
 * state-fidelity (state.clj:1164-1185) is |<a|b>|; |<a|b>|^2 is the overlap fidelity. */
int qsv_state_compare(const qsv_state* a, const qsv_state* b, double out[5]);

/* Counter-based uniform in [0,1) used by qsv_sample / MEASURE: splitmix64 of
 * (seed + (k+1) * 0x9E3779B97F4A7C15), top 53 bits. Exposed so callers can
 * reproduce shot streams. */
double qsv_uniform(uint64_t seed, uint64_t k);

/* ---- batched circuits (BatchJobBackend, application/backend.clj:136-165) --
 * Many independent circuits, each on its own state, applied in one persistent
 * launch. Instance i runs ops[op_offsets[i] .. op_offsets[i+1]) on a
 * num_qubits[i]-qubit |0..0> state; n_terms_i Pauli strings (Z-only) are
 * evaluated on the final state into energies[i] = sum_k coeff_k <P_k> (FP64,
 * term order). */
typedef struct qsv_batch_desc {
  uint64_t        n_instances;
  const int32_t*  num_qubits;     /* [n_instances]                                   */
  const uint64_t* op_offsets;     /* [n_instances+1] into ops                        */
  const qsv_op*   ops;
  const uint64_t* term_offsets;   /* [n_instances+1] into term_zmask / term_coeff    */
  const uint64_t* term_zmask;     /* Z-string as a qubit mask (bit q = Z on qubit q) */
  const double*   term_coeff;
  int32_t         precision;      /* QSV_C64 or QSV_C128                             */
} qsv_batch_desc;

int qsv_batch_expect_z(const qsv_batch_desc* desc, double* energies);

/* ---- noisy trajectories (the hardware simulator's shots) -------------------
 * execute-single-noisy-shot-with-trajectory repeated n_traj times
 * (src/org/soulspace/qclojure/adapter/backend/hardware_simulator.clj:80-104),
 * each trajectory on its own |0..0> state: every op, then the noise channel of
 * each site whose after_op is that op (noise.clj:53-103). A site's operator j
 * is chosen as apply-quantum-channel does (channel.clj:185-222): weights
 * w_j = max |k_j,rc|^2, the first j whose running sum exceeds the uniform u,
 * else the last one (a single operator always). It is applied to `qubit` and
 * the state renormalised (channel.clj:135-183). The measurement is one
 * measure-state draw (state.clj:869-913) -> outcomes[t] (basis index).
 * Uniforms: site k of trajectory t uses qsv_uniform(seed, t*QSV_TRAJ_SLOTS + k),
 * its measurement qsv_uniform(seed, t*QSV_TRAJ_SLOTS + QSV_TRAJ_SLOTS - 1).
 * final_states (optional) receives the first n_keep trajectories' normalised
 * final states, 2^n interleaved FP64 (re, im) each. Sites must be ordered by
 * after_op; mid-circuit MEASURE ops are rejected. */
#define QSV_MAX_KRAUS   4
#define QSV_TRAJ_SLOTS  (1ull << 20)
typedef struct qsv_kraus_site {
  uint64_t after_op;               /* index into ops                          */
  int32_t  qubit;                  /* reference qubit number                  */
  int32_t  n_kraus;                /* 1..QSV_MAX_KRAUS                        */
  double   m[QSV_MAX_KRAUS][8];    /* row-major 2x2 complex (re, im) each     */
} qsv_kraus_site;
int qsv_noisy_trajectories(int num_qubits, int precision, const qsv_op* ops, uint64_t n_ops,
                           const qsv_kraus_site* sites, uint64_t n_sites, uint64_t n_traj, uint64_t seed,
                           uint64_t* outcomes, uint64_t n_keep, double* final_states);

/* ---- serving: continuous batching of circuit requests ---------------------
 * The asynchronous job model of QuantumBackend (submit-circuit / job-status /
 * job-result, application/backend.clj:72-112; the reference runs each job in
 * its own future, ideal_simulator.clj:120-136) as a native request loop.
 * A scheduler thread admits queued requests in arrival order while the sum of
 * their state bytes (2^n * 16 or 8) fits the HBM budget, keeps admitted states
 * on the device, gathers every admitted request small enough for one
 * workgroup's LDS into a single batched launch (gates + shots), runs larger
 * ones concurrently on their own streams through fused passes, and records
 * submit / admit / done timestamps inside the library. */
typedef struct qsv_server qsv_server;

typedef struct qsv_server_config {
  uint64_t hbm_budget_bytes;  /* 0: 80% of the free device memory                 */
  int32_t  precision;         /* QSV_C64 or QSV_C128                               */
  int32_t  max_inflight;      /* admitted requests on the device at once (0: 32)   */
  uint32_t flags;             /* QSV_FLAG_* applied to every request               */
  int32_t  batch_small;       /* 1: batch LDS-sized requests into one launch       */
} qsv_server_config;

#define QSV_REQ_QUEUED    0
#define QSV_REQ_RUNNING   1
#define QSV_REQ_COMPLETED 2
#define QSV_REQ_FAILED    3
#define QSV_REQ_CANCELLED 4

typedef struct qsv_request_info {
  int32_t  status;            /* QSV_REQ_*                                          */
  int32_t  num_qubits;
  uint64_t shots;
  double   queue_ms;          /* submit -> admitted                                 */
  double   latency_ms;        /* submit -> results on the host                      */
  char     error[128];
  double   norm2;             /* completed: FP64 sum |a|^2 of the final state (the
                                 sampler's total); |norm2 - 1| is the request's
                                 norm drift. NaN otherwise.                          */
} qsv_request_info;

typedef struct qsv_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t queued_now, inflight_now, max_queue_depth;
  uint64_t small_batches, small_batched_requests;
  uint64_t hbm_budget_bytes, hbm_in_use_bytes;
  uint64_t cancelled;
} qsv_server_stats;

int qsv_server_create(const qsv_server_config* cfg, qsv_server** out);
/* Validates and decodes the ops (errors return immediately) and queues the request. */
int qsv_server_submit(qsv_server* srv, int num_qubits, const qsv_op* ops, uint64_t n_ops,
                      uint64_t shots, uint64_t seed, uint64_t* request_id);
/* Non-blocking status (QSV_E_NOTFOUND for unknown or already collected ids). */
int qsv_server_status(qsv_server* srv, uint64_t request_id, qsv_request_info* info);
/* Waits up to timeout_s for completion; on completion copies `shots` outcomes
 * (if outcomes != NULL) and forgets the request. Returns QSV_E_STATE on timeout. */
int qsv_server_wait(qsv_server* srv, uint64_t request_id, double timeout_s,
                    qsv_request_info* info, uint64_t* outcomes);
/* cancel-job (ideal_simulator.clj:154-163): a queued request leaves the queue
 * at once; an admitted one, whose launches are already on the device, runs to
 * its end and then finishes as QSV_REQ_CANCELLED without outcomes. *outcome =
 * QSV_CANCEL_DONE (:cancelled) or QSV_CANCEL_CANNOT (already completed or
 * failed: :cannot-cancel); unknown ids return QSV_E_NOTFOUND (:not-found).
 * qsv_server_wait on a cancelled request returns at once with its status. */
#define QSV_CANCEL_DONE   0
#define QSV_CANCEL_CANNOT 1
int qsv_server_cancel(qsv_server* srv, uint64_t request_id, int* outcome);
int qsv_server_stats_read(qsv_server* srv, qsv_server_stats* out);
/* Drains in-flight work, then frees every state and the scheduler thread. */
int qsv_server_destroy(qsv_server* srv);

/* ---- per-kernel timing (measurement support) ----------------------------
 * When enabled, qsv_apply_program brackets every kernel it launches with HIP
 * events on the state's own stream and accumulates (count, total ms,
 * algorithmic bytes) per kernel class. */
#define QSV_KCLASS_PAIR    0   /* dense 2x2 on a pair (optionally controlled) */
#define QSV_KCLASS_DIAG    1   /* diagonal phase                              */
#define QSV_KCLASS_PERM    2   /* swap-type permutation                       */
#define QSV_KCLASS_BLOCK   3   /* fused LDS-tiled pass                        */
#define QSV_KCLASS_OTHER   4
#define QSV_KCLASS_COUNT   5
typedef struct qsv_kernel_stat {
  uint64_t launches;
  double   total_ms;
  double   bytes;          /* algorithmic bytes over all launches */
} qsv_kernel_stat;
int qsv_timing_enable(qsv_state* s, int enable);
int qsv_timing_read(qsv_state* s, qsv_kernel_stat out[QSV_KCLASS_COUNT], int reset);

/* ---- fused-pass tuner ----------------------------------------------------
 * No counterpart in the reference (its simulator has no tunables). Times a
 * fused program of eight random layers (H, RX, RY or RZ on every qubit, then
 * CNOTs on a random matching; device time of its kernels, host planning
 * excluded) on a scratch state of `num_qubits` (16..30; 0 = 26) for every
 * candidate width of the contiguous index run a tile keeps (low_bits), and
 * makes the narrowest width within 1 % of the fastest the default for later
 * fused programs of that precision in this process. QSV_TILE_LOWBITS in the environment still overrides it. */
#define QSV_TUNE_MAX 8
typedef struct qsv_tune_result {
  int32_t  low_bits;               /* the chosen width                      */
  int32_t  n_candidates;
  int32_t  candidate[QSV_TUNE_MAX]; /* widths tried                         */
  double   ms[QSV_TUNE_MAX];        /* best time per layer for each         */
} qsv_tune_result;
int qsv_tune_fused(int precision, int num_qubits, qsv_tune_result* out);

/* ---- complex linear algebra on the device: the :hip math backend -----------
 * The MatrixAlgebra protocol and eigen-hermitian of
 * src/org/soulspace/qclojure/domain/math/protocols.clj:81-375, as implemented
 * by the reference's backends (math/fastmath/complex_linear_algebra.clj,
 * math/clojure/complex_linear_algebra.clj). A qsv_mat is a rows x cols
 * complex128 matrix in HBM, row-major, (re, im) interleaved; vectors are
 * n x 1. Results go to caller-created matrices of the right shape; calls are
 * synchronous and serialised on one library stream. */
typedef struct qsv_mat qsv_mat;
/* matrix->backend / vector->backend (protocols.clj:18-56): a zero-filled matrix,
 * then its contents from host memory; backend->matrix: the contents back. */
int qsv_mat_create(int64_t rows, int64_t cols, qsv_mat** out);
int qsv_mat_destroy(qsv_mat* m);
int qsv_mat_upload(qsv_mat* m, const double* host);        /* rows*cols (re, im) pairs */
int qsv_mat_download(const qsv_mat* m, double* host);
int qsv_mat_copy(const qsv_mat* src, qsv_mat* dst);        /* same element count */
/* shape (protocols.clj:106-114). */
int qsv_mat_shape(const qsv_mat* m, int64_t* rows, int64_t* cols);
/* A complex128 state's amplitudes as a 2^n x 1 matrix (no copy; the view must
 * not outlive the state). Every call that takes the view first waits for the
 * work queued on the state. No counterpart in the reference: it lets results
 * of the engine enter the backend without a host round trip. */
int qsv_mat_view_state(qsv_state* s, qsv_mat** out);
/* add, subtract, scale, negate (protocols.clj:116-160): C = alpha A + beta B
 * (B may be NULL: C = alpha A). */
int qsv_la_axpby(double alpha_re, double alpha_im, const qsv_mat* A,
                 double beta_re, double beta_im, const qsv_mat* B, qsv_mat* C);
/* hadamard-product (protocols.clj:213-222). */
int qsv_la_hadamard(const qsv_mat* A, const qsv_mat* B, qsv_mat* C);
/* matrix-multiply (protocols.clj:163-172): C = op(A) B, op(A) = A or, with
 * conj_transpose_a, A^H (used by unitary?, protocols.clj:328-337). */
int qsv_la_gemm(int conj_transpose_a, const qsv_mat* A, const qsv_mat* B, qsv_mat* C);
/* matrix-vector-product (protocols.clj:174-183). */
int qsv_la_gemv(const qsv_mat* A, const qsv_mat* x, qsv_mat* y);
/* inner-product (protocols.clj:185-197): sum conj(x_i) y_i with conj_x = 1. */
int qsv_la_dot(int conj_x, const qsv_mat* x, const qsv_mat* y, double out[2]);
/* norm2 (protocols.clj:273-282): sqrt(sum |x_i|^2). */
int qsv_la_norm2(const qsv_mat* x, double* out);
/* outer-product (protocols.clj:199-211): C_ij = x_i conj(y_j). */
int qsv_la_outer(const qsv_mat* x, const qsv_mat* y, qsv_mat* C);
/* kronecker-product (protocols.clj:224-237). */
int qsv_la_kron(const qsv_mat* A, const qsv_mat* B, qsv_mat* C);
/* transpose / conjugate-transpose (protocols.clj:239-261). */
int qsv_la_transpose(int conj, const qsv_mat* A, qsv_mat* C);
/* trace (protocols.clj:263-271). */
int qsv_la_trace(const qsv_mat* A, double out[2]);
/* hermitian?, diagonal?, unitary? (protocols.clj:306-337): the largest
 * element-wise deviation, which the caller compares with its tolerance:
 * max |A_ij - conj(A_ji)|, max_{i != j} |A_ij|, max |A_ij - delta_ij|. */
#define QSV_LA_DEV_HERMITIAN 0
#define QSV_LA_DEV_OFFDIAG   1
#define QSV_LA_DEV_IDENTITY  2
int qsv_la_deviation(int kind, const qsv_mat* A, double* out);
/* solve-linear-system / inverse (protocols.clj:284-304): Gauss-Jordan with
 * partial pivoting on [A | B] (B = NULL: the identity, X = A^-1); a pivot of
 * magnitude below 1e-14 fails with QSV_E_INVALID ("Matrix is singular"). */
int qsv_la_solve(const qsv_mat* A, const qsv_mat* B, qsv_mat* X);
/* eigen-hermitian (protocols.clj:363-373): eigenvalues (n doubles, in the
 * order of V's columns; the caller sorts) and eigenvectors as the columns of
 * V, by parallel cyclic Jacobi; sweeps (may be NULL) reports the sweeps used. */
int qsv_la_eigh(const qsv_mat* A, double* eigenvalues, qsv_mat* V, int* sweeps);
/* eigen-general (protocols.clj:375-387, fastmath/complex_linear_algebra.clj:792)
 * for a square complex A with n <= 64: eigenvalues as n (re, im) pairs (2n
 * doubles, in the order of V's columns; the caller sorts) and unit
 * eigenvectors as the columns of V, by Householder reduction to Hessenberg
 * form and the shifted QR algorithm in one workgroup. Larger n fails with
 * QSV_E_INVALID, a QR iteration that does not converge with QSV_E_DEVICE. */
int qsv_la_eig(const qsv_mat* A, double* eigenvalues, qsv_mat* V);
/* matrix-vector-product of a Kronecker product kept as its factors: x <-
 * (F_0 kron ... kron F_{n-1}) x, each factor a row-major 2x2 (re, im) or NULL
 * for the identity, factor k acting on reference qubit k. The shape the
 * reference builds in expand-gate-to-n-qubits (src/org/soulspace/qclojure/domain/gate.clj:323-353),
 * applied in O(2^n) per factor instead of through the 2^n x 2^n matrix. */
int qsv_la_kron_apply(const double* const* factors, int nfactors, qsv_mat* x);
/* cholesky-decomposition (src/org/soulspace/qclojure/domain/math/protocols.clj:431;
 * fastmath/complex_linear_algebra.clj:1104-1166): L lower triangular with
 * A = L L^H, from A's lower triangle; the caller checks hermitian? first.
 * QSV_E_INVALID "Hermitian matrix diagonal must be real" when a diagonal
 * residual has |imag| > eps, "Matrix is not positive definite" when its real
 * part is below eps. */
int qsv_la_cholesky(const qsv_mat* A, double eps, qsv_mat* L);
/* lu-decomposition (protocols.clj:403; fastmath/complex_linear_algebra.clj:965-1046):
 * P A = L U with partial pivoting (the first row of largest |a_ik|), L unit
 * lower, U upper; perm[k] = the row of A that is row k of P A (n entries).
 * QSV_E_INVALID "Matrix is singular" for a pivot below 1e-11. */
int qsv_la_lu(const qsv_mat* A, qsv_mat* L, qsv_mat* U, int64_t* perm);
/* qr-decomposition (protocols.clj:416; fastmath/complex_linear_algebra.clj:1048-1102):
 * modified Gram-Schmidt, A (m x n) = Q (m x n) R (n x n), R upper triangular
 * with real diagonal ||v_j||; a column whose residual norm is below 1e-12 gets
 * the unit vector e_j in Q (zero when j >= m). */
int qsv_la_qr(const qsv_mat* A, qsv_mat* Q, qsv_mat* R);

#ifdef __cplusplus
}
#endif
#endif /* QSV_H */
