/*
 * bqa_hip.h -- C ABI of libbqa_hip.so, the MI355X (gfx950) backend for
 * belief-propagation tensor-network quantum-annealing emulation.
 *
 * One job = one annealing instance resident on one GPU: degree-class node
 * tensors, two message buffers, Vidal-gauge lambdas and the simple-update
 * workspaces, all in device memory owned by the library.  The host drives
 * the reference's state machine (src/bqa/state.py) through phase entry points;
 * each phase replaces a group of the reference's batched backend calls
 * (src/bqa/backends.py) and runs as hand-written HIP kernels.
 *
 * Conventions
 *   - every function returns 0 on success or a BQA_ERR_* code; the message of
 *     the last failure on the calling thread is bqa_last_error();
 *   - arrays are plain host pointers; complex numbers are interleaved
 *     (re, im) doubles; layout arrays are int32 laid out [bond][node];
 *   - message position p is the reference's position (config_canonicalization.py:189-205):
 *     forward edges 0..E-1 in input order, reversed edges E..2E-1;
 *   - a job is not thread-safe; distinct jobs may run on distinct threads.
 */
#ifndef BQA_HIP_H
#define BQA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BQA_OK 0
#define BQA_ERR_INVALID 1      /* bad argument or shape   -> ValueError   */
#define BQA_ERR_MEMORY 2       /* device allocation failed -> MemoryError  */
#define BQA_ERR_HIP 3          /* HIP runtime error        -> RuntimeError */
#define BQA_ERR_NUMERIC 4      /* BP produced no finite distance (state.py:122) */

#define BQA_PREC_SINGLE 0      /* complex64 tensors and messages  */
#define BQA_PREC_DOUBLE 1      /* complex128 tensors and messages */

typedef struct bqa_job bqa_job;

/* Thread-local text of the last error (never NULL). */
const char* bqa_last_error(void);

/* Number of visible gfx950 devices (0 when none); *arch_ok=1 when device 0 is gfx950. */
int bqa_device_info(int* device_count, int* arch_ok, char* name_out, int name_len);

/*
 * Create a job from a compiled degree-grouped layout
 * (replaces the backend.make_from_list uploads of
 *  src/bqa/config/config_canonicalization.py:109-126).
 *
 *   nodes_number, edges_number  -- qubits n and directed messages 2E
 *   classes                     -- number of degree classes
 *   class_degree[c], class_size[c]
 *   node_ids      -- concat over classes of (n_c) ids
 *   input_pos, output_pos, lmbd_pos -- concat over classes of (d_c * n_c), [bond][node]
 *   node_ampls    -- concat over classes of (n_c) fields h
 *   edge_ampls    -- concat over classes of (d_c * n_c) couplings J, [bond][node]
 *   fwd_lhs, fwd_rhs, fwd_j -- the E forward edges (energy readout)
 *   max_bond_dim  -- chi_max (device buffers are sized for it)
 *   precision     -- BQA_PREC_SINGLE or BQA_PREC_DOUBLE
 */
int bqa_job_create(bqa_job** job, int32_t nodes_number, int32_t edges_number, int32_t classes,
                   const int32_t* class_degree, const int32_t* class_size,
                   const int32_t* node_ids, const int32_t* input_pos, const int32_t* output_pos,
                   const int32_t* lmbd_pos, const double* node_ampls, const double* edge_ampls,
                   const int32_t* fwd_lhs, const int32_t* fwd_rhs, const double* fwd_j,
                   int32_t max_bond_dim, int32_t precision);

int bqa_job_destroy(bqa_job* job);

/* Device bytes held by the job (admission accounting). */
int bqa_job_device_bytes(const bqa_job* job, uint64_t* bytes);

/* Bytes a job of this shape would hold, without creating it. */
int bqa_job_estimate_bytes(int32_t edges_number, int32_t classes, const int32_t* class_degree,
                           const int32_t* class_size, int32_t max_bond_dim, int32_t precision,
                           uint64_t* bytes);

/* |-> on every qubit, bond dimension 1, lambda = 1, messages [[1]]
 * (src/bqa/state.py:36-74). */
int bqa_init_state(bqa_job* job);

/*
 * Jacobi BP to convergence (src/bqa/state.py:97-124): dist = max|new-old| /
 * max|new+old| over all messages; stop before applying the update when
 * dist < bp_eps; otherwise msgs <- damping*msgs + (1-damping)*new; after
 * max_iters sweeps the messages are the last undamped iterate.
 * Outputs: *sweeps (sweeps run), *converged (0/1), *last_dist.
 * Returns BQA_ERR_NUMERIC when no sweep gave a finite distance.
 */
int bqa_run_bp(bqa_job* job, int32_t max_iters, double bp_eps, double damping,
               int32_t* sweeps, int32_t* converged, double* last_dist);

/*
 * ZZ layer with truncation (src/bqa/state.py:230-247): extended messages,
 * FP64 canonicalisers (SVD via one-sided Jacobi), the global bond dimension
 * chi' = min(chi_max, #{k: max_e lambda_e[k] >= pinv_eps}) and absorption of
 * the canonicalisers into the node tensors.  No gates, no gauge.
 * Outputs: *new_bond_dim, *trunc_err (norm of the discarded max-lambdas).
 */
int bqa_simple_update(bqa_job* job, double ztime, double pinv_eps,
                      int32_t* new_bond_dim, double* trunc_err);

/*
 * Vidal gauge (src/bqa/state.py:204-216): lambdas and bond canonicalisers of
 * the current chi x chi messages (FP64, the canonicaliser of bqa_simple_update,
 * state.py:171-200), every tensor bond contracted with the canonicaliser at its
 * input message position (backends.py:410-414).  The bond dimension and the
 * messages stay as they are; bqa_get_lambdas returns the (E, chi) lambdas.
 */
int bqa_set_vidal_gauge(bqa_job* job, double pinv_eps);

/*
 * Truncation of a Vidal-gauge state (src/bqa/state.py:160-168): the global
 * bond dimension chi' = min(chi_max, #{k: max_e lambda_e[k] >= pinv_eps})
 * (backends.py:297-303), every tensor bond and the lambdas cut to chi'.  The
 * reference leaves its messages at the old size; here they keep their leading
 * chi' x chi' block.  Outputs: *new_bond_dim, *trunc_err.
 */
int bqa_truncate_vidal_gauge(bqa_job* job, double pinv_eps, int32_t* new_bond_dim, double* trunc_err);

/* Local Z layer exp(-i ztime h Z) then X layer exp(-i xtime X) on every
 * qubit (src/bqa/state.py:142-156).  Pass 0 to skip either. */
int bqa_apply_local_gates(bqa_job* job, double ztime, double xtime);

/* Symmetric gauge (src/bqa/state.py:219-227): T <- T prod sqrt(lambda),
 * L2-normalised per node; messages <- diag(lambda) / sum(lambda). */
int bqa_set_symmetric_gauge(bqa_job* job);

/*
 * One Trotter step, fused (src/bqa/state.py:315-321): simple update, Z and X
 * layers, symmetric gauge and BP in one pass of kernels; outputs as above.
 */
int bqa_run_layer(bqa_job* job, double xtime, double ztime, double pinv_eps,
                  int32_t max_iters, double bp_eps, double damping,
                  int32_t* new_bond_dim, double* trunc_err,
                  int32_t* sweeps, int32_t* converged, double* last_dist);

/*
 * Per-qubit reduced density matrices and Bloch vectors
 * (src/bqa/state.py:77-94, backends.py:440-448, utils.py:23-27).
 *   rho_out   -- n * 2 * 2 complex (n*8 doubles), indexed by qubit id; may be NULL
 *   bloch_out -- n * 3 doubles (x, y, z); may be NULL
 */
int bqa_density(bqa_job* job, double* rho_out, double* bloch_out);

/* FP64 Ising energy of the rounded spins s_i = +1 iff <Z_i> > 0 of the last
 * bqa_density call: sum_i h_i s_i + sum_(ij) J_ij s_i s_j; spins_out (n int8) may be NULL. */
int bqa_energy(bqa_job* job, double* energy, int8_t* spins_out);

/*
 * Projective measurement of qubits (replaces Tensor.measure,
 * src/bqa/backends.py:309-310,729-734, as called by state.py:281-285):
 * for each k, the physical component 1 - outcomes[k] (outcome 0 = |0>,
 * 1 = |1>) of qubit qubit_ids[k]'s tensor is set to zero.  The reference also
 * divides the whole degree class by its Frobenius norm; every quantity read
 * afterwards (trace-normalised messages, densities, the per-node gauge
 * normalisation) is invariant to that scale, so it is not applied here.
 * The sampling loop itself (state.py:250-312) runs on the host.
 */
int bqa_project(bqa_job* job, int32_t count, const int32_t* qubit_ids, const int32_t* outcomes);

/*
 * Allocation-free variant of bqa_energy.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The rounded-spin
 * energy sum_i h_i s_i + sum_(ij) J_ij s_i s_j, s_i = +1 iff bloch[3 i + 2] > 0,
 * of caller-owned device arrays: bloch_dev (3 n doubles), fields_dev (n),
 * lhs_dev / rhs_dev / j_dev (the E forward edges); the FP64 result is written
 * to the device double *energy_dev.  Kernels run on `stream` (a hipStream_t;
 * NULL for the null stream) without a host synchronisation.  The block
 * partial sums go to a 4 KB static device scratch array shared by every call
 * on a GPU: calls that overlap on one GPU race.  n and E are at most 2^30.
 * Same kernels and summation order as bqa_energy, so the results are equal.
 */
int bqa_energy_noalloc(const double* bloch_dev, int32_t nodes_number, const double* fields_dev,
                       int32_t edges_number, const int32_t* lhs_dev, const int32_t* rhs_dev,
                       const double* j_dev, double* energy_dev, void* stream);

/* Current bond dimension. */
int bqa_bond_dim(const bqa_job* job, int32_t* chi);

/* State transfer (testing and checkpoint): complex128 host arrays.
 * tensors of class c: (n_c, 2, chi^d_c); messages: (2E, chi, chi); lambdas: (E, chi). */
int bqa_get_tensors(bqa_job* job, int32_t class_index, double* out);
int bqa_set_tensors(bqa_job* job, int32_t class_index, int32_t chi, const double* in);
int bqa_get_messages(bqa_job* job, double* out);
int bqa_set_messages(bqa_job* job, int32_t chi, const double* in);
int bqa_get_lambdas(bqa_job* job, double* out);
int bqa_set_lambdas(bqa_job* job, int32_t chi, const double* in);

/*
 * Exact state-vector validator on the GPU (replaces src/bqa/exact_sim.py:14-88,
 * whose simulator comes from the qem package): the same Trotterised circuit on
 * the full 2^n amplitude vector in FP64, for n <= 32 qubits (24 GB of device
 * memory at 30 qubits).  Qubit q is bit q of the basis index.
 *   bqa_sv_create    -- fields h (n), the E undirected edges (lhs, rhs, J); E <= 2048
 *   bqa_sv_init      -- |->^n (exact_sim.py:45-49)
 *   bqa_sv_run_layer -- exp(-i ztime h_q Z_q) and exp(-i ztime J_ab Z_a Z_b) on every
 *                       qubit and edge, then exp(-i xtime X_q) on every qubit (:51-63)
 *   bqa_sv_density   -- per-qubit reduced density matrices (n*8 doubles) and Bloch
 *                       vectors (n*3 doubles) as in bqa_density; either may be NULL (:69-70)
 *   bqa_sv_get_state -- the amplitudes (2^n complex, n <= 26), for testing
 */
typedef struct bqa_sv bqa_sv;
int bqa_sv_create(bqa_sv** sv, int32_t qubits, const double* fields, int32_t edges_number,
                  const int32_t* lhs, const int32_t* rhs, const double* couplings);
int bqa_sv_destroy(bqa_sv* sv);
int bqa_sv_init(bqa_sv* sv);
int bqa_sv_run_layer(bqa_sv* sv, double xtime, double ztime);
int bqa_sv_density(bqa_sv* sv, double* rho_out, double* bloch_out);
int bqa_sv_get_state(bqa_sv* sv, double* out);

/* Kernel timing of the BP sweep kernels, recorded with HIP events on the job's
 * stream while enabled: launches, total milliseconds, algorithmic bytes moved. */
int bqa_timing_enable(bqa_job* job, int32_t on);
int bqa_timing_read(bqa_job* job, int64_t* launches, double* total_ms, double* bytes);

/* Serving batch (no counterpart in the reference, which runs one job per process):
 * one Trotter step of several jobs, each as bqa_run_layer would run it.  Each job's
 * simple update runs on the job's own stream (extended messages and canonicalisers
 * of all jobs before the truncation rules' readbacks, then the local steps); BP runs
 * on the batch's stream once every job's local step is done.  Each BP sweep is one
 * launch per (precision, degree, chi) shape over the classes of every job that has
 * it (k_small / k_mfma8 shapes; a job's other classes keep their own launches), with
 * each job's convergence and iteration rules applied to its own status after every
 * chunk of sweeps.  A job's results equal those of bqa_run_layer on it alone.  The
 * jobs must be distinct, on the batch's device, with kernel timing off, and not in
 * use by another thread during the call; per-job arrays have `count` entries.
 * job_status[k] is BQA_ERR_NUMERIC where job k's BP gave no finite distance (the
 * error bqa_run_layer returns for it; its state has advanced as there) and 0
 * otherwise; any other failure fails the whole call. */
typedef struct bqa_batch bqa_batch;
int bqa_batch_create(bqa_batch** out, int32_t max_jobs);
int bqa_batch_destroy(bqa_batch* batch);
int bqa_batch_run_layer(bqa_batch* batch, int32_t count, bqa_job* const* jobs, const double* xtime,
                        const double* ztime, const double* pinv_eps, const int32_t* max_iters, const double* bp_eps,
                        const double* damping, int32_t* new_bond_dim, double* trunc_err, int32_t* sweeps,
                        int32_t* converged, double* last_dist, int32_t* job_status);

/* BP launch autotuner (no counterpart in the reference, which has one code path
 * per operation).  Each (device, precision, degree, chi, ceil(log2 n)) shape is
 * timed once per process, on the first sweep of a BP run that meets it, over its
 * launch variants (wave splits of a node and grid policies); later launches use
 * the fastest.  BQA_AUTOTUNE=0 keeps variant 0; BQA_BP_VARIANT=k forces variant k.
 * bqa_tuning_report writes the table as JSON text into buf when capacity suffices
 * and sets *needed to its size with the terminating zero; bqa_tuning_reset
 * forgets every choice. */
int bqa_tuning_report(char* buf, int64_t capacity, int64_t* needed);
int bqa_tuning_reset(void);

/* Graph-captured BP chunks (no counterpart in the reference, whose sweeps are
 * host-driven library calls, state.py:105-121).  Once a job has run BP eight
 * times in a row at one bond dimension, its sweeps run as replays of a HIP graph
 * holding four sweeps of every degree class; a graph is captured once per bond dimension, buffer set and
 * launch variants (thread-local capture mode) and replayed until the run
 * converges.  Kernel timing (bqa_timing_enable) and BQA_GRAPHS=0 keep the
 * launches eager, and so does a capture that a runtime call from outside this
 * library invalidates (counted as a failure; the job continues eagerly).  Counts
 * of captures, replays and failed captures so far: */
int bqa_job_graph_stats(const bqa_job* job, int64_t* captures, int64_t* launches, int64_t* failures);

/* Plain device-to-device copy of `bytes` bytes (HBM counter calibration). */
int bqa_copy_calibrate(uint64_t bytes, int32_t repeats, double* ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* BQA_HIP_H */
