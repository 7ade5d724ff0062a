/*
 * yaohip.h — C ABI of the MI355X (gfx950) state-vector / density-matrix backend.
 *
 * The entry points replace the dense simulation path of yao-rs. Every function
 * names the reference item it stands in for (paths are relative to the yao-rs
 * repository). Conventions carried over unchanged:
 *   - qubit loc 0 is the MOST significant bit of a basis index (bit = n-1-loc),
 *     src/instruct_qubit.rs:13-17;
 *   - a multi-target gate matrix maps gate-index bit k to targets[k]
 *     (src/instruct_qubit.rs:346-355);
 *   - a density matrix is row-major rho[row*2^n + col] (src/density_matrix.rs:9-12);
 *   - expectation values on a density matrix read Pauli sites as LSB-first bit
 *     positions (src/expect.rs:55-63); every other call is MSB-first.
 * States cross the boundary as interleaved (re, im) pairs of doubles.
 *
 * Every call returns a status (YH_OK = 0); yh_last_error() returns the message
 * of the last failure on the calling thread. Handles are not thread safe: like
 * the reference's `&mut` receivers, one handle is used by one thread at a time.
 * Each handle owns one HIP stream; calls are synchronous with respect to the host
 * unless stated otherwise.
 */
#ifndef YAOHIP_H
#define YAOHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: every reference panic / CircuitError maps to one ---- */
enum yh_status {
  YH_OK = 0,
  YH_ERR_INVALID_ARG = 1,          /* null pointer, bad enum, bad size           */
  YH_ERR_CONTROL_CONFIG_LEN = 2,   /* CircuitError::ControlConfigLengthMismatch  */
  YH_ERR_LOC_OUT_OF_RANGE = 3,     /* CircuitError::LocOutOfRange                */
  YH_ERR_OVERLAPPING_LOCS = 4,     /* CircuitError::OverlappingLocs              */
  YH_ERR_MATRIX_SIZE = 5,          /* CircuitError::MatrixSizeMismatch           */
  YH_ERR_NBITS_MISMATCH = 6,       /* apply.rs:128-132 register/circuit mismatch */
  YH_ERR_MEASURE_LOC = 7,          /* measure.rs:124-133 out of range/duplicate  */
  YH_ERR_CHANNEL_PARAM = 8,        /* noise.rs:90-176 parameter validation       */
  YH_ERR_HIP = 9,                  /* HIP runtime failure                         */
  YH_ERR_OUT_OF_MEMORY = 10,
  YH_ERR_UNSUPPORTED = 11,         /* valid in the reference, not on this backend */
  YH_ERR_DTYPE = 12,               /* operation not defined for the handle dtype  */
  YH_ERR_RESET_VALUE = 13,         /* measure.rs:256-261 reset value too large    */
  YH_ERR_NO_DEVICE = 14,
  YH_ERR_IO = 15,                  /* state file could not be opened, read or written */
  YH_ERR_FORMAT = 16               /* state file header rejected (state_io.rs:63-88) */
};

enum yh_dtype { YH_C128 = 0, YH_C64 = 1 };

/* Element kinds of a circuit (src/circuit.rs:33-37 CircuitElement). */
enum yh_elem { YH_ELEM_GATE = 0, YH_ELEM_CHANNEL = 1, YH_ELEM_ANNOTATION = 2 };

/* Gate kinds (src/gate.rs:7-35 enum Gate). */
enum yh_gate_kind {
  YH_GATE_X = 0, YH_GATE_Y, YH_GATE_Z, YH_GATE_H, YH_GATE_S, YH_GATE_T, YH_GATE_SWAP,
  YH_GATE_PHASE, YH_GATE_RX, YH_GATE_RY, YH_GATE_RZ, YH_GATE_SQRTX, YH_GATE_SQRTY,
  YH_GATE_SQRTW, YH_GATE_ISWAP, YH_GATE_FSIM, YH_GATE_CUSTOM
};

/* Channel kinds (src/noise.rs:16-60 enum NoiseChannel). params[] per kind:
 *  BITFLIP{p} PHASEFLIP{p} DEPOLARIZING{p, n (0 -> nt)} PAULI{px,py,pz}
 *  RESET{p0,p1} AMPLITUDE_DAMPING{gamma, excited_population}
 *  PHASE_DAMPING{gamma} PHASE_AMPLITUDE_DAMPING{amplitude, phase, excited_population}
 *  THERMAL_RELAXATION{t1, t2, time, excited_population}
 *  COHERENT{matrix} CUSTOM{nmat Kraus matrices back to back in `matrix`}        */
enum yh_channel_kind {
  YH_CH_BITFLIP = 0, YH_CH_PHASEFLIP, YH_CH_DEPOLARIZING, YH_CH_PAULI, YH_CH_RESET,
  YH_CH_AMPLITUDE_DAMPING, YH_CH_PHASE_DAMPING, YH_CH_PHASE_AMPLITUDE_DAMPING,
  YH_CH_THERMAL_RELAXATION, YH_CH_COHERENT, YH_CH_CUSTOM
};

/* Single-site operators (src/operator.rs:8-17 enum Op). */
enum yh_pauli { YH_OP_I = 0, YH_OP_X, YH_OP_Y, YH_OP_Z, YH_OP_P0, YH_OP_P1, YH_OP_PU, YH_OP_PD };

#define YH_MAX_TARGETS 12
#define YH_MAX_CTRLS 16

/* One circuit element: PositionedGate (src/circuit.rs:103-108), PositionedChannel
 * (:25-29) or PositionedAnnotation (:19-22). Plain data; `matrix` is only read
 * during the call that receives it. */
typedef struct yh_op {
  int32_t elem;                  /* enum yh_elem                                   */
  int32_t kind;                  /* enum yh_gate_kind or enum yh_channel_kind      */
  int32_t nt;                    /* number of target locs (gate) / locs (channel)  */
  int32_t nc;                    /* number of control locs                          */
  int32_t targets[YH_MAX_TARGETS];
  int32_t ctrls[YH_MAX_CTRLS];
  int32_t ctrl_vals[YH_MAX_CTRLS]; /* control_configs as 0/1                        */
  double params[4];
  const double *matrix;          /* CUSTOM gate: dim x dim row-major interleaved;
                                    COHERENT: one matrix; CUSTOM channel: nmat of them */
  int32_t nmat;                  /* CUSTOM channel: number of Kraus operators      */
  int32_t dim;                   /* matrix rows (CUSTOM gate / channel matrices)   */
  int32_t is_diag;               /* CUSTOM gate is_diagonal flag (gate.rs:31)       */
  int32_t nconf;                 /* control_configs length; must equal nc, else
                                    CircuitError::ControlConfigLengthMismatch     */
} yh_op;

typedef struct yh_reg yh_reg;    /* ArrayReg on the device   (register.rs:15-18)     */
typedef struct yh_dm yh_dm;      /* DensityMatrix on device  (density_matrix.rs:9-12) */

/* ---- library ---- */
const char *yh_last_error(void);
/* Selects device 0 and checks that it is a gfx950 agent. */
int yh_init(void);
int yh_device_name(char *buf, size_t len);
/* Device-to-device copy rate of a `bytes`-sized buffer (read + write, GB/s), the
 * fastest of `reps` copies timed with HIP events after one warm-up: the practical
 * HBM ceiling the streaming kernels are compared with. No reference counterpart. */
int yh_copy_bandwidth(size_t bytes, int reps, double *gbps);

/* ---- ArrayReg (src/register.rs) ---- */
int yh_reg_create(int nbits, int dtype, yh_reg **out);            /* zero_state :21-26; nbits <= 35 */
int yh_reg_destroy(yh_reg *reg);
int yh_reg_nbits(const yh_reg *reg, int *nbits);                  /* nqubits :87-89     */
int yh_reg_dtype(const yh_reg *reg, int *dtype);
int yh_reg_upload(yh_reg *reg, const double *amps, size_t n_amps); /* from_vec :65-69   */
int yh_reg_download(const yh_reg *reg, double *amps, size_t n_amps); /* state_vec :91  */
int yh_reg_clone(const yh_reg *src, yh_reg **out);                /* Clone / apply :150 */
int yh_reg_copy(yh_reg *dst, const yh_reg *src);
int yh_reg_norm(const yh_reg *reg, double *out);                  /* norm :98-104       */
int yh_reg_normalize(yh_reg *reg);                                /* normalize :106-113 */
int yh_reg_fidelity(const yh_reg *a, const yh_reg *b, double *out); /* fidelity :115-124 */
/* Device address of the amplitudes (interleaved re,im; double or float by dtype). */
int yh_reg_device_ptr(const yh_reg *reg, void **ptr);
/* Constructors (src/register.rs), each a new register filled on the device: */
/* product_state (:28-33): the one-hot state of basis index `index`, i.e. BitStr::onehot
 * (bitbasis/src/bitstr.rs:24-29); index < 2^nbits. */
int yh_reg_product_state(int nbits, int dtype, uint64_t index, yh_reg **out);
int yh_reg_uniform_state(int nbits, int dtype, yh_reg **out);       /* uniform_state :47-54 */
int yh_reg_ghz_state(int nbits, int dtype, yh_reg **out);           /* ghz_state :56-63     */
/* deterministic_state (:71-84): amplitude k = (cos 0.1k, sin 0.2k), normalised. */
int yh_reg_deterministic_state(int nbits, int dtype, yh_reg **out);
/* rand_state (:35-45) with the caller's stream: amplitude k = (u[2k] - 0.5,
 * u[2k+1] - 0.5), normalised; n_uniforms = 2 * 2^nbits f64 uniforms, the values
 * `rng.random::<f64>()` returns in the reference's loop order. */
int yh_reg_rand_state(int nbits, int dtype, const double *uniforms, size_t n_uniforms, yh_reg **out);

/* ---- state files: yao-state-v1 (yao-cli/src/state_io.rs) ---- */
/* write_state (state_io.rs:18-47): header line + little-endian (re, im) f64 pairs,
 * streamed from HBM in chunks (complex64 registers are widened). */
int yh_state_write(const yh_reg *reg, const char *path);
/* read_state_from_reader (state_io.rs:55-102) into a new register of `dtype`;
 * header checks and messages as the reference (YH_ERR_FORMAT), file errors
 * YH_ERR_IO. */
int yh_state_read(const char *path, int dtype, yh_reg **out);
/* Host-only: parse and check a state file's header; *num_qubits on success. */
int yh_state_read_header(const char *path, int *num_qubits);

/* ---- circuit application (src/apply.rs) ---- */
/* Validates the element list as Circuit::new does (circuit.rs:150-272) for a
 * qubit circuit of `nbits` sites; returns the matching CircuitError status. */
int yh_validate(int nbits, const yh_op *ops, size_t nops);
/* apply_inplace (apply.rs:123-147): gates in order, channels and annotations
 * skipped for a pure register. One kernel launch per gate (dispatch_arrayreg_gate
 * :7-120). */
int yh_apply(yh_reg *reg, int circuit_nbits, const yh_op *ops, size_t nops);
/* Same semantics as yh_apply (apply_inplace, apply.rs:123-147), with gates packed
 * into passes that each stream the state through HBM once.
 * max_fuse_qubits = 5 or 0 (default): gates fused into blocks of up to 5 qubits
 *   (32 x 32 complex matrices, or 16 x 16 for blocks on at most 4 qubits; FP64 on
 *   the host), each applied as an MFMA contraction (FP64 MFMA for complex128,
 *   FP32 MFMA for complex64). The planner keeps the plan with the lowest modelled
 *   launch time of several strategies and caches it by circuit structure (gate
 *   kinds and positions), so repeated calls only rebuild the block matrices.
 *   Blocks whose bits fit a 12-bit tile together with memory bits 0..3 share one
 *   HBM pass (up to 5 blocks applied from LDS, each complex product formed from
 *   three real ones); a block that fits with no other streams through HBM on its
 *   own (32 x 32, real-split product, four real ones).
 * max_fuse_qubits in 1..4: multi-gate LDS tile passes (12 qubits per pass, gates
 *   applied in groups of up to 4 qubits). */
int yh_apply_fused(yh_reg *reg, int circuit_nbits, const yh_op *ops, size_t nops,
                   int max_fuse_qubits);
/* Host-only plan of the 5-qubit block path: out[0] blocks, out[1] gates in blocks
 * (X split per target), out[2] gates run as their own launch. No device call. */
int yh_block_plan_stats(int nbits, const yh_op *ops, size_t nops, int64_t out[3]);
/* Host-only block plan, step by step: for step i, bits[5 i + j] is the memory bit
 * of row bit j of the block's 32 x 32 matrix, or -1 (five times) for a gate run as
 * its own launch. Fills at most `cap` steps; *nsteps receives the plan length.
 * No reference counterpart (planning introspection). */
int yh_block_plan(int nbits, const yh_op *ops, size_t nops, int32_t *bits, size_t cap, size_t *nsteps);
/* Host-only launch plan of the block path, pass by pass: dm = 0 for yh_apply_fused
 * (max_fuse_qubits 5) on a register of nbits qubits, dm = 1 for yh_dm_apply_blocks
 * on a density matrix of nbits qubits. out[0] block-kernel launches (one HBM pass
 * each), out[1] multi-block passes, out[2] blocks in them, out[3] of those applied
 * as 16 x 16 blocks (acting on at most 4 bits), out[4] multi-block passes whose
 * first block reads HBM directly, out[5] whose last block writes HBM directly,
 * out[6] elements run as their own launch. No device call; no reference
 * counterpart (planning introspection). */
int yh_block_pass_stats(int nbits, const yh_op *ops, size_t nops, int dm, int64_t out[7]);
/* Host-only plan of yh_apply_fused for a circuit on nbits qubits: out[0] HBM passes
 * of the tile kernel, out[1] gate groups, out[2] gates placed in tiles, out[3]
 * gates run as their own launch. No device call. */
int yh_plan_stats(int nbits, const yh_op *ops, size_t nops, int64_t out[4]);
/* Circuit::num_params (circuit.rs:308-316): trainable parameters of the gate
 * elements (Rx, Ry, Rz, Phase: 1; FSim: 2), in Circuit::parameters order. */
int yh_circuit_num_params(const yh_op *ops, size_t nops, size_t *n);
/* expect_grad (ad.rs:75-156): value = Re <psi|H|psi> with psi = circuit |reg>, and
 * its gradient with respect to every trainable parameter (ngrad must equal
 * yh_circuit_num_params), by one forward pass and an adjoint sweep on the device.
 * `reg` is not modified. H is flattened as for yh_expect (MSB-first sites). A
 * circuit with a noise channel returns YH_ERR_INVALID_ARG. */
int yh_expect_grad(yh_reg *reg, int circuit_nbits, const yh_op *ops, size_t nops, size_t nterms,
                   const double *coeffs, const int32_t *term_len, const int32_t *sites,
                   const int32_t *opcodes, double *value, double *grad, size_t ngrad);
/* dispatch_arrayreg_gate for one gate element (apply.rs:7). */
int yh_instruct(yh_reg *reg, const yh_op *op);
/* Host-side gate matrix, as Gate::matrix (gate.rs:63-68, 284-393); out holds
 * dim*dim interleaved complex, dim = 2 or 4. */
int yh_gate_matrix(const yh_op *op, double *out, int *dim);

/* ---- observables (src/expect.rs, src/operator.rs) ---- */
/* Operator polynomial, flattened: term t has coefficient (coeffs[2t], coeffs[2t+1])
 * and term_len[t] (site, op) pairs taken in order from sites[]/ops[].
 * OperatorString::new rules (operator.rs:27-33) are applied: sorted by site, I dropped. */
int yh_expect(const yh_reg *reg, size_t nterms, const double *coeffs,
              const int32_t *term_len, const int32_t *sites, const int32_t *ops,
              double out[2]);                                       /* expect_arrayreg :7 */

/* No reference counterpart: expect_arrayreg (expect.rs:7-11) for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). `amps` is a caller-owned device state of 2^nbits
 * amplitudes (interleaved re, im; double for YH_C128, float for YH_C64) on the
 * current device; terms as for yh_expect; `stream` a hipStream_t (NULL: the null
 * stream), synchronised before return. The reduction partials, results and at most
 * 256 terms live in the library's static device array (about 760 KiB per device,
 * held until the process ends); more terms return YH_ERR_INVALID_ARG. No device or
 * pinned memory is allocated. Every call on one GPU shares that array: calls that
 * overlap on one GPU race. Results equal yh_expect's (same kernels, same order). */
int yh_expect_noalloc(const void *amps, int nbits, int dtype, size_t nterms, const double *coeffs,
                      const int32_t *term_len, const int32_t *sites, const int32_t *ops, void *stream,
                      double out[2]);

/* ---- measurement (src/measure.rs) ---- */
/* probs (measure.rs:82-104): locs == NULL -> full distribution (2^n values),
 * else marginal over locs (2^nlocs values, output bit nlocs-1-i <- locs[i]). */
int yh_probs(const yh_reg *reg, const int32_t *locs, int nlocs, double *out);
/* Shot sampling as sample_from_probs (:140-150) with caller-supplied uniforms:
 * outcome[s] = first i with uniforms[s] < cumsum_i (sequential f64 cumsum), else
 * the last index. The distribution is computed once for all shots. */
int yh_sample(const yh_reg *reg, const int32_t *locs, int nlocs, const double *uniforms,
              size_t nshots, uint64_t *outcomes);
/* measure_with_postprocess (:243-274) with one uniform. post: 0 NoPostProcess,
 * 1 ResetTo(reset_val), 2 RemoveMeasured (the register shrinks to n-nlocs qubits).
 * outcome_bits receives nlocs bits, MSB-first over locs. */
int yh_measure(yh_reg *reg, const int32_t *locs, int nlocs, int post, uint64_t reset_val,
               double uniform, int32_t *outcome_bits);

/* ---- DensityMatrix (src/density_matrix.rs) ---- */
int yh_dm_create(int nbits, yh_dm **out);                          /* zero_state :19-21 */
int yh_dm_from_reg(const yh_reg *reg, yh_dm **out);                /* from_reg :23-35   */
/* mixed (density_matrix.rs:37-61): sum_j weights[j] |psi_j><psi_j| over nregs
 * complex128 registers of one size, accumulated per entry in register order. */
int yh_dm_mixed(const double *weights, const yh_reg *const *regs, size_t nregs, yh_dm **out);
/* density_matrix_from_reg (density_matrix.rs:266-272): the reduced density matrix
 * on the kept `locs` (read as bit positions, like partial_tr; the others are traced
 * out), computed from the amplitudes without forming the full 4^n matrix; entries
 * are summed in the reference's order. */
int yh_dm_from_reg_locs(const yh_reg *reg, const int32_t *locs, int nlocs, yh_dm **out);
int yh_dm_destroy(yh_dm *dm);
int yh_dm_nbits(const yh_dm *dm, int *nbits);
int yh_dm_upload(yh_dm *dm, const double *rho, size_t n_entries);
int yh_dm_download(const yh_dm *dm, double *rho, size_t n_entries);
/* Register::apply for DensityMatrix (:244-259): gate -> U rho U^dagger (:161-178),
 * channel -> sum_i K_i rho K_i^dagger (:181-224). One pass over rho per element on at
 * most 2 qubits; a channel on more qubits (or with more than 64 Kraus operators)
 * sums its Kraus branches as the reference does (a branch copy of rho per operator,
 * K on the row bits, conj(K) on the column bits, accumulated in order). */
int yh_dm_apply(yh_dm *dm, int circuit_nbits, const yh_op *ops, size_t nops);
/* Same result; each gate is fused with the 1-qubit channels that follow it on its
 * target qubits into one superoperator pass. */
int yh_dm_apply_fused(yh_dm *dm, int circuit_nbits, const yh_op *ops, size_t nops);
/* Same result through the 5-qubit block path: rho is a vector of 2n bits (row bits
 * above column bits), a gate is U (x) conj(U) on the row and column bits of its
 * qubits and a channel sum_k K_k (x) conj(K_k); these superoperators (at most 4
 * bits for 1- and 2-qubit elements) are fused into 5-bit blocks by the planner of
 * yh_apply_fused and applied with the same MFMA kernels, several blocks per HBM
 * pass where they share a 12-bit tile. Elements on more than 2 qubits run on their
 * own as in yh_dm_apply. */
int yh_dm_apply_blocks(yh_dm *dm, int circuit_nbits, const yh_op *ops, size_t nops);
int yh_dm_trace(const yh_dm *dm, double out[2]);                   /* trace :63-67      */
int yh_dm_purity(const yh_dm *dm, double *out);                    /* purity :69-78     */
/* partial_tr (density_matrix.rs:80-111): the reduced density matrix over the
 * kept qubits, as a new handle. `traced` are BIT positions (LSB-first, like the
 * reference's loc here); kept bit idx of the reduced index is the idx-th kept bit.
 * Entries are summed in the reference's order (bit-identical). */
int yh_dm_partial_tr(const yh_dm *dm, const int32_t *traced, int ntraced, yh_dm **out);
/* von_neumann_entropy (density_matrix.rs:113-122). The Hermitian eigenvalues are
 * computed on the host (Householder tridiagonalisation + implicit QL) for
 * nbits <= 10, else on the device with rocSOLVER's zheevd on a copy of rho (the
 * reference uses faer's eigensolver on the CPU). */
int yh_dm_entropy(const yh_dm *dm, double *out);
/* expect_dm (expect.rs:38-72): Pauli sites are LSB-first bit positions. */
int yh_dm_expect(const yh_dm *dm, size_t nterms, const double *coeffs,
                 const int32_t *term_len, const int32_t *sites, const int32_t *ops,
                 double out[2]);
/* probs for a DensityMatrix (measure.rs:50-61, 106-122). */
int yh_dm_probs(const yh_dm *dm, const int32_t *locs, int nlocs, double *out);

/* ---- noise channels (src/noise.rs) ---- */
/* kraus_operators (noise.rs:183-352): writes up to max_ops matrices of dim x dim
 * interleaved complex into out, the count into *nops and the dimension into *dim. */
int yh_channel_kraus(const yh_op *channel, double *out, int max_ops, int *nops, int *dim);

/* ---- measurement hooks (no reference counterpart) ----
 * Per-kernel-class launch timing with HIP events recorded on the handle's own
 * stream around every launch while enabled. Classes: */
enum yh_kclass {
  YH_K_1Q = 0,      /* dense 2x2 pair stream            */
  YH_K_DENSE = 1,   /* dense 2^k group, controls         */
  YH_K_DIAG = 2,    /* diagonal phase multiply           */
  YH_K_X = 3,       /* X pair swap                       */
  YH_K_SWAP = 4,    /* SWAP                               */
  YH_K_OOP = 5,     /* out-of-place dense (> 5 targets)  */
  YH_K_FUSED = 6,   /* fused 2^k block (MFMA)            */
  YH_K_TILE = 7,    /* multi-gate LDS tile pass          */
  YH_K_DM = 8,      /* density-matrix block pass         */
  YH_K_VQE = 9,     /* batched VQE persistent launch      */
  YH_K_TBLOCK = 10, /* several fused blocks per LDS tile  */
  YH_K_NCLASS = 11
};
int yh_reg_timing_enable(yh_reg *reg, int on);
int yh_reg_timing_read(const yh_reg *reg, int kclass, uint64_t *launches, double *total_ms);
int yh_dm_timing_enable(yh_dm *dm, int on);
int yh_dm_timing_read(const yh_dm *dm, int kclass, uint64_t *launches, double *total_ms);

/* ---- batched VQE (no reference counterpart: many ansatz instances per launch) ---- */
typedef struct yh_vqe_batch yh_vqe_batch;
/* Hardware-efficient ansatz from |0...0>: `depth` layers, each a single-qubit block on
 * every qubit followed by a CZ chain CZ(i, i+1), i < n-1 (CZ = control([i],[i+1],Z),
 * easybuild.rs:516). angles_per_qubit = 2: the block is Ry then Rz; 3: Rz, Ry, Rz
 * (general_u2, easybuild.rs:75-81). Instance b has nq[b] qubits and depth[b] layers;
 * its angles_per_qubit*nq*depth angles are angles[offset[b] ...] ordered layer,
 * qubit, then the block's gates in application order.
 * Energy: TFIM H = -J sum Z_i Z_{i+1} - h sum X_i, open boundary.
 * dtype YH_C64 keeps amplitudes in complex64, reduces in FP64 and returns
 * <psi|H|psi> / <psi|psi> (the FP32 roundings move the norm of the computed state,
 * which the division takes out); YH_C128 is the all-FP64 path and returns
 * <psi|H|psi> as expect.rs:7-24 does. The first call for a given list of (nq, depth) shapes times every
 * launch configuration on that batch and keeps the fastest for later calls with
 * the same shapes; the launch sequence is then replayed as a captured hipGraph. */
int yh_vqe_batch_create(int max_instances, int max_qubits, int dtype, yh_vqe_batch **out);
int yh_vqe_batch_destroy(yh_vqe_batch *b);
/* Fix the launch configuration instead of autotuning: tile_bits in {11, 12, 13}
 * (instances with more qubits stream through HBM in 2^tile_bits-amplitude work
 * items), blocks_per_cu workgroups per CU for the LDS-resident kernel,
 * resident_grid != 0 for step launches of as many workgroups as stay resident
 * (else one workgroup per work item), step_threads 256 or 512 (workgroup size of
 * the streaming kernel). tile_bits = 0 returns to autotuning. */
int yh_vqe_batch_configure(yh_vqe_batch *b, int tile_bits, int blocks_per_cu, int resident_grid,
                           int step_threads);
/* The configuration of the last yh_vqe_batch_energies call: tile_bits,
 * blocks_per_cu, resident_grid, step_threads. */
int yh_vqe_batch_config(const yh_vqe_batch *b, int32_t out[4]);
int yh_vqe_batch_energies(yh_vqe_batch *b, int ninst, const int32_t *nq,
                          const int32_t *depth, const int64_t *offset, const double *angles,
                          int angles_per_qubit, double J, double h, double *energies);
/* Same circuits, evaluated one instance per launch sequence (the plain per-circuit
 * GPU path through yh_apply + yh_expect), for comparison. */
int yh_vqe_plain_energies(int ninst, const int32_t *nq, const int32_t *depth,
                          const int64_t *offset, const double *angles, int angles_per_qubit,
                          double J, double h, int dtype, double *energies);

/* ---- MLP surrogate (no reference counterpart) ----
 * in_dim -> `hidden` layers of `width` (GELU) -> 1, y = x W^T + b per layer. FP32
 * master parameters; forward and backward run BF16 MFMA GEMMs with FP32
 * accumulation. Parameters are flat FP32, layer by layer: W (width x fan_in,
 * row-major) then b. Input matrices are row-major FP32 (n x in_dim).
 * Input encodings: YH_MLP_LINEAR feeds x as is (fan_in of layer 0 = in_dim,
 * in_dim % 16 == 0); YH_MLP_TRIG feeds (cos x_j, sin x_j) at features 2j, 2j+1
 * (fan_in = 2 in_dim, in_dim % 8 == 0), for inputs that are rotation angles.
 * width % 64 == 0, 1 <= hidden <= 16, max_batch % 64 == 0. */
typedef struct yh_mlp yh_mlp;
typedef enum { YH_MLP_LINEAR = 0, YH_MLP_TRIG = 1 } yh_mlp_encoding;
int yh_mlp_create(int in_dim, int encoding, int width, int hidden, int max_batch, uint64_t seed, yh_mlp **out);
int yh_mlp_destroy(yh_mlp *m);
int yh_mlp_num_params(const yh_mlp *m, size_t *n);
int yh_mlp_get_params(const yh_mlp *m, float *params, size_t n);
int yh_mlp_set_params(yh_mlp *m, const float *params, size_t n);
/* y = f(x) for n rows (any n; processed in chunks of max_batch). */
int yh_mlp_predict(yh_mlp *m, const float *x, size_t n, float *y);
/* The same for FP64 input rows, rounded to FP32 as they are staged (no FP32 copy
 * of the input on the caller's side). */
int yh_mlp_predict_f64(yh_mlp *m, const double *x, size_t n, float *y);
/* Mean-squared-error loss of one batch (n <= max_batch) and its gradient with
 * respect to every parameter (grads may be NULL), without updating. */
int yh_mlp_loss_grad(yh_mlp *m, const float *x, const float *y, int n, float *loss, float *grads);
/* Adam training: the n rows are uploaded once; each epoch visits them in a seeded
 * random order in batches of `batch` rows. epoch_loss[e] receives the mean batch
 * loss of epoch e (may be NULL). */
int yh_mlp_train(yh_mlp *m, const float *x, const float *y, size_t n, int batch, int epochs, float lr,
                 uint64_t seed, float *epoch_loss);

#ifdef __cplusplus
}
#endif
#endif /* YAOHIP_H */
