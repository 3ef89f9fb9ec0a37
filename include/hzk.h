/* hzk.h — C ABI of the MI355X-native alt_bn128 Groth16 prover (libhzk.so).
 *
 * The drop-in boundary for LibLightZK's operator and protocol API: every entry
 * point below names the reference interface it replaces (file:line in the
 * reference tree). Plain C types only; device buffers are `void*` device
 * pointers owned by the caller; results are valid after hzk_sync().
 *
 * Byte layouts (little-endian, Montgomery form, R = 2^256, fully reduced):
 *   Fr / Fq element   32 B  (= reference mont_t / libff bigint<4>)
 *   G1 affine         64 B  x | y                 infinity = 64 zero bytes
 *   G2 affine        128 B  x.c0 | x.c1 | y.c0 | y.c1   infinity = zero bytes
 *   G1 Jacobian       96 B  X | Y | Z              (reference g1_t, jacobian_t.cuh:11)
 *   G2 Jacobian      192 B  X.c0|X.c1|Y.c0|Y.c1|Z.c0|Z.c1 (reference write_to, alt_bn128_t.cuh:252-256)
 * MSM results are written normalised: (x, y, 1), or (0, 1, 0) for infinity
 * (libff's convention, alt_bn128_g1.cu:56-73).
 *
 * Return codes: 0 on success, a negative HZK_ERR_* otherwise; the message is
 * in hzk_last_error(). No entry point aborts the process.
 */
#ifndef HZK_H
#define HZK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HZK_OK 0
#define HZK_ERR_INVALID (-1)   /* bad argument */
#define HZK_ERR_NOMEM (-2)     /* device or host allocation failed */
#define HZK_ERR_HIP (-3)       /* HIP runtime error */
#define HZK_ERR_UNSUPPORTED (-4)

typedef struct hzk_ctx hzk_ctx;

/* batch flags */
#define HZK_BATCH_HOST_DEST 1u   /* MSM results written to host pointers */
#define HZK_BATCH_GRAPH 2u       /* capture the device stages into a hipGraph and replay it */
/* the window combine (part sums, Horner over the windows, normalisation) as a
 * device kernel, one lane per instance, so the call returns without waiting
 * for the device and, with HZK_BATCH_GRAPH, the whole MSM is one graph;
 * dest must be device memory. It ends the MSM later than the host combine
 * (about 255 dependent doublings on one lane: 1 ms G1, 3 ms G2 on MI355X,
 * hzk_roofline_probe kinds 4/5); the reference also ends its MSM on the
 * device (msm.cuh:226-227). hzk_msm_last_stats is not updated in this mode. */
#define HZK_BATCH_DEVICE_COMBINE 4u

/* ---- context (replaces the reference's implicit CUDA runtime state and legacy
 * default stream, include/device-arith/device-utils.cuh:3) ---- */
int hzk_ctx_create(int device, void* hip_stream /* nullable: own stream */, hzk_ctx** out);
void hzk_ctx_destroy(hzk_ctx* ctx);
const char* hzk_last_error(const hzk_ctx* ctx);
int hzk_sync(hzk_ctx* ctx);
/* wait for all work on the context's device, every stream (the reference's
 * operators run on the legacy default stream; its callers rely on that order) */
int hzk_device_sync(hzk_ctx* ctx);
void* hzk_ctx_stream(hzk_ctx* ctx);
const char* hzk_version(void);

/* ---- buffers (replaces xpu::vector<T> allocate/store/load/clear,
 * include/data-structure/xpu-vector.cuh:10-101) ---- */
int hzk_device_alloc(hzk_ctx* ctx, size_t bytes, void** out);
int hzk_device_free(hzk_ctx* ctx, void* ptr);
int hzk_memcpy_h2d(hzk_ctx* ctx, void* dst, const void* src, size_t bytes);   /* synchronous */
int hzk_memcpy_d2h(hzk_ctx* ctx, void* dst, const void* src, size_t bytes);   /* synchronous */
int hzk_memset(hzk_ctx* ctx, void* dst, int value, size_t bytes);              /* stream-ordered */
int hzk_mem_info(hzk_ctx* ctx, size_t* free_bytes, size_t* total_bytes);

/* ---- device field arithmetic unit entry (the reference exercises mont_t/fp2_t
 * through test kernels: test/fp2-check.cu:7-62, test/cpu-gpu-check.cu:15-57).
 * field: 0 Fq, 1 Fr, 2 Fq2. op: 0 mul, 1 add, 2 sub, 3 neg(a), 4 sqr(a),
 * 5 inv(a), 6 from_mont(a), 7 to_mont(a). a, b, out: device arrays of n
 * elements. ---- */
int hzk_field_op(hzk_ctx* ctx, int field, int op, const void* a, const void* b, void* out, size_t n);

/* ---- MSM (replaces multi_scalar_multiplication_g1/_g2,
 * include/operator/msm.cuh:210-248).
 * scalars: device Fr[n] (Montgomery); bases: device affine points (layout
 * HZK_BASES_AFFINE); dest: device pointer to one Jacobian point (96 / 192 B),
 * written normalised; dest may alias other buffers of the call.
 * win_siz: window bits c (0 = autotuned). ---- */
#define HZK_BASES_AFFINE 0u
/* the reference's pre_comp table (include/operator/msm.cuh:48-75): G1 96-B
 * Jacobian records X|Y|Z with Z = 1; G2 two 96-B lane records per point
 * [X.c0|Y.c0|Z.c0][X.c1|Y.c1|Z.c1] (alt_bn128_t.cuh:244-256); level-major.
 * The MSM reads level 0 (the first n points) only. */
#define HZK_BASES_REF_LEVELS 1u
/* a fixed-base level table built by hzk_bases_levels for window size win_siz
 * with all L = ceil(255 / win_siz) levels, point j*n + i = 2^(win_siz*j) * P_i:
 * every window then shares one bucket set (one bucket reduction instead of L),
 * the layout the reference's pre_comp levels serve (msm.cuh:98-115). Batched
 * MSMs may use tables of fewer levels (hzk_msm_desc.levels). */
#define HZK_BASES_AFFINE_LEVELS 2u
int hzk_msm_g1(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, uint32_t base_layout,
               size_t win_siz, void* dest);
int hzk_msm_g2(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, uint32_t base_layout,
               size_t win_siz, void* dest);

/* Batched MSMs of uneven sizes (no reference counterpart: the reference runs
 * one MSM per call). Every stage runs as one launch over the whole batch; with
 * HZK_BATCH_GRAPH the device stages are captured into a hipGraph on first use
 * of a batch shape (sizes and plans) and replayed afterwards. Each result is
 * bit-identical to the same instance run alone through hzk_msm_g1/_g2. */
typedef struct hzk_msm_desc {
    const void* scalars;   /* device Fr[n] */
    const void* bases;     /* device affine points */
    size_t n;
    void* dest;            /* Jacobian result (device, or host with HZK_BATCH_HOST_DEST) */
    uint32_t win_siz;      /* 0 = autotuned */
    uint32_t levels;       /* 0 or 1: affine bases; Lg > 1: bases is a level table of Lg levels
                              (hzk_bases_levels with win_siz), each bucket set serving Lg windows */
} hzk_msm_desc;
int hzk_msm_g1_batch(hzk_ctx* ctx, const hzk_msm_desc* d, size_t count, uint32_t flags);
int hzk_msm_g2_batch(hzk_ctx* ctx, const hzk_msm_desc* d, size_t count, uint32_t flags);

/* statistics of the last MSM on this context: {bucket entries (nonzero
 * digits), accumulation chunks, largest bucket, window bits c, windows} */
int hzk_msm_last_stats(hzk_ctx* ctx, uint32_t* out5);

/* ---- NTT over Fr (replaces number_theory_transformation,
 * include/operator/poly.cuh:65-73, same semantics): in place, natural order,
 *   coeffs[k] <- factor * sum_j coeffs[j] * w^(j k),  deg a power of two <= 2^26.
 * omega_powers: device table w^e for e < deg/2 (ntt_setup layout,
 * include/protocol/groth16-operator.cuh:6-32), or NULL for the library's own
 * table of the primitive deg-th root (inverse = 0) or its inverse (inverse = 1,
 * then scaled by 1/deg unless factor is given). factor: device Fr or NULL. */
int hzk_ntt(hzk_ctx* ctx, void* coeffs, size_t deg, const void* omega_powers, int inverse, const void* factor);
/* The reference's ntt_setup table for an M-point domain, written to host
 * memory as 6M + 2 Fr elements (replaces ntt_setup<FieldT, ntt_config*>,
 * include/protocol/groth16-operator.cuh:6-32): [0, M) w_M^i, [M, 2M) w_M^-i,
 * 2M: M^-1, [2M+1, 4M+1) w_2M^i, [4M+1, 6M+1) w_2M^-i, 6M+1: (4M)^-1 (the
 * value the reference writes there). M a power of two <= 2^26. Needs no GPU. */
int hzk_ntt_setup(size_t M, void* host_out);
/* coset NTT on g*<w>, g = 5 (alt_bn128 Fr multiplicative generator,
 * alt_bn128_init.cu:50): inverse = 0 evaluates the coefficients at g*w^k;
 * inverse = 1 interpolates from those evaluations back to coefficients.
 * No reference counterpart (the reference evaluates on a 2*deg domain
 * instead, include/protocol/groth16.cuh:176-185). */
int hzk_coset_ntt(hzk_ctx* ctx, void* coeffs, size_t deg, int inverse);

/* ---- R1CS SpMV (replaces batch_sparse_matrix_vector_multiplication,
 * include/operator/poly.cuh:5-30; CSR = the reference's spmat_descript,
 * include/data-structure/xpu-spmat.cuh:20-25: device arrays, row_ptr has `row`
 * entries and no sentinel, the last row ends at cnt). Az/Bz/Cz need not be
 * zeroed (the reference requires it, poly.cuh:8). */
typedef struct hzk_csr {
    const void* val;            /* Fr[cnt] */
    const uint64_t* col_idx;    /* [cnt] */
    const uint64_t* row_ptr;    /* [row] */
    size_t row, col, cnt;
} hzk_csr;
int hzk_spmv3(hzk_ctx* ctx, const hzk_csr* A, const hzk_csr* B, const hzk_csr* C, const void* z, void* Az, void* Bz,
              void* Cz);
/* A <- A o B - C over deg Fr elements (replaces polynomial_A_times_B_minus_C,
 * include/operator/poly.cuh:75-86) */
int hzk_pointwise_abc(hzk_ctx* ctx, size_t deg, void* A, const void* B, const void* C);

/* G1 MSM that allocates nothing (no counterpart in the reference): for callers
 * that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same result and layouts as
 * hzk_msm_g1 with base_layout HZK_BASES_AFFINE and win_siz 0 (dest: device
 * pointer to 96 bytes). Its scratch is a static device array of
 * HZK_NOALLOC_SCRATCH_MB MiB per GPU (mapped when a context is created, held
 * until the process ends); an MSM whose plan needs more (about n > 2^15
 * points) fails with HZK_ERR_UNSUPPORTED. All calls on one GPU share that
 * array: calls that overlap on one GPU, from any contexts, race. Host
 * bookkeeping uses the C++ heap and pageable copies; nothing is allocated on
 * the device or pinned. */
#define HZK_NOALLOC_SCRATCH_MB 128
int hzk_msm_g1_noalloc(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, void* dest);

/* Batched transforms of uneven sizes (no reference counterpart). kind: 0 NTT,
 * 1 inverse NTT (scaled by 1/deg), 2 coset NTT, 3 coset inverse NTT, all with
 * the library's tables; in place. With HZK_BATCH_GRAPH the launches are
 * captured into a hipGraph per batch shape and replayed. */
typedef struct hzk_ntt_desc {
    void* coeffs;
    size_t deg;
    int kind;
} hzk_ntt_desc;
int hzk_ntt_batch(hzk_ctx* ctx, const hzk_ntt_desc* d, size_t count, uint32_t flags);

/* ---- fixed-base generation (replaces fix_base_multi_scalar_multiplication_g1/_g2,
 * include/operator/msm.cuh:5-46, plus the affine normalisation of pre_comp_g1/_g2,
 * msm.cuh:48-75): out_affine[i] = k_i * G for device Fr scalars k (Montgomery);
 * scalars are not modified (the reference converts them in place, msm.cuh:16). */
int hzk_fixbase_g1(hzk_ctx* ctx, const void* scalars, size_t n, void* out_affine);
int hzk_fixbase_g2(hzk_ctx* ctx, const void* scalars, size_t n, void* out_affine);
/* convert n base points between layouts (HZK_BASES_AFFINE <-> HZK_BASES_REF_LEVELS,
 * level 0); group 1 = G1, 2 = G2 */
int hzk_bases_convert(hzk_ctx* ctx, int group, uint32_t from_layout, uint32_t to_layout, const void* src, void* dst,
                      size_t n);
/* replaces pre_comp_g1/_g2 (include/operator/msm.cuh:48-75): normalises the
 * `length` Jacobian records of level 0 in place (Z = 1; infinity -> (0, 0, 1)
 * as the reference leaves it) and writes levels j = 1..level-1 as
 * 2^(gap*j) * P_i at record j*length + i (G2: point j*length + i) */
int hzk_precomp(hzk_ctx* ctx, int group, void* points, size_t length, size_t gap, size_t level);
/* level table for window size c from n affine bases: out[j*n + i] =
 * 2^(c*j) * bases[i] for j < levels <= ceil(255 / c) (device, affine) */
int hzk_bases_levels(hzk_ctx* ctx, int group, int c, int levels, const void* bases, size_t n, void* out);

/* ---- seeded synthetic inputs (replaces the reference's std::random_device draws,
 * include/host-arith/field/bigint.cuh:172-182, and Fp_model::random_element,
 * include/host-arith/field/fp.cuh:789-816). kind: 0 uniform Fr, 1 sparse Fr
 * (70% in {0,1}), 2 uniform Fq; n elements of 32 B into host memory. */
int hzk_gen_field(uint64_t seed, uint64_t stream, size_t n, int kind, void* host_out);

/* ---- Groth16 prover (replaces the protocol class Groth16<...>,
 * include/protocol/groth16.cuh:18-223: the constructor + Setup(win_siz) become
 * hzk_g16_setup, Prove() becomes hzk_g16_prove, which returns the proof instead
 * of printing it, groth16.cuh:219-221). The statement is the synthetic
 * multiplication-gate circuit with M = 2^log_m constraint rows (circuit and
 * toxic waste from `seed`); the proving key stays resident in HBM. */
typedef struct hzk_g16 hzk_g16;
int hzk_g16_setup(hzk_ctx* ctx, uint64_t seed, int log_m, int num_public, hzk_g16** out);
void hzk_g16_free(hzk_g16* g);
/* {M, num_vars, num_public, num_gates, nnz A, nnz B, nnz C, resident device bytes of the key: CSR,
 * full and compacted query arrays, level tables and the plain prover's buffers} */
int hzk_g16_info(const hzk_g16* g, uint64_t* out8);
/* seeded satisfying witness z[num_vars] (host, Fr Montgomery; z[0] = 1) */
int hzk_g16_witness(const hzk_g16* g, uint64_t wseed, void* z_host);
/* prover randomness r | s (host, 2 x 32 B) for a witness seed */
int hzk_g16_rs(uint64_t wseed, void* rs_host);
/* proof = A (G1 affine 64 B) | B (G2 affine 128 B) | C (G1 affine 64 B), host memory */
int hzk_g16_prove(hzk_ctx* ctx, hzk_g16* g, const void* z_host, const void* rs_host, void* proof_host);
/* prove over fixed-base level tables of the key with `levels` levels per
 * bucket set (built on ctx; each MSM then reduces ceil(W / levels) bucket sets
 * instead of W), or over the plain key arrays (levels <= 1, the default).
 * Proofs are identical either way. */
int hzk_g16_set_levels(hzk_ctx* ctx, hzk_g16* g, int levels);
/* device pointer / count of a key array: 0 A1, 1 B1, 2 B2 (G2), 3 L1, 4 H1 */
int hzk_g16_key_array(const hzk_g16* g, int which, void** dev_ptr, size_t* count);
/* the reference's own statement: Groth16(input_scale) + Setup over the MatMul
 * R1CS of xpu::spmat (include/data-structure/xpu-spmat.cuh:33-65,
 * include/protocol/groth16.cuh:11-16, :39-71) for p x p matrices: k = 3p^2
 * public inputs (X, Y, X*Y), N = 3p^2 + p^3 variables besides z_0,
 * M = next_pow2(p^3 + p^2) rows, all coefficients one (input_scale p = 39, 63,
 * 100 give M = 2^16, 2^18, 2^20). hzk_g16_witness draws X and Y from the seed
 * and computes the partial products and X*Y forward. 2 <= p <= 255. */
int hzk_g16_setup_matmul(hzk_ctx* ctx, uint64_t seed, int p, hzk_g16** out);
/* proving key and R1CS to / from a file (no counterpart in the reference,
 * which regenerates random keys per run, groth16.cuh:84-122): the loaded key
 * proves byte-identically to the saved one. The load validates magic, version,
 * header consistency, file size, CSR bounds and a checksum of the whole file
 * and fails with HZK_ERR_INVALID naming the defect. */
int hzk_g16_save(hzk_ctx* ctx, const hzk_g16* g, const char* path);
int hzk_g16_load(hzk_ctx* ctx, const char* path, hzk_g16** out);

/* ---- proof server (no counterpart in the reference, which proves one
 * statement per process, test/groth16.cu:3-8): several circuits' proving keys
 * resident in HBM, requests submitted from any thread and proved by a worker
 * with continuous batching. mode 0: each step runs the SpMV + H stage of newly
 * admitted requests on one stream while the batched MSMs of the previous
 * step's requests run on another; mode 1: one request at a time, no overlap.
 * Admission is in arrival order while the requests' device footprint fits
 * mem_budget (0 = half the device memory still free when hzk_server_start
 * runs, after the circuits' keys are resident); a request that can never fit fails
 * with HZK_ERR_NOMEM. A failing stage fails only its requests. A proof is a
 * function of (circuit, z, r|s) alone: the same alone or in any mix. */
typedef struct hzk_server hzk_server;
int hzk_server_create(int device, size_t mem_budget, int max_batch, int mode, hzk_server** out);
/* number of pipelines (worker thread + two streams + MSM scratch each) sharing
 * the admission queue, 1..8, before hzk_server_start; default 3. In mode 1
 * only the first pipeline proves. */
int hzk_server_set_pipelines(hzk_server* s, int n);
void hzk_server_destroy(hzk_server* s);
const char* hzk_server_last_error(const hzk_server* s);
/* set up a circuit (as hzk_g16_setup) before hzk_server_start */
int hzk_server_add_circuit(hzk_server* s, uint64_t seed, int log_m, int num_public, int* circuit_id);
/* add a circuit from a proving-key file (hzk_g16_save), before hzk_server_start */
int hzk_server_add_circuit_file(hzk_server* s, const char* path, int* circuit_id);
/* prove the circuit over fixed-base level tables (hzk_g16_set_levels), before
 * hzk_server_start; proofs are identical either way */
int hzk_server_set_levels(hzk_server* s, int circuit_id, int levels);
/* the circuit's hzk_g16 (witness generation, info); owned by the server */
int hzk_server_circuit(hzk_server* s, int circuit_id, hzk_g16** out);
/* device bytes one request of the circuit is charged against the budget */
int hzk_server_footprint(hzk_server* s, int circuit_id, size_t* bytes);
int hzk_server_start(hzk_server* s);
/* switch between mode 0 and 1 while no request is pending or in flight */
int hzk_server_set_mode(hzk_server* s, int mode);
/* z: num_vars x 32 B host (copied), rs: r | s host (copied) */
int hzk_server_submit(hzk_server* s, int circuit_id, const void* z_host, const void* rs_host, uint64_t* request_id);
/* 1 done (proof_out = 256 B proof, latency from submit to completion),
 * 0 still running after timeout_s, < 0 failed; a finished request's id is
 * released by the call that returns its result */
int hzk_server_wait(hzk_server* s, uint64_t request_id, double timeout_s, void* proof_out, double* latency_s);
/* {steps, completed, failed, max in flight, mean requests per MSM stage,
 *  budget bytes, bytes admitted} */
int hzk_server_stats(hzk_server* s, double* out7);
/* make the n-th following MSM scratch allocation fail (-1 = off) */
int hzk_server_inject_failure(hzk_server* s, long after_n);

/* ---- timing (replaces the reference's Timer phase printouts,
 * include/timer/timer.cuh:5-12). When enabled, each kernel launch is
 * bracketed by HIP events on the context stream; totals are per kernel name. */
int hzk_timing_enable(hzk_ctx* ctx, int enable);
int hzk_timing_reset(hzk_ctx* ctx);
/* total milliseconds and launch count of kernel `name` since the last reset */
int hzk_timing_get(hzk_ctx* ctx, const char* name, double* ms, long* launches);
/* newline-separated "name ms launches" report; returns bytes needed */
size_t hzk_timing_report(hzk_ctx* ctx, char* buf, size_t cap);

/* ---- roofline denominators (no counterpart in the reference: the measured
 * peaks that achieved rates are divided by). kind: 0 v_mad_u64_u32 per second,
 * 1 Fq Montgomery products per second, 2 32-bit VALU adds per second,
 * 3 HBM copy bytes per second (read + write), 4 / 5 G1 / G2 point doublings
 * per second along one lane's dependent chain on a single wave (the latency
 * a device-side window combine of about 255 doublings would pay). */
int hzk_roofline_probe(hzk_ctx* ctx, int kind, double* value);

/* ---- MSM autotuning (no counterpart in the reference, which hard-wires
 * c = 12 and its launch geometry, include/operator/msm.cuh:121, :225, :245).
 * A plan per group (1 = G1, 2 = G2) and size class log_n = floor(log2 n):
 * window bits c, accumulation chunk length K (entries per lane, a multiple of
 * 4) and bucket-reduction run length g (a power of two <= 2^(c-1)). Every MSM
 * on the context, single or in a narrow batch, uses the class's plan; results
 * do not depend on the plan. c = 0 clears the entry (built-in table or
 * heuristic). */
int hzk_msm_tune_set(hzk_ctx* ctx, int group, int log_n, int c, uint32_t K, uint32_t g);
/* the plan an MSM of 2^log_n points would use now */
int hzk_msm_tune_get(hzk_ctx* ctx, int group, int log_n, int* c, uint32_t* K, uint32_t* g);
/* time candidate plans for classes log_min..log_max on seeded synthetic
 * inputs (coordinate descent over c, K, g; at most budget_ms per class) and
 * store the fastest in the context's table */
int hzk_msm_autotune(hzk_ctx* ctx, int group, int log_min, int log_max, double budget_ms);
/* The same for the throughput regime: the plans used by the instances of a
 * wide batch (two or more MSMs whose digit count under the single-MSM plans
 * reaches 2^26, or at least 8 G1 / 16 G2 MSMs with 2^24 digits), which fills
 * the GPU without help from the plan; typically
 * fewer window bits (fewer buckets to reduce), longer chunks and runs. The
 * autotuner times each class as a batch of 2^26 / (16 n) instances (2..256). */
int hzk_msm_batch_tune_set(hzk_ctx* ctx, int group, int log_n, int c, uint32_t K, uint32_t g);
int hzk_msm_batch_tune_get(hzk_ctx* ctx, int group, int log_n, int* c, uint32_t* K, uint32_t* g);
int hzk_msm_batch_autotune(hzk_ctx* ctx, int group, int log_min, int log_max, double budget_ms);

/* ---- tuning / test hooks ---- */
/* force MSM window bits and bucket-reduction segment size (0 = automatic) */
int hzk_set_tuning(hzk_ctx* ctx, int window_bits, int segment);
/* MSM plan regime: -1 by batch width (default), 0 single-MSM plans, 1 wide-batch plans */
int hzk_set_msm_regime(hzk_ctx* ctx, int regime);
/* make the n-th following scratch allocation fail with HZK_ERR_NOMEM (-1 = off) */
int hzk_inject_alloc_failure(hzk_ctx* ctx, long after_n);

#ifdef __cplusplus
}
#endif

#endif /* HZK_H */
