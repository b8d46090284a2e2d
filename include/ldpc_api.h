/*
 * ldpc_api.h -- C ABI of the MI355X QC-LDPC decoder library (libnrldpc.so).
 *
 * Plain C: opaque context, status codes, caller-owned buffers, plain pointers and sizes.
 * Device pointers are HIP device memory; `stream` is a hipStream_t passed as void*
 * (NULL = the default stream).  Every *_async call only enqueues work on `stream`.
 * A context may be used from several host threads as long as calls on it are not
 * concurrent with ldpc_ctx_destroy / ldpc_set_launch / ldpc_set_launch_layout / ldpc_autotune /
 * ldpc_set_layout_model / ldpc_noms_set_weights.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference checkout).
 */
#ifndef NRLDPC_LDPC_API_H
#define NRLDPC_LDPC_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------- status codes */
#define LDPC_OK 0
#define LDPC_EINVAL (-1)       /* bad argument: unsupported (kind, Z, rows), sizes, NULLs */
#define LDPC_ENOMEM (-2)       /* allocation failed or batch larger than the context allows */
#define LDPC_EHIP (-3)         /* HIP runtime error; see ldpc_last_error() */
#define LDPC_ENODEV (-4)       /* no usable GPU */
#define LDPC_EUNSUPPORTED (-5) /* valid request that this build does not implement */

/* ----------------------------------------------------------------- codes */
#define LDPC_CODE_BG1 1        /* 5G NR base graph 1: 46 x 68, K = 22 Z */
#define LDPC_CODE_BG2 2        /* 5G NR base graph 2: 42 x 52, K = 10 Z */
#define LDPC_CODE_QC12X24 3    /* 12 x 24 QC example code (802.11n rate-1/2 prototype), K = 12 Z */

#define LDPC_FLAG_EARLY_EXIT 1u /* stop a codeword once its hard decision has zero syndrome */
#define LDPC_FLAG_FALLBACK_PLAIN 2u /* learned decode: re-decode syndrome failures with plain min-sum */
#define LDPC_FLAG_LEARNED 4u    /* mixed batches: learned weights where set, plain min-sum elsewhere */

typedef struct ldpc_ctx ldpc_ctx;

typedef struct {
    int kind, Z, ils;  /* code kind, lifting size, TS 38.212 lifting-set index */
    int rows, cols;    /* base-graph rows / columns in use */
    int kb;            /* information block columns */
    int punct;         /* punctured leading block columns (2 for NR, 0 for the 12x24 code) */
    int n_vn;          /* decoder columns = cols * Z (LLR entries per codeword) */
    int K;             /* information bits = kb * Z (incl. punctured) */
    int N;             /* transmitted bits = (cols - punct) * Z */
    int hd_bytes;      /* packed hard-decision bytes per codeword = ceil(n_vn / 8) */
    int nnz;           /* block edges in use */
    int cw_per_block;  /* codewords per workgroup used by the decoder launch */
    int nnz_full;      /* block edges of the whole base graph (learned-weight row length) */
} ldpc_code_info;

const char* ldpc_strerror(int status);
/* Message of the last failing call on this host thread ("" if none). */
const char* ldpc_last_error(void);
int ldpc_device_count(int* count);

/* Create a decoder context for one code on one device.
 * Replaces the one-code-per-process setup of the reference:
 *   setup_ldpc                         src/nr_ldpc.cpp:31-230
 *   put_h_compact_and_params_in_constant_memory  src/setup_exp.cu:286-312
 *   allocate_constant_memory / put_params_in_constant_mem  src/ldpc_kernels.cu:29-57
 * m_rows: kept base-graph parity rows (rate matching by row truncation); <= 0 keeps all. */
int ldpc_ctx_create(ldpc_ctx** out, int kind, int Z, int m_rows, int device);
int ldpc_ctx_destroy(ldpc_ctx* ctx);
int ldpc_ctx_info(const ldpc_ctx* ctx, ldpc_code_info* info);

/* Decode n_cw codewords with flooding scaled min-sum (alpha = 0.75), n_iter >= 1 iterations
 * (LDPC_EINVAL otherwise; n_cw == 0 launches nothing and returns LDPC_OK).
 * Replaces perform_gpu_ldpc_decoding<T>  src/gpu_decoder_top.cu:16-52 and the kernels
 * cnp_kernel_1st_iter / cnp_kernel / vnp_kernel_normal / vnp_kernel_last_iter /
 * pack_hard_decision  src/ldpc_kernels.cu:72-500.
 *   llr        device, [n_cw][n_vn] float32 channel LLRs (punctured columns = 0)
 *   hd_packed  device, [n_cw][hd_bytes]; bit k of byte b = hard decision of column 8b+k
 *              (1 when APP <= 0), the bit order of unpack_char src/host_helpers.cpp:60-64
 *   app_out    device, [n_cw][n_vn] posterior LLRs, or NULL
 *   iters_out  device, [n_cw] iterations run, or NULL
 *   flags      LDPC_FLAG_EARLY_EXIT or 0 */
int ldpc_decode_async(ldpc_ctx* ctx, void* stream, const float* llr, int n_cw, int n_iter,
                      unsigned flags, uint8_t* hd_packed, float* app_out, int* iters_out);

/* The reference's HALF_PRC build: binary16 channel LLRs (IEEE bit patterns, [n_cw][n_vn]) and
 * R, dt and the posterior between iterations stored in binary16 where the reference stores them
 * (src/ldpc_kernels.cu HALF_PRC branches: :158-160, :258-266, :304-306, :351, :385-386), all
 * arithmetic in single precision; outputs as for ldpc_decode_async (app_out: the final
 * single-precision posterior).  It runs in the launch layout a single-precision decode of the
 * same batch would use (ldpc_set_launch_layout). */
int ldpc_decode_half_async(ldpc_ctx* ctx, void* stream, const uint16_t* llr_half, int n_cw, int n_iter,
                           unsigned flags, uint8_t* hd_packed, float* app_out, int* iters_out);
/* The reference's CHAR_PRC build (its default, src/globals.h:12-13): signed-char channel LLRs
 * ([n_cw][n_vn], as quantize(0.5, x, 8) makes them, src/host_helpers.cpp:46-58, used as the LLR
 * values), R and dt stored as signed chars and the posterior between iterations clamped to
 * [-128, 127] (src/ldpc_kernels.cu CHAR_PRC branches: :173-175, :258-264, :304-306, :387-390), a
 * saturated posterior (127 or -128) entering the check node without R_old (:260-262).  float ->
 * char conversions truncate toward zero; where the value lies outside [-128, 127], which C++
 * leaves undefined, they saturate.  Outputs as for ldpc_decode_async (app_out: the final
 * posterior before clamping). */
int ldpc_decode_char_async(ldpc_ctx* ctx, void* stream, const int8_t* llr_i8, int n_cw, int n_iter,
                           unsigned flags, uint8_t* hd_packed, float* app_out, int* iters_out);
/* int8 LLRs (as the reference's quantize(D, x, 8) makes them, src/host_helpers.cpp:46-58) ->
 * float LLRs out[i] = in[i] * scale (scale = D), n values. */
int ldpc_llr_from_int8_async(void* stream, const int8_t* in, size_t n, float scale, float* out);
/* float LLRs -> binary16 bit patterns (round to nearest even), n values; the reference's
 * reduce_precision<half>, src/setup_exp.cu:33-38, on the device. */
int ldpc_llr_to_half_async(void* stream, const float* llr, size_t n, uint16_t* out);

/* Launch shape of plain and learned decodes: an LDS layout and codewords per workgroup.
 *   LDPC_LAYOUT_DUP   per codeword a read copy of the posteriors with every sub-column stored
 *                     twice (the circulant read never wraps) and an accumulator, the degree-1
 *                     (extension parity) posteriors in registers; up to 512 threads per workgroup,
 *                     one workgroup per cw_per_block codewords
 *   LDPC_LAYOUT_SLIM  a single read copy and the accumulator, the degree-1 posteriors in a device
 *                     buffer of the context (one slot per CU); up to 768 threads (three
 *                     wavefronts per SIMD), a persistent grid of one workgroup per CU
 * Both give identical outputs.  LDPC_LAYOUT_AUTO (cw_per_block 0) chooses per batch the layout
 * with the shorter modelled time: ceil(n_cw / codewords per round) rounds of a measured round time,
 * each layout at its fastest codewords per workgroup, from the table built into the library
 * (measured on MI355X) or from ldpc_autotune.  cw_per_block 0 with DUP or SLIM takes that layout's
 * tuned count.  Decodes in the slim layout on one context run one after another on the GPU,
 * whatever stream issues them (under stream capture that order is not recorded: graphs sharing one
 * context must not run concurrently).  A set-up call: it may allocate the slim layout's buffer.
 * No reference counterpart: the reference fixes n_cw_per_mcw at build time
 * (setup_gpu_grid src/setup_exp.cu:131-166). */
#define LDPC_LAYOUT_AUTO 0
#define LDPC_LAYOUT_DUP 1
#define LDPC_LAYOUT_SLIM 2
int ldpc_set_launch_layout(ldpc_ctx* ctx, int layout, int cw_per_block);

/* Codewords per workgroup of the decode launch: counts that fit the dup layout select it, larger
 * ones the slim layout (at BG1 Z=384: 1 dup, 2 slim); 0 restores the per-batch choice
 * (LDPC_LAYOUT_AUTO).  Same reference note as ldpc_set_launch_layout. */
int ldpc_set_launch(ldpc_ctx* ctx, int cw_per_block);

/* The launch shape a plain or learned decode of n_cw codewords on ctx uses now: layout (may be
 * NULL), codewords per workgroup and workgroups.  No reference counterpart (setup_gpu_grid
 * src/setup_exp.cu:131-166 computes the reference's fixed grid). */
int ldpc_launch_layout(const ldpc_ctx* ctx, int n_cw, int* layout, int* cw_per_block, int* workgroups);
int ldpc_launch_shape(const ldpc_ctx* ctx, int n_cw, int* cw_per_block, int* workgroups);

/* Autotuner of the launch shape: times every candidate (layout, codewords per workgroup) on whole
 * rounds of the device (every CU busy; n_iter fixed iterations, no early exit, random channel
 * LLRs), installs the fastest of each layout as ctx's cost model and returns to the per-batch
 * choice (LDPC_LAYOUT_AUTO).  Candidates: per layout each count that fills one more wavefront and
 * the largest that fits.  Candidate k (k < max_candidates) is reported as cand_layout[k],
 * cand_cpb[k], cand_cw_per_cu[k] (codewords per CU in a round) and cand_round_us[k] (microseconds
 * per round, the median of three launches); any of them may be NULL.  A set-up call: it allocates
 * up to 1 GiB of temporaries and synchronises the device.  No reference counterpart (the reference
 * fixes the grid at build time, src/setup_exp.cu:131-166). */
int ldpc_autotune(ldpc_ctx* ctx, int n_iter, int max_candidates, int* cand_layout, int* cand_cpb,
                  int* cand_cw_per_cu, float* cand_round_us, int* n_candidates);

/* The context's launch cost model of one layout (LDPC_LAYOUT_DUP or LDPC_LAYOUT_SLIM): the
 * codewords per workgroup that layout launches with, the codewords per CU in one round and the
 * round time in microseconds, as installed from the table built into the library, by
 * ldpc_autotune or by ldpc_set_layout_model.  cw_per_block 0 means the layout is not in the model
 * (the slim layout of a code the table has no slim entry for).  ldpc_set_layout_model installs a
 * model measured earlier -- an autotune cache a caller keeps across processes -- without timing
 * again, and returns to the per-batch choice (LDPC_LAYOUT_AUTO) as ldpc_autotune does; it refuses
 * counts beyond the layout's workgroup or LDS limit, cw_per_cu < 1 and round times that are not
 * positive and finite, and takes cw_per_block 0 only for the slim layout (which drops it from the
 * model).  A set-up call: a slim model may allocate the slim layout's buffer.  No reference
 * counterpart (the reference fixes the grid at build time, src/setup_exp.cu:131-166). */
int ldpc_get_layout_model(const ldpc_ctx* ctx, int layout, int* cw_per_block, int* cw_per_cu, float* round_us);
int ldpc_set_layout_model(ldpc_ctx* ctx, int layout, int cw_per_block, int cw_per_cu, float round_us);

/* Repack per-codeword packed decisions ([n_cw][hd_bytes], as ldpc_decode_async writes them) into
 * one bit stream over the concatenated [n_cw][n_vn] columns: bit g of out (byte g/8, bit g%8) is
 * column g % n_vn of codeword g / n_vn; out holds ceil(n_cw * n_vn / 8) bytes.  This is the
 * layout of pack_hard_decision src/ldpc_kernels.cu:482-500 (identical to the per-codeword
 * layout when n_vn % 8 == 0). */
int ldpc_hd_to_stream_async(ldpc_ctx* ctx, void* stream, const uint8_t* hd_packed, int n_cw, uint8_t* out);

/* Syndrome weight (unsatisfied checks) of packed hard decisions, one int per codeword.
 * No reference counterpart (early termination is a TODO in the reference, README.md:193). */
int ldpc_syndrome_async(ldpc_ctx* ctx, void* stream, const uint8_t* hd_packed, int n_cw,
                        int* weight_out);

/* ----------------------------------------------------------------- host-side code utilities */

/* Raw base-matrix table (rows x cols ints, -1 = zero block) for (kind, ils).
 * Replaces the h_base_{1,2}_i{0..7} globals  src/h_matrices/h_bg_*.h (with the table
 * defects of BG1 i_LS 2..7 repaired). */
int ldpc_base_matrix(int kind, int ils, int* out, int out_len);

/* TS 38.212 parameter derivation for payload A and nominal rate.
 * Replaces setup_ldpc  src/nr_ldpc.cpp:31-230.  out[10] = {bg, Z, K, N, H_rows, H_cols,
 * ils, N_before_puncture, n_nz_in_row, n_nz_in_col}. */
int ldpc_setup_params(int A, float rate, int* out);

/* Systematic encoder on the host: info [n_cw][K] bits (one byte each) -> codeword
 * [n_cw][n_vn] bits (one byte each, punctured columns included) with H c = 0.
 * Replaces ldpc_encoder_orig  src/nr_ldpc.cpp:655-913 (which, for BG2 at full rate, leaves
 * the last 20 parity rows unencoded; this encoder solves every kept row). */
int ldpc_encode_host(const ldpc_ctx* ctx, const uint8_t* info, int n_cw, uint8_t* cw);

/* The same systematic encoder on the device: info (device, [n_cw][K] bits, one byte each; only
 * bit 0 of each byte is read) -> cw (device, [n_cw][n_vn] bits, one byte each), equal to
 * ldpc_encode_host's output.  One workgroup per codeword runs the context's encoding program
 * (the closed form of the quasi-cyclic parity equations) in LDS.
 * Replaces the encoding step of transmit  src/encoding_tools.cpp:67-83, i.e. ldpc_encoder_orig
 * src/nr_ldpc.cpp:655-913 with the generator shifts of src/Gen_shift_value.h, run per codeword on
 * the host by generate_new_codewords  src/setup_exp.cu:168-196. */
int ldpc_encode_async(ldpc_ctx* ctx, void* stream, const uint8_t* info, int n_cw, uint8_t* cw);

/* Generator shift lists regenerated from H for (kind, Z): fills no_shift[rows*kb] and
 * pointer_shift[rows*kb] (as the OAI tables lay them out) and returns the number of shift
 * values written to gen_shift (at most gen_len), or a negative status.
 * Replaces choose_generator_matrix  src/nr_ldpc.cpp:234-652 and the tables of
 * src/Gen_shift_value.h. */
int ldpc_generator_shifts(int kind, int Z, short* no_shift, short* pointer_shift, short* gen_shift,
                          int gen_len);

/* ----------------------------------------------------------------- learned decoder
 * Learned offset/normalized min-sum: R = alpha[t][e] * sign * max(m - beta[t][e], 0), one
 * (alpha, beta) per iteration t and base-graph edge e (row-major edge order of the whole base
 * graph, nnz_full per iteration).  alpha = 0.75, beta = 0 reproduces the plain decoder bit for
 * bit.  No reference counterpart (the reference hard-codes alpha = 0.75,
 * src/ldpc_kernels.cu:164,295). */

/* Set the weights for n_iter iterations: host array [n_iter][nnz_full][2] = (alpha, beta). */
int ldpc_noms_set_weights(ldpc_ctx* ctx, int n_iter, const float* weights);
/* Copy the current weights to the host ([n_iter][nnz_full][2]); *n_iter receives T. */
int ldpc_noms_get_weights(ldpc_ctx* ctx, int* n_iter, float* weights);

/* Decode with the learned weights (n_iter must equal the weights' iteration count).  With
 * LDPC_FLAG_FALLBACK_PLAIN, codewords whose learned hard decision fails the syndrome check are
 * decoded again with plain min-sum (n_iter iterations); the plain outputs (hard decisions, APP,
 * iterations) replace the learned ones for the codewords whose plain hard decision passes the
 * syndrome check.  Outputs as for ldpc_decode_async.  The fallback uses the context's scratch
 * memory: fallback calls (and ldpc_noms_train_grad_async) on one context run one after another on
 * the GPU, whatever thread or stream issues them (each call's stream waits for the previous
 * call's use of the scratch).  Under stream capture that wait is not recorded; graphs that
 * share one context must then not be launched concurrently. */
int ldpc_decode_learned_async(ldpc_ctx* ctx, void* stream, const float* llr, int n_cw, int n_iter,
                              unsigned flags, uint8_t* hd_packed, float* app_out, int* iters_out);

/* Size of the fixed device array that ldpc_decode_learned_noalloc_async uses for its
 * temporaries: one array per GPU, held from the library's first use of that GPU until the
 * process ends. */
#define LDPC_NOALLOC_SCRATCH_BYTES (32u << 20)

/* ldpc_decode_learned_async that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, outputs and results as
 * ldpc_decode_learned_async; with LDPC_FLAG_FALLBACK_PLAIN the fallback's temporaries live in the
 * fixed per-GPU array (LDPC_NOALLOC_SCRATCH_BYTES) instead of the context's grow-only scratch, and
 * the call makes no hipMalloc or hipFree and creates no event (a slim-layout decode at BG1 Z=384
 * uses the degree-1 buffer and ordering event its context allocated when it was created; see
 * ldpc_set_launch).  Returns LDPC_ENOMEM, enqueueing nothing, when
 * the batch needs more than that array (per codeword about 16 bytes + hd_bytes, plus n_vn floats
 * with app_out: at BG1 Z=384 up to about 10,200 codewords without app_out, 310 with it).  Every
 * call on one GPU shares that GPU's array, whatever context or stream issues it, and finds it
 * holding the previous call's contents (nothing in it needs clearing): calls that overlap on one
 * GPU race and give wrong results, so the caller orders them. */
int ldpc_decode_learned_noalloc_async(ldpc_ctx* ctx, void* stream, const float* llr, int n_cw, int n_iter,
                                      unsigned flags, uint8_t* hd_packed, float* app_out, int* iters_out);

/* Workspace bytes for ldpc_noms_train_grad_async with n_cw codewords per call. */
size_t ldpc_noms_workspace_bytes(const ldpc_ctx* ctx, int n_cw, int n_iter);

/* One gradient evaluation on the device: learned forward with training records, mean BCE of
 * sigmoid(-APP) against cw_bits, hand-written backward.  grad [n_iter][nnz_full][2] and loss[1]
 * (device float) are ACCUMULATED (clear them first).  workspace: device, at least
 * ldpc_noms_workspace_bytes(ctx, n_cw, n_iter) bytes. */
int ldpc_noms_train_grad_async(ldpc_ctx* ctx, void* stream, const float* llr, const uint8_t* cw_bits,
                               int n_cw, int n_iter, float* grad, float* loss, void* workspace,
                               size_t workspace_bytes);
/* ldpc_noms_train_grad_async with a loss weight per codeword: cw_weight (device, [n_cw], or NULL
 * for all 1) multiplies each codeword's BCE terms in loss and gradient (the mean is still taken
 * over n_cw codewords).  Used to weight the training frames, e.g. by their Eb/N0.  No reference
 * counterpart. */
int ldpc_noms_train_grad_weighted_async(ldpc_ctx* ctx, void* stream, const float* llr, const uint8_t* cw_bits,
                                        const float* cw_weight, int n_cw, int n_iter, float* grad, float* loss,
                                        void* ws, size_t ws_bytes);

/* Adam update of the context's weights in place from grad [n_iter][nnz_full][2]; m and v are
 * caller-owned device moment buffers of the same size (zero before step 1). */
int ldpc_noms_adam_async(ldpc_ctx* ctx, void* stream, const float* grad, float* m, float* v, int step,
                         float lr);

/* ----------------------------------------------------------------- on-device channel */

/* Noise standard deviation for Eb/N0 (dB) and nominal rate.
 * Replaces compute_noise_stddev  src/encoding_tools.cpp:27-38. */
float ldpc_noise_stddev(float ebn0_db, float rate);

/* Random codewords and channel LLRs generated on the device with Philox4x32-10:
 * info bits from (seed, stream_base + i), systematic encoding, BPSK (0 -> +1),
 * y = x + sigma * n, LLR = (2 y) / (sigma^2), punctured columns 0.
 * sigma: one value per codeword (device array) or, if sigma_dev is NULL, sigma_all.
 *   cw_bits   device [n_cw][n_vn] bytes (may be NULL)   llr  device [n_cw][n_vn] floats
 * Replaces the host chain generate_new_codewords -> transmit -> calculate_llr
 *   src/setup_exp.cu:168-196, src/encoding_tools.cpp:13-99. */
int ldpc_channel_async(ldpc_ctx* ctx, void* stream, uint64_t seed, uint64_t stream_base, int n_cw,
                       const float* sigma_dev, float sigma_all, uint8_t* cw_bits, float* llr);

/* Bit / frame errors of decoded codewords against cw_bits over the first K columns
 * (info bits incl. punctured ones), accumulated into counts[2] = {bit_errors, frame_errors}
 * (device int64).  Replaces cuda_error_check  src/encoding_tools.cpp:112-155. */
int ldpc_count_errors_async(ldpc_ctx* ctx, void* stream, const uint8_t* hd_packed,
                            const uint8_t* cw_bits, int n_cw, long long* counts);

/* ----------------------------------------------------------------- mixed batches
 * Codewords of different base graphs, lifting sizes and rates decoded by ONE persistent launch
 * with per-codeword early exit.  No reference counterpart: the reference decodes one code per
 * process (the code lives in __constant__ memory, src/ldpc_kernels.cu:29-57).
 *
 * Usage: create once, plan a batch from host descriptors, then decode (directly, or through a
 * captured HIP graph that is replayed for every later plan of the same object). */
typedef struct ldpc_mixed ldpc_mixed;

typedef struct {
    int kind;              /* LDPC_CODE_BG1 or LDPC_CODE_BG2 */
    int Z;                 /* lifting size (TS 38.212 set) */
    int m_rows;            /* parity rows kept: 4..46 (BG1), 4..42 (BG2); rate kb/(kb + m_rows - 2) */
    int reserved;          /* 0 */
    long long llr_offset;  /* float index of the codeword's (kb + m_rows) * Z channel LLRs; app_out
                              uses the same offsets */
    long long hd_offset;   /* byte offset of its ceil((kb + m_rows) * Z / 8) packed hard decisions */
} ldpc_mixed_cw;

/* device < 0: host-only object (planning and plan inspection, no decoding).  max_cw bounds the
 * codewords of one plan. */
int ldpc_mixed_create(ldpc_mixed** out, int device, int max_cw);
int ldpc_mixed_destroy(ldpc_mixed* p);

/* Validate and plan a batch (host work: grouping by (kind, Z), tiling by codewords per
 * workgroup, longest-first order).  Waits until the previous plan's upload has been consumed. */
int ldpc_mixed_plan(ldpc_mixed* p, const ldpc_mixed_cw* cws, int n_cw);

/* Plan summary: tiles, launch threads, grid, LDS bytes; any pointer may be NULL. */
int ldpc_mixed_plan_info(const ldpc_mixed* p, int* n_tiles, int* threads, int* grid, int* lds_bytes);
/* Tiles of the current plan: out[t*6 + 0..5] = kind, Z, m_rows, cpb, first, count; the codeword
 * order (descriptor index per planned slot) goes to order[n_cw] when non-NULL.  Returns the tile
 * count (or a negative status). */
int ldpc_mixed_tiles(const ldpc_mixed* p, int* out, int max_tiles, int* order);

/* Decode the planned batch: llr / hd / app_out are device buffers addressed by the descriptors'
 * offsets; iters_out[i] (device, may be NULL) is the iteration count of descriptor i.
 * Launches of one object (decodes and replays) share its device plan and work counter: each one
 * waits on the GPU for the previous one, whatever stream either runs on (except on a stream under
 * capture, where the caller must order them).  Objects are independent: use one per stream for
 * concurrent decodes.
 * flags: LDPC_FLAG_EARLY_EXIT; LDPC_FLAG_LEARNED (codes with weights set use them, the others
 * plain min-sum); with LDPC_FLAG_LEARNED, LDPC_FLAG_FALLBACK_PLAIN re-decodes, inside the same
 * launch, the learned codewords whose final hard decision fails the syndrome check with plain
 * min-sum and keeps the plain result where it passes (as ldpc_decode_learned_async). */
int ldpc_mixed_decode_async(ldpc_mixed* p, void* stream, const float* llr, int n_iter, unsigned flags,
                            uint8_t* hd, float* app_out, int* iters_out);

/* Build a HIP graph owned by p: the plan upload and the decode (as ldpc_mixed_decode_async).
 * Each replay enqueues, on its stream, an optional host->device copy of llr_host into llr (at
 * most llr_floats floats), the graph, and an optional device->host copy of hd into hd_host (at
 * most hd_bytes); the two batch copies are stream copies (DMA engines) so that replays on
 * different streams overlap them with each other's decodes.  Host buffers must be pinned
 * (hipHostMalloc).  llr_floats and hd_bytes are the sizes of llr in elements of the input type
 * (ldpc_mixed_set_input; app_out holds as many floats) and of hd in bytes; they are required with
 * a host buffer and optional (0) without one.
 * Each replay decodes the plan current at that time, into the same buffers, and cuts the copies
 * to the LLR floats and hard-decision bytes that plan addresses; it returns LDPC_EINVAL, and
 * enqueues nothing, when that plan addresses more than the sizes given here or holds learned
 * weights for another iteration count than n_iter.  The graph
 * keeps the buffer addresses given here: every one of them (llr, hd, app_out, iters_out,
 * llr_host, hd_host) must stay allocated, and be used for nothing else while a replay runs, until
 * the last replay has finished or p is destroyed.  `stream` is unused (the graph runs on the
 * replay's stream). */
int ldpc_mixed_capture(ldpc_mixed* p, void* stream, const float* llr, int n_iter, unsigned flags, uint8_t* hd,
                       float* app_out, int* iters_out, const float* llr_host, size_t llr_floats,
                       uint8_t* hd_host, size_t hd_bytes);
int ldpc_mixed_replay(ldpc_mixed* p, void* stream);

/* Element type of the channel LLRs of later ldpc_mixed_decode_async / ldpc_mixed_capture calls on p
 * (a capture keeps the type current at capture time).  The descriptors' llr_offset and the
 * capture's llr_floats then count elements of that type; app_out stays float at the same offsets.
 * The decode runs in single precision on the values converted on load:
 *   LDPC_LLR_F32  float (the default)
 *   LDPC_LLR_I8   signed char q, LLR = q * scale (quantize(D, x, 8) values with scale = D,
 *                 src/host_helpers.cpp:46-58); scale must be positive and finite
 *   LDPC_LLR_F16  IEEE binary16 bit patterns (the reference's reduce_precision<half>,
 *                 src/setup_exp.cu:33-38)
 * The narrower types cut the host-to-device bytes of a batch by 4x / 2x.  A narrow batch is
 * converted by one small kernel into a float buffer the object owns (grown, with an allocation,
 * by the first launch or capture that needs more; a capture sizes it to llr_floats, which narrow
 * inputs require) and decoded from there.  No reference counterpart for mixed batches. */
#define LDPC_LLR_F32 0
#define LDPC_LLR_I8 1
#define LDPC_LLR_F16 2
int ldpc_mixed_set_input(ldpc_mixed* p, int format, float scale);

/* Resident workgroups of the persistent grid of later launches and captures of p (0 = every
 * workgroup the device holds at once, the default; larger values are cut to it).  Objects that
 * decode concurrently on different streams can each take their share of the device instead of
 * contending with full-device grids. */
int ldpc_mixed_set_grid(ldpc_mixed* p, int workgroups);

/* Learned weights of one (kind, Z) for LDPC_FLAG_LEARNED decodes: host [n_iter][nnz_full][2]
 * (rows beyond a codeword's parity-row count are unused).  Plan again after setting them. */
int ldpc_mixed_set_weights(ldpc_mixed* p, int kind, int Z, int n_iter, const float* weights);

/* Launch layout of every tile of the object's later plans (the single-code layouts of
 * ldpc_set_launch_layout): LDPC_LAYOUT_DUP, 512-thread workgroups with the degree-1 posteriors in
 * registers; LDPC_LAYOUT_SLIM, 768-thread workgroups (three wavefronts per SIMD) with a single LDS
 * read copy and the degree-1 posteriors in a device buffer of the object (one 129 KB slot per
 * resident workgroup, allocated here); LDPC_LAYOUT_AUTO, the library's default.  Codewords per
 * workgroup are kept per layout.  Discards the current plan and any captured graph (capture again).
 * A set-up call: it may allocate and waits for the object's previous launch. */
int ldpc_mixed_set_layout(ldpc_mixed* p, int layout);
int ldpc_mixed_get_layout(const ldpc_mixed* p, int* layout);
/* Layout autotuner for a representative batch: plans cws (n_cw descriptors addressing llr and hd,
 * device buffers as for ldpc_mixed_decode_async), times both layouts on it with the caller's
 * iteration count and flags (the mean of three launches after two warm-ups,
 * microseconds per batch to us_dup / us_slim, either may be NULL) and keeps the faster
 * (best_layout).  The early-exit decode of a realistic batch and the fixed-iteration decode can
 * prefer different layouts: larger slim tiles run until their slowest codeword stops.  A set-up
 * call: it synchronises the device, discards the plan and any captured graph. */
int ldpc_mixed_autotune_layout(ldpc_mixed* p, const ldpc_mixed_cw* cws, int n_cw, const float* llr, int n_iter,
                               unsigned flags, uint8_t* hd, float* us_dup, float* us_slim, int* best_layout);

/* Codewords per workgroup (LDS codeword slots) of (kind, Z) in the object's current layout;
 * 0 = default (the most that fit).
 * ldpc_mixed_autotune times the candidate counts (each count that fills one more wavefront,
 * 1 and the maximum) on n_cw full-rate codewords for n_iter iterations, keeps the fastest, and
 * writes candidate k's count and microseconds per codeword to cand_cpb[k], us_per_cw[k] (k <
 * max_candidates; either may be NULL).  It discards the current plan. */
int ldpc_mixed_set_cpb(ldpc_mixed* p, int kind, int Z, int cpb);
int ldpc_mixed_get_cpb(ldpc_mixed* p, int kind, int Z, int* cpb);
int ldpc_mixed_autotune(ldpc_mixed* p, int kind, int Z, int n_iter, int n_cw, int* cand_cpb, float* us_per_cw,
                        int max_candidates, int* n_candidates, int* best_cpb);
/* Tuning cache: text lines "kind Z cpb" (dup layout) and "kind Z cpb slim" (slim layout). */
int ldpc_mixed_save_tuning(const ldpc_mixed* p, const char* path);
int ldpc_mixed_load_tuning(ldpc_mixed* p, const char* path);

/* ----------------------------------------------------------------- transport blocks
 * TS 38.212 around the decoder: CRC (5.1), code-block segmentation with C > 1 (5.2.2), filler
 * bits (5.3.2), rate matching for LDPC (5.4.2: bit selection from the full circular buffer,
 * I_LBRM = 0, one layer, and bit interleaving).  No reference counterpart: the reference
 * handles one code block (setup_ldpc rejects C > 1, src/nr_ldpc.cpp:78-81) and no rate
 * matching. */
#define LDPC_CRC16 0   /* gCRC16 */
#define LDPC_CRC24A 1  /* gCRC24A */
#define LDPC_CRC24B 2  /* gCRC24B */

/* CRC parity of bits[0..n) (one bit per byte, bit 0 first), register initialised to 0: the
 * parity bits p_0..p_{L-1} are bits L-1..0 of *crc. */
int ldpc_crc(int type, const uint8_t* bits, long long n, uint32_t* crc);

typedef struct ldpc_tb ldpc_tb;
typedef struct {
    int A, L_tb, B;     /* payload bits, TB CRC bits (16 or 24), A + L_tb */
    int bg, C, L_cb;    /* base graph, code blocks, code-block CRC bits (0 or 24) */
    int K_prime, K, Z;  /* bits per block before filler (incl. CRC), kb*Z, lifting size */
    int F, N, n_vn;     /* filler bits per block, circular buffer (cols-2)*Z, decoder columns */
    int G, Qm, rv, k0;  /* rate matching: coded bits, modulation order, redundancy version, k0 */
} ldpc_tb_info;

/* Configure a transport block of A bits at nominal rate `rate` (base graph per TS 38.212 7.2.2
 * when bg = 0), rate-matched to G bits (multiple of Qm) with redundancy version rv.
 * device < 0: host only (encoding); device >= 0 also uploads the de-rate-matching map. */
int ldpc_tb_create(ldpc_tb** out, int A, float rate, int bg, int G, int Qm, int rv, int device);
int ldpc_tb_destroy(ldpc_tb* tb);
/* Parameters, and the rate-matched length E_r of code blocks r < max_e into E (nullable). */
int ldpc_tb_get_info(const ldpc_tb* tb, ldpc_tb_info* info, int* E, int max_e);
/* Host: payload a[A] (one bit per byte) -> transmitted bits f[G] (CRC, segmentation, LDPC
 * encoding, rate matching, interleaving); cw[C][n_vn] (nullable) receives the code blocks. */
int ldpc_tb_encode_host(const ldpc_tb* tb, const uint8_t* a, uint8_t* f, uint8_t* cw);
/* Device: received LLRs llr_in[G] (in transmitted order) -> decoder LLRs llr_cb[C][n_vn]:
 * repetitions summed in bit-selection order, filler bits +10000 (known 0), punctured and
 * unsent columns 0.  Decode llr_cb with a context of ldpc_tb_decoder_context. */
int ldpc_tb_derate_async(const ldpc_tb* tb, void* stream, const float* llr_in, float* llr_cb);
/* Device: packed decisions hd[C][ceil(n_vn/8)] -> cb_ok[C] (CRC24B of each block; 1 when
 * C = 1) and, if tb_bits is not NULL, the B transport-block bits (payload + TB CRC, one
 * per byte); check the TB CRC with ldpc_crc over them (remainder 0). */
int ldpc_tb_check_async(const ldpc_tb* tb, void* stream, const uint8_t* hd, int* cb_ok, uint8_t* tb_bits);
/* A decoder context for the transport block's code (bg, Z, all rows) on `device`. */
int ldpc_tb_decoder_context(const ldpc_tb* tb, int device, ldpc_ctx** out);

#ifdef __cplusplus
}
#endif
#endif
