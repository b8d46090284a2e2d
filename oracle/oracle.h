/*
 * oracle.h -- CPU restatement of the reference decoder, for tests and the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or calls this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as
 * the checker.  Every routine cites the reference file:line it restates.
 *
 * Pinning (what checks each routine against something other than itself) is listed per
 * function below and in the header comment of oracle.c.
 */
#ifndef NRLDPC_ORACLE_H
#define NRLDPC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Code kinds: 5G NR base graph 1, base graph 2, and the 12x24 QC example code. */
enum { OR_BG1 = 1, OR_BG2 = 2, OR_WIFI = 3 };

typedef struct {
    int kind, Z, ils;        /* ils: lifting-set index (NR only) */
    int rows, cols;          /* rows/columns in use (after parity-row truncation) */
    int rows_full, cols_full;
    int kb;                  /* information block columns */
    int *V;                  /* rows_full*cols_full, V mod Z or -1 */
} oracle_code;

/* Build the code description; m_rows <= 0 keeps every row.  Returns 0 on success. */
int oracle_code_init(oracle_code *c, int kind, int Z, int m_rows);
void oracle_code_free(oracle_code *c);
/* Shipped-table values V (not reduced mod Z) for (kind, ils): rows_full*cols_full ints. */
int oracle_base_matrix(int kind, int ils, int *out);
int oracle_lifting_set(int Z);  /* TS 38.212 Table 5.3.2-1 set index, -1 if Z is not a lifting size */

/* setup_ldpc parameter derivation, src/nr_ldpc.cpp:31-230.  out[] = {bg, Z, K, N, H_rows,
 * H_cols, ils, N_before_puncture, n_nz_in_row, n_nz_in_col}. */
void oracle_setup_ldpc(int A, float rate, int *out);

/* Flooding scaled min-sum with the reference schedule (src/ldpc_kernels.cu:72-468,
 * src/gpu_decoder_top.cu:16-52).  llr: [n_cw][cols*Z].  hd: [n_cw][cols*Z] bytes (0/1).
 * app (nullable): posterior LLR after the last iteration.  iters (nullable): iterations run.
 * early_exit: stop a codeword after the first iteration whose hard decision has zero
 * syndrome (no reference counterpart; the reference always runs n_iter).
 * threads: worker threads (<=0: one). */
void oracle_decode(const oracle_code *c, const float *llr, int n_cw, int n_iter, int early_exit,
                   uint8_t *hd, float *app, int *iters, int threads);

/* The reference's HALF_PRC build: LLRs, R, dt and the APP between iterations stored as IEEE
 * binary16, arithmetic in single precision (src/ldpc_kernels.cu, HALF_PRC branches). */
void oracle_decode_half(const oracle_code *c, const uint16_t *llr, int n_cw, int n_iter, int early_exit,
                        uint8_t *hd, float *app, int *iters, int threads);
/* The reference's CHAR_PRC build (its default): signed-char LLRs (quantize(0.5, x, 8)), R, dt and
 * the APP between iterations, arithmetic in single precision; float -> char conversions truncate
 * toward zero and saturate out of range (undefined in C++).  Returns the number of out-of-range
 * conversions, the cases where the reference's own result is not defined. */
long long oracle_decode_char(const oracle_code *c, const int8_t *llr, int n_cw, int n_iter, int early_exit,
                             uint8_t *hd, float *app, int *iters, int threads);
uint16_t oracle_float_to_half(float f);  /* round to nearest even */
float oracle_half_to_float(uint16_t h);

/* Learned offset/normalized min-sum: R = alpha[t][e] * sign * max(m - beta[t][e], 0) with one
 * (alpha, beta) per (iteration, base-graph edge in use, [n_iter][nnz]); alpha = 0.75, beta = 0
 * is the plain decoder above.  Single precision, same schedule as oracle_decode (the
 * reference's fixed-iteration loop, src/gpu_decoder_top.cu:25-47, with the constant 0.75 of
 * src/ldpc_kernels.cu:164,295 replaced by the weights), including early_exit, iters and
 * threads as there. */
void oracle_decode_noms(const oracle_code *c, const float *llr, int n_cw, int n_iter, int early_exit,
                        const float *alpha, const float *beta, uint8_t *hd, float *app, int *iters,
                        int threads);

/* Double-precision NOMS forward + mean BCE loss over all cols*Z bits (target bit x, soft
 * output p = sigmoid(-APP)), and its analytic gradient w.r.t. alpha and beta
 * ([n_iter][nnz] each).  grad_* may be NULL.  Returns the loss. */
double oracle_noms_loss_grad(const oracle_code *c, const float *llr, const uint8_t *bits,
                             int n_cw, int n_iter, const double *alpha, const double *beta,
                             double *grad_alpha, double *grad_beta);
/* The same with a loss weight per codeword (cw_weight [n_cw], NULL = all 1): codeword w's BCE
 * terms and their gradient are multiplied by cw_weight[w]; the mean is still over n_cw. */
double oracle_noms_loss_grad_weighted(const oracle_code *c, const float *llr, const uint8_t *bits,
                                      const double *cw_weight, int n_cw, int n_iter, const double *alpha,
                                      const double *beta, double *grad_alpha, double *grad_beta);

/* Syndrome weight of hard decisions hd ([cols*Z] bytes) for one codeword. */
int oracle_syndrome(const oracle_code *c, const uint8_t *hd);

/* Channel helpers restating src/encoding_tools.cpp. */
float oracle_noise_stddev(float ebn0_db, float rate);                        /* :27-38 */
void oracle_llr(const float *recv, int n_symbols, float sigma, int z, float *llr);  /* :85-99 */

/* Philox4x32-10 (Salmon et al., SC'11) and the derived Gaussian stream. */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Noise sample n for (seed, stream, index) exactly as the device generator defines it. */
float oracle_gauss(uint64_t seed, uint64_t stream, uint64_t index);
uint32_t oracle_info_bit_word(uint64_t seed, uint64_t stream, uint64_t word);

#ifdef __cplusplus
}
#endif
#endif
