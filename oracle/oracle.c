/*
 * oracle.c -- CPU restatement of the reference decode path (test infrastructure only).
 *
 * This file is the checker, never the product: the product library does not link or call
 * it, and the product path fails loudly when its HIP build is missing.  It is compiled by
 * oracle/Makefile into oracle/_build/liboracle.so and used by tests/ (as the parity oracle),
 * __graft_entry__.smoke() and bench.py (the CPU baseline).
 *
 * What pins each part (beyond self-consistency):
 *   tables .............. sha256 digests of the shipped and corrected base matrices, and
 *                         H.c = 0 for every (BG, Z) with the product encoder (tests).
 *   setup_ldpc .......... closed-form cases from the reference's own comments / main()
 *                         (A=8424, rate 1/3 -> BG1, Z=384, i_LS=1; A=2064, rate 0.2 -> BG2, Z=208).
 *   decode .............. hand-computed check-node closed forms (ties, zeros, the 1000.0f
 *                         clamp), a dense-matrix numpy restatement on small codes, noiseless
 *                         codewords, and the fixed-point property of zero syndrome (tests).
 *   NOMS ................ alpha=0.75/beta=0 equals the plain decoder bit for bit; the analytic
 *                         gradient matches double-precision finite differences (tests).
 *   Philox4x32-10 ....... the Random123 known-answer vectors (tests).
 *   Gaussian transform .. moment tests; the device generator must equal it bit for bit.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int row, col, v[8]; } nrldpc_bg_entry;
typedef struct { int row, col, v; } nrldpc_qc_entry;
#define NRLDPC_TABLE_QUAL static const
#include "../nr_gpu_ldpc_amd/csrc/bg_tables.inc"

/* ---------------------------------------------------------------- code description */

int oracle_lifting_set(int Z) {
    /* TS 38.212 Table 5.3.2-1: Z = a * 2^j, a in {2,3,5,7,9,11,13,15}, Z <= 384
     * (the reference builds the same table at src/nr_ldpc.cpp:91-106). */
    static const int base[8] = {2, 3, 5, 7, 9, 11, 13, 15};
    for (int i = 0; i < 8; i++)
        for (int z = base[i]; z <= 384; z *= 2)
            if (z == Z) return i;
    return -1;
}

int oracle_base_matrix(int kind, int ils, int *out) {
    int rows, cols;
    if (kind == OR_BG1) { rows = 46; cols = 68; }
    else if (kind == OR_BG2) { rows = 42; cols = 52; }
    else if (kind == OR_WIFI) { rows = 12; cols = 24; }
    else return -1;
    for (int i = 0; i < rows * cols; i++) out[i] = -1;
    if (kind == OR_BG1)
        for (int n = 0; n < NRLDPC_BG1_NNZ; n++) out[kBG1Entries[n].row * cols + kBG1Entries[n].col] = kBG1Entries[n].v[ils];
    else if (kind == OR_BG2)
        for (int n = 0; n < NRLDPC_BG2_NNZ; n++) out[kBG2Entries[n].row * cols + kBG2Entries[n].col] = kBG2Entries[n].v[ils];
    else
        for (int n = 0; n < NRLDPC_WIFI_NNZ; n++) out[kWifiEntries[n].row * cols + kWifiEntries[n].col] = kWifiEntries[n].v;
    return 0;
}

int oracle_code_init(oracle_code *c, int kind, int Z, int m_rows) {
    memset(c, 0, sizeof(*c));
    c->kind = kind;
    c->Z = Z;
    if (kind == OR_BG1) { c->rows_full = 46; c->cols_full = 68; c->kb = 22; }
    else if (kind == OR_BG2) { c->rows_full = 42; c->cols_full = 52; c->kb = 10; }
    else if (kind == OR_WIFI) { c->rows_full = 12; c->cols_full = 24; c->kb = 12; }
    else return -1;
    if (Z < 1 || Z > 384) return -1;
    if (kind == OR_WIFI) {
        c->ils = 0;
        c->rows = 12;
    } else {
        c->ils = oracle_lifting_set(Z);
        if (c->ils < 0) return -1;
        c->rows = (m_rows <= 0) ? c->rows_full : m_rows;
        if (c->rows < 4 || c->rows > c->rows_full) return -1;
    }
    /* Columns in use: info + core + one extension column per kept extension row. */
    c->cols = (kind == OR_WIFI) ? 24 : c->kb + c->rows;
    c->V = (int *)malloc(sizeof(int) * c->rows_full * c->cols_full);
    oracle_base_matrix(kind, c->ils, c->V);
    for (int i = 0; i < c->rows_full * c->cols_full; i++)
        if (c->V[i] >= 0) c->V[i] %= Z; /* TS 38.212 5.3.2: P = V mod Z */
    return 0;
}

void oracle_code_free(oracle_code *c) { free(c->V); c->V = NULL; }

/* ---------------------------------------------------------------- setup_ldpc */

void oracle_setup_ldpc(int A, float rate, int *out) {
    /* src/nr_ldpc.cpp:37-112 */
    const int l = (A > 3824) ? 24 : 16;
    const int B = A + l;
    int bg, K_cb, K_b;
    if ((A <= 292) | ((A < 3824) & (rate <= 0.67)) | (rate <= 0.25)) {
        bg = 2;
        K_cb = 3840;
        K_b = (B > 640) ? 10 : (B > 560) ? 9 : (B > 192) ? 8 : 6;
    } else {
        bg = 1;
        K_cb = 8448;
        K_b = 22;
    }
    int C, B_prime;
    if (B <= K_cb) { C = 1; B_prime = B; }
    else { C = (int)ceil(B / (K_cb - 24)); B_prime = B + C * 24; }
    const int K_prime = B_prime / C;
    const int Z_min = K_prime / K_b; /* integer floor, as the reference computes it */
    int Z_c = 1000, ils = -1;
    static const int base[8] = {2, 3, 5, 7, 9, 11, 13, 15};
    for (int i = 0; i < 8; i++) {
        int t = base[i];
        for (int k = 0; k < 8; k++) {
            const int z = (t < 385) ? t : 0;
            t *= 2;
            if ((z >= Z_min) & (z < Z_c)) { Z_c = z; ils = i; }
        }
    }
    const int K = (bg == 1) ? 22 * Z_c : 10 * Z_c;
    const int N = (bg == 1) ? 66 * Z_c : 50 * Z_c;
    const int H_rows = (bg == 1) ? 46 : 42, H_cols = (bg == 1) ? 68 : 52;
    /* src/nr_ldpc.cpp:188-228: row/column degree maxima of the chosen table */
    int *H = (int *)malloc(sizeof(int) * H_rows * H_cols);
    oracle_base_matrix(bg == 1 ? OR_BG1 : OR_BG2, ils, H);
    int mr = 0, mc = 0;
    for (int i = 0; i < H_rows; i++) {
        int n = 0;
        for (int j = 0; j < H_cols; j++) n += H[i * H_cols + j] != -1;
        if (n > mr) mr = n;
    }
    for (int j = 0; j < H_cols; j++) {
        int n = 0;
        for (int i = 0; i < H_rows; i++) n += H[i * H_cols + j] != -1;
        if (n > mc) mc = n;
    }
    free(H);
    out[0] = bg; out[1] = Z_c; out[2] = K; out[3] = N; out[4] = H_rows; out[5] = H_cols;
    out[6] = ils; out[7] = N + 2 * Z_c; out[8] = mr; out[9] = mc;
}

/* ---------------------------------------------------------------- decoder */

typedef struct {
    int n_edges;  /* block edges in use */
    int *erow, *ecol, *eshift;        /* per block edge, row-major ascending order */
    int *row_start;                   /* rows+1 */
    int *col_edges, *col_start;       /* per column: block edges in ascending block row */
} edge_list;

static void edges_build(const oracle_code *c, edge_list *el) {
    /* Compact row list: nonzeros of each block row in ascending block column
     * (src/setup_exp.cu:241-250); column list in ascending block row (:261-271). */
    int n = 0;
    for (int i = 0; i < c->rows; i++)
        for (int j = 0; j < c->cols; j++) n += c->V[i * c->cols_full + j] >= 0;
    el->n_edges = n;
    el->erow = malloc(sizeof(int) * n);
    el->ecol = malloc(sizeof(int) * n);
    el->eshift = malloc(sizeof(int) * n);
    el->row_start = malloc(sizeof(int) * (c->rows + 1));
    int e = 0;
    for (int i = 0; i < c->rows; i++) {
        el->row_start[i] = e;
        for (int j = 0; j < c->cols; j++) {
            const int v = c->V[i * c->cols_full + j];
            if (v >= 0) { el->erow[e] = i; el->ecol[e] = j; el->eshift[e] = v; e++; }
        }
    }
    el->row_start[c->rows] = e;
    el->col_start = malloc(sizeof(int) * (c->cols + 1));
    el->col_edges = malloc(sizeof(int) * n);
    int k = 0;
    for (int j = 0; j < c->cols; j++) {
        el->col_start[j] = k;
        for (int i = 0; i < c->rows; i++)
            for (int q = el->row_start[i]; q < el->row_start[i + 1]; q++)
                if (el->ecol[q] == j) el->col_edges[k++] = q;
    }
    el->col_start[c->cols] = k;
}

static void edges_free(edge_list *el) {
    free(el->erow); free(el->ecol); free(el->eshift); free(el->row_start);
    free(el->col_edges); free(el->col_start);
}

int oracle_syndrome(const oracle_code *c, const uint8_t *hd) {
    const int Z = c->Z;
    int w = 0;
    for (int i = 0; i < c->rows; i++)
        for (int r = 0; r < Z; r++) {
            int p = 0;
            for (int j = 0; j < c->cols; j++) {
                const int v = c->V[i * c->cols_full + j];
                if (v >= 0) p ^= hd[j * Z + (r + v) % Z] & 1;
            }
            w += p;
        }
    return w;
}

/* One codeword of (learned) scaled min-sum in single precision.  alpha/beta == NULL selects
 * the reference constant 0.75f and no offset, computed exactly as the reference writes it. */
/* IEEE binary16 conversions (round to nearest, ties to even), written out from the format
 * definition: what __float2half / __half2float do for the reference's HALF_PRC build. */
uint16_t oracle_float_to_half(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) /* inf, nan (quiet, payload kept where it fits) */
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520: rounds to inf */
    if (ax >= 0x38800000u) {                                  /* normal half */
        uint32_t m = ax + 0xfffu + ((ax >> 13) & 1u);
        m -= (uint32_t)(127 - 15) << 23;
        return (uint16_t)(sign | (m >> 13));
    }
    if (ax < 0x33000000u) return (uint16_t)sign; /* below half the smallest subnormal */
    const uint32_t e = ax >> 23;
    const uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - e; /* value = q * 2^-24 */
    uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), halfway = 1u << (shift - 1u);
    if (rem > halfway || (rem == halfway && (q & 1u))) q++;
    return (uint16_t)(sign | q);
}

float oracle_half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t x;
    if (e == 31) x = sign | 0x7f800000u | (m << 13);
    else if (e != 0) x = sign | ((e + 112u) << 23) | (m << 13);
    else if (m == 0) x = sign;
    else { /* subnormal: normalise */
        int k = 0;
        uint32_t mm = m;
        while (!(mm & 0x400u)) { mm <<= 1; k++; }
        x = sign | ((uint32_t)(113 - k) << 23) | ((mm & 0x3ffu) << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

static inline float rh(float x) { return oracle_half_to_float(oracle_float_to_half(x)); }

/* float -> signed char of the CHAR_PRC build: truncation toward zero; values outside
 * [-128, 127] (undefined behaviour in C++) saturate, and *oor counts them. */
static inline float to_char(float x, long long *oor) {
    float t = truncf(x);
    if (t > 127.0f) { t = 127.0f; (*oor)++; }
    else if (t < -128.0f) { t = -128.0f; (*oor)++; }
    return t;
}

enum { PREC_FLOAT = 0, PREC_HALF = 1, PREC_CHAR = 2 };

/* prec: PREC_HALF is the HALF_PRC build, where R, dt and the APP between iterations are stored
 * as binary16 (src/ldpc_kernels.cu:158-160,266,304-306,351,385-386) and all arithmetic is single
 * precision; the last iteration's APP is not stored (:450-467).  PREC_CHAR is the CHAR_PRC build
 * (the reference's default, src/globals.h:12-13): LLRs, R, dt and the APP between iterations are
 * signed chars, R and dt converted with to_char (:173-175, :304-306), the APP clamped to
 * [-128, 127] (:385-390), and a saturated APP (127 or -128) enters the check node unchanged
 * (:258-264); *oor counts the out-of-range conversions. */
static void decode_one(const oracle_code *c, const edge_list *el, const float *llr, int n_iter,
                       int early_exit, const float *alpha, const float *beta, float *R, float *dt,
                       float *app, uint8_t *hd, int *iters_out, int prec, long long *oor) {
    const int half_store = prec == PREC_HALF, char_store = prec == PREC_CHAR;
    const int Z = c->Z, nvn = c->cols * Z;
    memcpy(app, llr, sizeof(float) * nvn);
    memset(R, 0, sizeof(float) * el->n_edges * Z);
    int it;
    for (it = 0; it < n_iter; it++) {
        /* Check-node processing over every expanded row: cnp_kernel_1st_iter /
         * cnp_kernel, src/ldpc_kernels.cu:72-181 and :198-314. */
        for (int i = 0; i < c->rows; i++) {
            const int e0 = el->row_start[i], d = el->row_start[i + 1] - e0;
            for (int r = 0; r < Z; r++) {
                float rmin1 = 1000.0f, rmin2 = 1000.0f, sign = 1.0f, Qv[32];
                int q_sign = 0, idx_min = 0;
                for (int k = 0; k < d; k++) {
                    const int e = e0 + k;
                    int s = r + el->eshift[e];
                    if (s >= Z) s -= Z;
                    const int col = el->ecol[e] * Z + s;
                    /* iteration 0 reads the channel LLR (:138); later ones APP - R_old (:266-267) */
                    float Q = (it == 0) ? app[col] : app[col] - R[e * Z + r];
                    if (char_store && (app[col] == 127.0f || app[col] == -128.0f)) Q = app[col];
                    const float Q_abs = fabsf(Q);
                    const int sq = Q < 0;
                    sign = sign * (float)(1 - sq * 2);
                    q_sign |= sq << k;
                    if (Q_abs < rmin1) { rmin2 = rmin1; rmin1 = Q_abs; idx_min = k; }
                    else if (Q_abs < rmin2) { rmin2 = Q_abs; }
                    Qv[k] = Q;
                }
                (void)Qv;
                for (int k = 0; k < d; k++) {
                    const int e = e0 + k;
                    const int sq = 1 - 2 * ((q_sign >> k) & 1);
                    const float m = (k != idx_min) ? rmin1 : rmin2;
                    float Rn;
                    if (!alpha) {
                        Rn = 0.75f * sign * (float)sq * m; /* :164, :295 */
                    } else {
                        const int b = (it * el->n_edges) + e; /* (iteration, base edge) */
                        float off = m - beta[b];
                        if (!(off > 0.0f)) off = 0.0f;
                        Rn = sign * (float)sq * (alpha[b] * off);
                    }
                    const float Ro = (it == 0) ? 0.0f : R[e * Z + r];
                    if (char_store) {
                        const float Rc = to_char(Rn, oor);
                        dt[e * Z + r] = (it == 0) ? Rc : to_char(Rc - Ro, oor); /* :173-174, :304 */
                        R[e * Z + r] = Rc;
                        continue;
                    }
                    const float d = (it == 0) ? Rn : Rn - Ro; /* :177, :308-310 */
                    dt[e * Z + r] = half_store ? rh(d) : d;
                    R[e * Z + r] = half_store ? rh(Rn) : Rn;
                }
            }
        }
        /* Variable-node processing: APP += sum of dt over incident rows in ascending block
         * row (vnp_kernel_normal / vnp_kernel_last_iter, src/ldpc_kernels.cu:329-468). */
        for (int j = 0; j < c->cols; j++) {
            for (int s = 0; s < Z; s++) {
                float a = app[j * Z + s];
                for (int q = el->col_start[j]; q < el->col_start[j + 1]; q++) {
                    const int e = el->col_edges[q];
                    int sf = s - el->eshift[e];
                    if (sf < 0) sf += Z;
                    a += dt[e * Z + sf];
                }
                if (char_store && it + 1 < n_iter) a = a > 127.0f ? 127.0f : (a < -128.0f ? -128.0f : a);
                app[j * Z + s] = (half_store && it + 1 < n_iter) ? rh(a) : a;
            }
        }
        /* Hard decision: APP > 0 -> 0, else 1 (src/ldpc_kernels.cu:463-467). */
        for (int n = 0; n < nvn; n++) hd[n] = app[n] > 0.0f ? 0 : 1;
        if (early_exit && oracle_syndrome(c, hd) == 0) { it++; break; }
    }
    if (iters_out) *iters_out = it;
}

typedef struct {
    const oracle_code *c;
    const edge_list *el;
    const float *llr;
    int n_cw, n_iter, early_exit, tid, nthreads;
    const float *alpha, *beta;
    uint8_t *hd;
    float *app;
    int *iters;
    int prec;
    long long oor;  /* out-of-range char conversions (PREC_CHAR) */
} dec_job;

static void *dec_worker(void *arg) {
    dec_job *jb = (dec_job *)arg;
    const int Z = jb->c->Z, nvn = jb->c->cols * Z;
    float *R = malloc(sizeof(float) * jb->el->n_edges * Z);
    float *dt = malloc(sizeof(float) * jb->el->n_edges * Z);
    float *app = malloc(sizeof(float) * nvn);
    /* static partition of codewords over workers */
    const int per = (jb->n_cw + jb->nthreads - 1) / jb->nthreads;
    const int b = jb->tid * per, e = (b + per < jb->n_cw) ? b + per : jb->n_cw;
    for (int w = b; w < e; w++) {
        decode_one(jb->c, jb->el, jb->llr + (size_t)w * nvn, jb->n_iter, jb->early_exit, jb->alpha,
                   jb->beta, R, dt, app, jb->hd + (size_t)w * nvn, jb->iters ? jb->iters + w : NULL,
                   jb->prec, &jb->oor);
        if (jb->app) memcpy(jb->app + (size_t)w * nvn, app, sizeof(float) * nvn);
    }
    free(R); free(dt); free(app);
    return NULL;
}

static long long decode_many(const oracle_code *c, const float *llr, int n_cw, int n_iter, int early_exit,
                              const float *alpha, const float *beta, uint8_t *hd, float *app, int *iters,
                              int threads, int prec) {
    edge_list el;
    edges_build(c, &el);
    if (threads < 1) threads = 1;
    if (threads > n_cw) threads = n_cw > 0 ? n_cw : 1;
    pthread_t th[256];
    dec_job jobs[256];
    if (threads > 256) threads = 256;
    for (int t = 0; t < threads; t++) {
        dec_job j = {c, &el, llr, n_cw, n_iter, early_exit, t, threads, alpha, beta, hd, app, iters, prec, 0};
        jobs[t] = j;
        if (threads == 1) dec_worker(&jobs[t]);
        else pthread_create(&th[t], NULL, dec_worker, &jobs[t]);
    }
    if (threads > 1)
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    edges_free(&el);
    long long oor = 0;
    for (int t = 0; t < threads; t++) oor += jobs[t].oor;
    return oor;
}

void oracle_decode(const oracle_code *c, const float *llr, int n_cw, int n_iter, int early_exit,
                   uint8_t *hd, float *app, int *iters, int threads) {
    decode_many(c, llr, n_cw, n_iter, early_exit, NULL, NULL, hd, app, iters, threads, PREC_FLOAT);
}

void oracle_decode_half(const oracle_code *c, const uint16_t *llr, int n_cw, int n_iter, int early_exit,
                        uint8_t *hd, float *app, int *iters, int threads) {
    const size_t n = (size_t)n_cw * c->cols * c->Z;
    float *f = malloc(sizeof(float) * (n ? n : 1));
    for (size_t i = 0; i < n; i++) f[i] = oracle_half_to_float(llr[i]);
    decode_many(c, f, n_cw, n_iter, early_exit, NULL, NULL, hd, app, iters, threads, PREC_HALF);
    free(f);
}

long long oracle_decode_char(const oracle_code *c, const int8_t *llr, int n_cw, int n_iter, int early_exit,
                             uint8_t *hd, float *app, int *iters, int threads) {
    const size_t n = (size_t)n_cw * c->cols * c->Z;
    float *f = malloc(sizeof(float) * (n ? n : 1));
    for (size_t i = 0; i < n; i++) f[i] = (float)llr[i];
    const long long oor = decode_many(c, f, n_cw, n_iter, early_exit, NULL, NULL, hd, app, iters, threads, PREC_CHAR);
    free(f);
    return oor;
}

void oracle_decode_noms(const oracle_code *c, const float *llr, int n_cw, int n_iter, int early_exit,
                        const float *alpha, const float *beta, uint8_t *hd, float *app, int *iters,
                        int threads) {
    decode_many(c, llr, n_cw, n_iter, early_exit, alpha, beta, hd, app, iters, threads, PREC_FLOAT);
}

/* ---------------------------------------------------------------- NOMS in double + gradient */

double oracle_noms_loss_grad(const oracle_code *c, const float *llr, const uint8_t *bits, int n_cw,
                             int n_iter, const double *alpha, const double *beta, double *grad_alpha,
                             double *grad_beta) {
    return oracle_noms_loss_grad_weighted(c, llr, bits, NULL, n_cw, n_iter, alpha, beta, grad_alpha, grad_beta);
}

double oracle_noms_loss_grad_weighted(const oracle_code *c, const float *llr, const uint8_t *bits,
                                      const double *cw_weight, int n_cw, int n_iter, const double *alpha,
                                      const double *beta, double *grad_alpha, double *grad_beta) {
    /* Forward in the non-incremental form APP_t = L + sum_e R_t(e), Q_t(e) = APP_{t-1} - R_{t-1}(e),
     * which equals the incremental update in exact arithmetic.  Per row and iteration the
     * arg-min (idx1) and second arg-min (idx2) are kept for the backward pass. */
    edge_list el;
    edges_build(c, &el);
    const int Z = c->Z, nvn = c->cols * Z, E = el.n_edges;
    const size_t ER = (size_t)E * Z;
    double *R = malloc(sizeof(double) * ER * (n_iter + 1));      /* R_t for t = 0..T */
    double *Qm = malloc(sizeof(double) * ER * n_iter);            /* Q_t (t = 1..T) */
    int *i1 = malloc(sizeof(int) * (size_t)c->rows * Z * n_iter);
    int *i2 = malloc(sizeof(int) * (size_t)c->rows * Z * n_iter);
    double *app = malloc(sizeof(double) * nvn);
    double *gapp = malloc(sizeof(double) * nvn);
    double *gR = malloc(sizeof(double) * ER);
    double *gQn = malloc(sizeof(double) * ER); /* dLoss/dQ_{t+1}, carried backwards */
    double loss = 0.0;
    if (grad_alpha) memset(grad_alpha, 0, sizeof(double) * n_iter * E);
    if (grad_beta) memset(grad_beta, 0, sizeof(double) * n_iter * E);
    for (int w = 0; w < n_cw; w++) {
        const float *L = llr + (size_t)w * nvn;
        const uint8_t *x = bits + (size_t)w * nvn;
        const double wcw = cw_weight ? cw_weight[w] : 1.0; /* the codeword's loss weight */
        memset(R, 0, sizeof(double) * ER);
        for (int n = 0; n < nvn; n++) app[n] = L[n];
        for (int t = 1; t <= n_iter; t++) {
            double *Rp = R + ER * (t - 1), *Rt = R + ER * t, *Qt = Qm + ER * (t - 1);
            for (int i = 0; i < c->rows; i++) {
                const int e0 = el.row_start[i], d = el.row_start[i + 1] - e0;
                for (int r = 0; r < Z; r++) {
                    double m1 = 1000.0, m2 = 1000.0;
                    int a1 = -1, a2 = -1, neg = 0;
                    for (int k = 0; k < d; k++) {
                        const int e = e0 + k;
                        const int col = el.ecol[e] * Z + (r + el.eshift[e]) % Z;
                        const double q = app[col] - Rp[e * Z + r];
                        Qt[e * Z + r] = q;
                        const double aq = fabs(q);
                        neg ^= q < 0;
                        if (aq < m1) { m2 = m1; a2 = a1; m1 = aq; a1 = k; }
                        else if (aq < m2) { m2 = aq; a2 = k; }
                    }
                    i1[((size_t)(t - 1) * c->rows + i) * Z + r] = a1;
                    i2[((size_t)(t - 1) * c->rows + i) * Z + r] = a2;
                    for (int k = 0; k < d; k++) {
                        const int e = e0 + k;
                        const double q = Qt[e * Z + r];
                        const double s = ((neg ^ (q < 0)) ? -1.0 : 1.0);
                        const double m = (k != a1) ? m1 : m2;
                        const int b = (t - 1) * E + e;
                        const double off = m - beta[b];
                        Rt[e * Z + r] = s * alpha[b] * (off > 0 ? off : 0.0);
                    }
                }
            }
            for (int n = 0; n < nvn; n++) app[n] = L[n];
            for (int e = 0; e < E; e++)
                for (int r = 0; r < Z; r++)
                    app[el.ecol[e] * Z + (r + el.eshift[e]) % Z] += Rt[e * Z + r];
        }
        /* BCE on p = sigmoid(-APP): -[x log p + (1-x) log(1-p)], mean over bits */
        for (int n = 0; n < nvn; n++) {
            const double a = app[n];
            const double sp_pos = (a > 0) ? a + log1p(exp(-a)) : log1p(exp(a)); /* softplus(a) */
            const double sp_neg = sp_pos - a;                                     /* softplus(-a) */
            loss += wcw * (x[n] ? sp_pos : sp_neg);
            /* dloss/da = x*sigmoid(a) - (1-x)*sigmoid(-a) */
            const double sg = 1.0 / (1.0 + exp(-a));
            gapp[n] = wcw * (x[n] ? sg : (sg - 1.0)) / ((double)nvn * n_cw);
        }
        if (!grad_alpha && !grad_beta) continue;
        memset(gQn, 0, sizeof(double) * ER);
        for (int t = n_iter; t >= 1; t--) {
            double *Rt = R + ER * t, *Qt = Qm + ER * (t - 1);
            /* dR_t(e) = gAPP_t[col(e)] - gQ_{t+1}(e) */
            for (int e = 0; e < E; e++)
                for (int r = 0; r < Z; r++)
                    gR[e * Z + r] = gapp[el.ecol[e] * Z + (r + el.eshift[e]) % Z] - gQn[e * Z + r];
            /* check-node backward: gradient to alpha, beta and to Q_t at the two arg-mins */
            memset(gQn, 0, sizeof(double) * ER);
            for (int i = 0; i < c->rows; i++) {
                const int e0 = el.row_start[i], d = el.row_start[i + 1] - e0;
                for (int r = 0; r < Z; r++) {
                    const int a1 = i1[((size_t)(t - 1) * c->rows + i) * Z + r];
                    const int a2 = i2[((size_t)(t - 1) * c->rows + i) * Z + r];
                    int neg = 0;
                    for (int k = 0; k < d; k++) neg ^= Qt[(e0 + k) * Z + r] < 0;
                    for (int k = 0; k < d; k++) {
                        const int e = e0 + k;
                        const int src = (k != a1) ? a1 : a2; /* edge whose |Q| is m */
                        /* src < 0: the 1000.0 start value is the minimum, a constant m */
                        const double m = (src < 0) ? 1000.0 : fabs(Qt[(e0 + src) * Z + r]);
                        const int b = (t - 1) * E + e;
                        const double off = m - beta[b];
                        const double s = ((neg ^ (Qt[e * Z + r] < 0)) ? -1.0 : 1.0);
                        const double g = gR[e * Z + r];
                        if (grad_alpha) grad_alpha[b] += g * s * (off > 0 ? off : 0.0);
                        if (off > 0) {
                            if (grad_beta) grad_beta[b] -= g * s * alpha[b];
                            if (src >= 0) {
                                const double qs = Qt[(e0 + src) * Z + r];
                                gQn[(e0 + src) * Z + r] += g * s * alpha[b] * (qs < 0 ? -1.0 : 1.0);
                            }
                        }
                    }
                }
            }
            (void)Rt;
            /* gAPP_{t-1}[col] = sum_e gQ_t(e); Q_t also depends on -R_{t-1}: folded into gR above */
            memset(gapp, 0, sizeof(double) * nvn);
            for (int e = 0; e < E; e++)
                for (int r = 0; r < Z; r++)
                    gapp[el.ecol[e] * Z + (r + el.eshift[e]) % Z] += gQn[e * Z + r];
        }
    }
    free(R); free(Qm); free(i1); free(i2); free(app); free(gapp); free(gR); free(gQn);
    edges_free(&el);
    return loss / ((double)nvn * n_cw);
}

/* ---------------------------------------------------------------- channel */

float oracle_noise_stddev(float ebn0_db, float rate) {
    /* src/encoding_tools.cpp:27-38.  The unqualified pow/sqrt calls on float arguments are
     * bound to the double-precision libm routines, which is how a host C++ compiler without
     * extra global overloads resolves them. */
    const float energy_per_bit = 1 / rate;
    const float eb_n0_db = ebn0_db;
    const double eb_n0 = pow(10.0f, eb_n0_db / 10);
    const float n0 = energy_per_bit / eb_n0;
    const float var = n0 / 2;
    return (float)sqrt(var);
}

void oracle_llr(const float *recv, int n_symbols, float sigma, int z, float *llr) {
    /* src/encoding_tools.cpp:85-99 */
    const int two_cols = 2 * z;
    for (int i = 0; i < n_symbols; i++)
        llr[i] = (i < two_cols) ? 0.0f : (recv[i - two_cols] * 2) / (sigma * sigma);
}

/* ---------------------------------------------------------------- Philox4x32-10 */

static inline uint32_t mulhilo(uint32_t a, uint32_t b, uint32_t *hi) {
    const uint64_t p = (uint64_t)a * b;
    *hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    /* Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11:
     * multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments 0x9E3779B9 / 0xBB67AE85. */
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, hi1;
        const uint32_t lo0 = mulhilo(0xD2511F53u, c0, &hi0);
        const uint32_t lo1 = mulhilo(0xCD9E8D57u, c2, &hi1);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Domain tags keep the info-bit and noise streams of one codeword independent. */
#define TAG_BITS 0x42495453u
#define TAG_NOISE 0x4E4F4953u

static void philox_block(uint64_t seed, uint64_t stream, uint64_t block, uint32_t tag, uint32_t out[4]) {
    const uint32_t ctr[4] = {(uint32_t)block, (uint32_t)(block >> 32), (uint32_t)stream,
                             (uint32_t)(stream >> 32) ^ tag};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    oracle_philox4x32_10(ctr, key, out);
}

uint32_t oracle_info_bit_word(uint64_t seed, uint64_t stream, uint64_t word) {
    uint32_t o[4];
    philox_block(seed, stream, word >> 2, TAG_BITS, o);
    return o[word & 3];
}

/* Elementary functions from basic IEEE operations only, evaluated in a fixed order, so that a
 * device implementation compiled without contraction returns identical bits. */
static float nr_log(float u) { /* u in (0, 1] */
    union { float f; uint32_t i; } b = {u};
    int e = (int)((b.i >> 23) & 0xff) - 127;
    b.i = (b.i & 0x007fffffu) | 0x3f800000u; /* m in [1, 2) */
    float m = b.f;
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 0.0909090909f;                 /* 1/11 */
    p = p * s2 + 0.111111111f;               /* 1/9 */
    p = p * s2 + 0.142857143f;               /* 1/7 */
    p = p * s2 + 0.2f;
    p = p * s2 + 0.333333333f;
    p = p * s2 + 1.0f;
    const float lm = 2.0f * s * p;
    return lm + (float)e * 0.693147181f;
}

static void nr_sincos_2pi(float u, float *sn, float *cs) { /* angle 2*pi*u, u in [0, 1) */
    const float v = u * 4.0f;
    const int q = (int)v;
    const float f = v - (float)q;
    const float x = f * 1.57079633f;
    const float x2 = x * x;
    float s = -2.50521084e-8f;               /* -1/11! */
    s = s * x2 + 2.75573192e-6f;
    s = s * x2 - 1.98412698e-4f;
    s = s * x2 + 8.33333333e-3f;
    s = s * x2 - 1.66666667e-1f;
    s = s * x2 + 1.0f;
    s = s * x;
    float c = 2.08767570e-9f;                /* 1/12! */
    c = c * x2 - 2.75573192e-7f;
    c = c * x2 + 2.48015873e-5f;
    c = c * x2 - 1.38888889e-3f;
    c = c * x2 + 4.16666667e-2f;
    c = c * x2 - 0.5f;
    c = c * x2 + 1.0f;
    switch (q & 3) {
        case 0: *sn = s; *cs = c; break;
        case 1: *sn = c; *cs = -s; break;
        case 2: *sn = -s; *cs = -c; break;
        default: *sn = -c; *cs = s; break;
    }
}

float oracle_gauss(uint64_t seed, uint64_t stream, uint64_t index) {
    /* Box-Muller on one Philox block: (o0, o1) -> normals 4b, 4b+1; (o2, o3) -> 4b+2, 4b+3.
     * Uniforms use the top 24 bits: u = (x >> 8) * 2^-24 + 2^-25, in (0, 1). */
    uint32_t o[4];
    philox_block(seed, stream, index >> 2, TAG_NOISE, o);
    const int pair = (int)((index >> 1) & 1);
    const float u1 = (float)(o[2 * pair] >> 8) * 5.9604644775390625e-8f + 2.98023223876953125e-8f;
    const float u2 = (float)(o[2 * pair + 1] >> 8) * 5.9604644775390625e-8f;
    const float rad = sqrtf(-2.0f * nr_log(u1));
    float sn, cs;
    nr_sincos_2pi(u2, &sn, &cs);
    return (index & 1) ? rad * sn : rad * cs;
}
