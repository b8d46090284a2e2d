// learned.h -- launch interface of the learned-decoder training kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nrldpc {

struct BackwardArgs {
    const float* app;      // [n_cw][n_vn] posterior after the last iteration (forward output)
    const uint8_t* bits;   // [n_cw][n_vn] transmitted codeword bits
    const float* rec_m1;   // forward training records [n_cw][n_iter][rows_full][Z]
    const float* rec_m2;
    const uint32_t* rec_w;
    const float2* weights; // [(n_iter+1)][nnz] (alpha, beta), row 0 unused
    const int* edge_col;   // per edge (row-major, rows in use)
    const int* edge_vm;
    float* grad;           // [n_iter][nnz][2], accumulated
    float* loss;           // accumulated mean BCE
    float* gapp;           // workspace [n_wg][2][n_vn]
    const float* cw_weight;  // [n_cw] weight of each codeword's loss terms, or null (all 1)
    int n_cw, n_iter, Z, m_rows, n_vn, nnz;
    float inv_norm;        // 1 / (n_vn * n_cw)
};

struct AdamArgs {
    float* w;              // parameters (alpha, beta interleaved), n values
    const float* grad;
    float* m;
    float* v;
    int n;
    float lr, b1, b2, eps, b1t, b2t;  // b1t = b1^step, b2t = b2^step
};

// Learned-decoder fallback: copy the plain re-decode of codeword cw into the outputs when the
// learned result failed its syndrome check (w_learned != 0) and the plain one passes (w_plain == 0).
struct MergeArgs {
    const int* w_learned;
    const int* w_plain;
    const uint8_t* hd_plain;
    const float* app_plain;  // NULL when no APP output is requested
    const int* iters_plain;
    uint8_t* hd;
    float* app;
    int* iters;              // may be NULL
    int n_cw, hd_bytes, n_vn;
};

int fallback_merge_launch(const MergeArgs& a, hipStream_t s);
int backward_launch(int kind, const BackwardArgs& a, int n_wg, hipStream_t s);
int adam_launch(const AdamArgs& a, hipStream_t s);

}  // namespace nrldpc
