// channel.h -- launch interface of the channel, syndrome and error-count kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nrldpc {

struct ChannelArgs {
    uint64_t seed, stream_base;
    int Z, n_vn, K, punct;
    int n_steps;
    const int* step_target;  // [n_steps]
    const int* step_start;   // [n_steps + 1]
    const int* term_col;
    const int* term_shift;
    const float* sigma_dev;  // [n_cw] or null
    float sigma_all;
    const uint8_t* info_in;  // [n_cw][K] information bits, one byte each, or null: Philox bits
    uint8_t* cw_bits;        // [n_cw][n_vn] or null
    float* llr;              // [n_cw][n_vn], or null: encode only (no channel)
};

struct SyndromeArgs {
    const uint8_t* hd;
    int hd_bytes, Z, rows;
    const int* row_start;  // [rows + 1]
    const int* edge_col;   // per edge, row-major
    const int* edge_vm;
    int* weight;
    int n_edges;           // row_start[rows]
    int cols;              // block columns in use
};

struct ErrorArgs {
    const uint8_t* hd;
    const uint8_t* cw_bits;
    int hd_bytes, n_vn, K;
    long long* counts;
};

// Per-codeword packed decisions -> one bit stream over the concatenated [cw][n_vn] columns.
struct StreamPackArgs {
    const uint8_t* hd;      // [n_cw][hd_bytes]
    uint8_t* out;           // ceil(n_cw * n_vn / 8) bytes
    int hd_bytes, n_vn;
    long long n_bits;       // n_cw * n_vn
};

int channel_launch(const ChannelArgs& a, int n_cw, hipStream_t s);
int stream_pack_launch(const StreamPackArgs& a, hipStream_t s);
// int8 LLRs -> float, scaled.
int from_int8_launch(const int8_t* in, size_t n, float scale, float* out, hipStream_t s);
// float -> IEEE binary16 bit patterns, round to nearest even.
int to_half_launch(const float* in, size_t n, uint16_t* out, hipStream_t s);
// Ascending indices of the codewords whose weight is nonzero, and their count (device).
int compact_failures_launch(const int* weight, int n_cw, int* list, int* count, hipStream_t s);
int syndrome_launch(const SyndromeArgs& a, int n_cw, hipStream_t s);
int count_errors_launch(const ErrorArgs& a, int n_cw, hipStream_t s);

}  // namespace nrldpc
