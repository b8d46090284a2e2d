// nr_gpu_ldpc.hpp -- the reference's C++ host API for the decode path, over libnrldpc.so.
//
// Same names, argument meaning and call sequence as the reference's host interface, so a
// program written against it (setup_ldpc -> setup_host -> setup_dev_memory -> setup_gpu_grid ->
// put_h_compact_and_params_in_constant_memory -> generate_new_codewords ->
// perform_gpu_ldpc_decoding -> unpack_char / cuda_error_check) builds against this header and
// links libnrldpc.so instead.  Every GPU operation goes through the C ABI of ldpc_api.h.
//
//   reference                         declared at
//   perform_gpu_ldpc_decoding<T>      src/gpu_decoder_top.h:13-18
//   setup_host / setup_dev_memory /   src/setup_exp.h:13-30
//     setup_gpu_grid / generate_new_codewords / put_h_compact_and_params_in_constant_memory
//   setup_ldpc / ldpc_encoder_orig    src/nr_ldpc.h:14-20
//   allocate_constant_memory /        src/ldpc_kernels.h:15-27
//     put_params_in_constant_mem, fill_in_compact_matrices (src/setup_exp.cu:221-284)
//   compute_noise_stddev / transmit / calculate_llr / error_check / cuda_error_check
//                                     src/encoding_tools.h:12-28
//   quantize / unpack_char            src/host_helpers.cpp:46-64
//   types                             src/types.h:16-142
//
// Differences, all deliberate:
//   * T = float (FP32 decode), T = half (the reference's HALF_PRC build: LLRs rounded to binary16
//     by __float2half_rn as its reduce_precision<half> does, decoded by ldpc_decode_half_async)
//     and T = char (its default CHAR_PRC build: LLRs quantised by quantize(0.5, x, 8), decoded by
//     ldpc_decode_char_async) are instantiated.
//   * dev_mem::dt and dev_mem::r stay NULL: the decoder keeps its check-node state on chip.
//   * Errors: HIP failures print the library's message and exit(EXIT_FAILURE), as the
//     reference's checkCudaErrors does; invalid configurations print and abort().
//   * The code description lives in a process-wide decoder context (one code per process, as
//     the reference's __constant__ memory); put_h_compact_and_params_in_constant_memory
//     replaces it.
//   * Packed hard decisions are exact for every Z (the reference's packer assumes Z % 8 == 0);
//     host_mem::mem_size_hd_cuda_packed is ceil(bits / 8).
//   * ldpc_encoder_orig(gen_code = 1) (OAI code generation to files) is not provided.
//   * perform_gpu_ldpc_decoding with n_iterations < 1 is an error (the reference launches no
//     kernel then and copies back whatever the device's hard-decision buffer holds).
//   * setup_host aborts with a message when a stream's LLR bytes exceed INT_MAX (the
//     host_mem size fields are int, as in the reference, whose sizes wrap there instead).
#ifndef NRLDPC_NR_GPU_LDPC_HPP
#define NRLDPC_NR_GPU_LDPC_HPP

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "ldpc_api.h"

struct error_result {
    int bit_error;
    int n_bits;
    int frame_error;
    int n_frames;
};
inline error_result operator+(const error_result& a, const error_result& b) {
    return {a.bit_error + b.bit_error, a.n_bits + b.n_bits, a.frame_error + b.frame_error, a.n_frames + b.n_frames};
}

struct h_element {
    char x;
    char y;
    short value;
    char valid;
};

struct ldpc_params {
    int A;                          // payload size before CRC
    float rate;                     // nominal code rate
    int bg_index;                   // 1 or 2 (NR base graph); 3 = the 12 x 24 QC code (extension)
    int Z;                          // lifting size
    int K;                          // information bits per code block (incl. punctured)
    int N;                          // transmitted bits
    int H_rows;                     // base graph rows
    int H_cols;                     // base graph columns
    int* H;                         // raw base matrix [H_rows][H_cols], -1 = zero block
    int n_nz_in_row;                // largest row degree
    int n_nz_in_col;                // largest column degree
    char* h_element_count_per_row;  // row degrees
    char* h_element_count_per_col;  // column degrees
    int n_iterations;
    int N_before_puncture;          // decoder columns = H_cols * Z
};

struct params {
    float snr_min;
    float snr_max;
    float snr_step;
    int n_codewords;
    int n_info_bits;
    float code_rate;
    int n_cw_per_mcw;
    int n_mcw;
    int n_streams;
};

template <typename T>
struct host_mem {
    int* info_bin;
    unsigned long mem_size_infobits;
    int* cwds;
    float* modulated_cwds;
    unsigned long mem_size_cw;
    float* llr;
    unsigned long mem_size_llr;
    float* recv;
    T* llr_quan;
    // per-stream pinned staging
    int** info_bin_cuda;
    T** llr_cuda;
    int mem_size_llr_cuda;
    char** hd_cuda;
    int mem_size_hd_cuda;
    char** hd_packed_cuda;
    int mem_size_hd_cuda_packed;
};

template <typename T>
struct dev_mem {
    hipStream_t* streams;
    T** llr;
    T** dt;                      // NULL (no message arrays in HBM)
    T** r;                       // NULL
    char** hard_decision;        // per-codeword packed decisions (used when the stream packing differs)
    char** hard_decision_packed; // decisions packed over the concatenated [cw][N_before_puncture] stream
};

struct gpu_result {
    float best_latency;
    float cpu_run_time;
    int n_bits_total;
    int n_bit_errors;
    int n_frames;
    int n_frame_errors;
};
inline gpu_result operator+(const gpu_result& a, const gpu_result& b) {
    return {a.best_latency < b.best_latency ? a.best_latency : b.best_latency, a.cpu_run_time + b.cpu_run_time,
            a.n_bits_total + b.n_bits_total, a.n_bit_errors + b.n_bit_errors, a.n_frames + b.n_frames,
            a.n_frame_errors + b.n_frame_errors};
}

// Geometry of one decode call.  The reference's fields keep their meaning for the batch shape
// (cnp_blocks.y = macro-codewords, cnp_threads.y = codewords per macro-codeword); the fused
// launch's own shape is in cnp_threads_per_block (lanes) and shared_r_cache_size (LDS bytes).
struct cuda_grid {
    dim3 cnp_blocks;
    dim3 cnp_threads;
    dim3 vnp_blocks;
    dim3 vnp_threads;
    dim3 pack_blocks;
    dim3 pack_threads;
    dim3 early_blocks;
    dim3 early_threads;
    unsigned long shared_r_cache_size;
    unsigned int cnp_threads_per_block;
};

struct kernel_params {
    unsigned int z;
    unsigned int n_cw_per_mcw;
    unsigned int n_total_vn;
    unsigned int n_total_cn;
    unsigned int n_blk_col;
    unsigned int threads_per_block;
};

// ---- code setup and host-side channel (src/nr_ldpc.h, src/encoding_tools.h, src/host_helpers.cpp)
ldpc_params setup_ldpc(int n_bits, float code_rate, int n_iterations);
// Extension: the 12 x 24 QC example code at lifting size Z (no puncturing).
ldpc_params setup_ldpc_qc12x24(int Z, int n_iterations);
int ldpc_encoder_orig(int* test_input, int* channel_input, short block_length, short BG, unsigned char gen_code);
short* choose_generator_matrix(short BG, short Zc);
float compute_noise_stddev(float snr, float code_rate, int modulation = 2);
void transmit(int* info_bin, int* codeword, float* modulate_cwds, unsigned long mem_size_codeword, float* recv,
              int n_info_bits, int n_coded_bits, float sigma, int bg_index);
void calculate_llr(float llr[], float recv[], int n_symbols, float sigma, int z);
int error_check(float trans[], float recv[], int n_symbols);
error_result cuda_error_check(int info[], char hard_decision[], int n_total_cws, int codeword_len, int info_length);
signed char quantize(double D, double x, unsigned char B);
void unpack_char(char* packed_input, char* unpacked_output, int size);

// ---- memory, grid and the decode entry point (src/setup_exp.h, src/gpu_decoder_top.h)
template <typename T>
host_mem<T> setup_host(const params* p, unsigned int n_info_bits, unsigned int n_cw_bits);
template <typename T>
dev_mem<T> setup_dev_memory(const params* p, const ldpc_params* ldpc, const host_mem<T>* host);
cuda_grid setup_gpu_grid(const ldpc_params* ldpc, int n_cw_per_mcw, int n_mcw);
template <typename T>
void generate_new_codewords(const params* p, host_mem<T>* h, const ldpc_params* l, int codewords_in_iteration,
                            float sigma);
void put_h_compact_and_params_in_constant_memory(const ldpc_params* ldpc, const params* p, const cuda_grid* g);
template <typename T>
void perform_gpu_ldpc_decoding(dev_mem<T>* d, host_mem<T>* h, const cuda_grid* g, int n_iterations,
                               int stream_index);

// ---- the code description (src/ldpc_kernels.h:15-27, src/setup_exp.cu:221-284)
// Compact lists of ldpc->H: h_compact1[k][i] = k-th nonzero block of block row i (ascending
// column), h_compact2[k][j] = k-th nonzero block of block column j (ascending row), unused slots
// {0, 0, -1, 0}.  The array extents are the caller's (the reference sizes them with its
// H_COMPACT* macros); the row-major cores take them explicitly.
void nr_gpu_ldpc_fill_compact(h_element* h1, int h1_deg, int h1_rows, h_element* h2, int h2_deg, int h2_cols,
                              const ldpc_params* ldpc);
void nr_gpu_ldpc_load_compact(const h_element* h1, int h1_deg, int h1_rows, const h_element* h2, int h2_deg,
                              int h2_cols, const ldpc_params* ldpc);
template <int D1, int R1, int D2, int C2>
void fill_in_compact_matrices(h_element (&h_compact1)[D1][R1], h_element (&h_compact2)[D2][C2],
                              const ldpc_params* ldpc) {
    nr_gpu_ldpc_fill_compact(&h_compact1[0][0], D1, R1, &h_compact2[0][0], D2, C2, ldpc);
}
// Makes the code the lists describe (they must match ldpc->H exactly) the process-wide code;
// the memory sizes are accepted for the reference's signature and not needed.
template <int D1, int R1, int D2, int C2>
void allocate_constant_memory(h_element (&h_compact1)[D1][R1], h_element (&h_compact2)[D2][C2],
                              const ldpc_params* ldpc, int /*mem_size_h_compact1*/, int /*mem_size_h_compact2*/) {
    nr_gpu_ldpc_load_compact(&h_compact1[0][0], D1, R1, &h_compact2[0][0], D2, C2, ldpc);
}
// Checks the kernel parameters against the process-wide code (aborts on a mismatch); the fused
// decoder derives its launch shape itself, so threads_per_block is not used.
void put_params_in_constant_mem(unsigned int z, unsigned int n_cw_per_mcw, unsigned int n_total_vn,
                                unsigned int n_total_cn, unsigned int n_blk_col, unsigned int threads_per_block);

// The C entry point the reference proposes for linking into OAI-style C projects
// (Linking_With_Other_Projects.md:19-25): decode one batch of the process-wide code and geometry
// (set by put_h_compact_and_params_in_constant_memory: n_mcw * n_cw_per_mcw codewords,
// ldpc_params::n_iterations iterations).  inputs: int8 LLRs [n_cw][N_before_puncture] in host
// memory, quantised as quantize(0.5, llr, 8) does; they are decoded with the reference's
// CHAR_PRC schedule, its default build (ldpc_decode_char_async).  outputs: hard decisions packed
// over the concatenated [n_cw][N_before_puncture] bit stream (bit g of byte g/8),
// ceil(n_cw * N_before_puncture / 8) bytes.  Synchronous.
extern "C" void ldpc_gpu_decoder(char* inputs, char* outputs);

// The process-wide decoder context created by put_h_compact_and_params_in_constant_memory
// (NULL before), for callers that mix this API with the C ABI.
ldpc_ctx* nr_gpu_ldpc_context();

extern template host_mem<float> setup_host<float>(const params*, unsigned int, unsigned int);
extern template dev_mem<float> setup_dev_memory<float>(const params*, const ldpc_params*, const host_mem<float>*);
extern template void generate_new_codewords<float>(const params*, host_mem<float>*, const ldpc_params*, int, float);
extern template void perform_gpu_ldpc_decoding<float>(dev_mem<float>*, host_mem<float>*, const cuda_grid*, int,
                                                      int);
extern template host_mem<__half> setup_host<__half>(const params*, unsigned int, unsigned int);
extern template dev_mem<__half> setup_dev_memory<__half>(const params*, const ldpc_params*, const host_mem<__half>*);
extern template void generate_new_codewords<__half>(const params*, host_mem<__half>*, const ldpc_params*, int, float);
extern template void perform_gpu_ldpc_decoding<__half>(dev_mem<__half>*, host_mem<__half>*, const cuda_grid*, int,
                                                       int);
extern template host_mem<char> setup_host<char>(const params*, unsigned int, unsigned int);
extern template dev_mem<char> setup_dev_memory<char>(const params*, const ldpc_params*, const host_mem<char>*);
extern template void generate_new_codewords<char>(const params*, host_mem<char>*, const ldpc_params*, int, float);
extern template void perform_gpu_ldpc_decoding<char>(dev_mem<char>*, host_mem<char>*, const cuda_grid*, int, int);

#endif
