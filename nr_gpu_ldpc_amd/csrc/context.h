// context.h -- the opaque ldpc_ctx of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/ldpc_api.h"
#include "code.h"
#include "decoder.h"
#include "encoder.h"

struct ldpc_ctx {
    int device = 0;
    nrldpc::CodeDesc code;
    nrldpc::EncodeProgram prog;
    int cus = 0;                   // compute units of the device
    // Launch shape of plain and learned decodes: the LDS layout (nrldpc::kLayoutDup / kLayoutSlim)
    // is fixed by ldpc_set_launch / ldpc_set_launch_layout (layout_set != 0) or chosen per batch
    // with the cost models m_dup and m_slim (the built-in tuning table, or ldpc_autotune).
    int cpb = 1, cpb_default = 1;  // codewords per workgroup of the dup layout
    int slim_cpb = 0;              // codewords per workgroup of the slim layout (0: not in use)
    int layout_set = 0;
    nrldpc::LayoutModel m_dup, m_slim;
    // Slim layout: the degree-1 APP slots of its persistent grid (one per CU, room for
    // d1_cpb codewords each) and the event that orders the launches sharing them.
    float* d_d1 = nullptr;
    int d1_cpb = 0;
    hipEvent_t d1_done = nullptr;
    std::mutex d1_mu;
    // device tables
    nrldpc::EdgeParam* d_etab = nullptr;
    int* d_row_start = nullptr;
    int* d_edge_col = nullptr;
    int* d_edge_vm = nullptr;
    int* d_step_target = nullptr;
    int* d_step_start = nullptr;
    int* d_term_col = nullptr;
    int* d_term_shift = nullptr;
    // learned offset/normalized min-sum weights: [(n_iter + 1)][nnz_full] float2, row 0 zero
    float2* d_weights = nullptr;
    int weights_iters = 0;
    // grow-only scratch for internal temporaries (fallback syndrome weights, training HD)
    void* d_scratch = nullptr;
    size_t scratch_bytes = 0;
    hipEvent_t scratch_done = nullptr;  // recorded after each use of d_scratch
    std::mutex mu;
};

namespace nrldpc {

int nnz_full(int kind);
int n_nd(int kind);
int default_cpb(const CodeDesc& c);
// Per-edge LDS offsets of the any-Z decoder for code c (row-major edge order, whole base graph).
void build_edge_params(const CodeDesc& c, std::vector<EdgeParam>* out);
// Whether a plain or learned decode of n_cw codewords on x runs the slim layout.
bool slim_chosen(const ldpc_ctx* x, int n_cw);
// Slim-layout resources for cpb codewords per workgroup (degree-1 slots, ordering event); a set-up
// call: it may allocate, after the launches using the previous buffer have finished.
int ensure_slim(ldpc_ctx* x, int cpb);
// ldpc_autotune: times every candidate launch shape and installs the fastest of each layout.
int autotune_layouts(ldpc_ctx* x, int n_iter, int max_candidates, int* cand_layout, int* cand_cpb,
                     int* cand_cw_per_cu, float* cand_round_us, int* n_candidates);
// Shared decode path; weights != null selects the learned decoder with [(n_iter+1)][nnz] weights.
int decode_with(ldpc_ctx* x, void* stream, const float* llr, int n_cw, int n_iter, unsigned flags,
                uint8_t* hd, float* app_out, int* iters_out, const float2* weights, const int* list,
                const int* list_count, int* fail_out = nullptr);
// The slim-layout launch of a (a.n_cw codewords) on x's slots, ordered after x's previous one.
int launch_slim(ldpc_ctx* x, DecodeArgs& a, bool learned, bool half, hipStream_t s);
int decode_half(ldpc_ctx* x, void* stream, const uint16_t* llr, int n_cw, int n_iter, unsigned flags, uint8_t* hd,
                float* app_out, int* iters_out);
int decode_char(ldpc_ctx* x, void* stream, const int8_t* llr, int n_cw, int n_iter, unsigned flags, uint8_t* hd,
                float* app_out, int* iters_out);

}  // namespace nrldpc
