// decode.cpp -- host side of the decode entry points.
#include <hip/hip_runtime.h>

#include "context.h"
#include "decoder.h"
#include "status.h"

namespace nrldpc {

int decode_with(ldpc_ctx* x, void* stream, const float* llr, int n_cw, int n_iter, unsigned flags,
                uint8_t* hd, float* app_out, int* iters_out, const float2* weights, const int* list,
                const int* list_count, int* fail_out) {
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(x->device));
    const CodeDesc& c = x->code;
    DecodeArgs a{};
    a.llr = llr;
    a.hd = hd;
    a.app_out = app_out;
    a.iters_out = iters_out;
    a.etab = x->d_etab;
    a.weights = weights;
    a.n_cw = n_cw;
    a.Z = c.Z;
    a.m_rows = c.rows;
    a.n_vn = c.n_vn();
    a.hd_bytes = (c.n_vn() + 7) / 8;
    a.n_iter = n_iter;
    a.cpb = x->cpb;
    a.early_exit = (flags & LDPC_FLAG_EARLY_EXIT) ? 1 : 0;
    a.lds_cw_bytes = (int)decoder_lds_bytes_per_cw(c.kind, c.Z);
    a.nnz = nnz_full(c.kind);
    a.list = list;
    a.list_count = list_count;
    a.fail_out = fail_out;
    return decoder_launch(c.kind, weights != nullptr, a, static_cast<hipStream_t>(stream));
}

int decode_half(ldpc_ctx* x, void* stream, const uint16_t* llr, int n_cw, int n_iter, unsigned flags, uint8_t* hd,
                float* app_out, int* iters_out) {
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(x->device));
    const CodeDesc& c = x->code;
    DecodeArgs a{};
    a.llr_half = llr;
    a.hd = hd;
    a.app_out = app_out;
    a.iters_out = iters_out;
    a.etab = x->d_etab;
    a.n_cw = n_cw;
    a.Z = c.Z;
    a.m_rows = c.rows;
    a.n_vn = c.n_vn();
    a.hd_bytes = c.hd_bytes();
    a.n_iter = n_iter;
    a.cpb = x->cpb;
    a.early_exit = (flags & LDPC_FLAG_EARLY_EXIT) ? 1 : 0;
    a.lds_cw_bytes = (int)decoder_lds_bytes_per_cw(c.kind, c.Z);
    a.nnz = nnz_full(c.kind);
    return decoder_launch_half(c.kind, a, static_cast<hipStream_t>(stream));
}

int decode_char(ldpc_ctx* x, void* stream, const int8_t* llr, int n_cw, int n_iter, unsigned flags, uint8_t* hd,
                float* app_out, int* iters_out) {
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(x->device));
    const CodeDesc& c = x->code;
    DecodeArgs a{};
    a.llr_i8 = llr;
    a.hd = hd;
    a.app_out = app_out;
    a.iters_out = iters_out;
    a.etab = x->d_etab;
    a.n_cw = n_cw;
    a.Z = c.Z;
    a.m_rows = c.rows;
    a.n_vn = c.n_vn();
    a.hd_bytes = c.hd_bytes();
    a.n_iter = n_iter;
    a.cpb = decoder_pair_slots(c.kind, c.Z);
    a.early_exit = (flags & LDPC_FLAG_EARLY_EXIT) ? 1 : 0;
    a.lds_cw_bytes = (int)decoder_pair_lds_bytes(c.kind, c.Z);
    a.nnz = nnz_full(c.kind);
    return decoder_launch_char(c.kind, a, static_cast<hipStream_t>(stream));
}

}  // namespace nrldpc
