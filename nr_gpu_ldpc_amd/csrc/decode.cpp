// decode.cpp -- host side of the decode entry points.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "context.h"
#include "decoder.h"
#include "status.h"

namespace nrldpc {

bool slim_chosen(const ldpc_ctx* x, int n_cw) {
    if (x->slim_cpb <= 0 || x->d1_done == nullptr) return false;  // slim resources not set up
    if (x->layout_set) return x->layout_set == kLayoutSlim;
    return layout_for_batch(x->m_dup, x->m_slim, n_cw, x->cus) == kLayoutSlim;
}

int ensure_slim(ldpc_ctx* x, int cpb) {
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    if (cpb < 1 || cpb > decoder_slim_cpb_max(x->code.kind, x->code.Z) || x->cus < 1) {
        set_last_error("slim layout: codewords per workgroup beyond the workgroup size or LDS limit");
        return LDPC_EINVAL;
    }
    NRLDPC_HIP(hipSetDevice(x->device));
    if (!x->d1_done) NRLDPC_HIP(hipEventCreateWithFlags(&x->d1_done, hipEventDisableTiming));
    const size_t slot = decoder_slim_slot_bytes(x->code.kind, x->code.Z, cpb);
    if (slot > 0 && (x->d_d1 == nullptr || cpb > x->d1_cpb)) {
        std::lock_guard<std::mutex> lk(x->d1_mu);
        NRLDPC_HIP(hipEventSynchronize(x->d1_done));  // launches on the old slots have finished
        if (x->d_d1) NRLDPC_HIP(hipFree(x->d_d1));
        x->d_d1 = nullptr;
        x->d1_cpb = 0;
        NRLDPC_HIP(hipMalloc(reinterpret_cast<void**>(&x->d_d1), (size_t)x->cus * slot));
        x->d1_cpb = cpb;
    }
    x->slim_cpb = cpb;
    return LDPC_OK;
}

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
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // The slim layout: explicitly chosen (ldpc_set_launch with its codewords per workgroup), or
    // by batch size when the launch shape is left to the context.
    if (!slim_chosen(x, n_cw)) return decoder_launch(c.kind, weights != nullptr, a, s);
    return launch_slim(x, a, weights != nullptr, false, s);
}

int launch_slim(ldpc_ctx* x, DecodeArgs& a, bool learned, bool half, hipStream_t s) {
    const CodeDesc& c = x->code;
    const int slim_cpb = x->slim_cpb;
    a.cpb = slim_cpb;
    a.lds_cw_bytes = (int)decoder_slim_lds_bytes_per_cw(c.kind, c.Z);
    a.d1_scratch = x->d_d1;
    a.d1_stride = (int)(decoder_slim_slot_bytes(c.kind, c.Z, slim_cpb) / sizeof(float));
    const int grid = std::min((a.n_cw + slim_cpb - 1) / slim_cpb, x->cus);
    // Launches on one context share its slots: each waits for the previous one (on any stream).
    // Under stream capture the wait is not recorded: the captured graph orders its own nodes,
    // and graphs that share one context must not run concurrently.
    std::lock_guard<std::mutex> lk(x->d1_mu);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    NRLDPC_HIP(hipStreamIsCapturing(s, &cs));
    const bool capturing = cs != hipStreamCaptureStatusNone;
    if (!capturing) NRLDPC_HIP(hipStreamWaitEvent(s, x->d1_done, 0));
    const int st = decoder_launch_slim(c.kind, learned, a, grid, s, half);
    if (!capturing) NRLDPC_HIP(hipEventRecord(x->d1_done, s));
    return st;
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
    // the same layout choice as the single-precision decode of this batch
    if (slim_chosen(x, n_cw)) return launch_slim(x, a, false, true, static_cast<hipStream_t>(stream));
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
