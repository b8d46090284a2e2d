// noms_api.cpp -- C ABI of the learned offset/normalized min-sum decoder.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ldpc_api.h"
#include "channel.h"
#include "context.h"
#include "decoder.h"
#include "learned.h"
#include "status.h"

namespace nrldpc {

// Grow-only device scratch owned by the context (fallback syndrome weights and re-decode
// outputs, the training forward pass's hard decisions).  One use at a time: a ScratchUse holds
// the context mutex (host threads) while its call enqueues work, makes the call's stream wait
// for the previous use's last kernel (the event x->scratch_done, so calls on different streams
// are serialised on the GPU), and records that event on the call's stream when it ends, error
// paths included.  Growing waits for the previous use to finish before freeing the old buffer.
// Under stream capture the event is neither waited on nor recorded: the captured graph orders
// its own nodes, and graphs that share one context's scratch must not run concurrently.
class ScratchUse {
  public:
    ScratchUse(ldpc_ctx* x, hipStream_t s) : x_(x), s_(s), lk_(x->mu) {}
    ScratchUse(const ScratchUse&) = delete;
    ScratchUse& operator=(const ScratchUse&) = delete;
    ~ScratchUse() {
        if (!began_ || capturing_) return;
        if (!x_->scratch_done &&
            hipEventCreateWithFlags(&x_->scratch_done, hipEventDisableTiming) != hipSuccess) {
            x_->scratch_done = nullptr;
            return;
        }
        (void)hipEventRecord(x_->scratch_done, s_);
    }
    int acquire(size_t bytes, void** out) {
        if (x_->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
        NRLDPC_HIP(hipSetDevice(x_->device));  // the scratch lives on the context's device
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        NRLDPC_HIP(hipStreamIsCapturing(s_, &cs));
        capturing_ = cs != hipStreamCaptureStatusNone;
        began_ = true;
        if (x_->scratch_done && !capturing_) NRLDPC_HIP(hipStreamWaitEvent(s_, x_->scratch_done, 0));
        if (x_->scratch_bytes < bytes) {
            if (x_->scratch_done) NRLDPC_HIP(hipEventSynchronize(x_->scratch_done));
            if (x_->d_scratch) NRLDPC_HIP(hipFree(x_->d_scratch));
            x_->d_scratch = nullptr;
            x_->scratch_bytes = 0;
            NRLDPC_HIP(hipMalloc(&x_->d_scratch, bytes));
            x_->scratch_bytes = bytes;
        }
        *out = x_->d_scratch;
        return LDPC_OK;
    }

  private:
    ldpc_ctx* x_;
    hipStream_t s_;
    std::lock_guard<std::mutex> lk_;
    bool began_ = false, capturing_ = false;
};

static size_t records_per_cw(const ldpc_ctx* x, int n_iter) {
    const int rows_full = x->code.rows_full;
    return (size_t)n_iter * rows_full * x->code.Z;
}

// Layout of the learned fallback's scratch for n_cw codewords (byte offsets, 256-aligned).
struct FallbackLayout {
    size_t w2, it, list, cnt, hd, app, total;
};
static FallbackLayout fallback_layout(const CodeDesc& c, int n_cw, bool with_app) {
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    FallbackLayout l;
    l.w2 = up((size_t)n_cw * sizeof(int));
    l.it = l.w2 + up((size_t)n_cw * sizeof(int));
    l.list = l.it + up((size_t)n_cw * sizeof(int));
    l.cnt = l.list + up((size_t)n_cw * sizeof(int));
    l.hd = l.cnt + 256;
    l.app = l.hd + up((size_t)n_cw * c.hd_bytes());
    l.total = l.app + (with_app ? (size_t)n_cw * c.n_vn() * sizeof(float) : 0);
    return l;
}

// Learned decode with plain min-sum fallback, its temporaries in `sp` (device, l.total bytes).
// Plain min-sum re-decodes the codewords whose learned hard decision fails the syndrome check;
// its outputs replace the learned ones only where the plain hard decision passes the check (a
// learned output with correct information bits but an unsatisfied parity check is kept when
// plain min-sum does no better).  Every scratch word the kernels read is written earlier in the
// same call (failure flags of every codeword, the compacted list and its count, the re-decode's
// outputs of every listed codeword), so the scratch needs no clearing.
static int decode_fallback(ldpc_ctx* x, void* stream, const float* llr, int n_cw, int n_iter, unsigned flags,
                           uint8_t* hd, float* app_out, int* iters_out, const FallbackLayout& l, char* sp) {
    const CodeDesc& c = x->code;
    int* w1 = reinterpret_cast<int*>(sp);
    int* w2 = reinterpret_cast<int*>(sp + l.w2);
    int* it2 = reinterpret_cast<int*>(sp + l.it);
    int* list = reinterpret_cast<int*>(sp + l.list);
    int* cnt = reinterpret_cast<int*>(sp + l.cnt);
    uint8_t* hd2 = reinterpret_cast<uint8_t*>(sp + l.hd);
    float* app2 = app_out ? reinterpret_cast<float*>(sp + l.app) : nullptr;
    const hipStream_t hs = static_cast<hipStream_t>(stream);
    int st;
    // both decodes check the parity of their output after the last iteration (w1, w2: 1 = fails)
    if ((st = decode_with(x, stream, llr, n_cw, n_iter, flags & LDPC_FLAG_EARLY_EXIT, hd, app_out, iters_out,
                          x->d_weights, nullptr, nullptr, w1)))
        return st;
    // the failing codewords, compacted, so that the re-decode's workgroups are all busy ones
    if ((st = compact_failures_launch(w1, n_cw, list, cnt, hs))) return st;
    if ((st = decode_with(x, stream, llr, n_cw, n_iter, flags & LDPC_FLAG_EARLY_EXIT, hd2, app2, it2, nullptr,
                          list, cnt, w2)))
        return st;
    MergeArgs m{w1, w2, hd2, app2, it2, hd, app_out, iters_out, n_cw, c.hd_bytes(), c.n_vn()};
    return fallback_merge_launch(m, hs);
}

// ---- fixed scratch of ldpc_decode_learned_noalloc_async
// One array per GPU, loaded with the library's code object and held until the process ends.  In
// host code `scratch` only names the symbol: its device address on each GPU comes from
// hipGetSymbolAddress (called once per device, with that device current) and is what the kernels
// are given.
static __device__ __align__(256) unsigned char scratch[LDPC_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be the current device).
static int noalloc_scratch(int device, char** out) {
    if (device < 0 || device >= kMaxDevices) {
        set_last_error("device index beyond the fixed-scratch table");
        return LDPC_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return LDPC_EHIP;
        }
        g_scratch_ptr[device] = ptr;
    }
    *out = static_cast<char*>(g_scratch_ptr[device]);
    return LDPC_OK;
}

}  // namespace nrldpc

using namespace nrldpc;

extern "C" {

int ldpc_noms_set_weights(ldpc_ctx* x, int n_iter, const float* w) {
    if (!x || n_iter <= 0 || !w) return LDPC_EINVAL;
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(x->device));
    const int nnz = nnz_full(x->code.kind);
    std::vector<float> h((size_t)(n_iter + 1) * nnz * 2, 0.f);
    std::copy(w, w + (size_t)n_iter * nnz * 2, h.begin() + (size_t)nnz * 2);
    std::lock_guard<std::mutex> lk(x->mu);
    // Learned decodes, fallbacks and Adam updates on any stream may still read or write the
    // table: replace it only once the device is idle.
    NRLDPC_HIP(hipDeviceSynchronize());
    if (x->d_weights && x->weights_iters != n_iter) {
        NRLDPC_HIP(hipFree(x->d_weights));
        x->d_weights = nullptr;
    }
    if (!x->d_weights) NRLDPC_HIP(hipMalloc(reinterpret_cast<void**>(&x->d_weights), h.size() * sizeof(float)));
    NRLDPC_HIP(hipMemcpy(x->d_weights, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    x->weights_iters = n_iter;
    return LDPC_OK;
}

int ldpc_noms_get_weights(ldpc_ctx* x, int* n_iter, float* w) {
    if (!x || !n_iter) return LDPC_EINVAL;
    *n_iter = x->weights_iters;
    if (!x->d_weights) return LDPC_EINVAL;
    if (w) {
        const int nnz = nnz_full(x->code.kind);
        NRLDPC_HIP(hipSetDevice(x->device));
        NRLDPC_HIP(hipMemcpy(w, reinterpret_cast<float*>(x->d_weights) + (size_t)nnz * 2,
                             (size_t)x->weights_iters * nnz * 2 * sizeof(float), hipMemcpyDeviceToHost));
    }
    return LDPC_OK;
}

int ldpc_decode_learned_async(ldpc_ctx* x, void* stream, const float* llr, int n_cw, int n_iter, unsigned flags,
                              uint8_t* hd, float* app_out, int* iters_out) {
    if (!x || !llr || !hd || n_cw < 0) return LDPC_EINVAL;
    if (!x->d_weights || n_iter != x->weights_iters) {
        set_last_error("learned weights missing or for a different iteration count");
        return LDPC_EINVAL;
    }
    if (n_cw == 0) return LDPC_OK;
    const bool fb = (flags & LDPC_FLAG_FALLBACK_PLAIN) != 0;
    if (!fb)
        return decode_with(x, stream, llr, n_cw, n_iter, flags & LDPC_FLAG_EARLY_EXIT, hd, app_out, iters_out,
                           x->d_weights, nullptr, nullptr);
    const FallbackLayout l = fallback_layout(x->code, n_cw, app_out != nullptr);
    void* scratch_mem = nullptr;
    ScratchUse use(x, static_cast<hipStream_t>(stream));
    int st = use.acquire(l.total, &scratch_mem);
    if (st) return st;
    return decode_fallback(x, stream, llr, n_cw, n_iter, flags, hd, app_out, iters_out, l,
                           static_cast<char*>(scratch_mem));
}

int ldpc_decode_learned_noalloc_async(ldpc_ctx* x, void* stream, const float* llr, int n_cw, int n_iter,
                                      unsigned flags, uint8_t* hd, float* app_out, int* iters_out) {
    if (!x || !llr || !hd || n_cw < 0) return LDPC_EINVAL;
    if (!x->d_weights || n_iter != x->weights_iters) {
        set_last_error("learned weights missing or for a different iteration count");
        return LDPC_EINVAL;
    }
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    if (n_cw == 0) return LDPC_OK;
    if (!(flags & LDPC_FLAG_FALLBACK_PLAIN))  // no temporaries at all
        return decode_with(x, stream, llr, n_cw, n_iter, flags & LDPC_FLAG_EARLY_EXIT, hd, app_out, iters_out,
                           x->d_weights, nullptr, nullptr);
    const FallbackLayout l = fallback_layout(x->code, n_cw, app_out != nullptr);
    if (l.total > (size_t)LDPC_NOALLOC_SCRATCH_BYTES) {
        set_last_error("allocation-free learned decode: " + std::to_string(n_cw) + " codewords need " +
                       std::to_string(l.total) + " bytes of scratch, beyond the fixed " +
                       std::to_string((size_t)LDPC_NOALLOC_SCRATCH_BYTES));
        return LDPC_ENOMEM;
    }
    NRLDPC_HIP(hipSetDevice(x->device));
    char* sp = nullptr;
    int st = noalloc_scratch(x->device, &sp);
    if (st) return st;
    return decode_fallback(x, stream, llr, n_cw, n_iter, flags, hd, app_out, iters_out, l, sp);
}

size_t ldpc_noms_workspace_bytes(const ldpc_ctx* x, int n_cw, int n_iter) {
    if (!x || n_cw <= 0 || n_iter <= 0) return 0;
    const size_t rec = records_per_cw(x, n_iter) * (size_t)n_cw;
    const size_t app = (size_t)n_cw * x->code.n_vn();
    const size_t n_wg = (size_t)std::min(n_cw, 1024);
    const size_t gapp = n_wg * 2 * (size_t)x->code.n_vn();
    return (rec * 3 + app + gapp) * sizeof(float) + 256;
}

int ldpc_noms_train_grad_async(ldpc_ctx* x, void* stream, const float* llr, const uint8_t* cw_bits, int n_cw,
                               int n_iter, float* grad, float* loss, void* ws, size_t ws_bytes) {
    return ldpc_noms_train_grad_weighted_async(x, stream, llr, cw_bits, nullptr, n_cw, n_iter, grad, loss, ws,
                                               ws_bytes);
}

int ldpc_noms_train_grad_weighted_async(ldpc_ctx* x, void* stream, const float* llr, const uint8_t* cw_bits,
                                        const float* cw_weight, int n_cw, int n_iter, float* grad, float* loss,
                                        void* ws, size_t ws_bytes) {
    if (!x || !llr || !cw_bits || !grad || !loss || !ws || n_cw <= 0) return LDPC_EINVAL;
    if (!x->d_weights || n_iter != x->weights_iters) {
        set_last_error("learned weights missing or for a different iteration count");
        return LDPC_EINVAL;
    }
    if (ws_bytes < ldpc_noms_workspace_bytes(x, n_cw, n_iter)) {
        set_last_error("training workspace too small");
        return LDPC_ENOMEM;
    }
    NRLDPC_HIP(hipSetDevice(x->device));
    const CodeDesc& c = x->code;
    const size_t rec = records_per_cw(x, n_iter) * (size_t)n_cw;
    float* p = static_cast<float*>(ws);
    float* rec_m1 = p;
    float* rec_m2 = p + rec;
    uint32_t* rec_w = reinterpret_cast<uint32_t*>(p + 2 * rec);
    float* app = p + 3 * rec;
    float* gapp = app + (size_t)n_cw * c.n_vn();
    std::vector<uint8_t> dummy;
    // forward with records; hard decisions go to the tail of the workspace-free hd buffer
    DecodeArgs a{};
    a.llr = llr;
    a.app_out = app;
    a.etab = x->d_etab;
    a.weights = x->d_weights;
    a.n_cw = n_cw;
    a.Z = c.Z;
    a.m_rows = c.rows;
    a.n_vn = c.n_vn();
    a.hd_bytes = (c.n_vn() + 7) / 8;
    a.n_iter = n_iter;
    a.cpb = 1;
    a.lds_cw_bytes = (int)decoder_lds_bytes_per_cw(c.kind, c.Z);
    a.nnz = nnz_full(c.kind);
    a.rec_m1 = rec_m1;
    a.rec_m2 = rec_m2;
    a.rec_w = rec_w;
    void* hd_scratch = nullptr;
    ScratchUse use(x, static_cast<hipStream_t>(stream));
    int st = use.acquire((size_t)n_cw * a.hd_bytes, &hd_scratch);
    if (st) return st;
    a.hd = static_cast<uint8_t*>(hd_scratch);
    if ((st = decoder_launch_train(c.kind, a, static_cast<hipStream_t>(stream)))) return st;
    BackwardArgs b{};
    b.app = app;
    b.bits = cw_bits;
    b.rec_m1 = rec_m1;
    b.rec_m2 = rec_m2;
    b.rec_w = rec_w;
    b.weights = x->d_weights;
    b.edge_col = x->d_edge_col;
    b.edge_vm = x->d_edge_vm;
    b.grad = grad;
    b.loss = loss;
    b.gapp = gapp;
    b.n_cw = n_cw;
    b.n_iter = n_iter;
    b.Z = c.Z;
    b.m_rows = c.rows;
    b.n_vn = c.n_vn();
    b.nnz = nnz_full(c.kind);
    b.inv_norm = 1.0f / ((float)c.n_vn() * (float)n_cw);
    b.cw_weight = cw_weight;
    return backward_launch(c.kind, b, std::min(n_cw, 1024), static_cast<hipStream_t>(stream));
}

int ldpc_noms_adam_async(ldpc_ctx* x, void* stream, const float* grad, float* m, float* v, int step, float lr) {
    if (!x || !grad || !m || !v || step < 1 || !x->d_weights) return LDPC_EINVAL;
    NRLDPC_HIP(hipSetDevice(x->device));
    const int nnz = nnz_full(x->code.kind);
    AdamArgs a{};
    a.w = reinterpret_cast<float*>(x->d_weights) + (size_t)nnz * 2;
    a.grad = grad;
    a.m = m;
    a.v = v;
    a.n = x->weights_iters * nnz * 2;
    a.lr = lr;
    a.b1 = 0.9f;
    a.b2 = 0.999f;
    a.eps = 1e-8f;
    a.b1t = (float)std::pow(0.9, step);
    a.b2t = (float)std::pow(0.999, step);
    return adam_launch(a, static_cast<hipStream_t>(stream));
}

}  // extern "C"
