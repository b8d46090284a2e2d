// tune.cpp -- autotuner of the single-code decode launch: LDS layout x codewords per workgroup.
//
// For one context (base graph, Z, kept rows) every candidate launch shape is timed on whole
// rounds of the device (every CU busy with its full share of codewords) at a fixed iteration
// count without early exit, so the work is data independent.  Candidates, per layout:
//   dup   each codewords-per-workgroup count that fills one more wavefront, up to kMaxBlock
//         threads and the LDS (a round: every CU holding as many workgroups as fit at once)
//   slim  the same up to kSlimMaxBlock threads (a round: one workgroup per CU)
// The fastest shape of each layout (time per codeword of a round) becomes the context's cost
// model, and every later batch takes the layout with the shorter modelled time
// (layout_for_batch): ceil(n / codewords per round) rounds of the measured round time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/ldpc_api.h"
#include "context.h"
#include "decoder.h"
#include "status.h"

namespace nrldpc {

namespace {

struct Cand {
    int layout, cpb, per_cu;
    float round_us;
};

// Device buffers and events of one tuning run, released on every path out.
struct TuneRes {
    float* llr = nullptr;
    uint8_t* hd = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~TuneRes() {
        if (s) (void)hipStreamSynchronize(s);
        if (llr) (void)hipFree(llr);
        if (hd) (void)hipFree(hd);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    }
};

constexpr int kRounds = 4;                       // rounds per timed launch
constexpr int kReps = 3;                         // timed launches per candidate (median)
constexpr size_t kMaxLlrBytes = size_t(1) << 30;  // LLR buffer cap of a tuning run

}  // namespace

int autotune_layouts(ldpc_ctx* x, int n_iter, int max_candidates, int* cand_layout, int* cand_cpb,
                     int* cand_cw_per_cu, float* cand_round_us, int* n_candidates) {
    const int kind = x->code.kind, Z = x->code.Z;
    if (x->cus < 1) { set_last_error("autotune: the device reports no compute units"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(x->device));
    std::vector<Cand> cand;
    auto add = [&](int layout, int cpb) {
        for (const Cand& c : cand)
            if (c.layout == layout && c.cpb == cpb) return;
        cand.push_back(Cand{layout, cpb, 0, 0.f});
    };
    const int dmax = decoder_dup_cpb_max(kind, Z), smax = decoder_slim_cpb_max(kind, Z);
    if (dmax > 0) {
        for (int k = 1; k <= kMaxBlock / 64; k++) add(kLayoutDup, std::max(1, std::min(dmax, 64 * k / Z)));
        add(kLayoutDup, dmax);
    }
    if (smax > 0) {
        for (int k = 1; k <= kSlimMaxBlock / 64; k++) add(kLayoutSlim, std::max(1, std::min(smax, 64 * k / Z)));
        add(kLayoutSlim, smax);
    }
    if (cand.empty()) { set_last_error("no launch shape fits this code"); return LDPC_EINVAL; }
    const size_t lds_dup = decoder_lds_bytes_per_cw(kind, Z);
    for (Cand& c : cand) {
        if (c.layout == kLayoutDup) {
            const int occ = decoder_dup_occupancy(kind, Z, false, c.cpb * Z, (size_t)c.cpb * lds_dup + 16);
            c.per_cu = c.cpb * std::max(1, occ);
        } else {
            c.per_cu = c.cpb;
        }
    }
    // codewords of a timed launch: kRounds rounds, fewer where the LLRs would exceed the cap
    const size_t n_vn = (size_t)x->code.n_vn(), hd_bytes = (size_t)x->code.hd_bytes();
    const long long cap_cw = (long long)(kMaxLlrBytes / (n_vn * sizeof(float)));
    auto rounds_of = [&](const Cand& c) {
        const long long per_round = (long long)x->cus * c.per_cu;
        return (int)std::max<long long>(1, std::min<long long>(kRounds, cap_cw / per_round));
    };
    long long n_max = 0;
    for (const Cand& c : cand) n_max = std::max(n_max, (long long)rounds_of(c) * x->cus * c.per_cu);
    if (n_max > INT32_MAX || n_max > cap_cw) {
        set_last_error("autotune: a round of this code exceeds the tuning buffer");
        return LDPC_ENOMEM;
    }
    TuneRes res;
    if (hipMalloc(reinterpret_cast<void**>(&res.llr), (size_t)n_max * n_vn * sizeof(float)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&res.hd), (size_t)n_max * hd_bytes) != hipSuccess) {
        set_last_error("autotune: device allocation failed");
        return LDPC_ENOMEM;
    }
    NRLDPC_HIP(hipStreamCreateWithFlags(&res.s, hipStreamNonBlocking));
    NRLDPC_HIP(hipEventCreate(&res.e0));
    NRLDPC_HIP(hipEventCreate(&res.e1));
    const float rate = (float)x->code.K() / (float)x->code.N();
    int st = ldpc_channel_async(x, res.s, 1, 0, (int)n_max, nullptr, ldpc_noise_stddev(1.0f, rate), nullptr, res.llr);
    if (st) return st;
    NRLDPC_HIP(hipStreamSynchronize(res.s));

    // the context's launch state comes back unchanged if a candidate fails
    struct Restore {
        ldpc_ctx* x;
        int layout_set, cpb, slim_cpb;
        bool done = false;
        ~Restore() {
            if (!done) { x->layout_set = layout_set; x->cpb = cpb; x->slim_cpb = slim_cpb; }
        }
    } restore{x, x->layout_set, x->cpb, x->slim_cpb};
    for (Cand& c : cand) {
        if (c.layout == kLayoutDup) {
            x->cpb = c.cpb;
            x->layout_set = kLayoutDup;
        } else {
            if ((st = ensure_slim(x, c.cpb))) return st;
            x->layout_set = kLayoutSlim;
        }
        const int rounds = rounds_of(c);
        const int n = rounds * x->cus * c.per_cu;
        if ((st = decode_with(x, res.s, res.llr, n, n_iter, 0, res.hd, nullptr, nullptr, nullptr, nullptr, nullptr)))
            return st;
        std::vector<float> ms;
        for (int r = 0; r < kReps; r++) {
            NRLDPC_HIP(hipEventRecord(res.e0, res.s));
            if ((st = decode_with(x, res.s, res.llr, n, n_iter, 0, res.hd, nullptr, nullptr, nullptr, nullptr,
                                  nullptr)))
                return st;
            NRLDPC_HIP(hipEventRecord(res.e1, res.s));
            NRLDPC_HIP(hipEventSynchronize(res.e1));
            float t = 0.f;
            NRLDPC_HIP(hipEventElapsedTime(&t, res.e0, res.e1));
            ms.push_back(t);
        }
        std::sort(ms.begin(), ms.end());
        c.round_us = ms[ms.size() / 2] * 1000.f / (float)rounds;
    }

    // the fastest shape of each layout per codeword of a round
    const Cand* best[3] = {nullptr, nullptr, nullptr};
    for (const Cand& c : cand) {
        const Cand*& b = best[c.layout];
        if (!b || c.round_us / c.per_cu < b->round_us / b->per_cu) b = &c;
    }
    LayoutModel m_dup, m_slim;
    if (best[kLayoutDup]) m_dup = LayoutModel{best[kLayoutDup]->cpb, best[kLayoutDup]->per_cu, best[kLayoutDup]->round_us};
    if (best[kLayoutSlim]) {
        m_slim = LayoutModel{best[kLayoutSlim]->cpb, best[kLayoutSlim]->per_cu, best[kLayoutSlim]->round_us};
        if ((st = ensure_slim(x, m_slim.cpb))) return st;
    } else {
        x->slim_cpb = 0;
    }
    x->m_dup = m_dup;
    x->m_slim = m_slim;
    if (x->m_dup.cpb > 0) x->cpb_default = x->m_dup.cpb;
    x->cpb = x->cpb_default;
    x->layout_set = 0;
    restore.done = true;
    const int n_out = std::min<int>((int)cand.size(), max_candidates);
    for (int k = 0; k < n_out; k++) {
        if (cand_layout) cand_layout[k] = cand[k].layout;
        if (cand_cpb) cand_cpb[k] = cand[k].cpb;
        if (cand_cw_per_cu) cand_cw_per_cu[k] = cand[k].per_cu;
        if (cand_round_us) cand_round_us[k] = cand[k].round_us;
    }
    if (n_candidates) *n_candidates = (int)cand.size();
    return LDPC_OK;
}

}  // namespace nrldpc
