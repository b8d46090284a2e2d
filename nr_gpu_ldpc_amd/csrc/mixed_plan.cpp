// mixed_plan.cpp -- planner, autotuner, graph capture and C ABI of the persistent mixed-batch
// decoder (kernel in mixed.hip).
//
// Plan of a batch (host):
//   1. group the codewords by (base graph, Z) and sort each group by parity rows, largest first;
//   2. cut every group into tiles of cpb codewords (cpb per (base graph, Z): tuned or default);
//      a tile runs with the largest row count of its codewords, the others mask the extra rows;
//   3. order the tiles longest first (estimated by SIMD-rounds x edges), so the work queue of
//      the persistent grid ends with the short tiles.
// The plan lives in pinned staging memory and is copied to the device on the decode stream;
// the kernel reads the tile count from the device copy, so one captured graph serves every
// later plan of the same object.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ldpc_api.h"
#include "code.h"
#include "context.h"
#include "decoder.h"
#include "status.h"

namespace nrldpc {

constexpr int kMixedThreads = 512;  // workgroup size of the dup-layout mixed kernel
// layout of a new object (and of LDPC_LAYOUT_AUTO)
constexpr int kMixedDefaultLayout = kLayoutDup;
constexpr size_t kHdrBytes = 256;

struct MixedCode {
    CodeDesc code;                 // whole base graph
    EdgeParam* d_etab = nullptr;
    // per layout of the object (index 0: dup, 1: slim): LDS bytes per codeword slot and codewords
    // per workgroup (cpb 0: the default, the most that fit)
    int lds_cw[2] = {0, 0};
    int cpb_max[2] = {1, 1}, cpb_default[2] = {1, 1}, cpb[2] = {0, 0};
    float2* d_weights = nullptr;   // trained weights [(T+1)][nnz_full], row 0 zero
    int weights_iters = 0;
};

static int code_key(int kind, int Z) { return kind * 1024 + Z; }

static int edges_in_rows(int kind, int m) {
    int n = 0;
    if (kind == kBG1) { for (const auto& e : kBG1Entries) n += e.row < m; }
    else { for (const auto& e : kBG2Entries) n += e.row < m; }
    return n;
}

}  // namespace nrldpc

using namespace nrldpc;

struct ldpc_mixed {
    int device = -1;
    int max_cw = 0;
    // launch layout of every tile (nrldpc::kLayoutDup: 512-thread workgroups, degree-1 APPs in
    // registers; kLayoutSlim: 768-thread workgroups at three wavefronts per SIMD, single LDS read
    // copy, degree-1 APPs in d_d1)
    int layout = kLayoutDup;
    float* d_d1 = nullptr;             // slim: kMixedSlimD1Stride floats per resident workgroup
    int d1_slots = 0;
    int n_cu = 256, grid = 256, threads = kMixedThreads;
    int grid_full = 256;               // resident workgroups of the whole chip (the default grid)
    int llr_fmt = kLlrF32;             // element type of the llr buffers of later launches
    float llr_scale = 1.0f;            // kLlrI8: LLR = q * llr_scale
    float* d_conv = nullptr;           // narrow inputs converted to float (grow-only)
    size_t conv_cap = 0;               // floats d_conv holds
    int lds = 0;                       // dynamic LDS bytes per workgroup
    std::map<int, MixedCode> codes;
    unsigned char* h_plan = nullptr;   // pinned staging: [header][tiles][codewords]
    unsigned char* d_plan = nullptr;
    size_t tiles_off = kHdrBytes, cws_off = 0, plan_bytes = 0;
    unsigned* d_queue = nullptr;
    int n_tiles = 0, n_cw = 0;
    std::vector<int> order, tiles6;
    // iteration count of the learned weights in the plan's tiles: 0 none, -1 the plan mixes
    // codes whose weights are for different iteration counts
    int plan_trained_T = 0;
    bool planned = false;
    // recorded on the stream of each launch (decode or replay) after its last use of the plan,
    // the work counter and the caller's buffers; the next launch of this object waits for it
    hipEvent_t uploaded = nullptr;
    bool upload_pending = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // memcpy nodes of the graph (null when absent) and what they copy at most
    hipGraphNode_t n_head = nullptr, n_body = nullptr;
    const float* g_llr_host = nullptr;
    float* g_llr = nullptr;
    uint8_t* g_hd = nullptr;
    uint8_t* g_hd_host = nullptr;
    size_t g_llr_floats = 0, g_hd_bytes = 0;  // buffer sizes given at capture (0: not given)
    int g_n_iter = 0;                         // iteration count and flags of the captured decode
    unsigned g_flags = 0;
    size_t g_elem = 4;                        // bytes per LLR element of the captured decode
    long long plan_llr_extent = 0, plan_hd_extent = 0;  // floats / bytes the current plan touches
    std::mutex mu;
};

namespace nrldpc {

static int mixed_get_code(ldpc_mixed* p, int kind, int Z, MixedCode** out) {
    auto it = p->codes.find(code_key(kind, Z));
    if (it != p->codes.end()) { *out = &it->second; return LDPC_OK; }
    if (kind != kBG1 && kind != kBG2) {
        set_last_error("mixed batches take base graph 1 or 2");
        return LDPC_EINVAL;
    }
    MixedCode mc;
    std::string err;
    int st = make_code(kind, Z, 0, &mc.code, &err);
    if (st) { set_last_error(err); return st; }
    const int slots = p->lds - 32;  // workgroup flag and queue word above the codeword slots
    for (int L = 0; L < 2; L++) {
        const int threads = L ? kSlimMaxBlock : kMixedThreads;
        mc.lds_cw[L] = (int)(L ? decoder_slim_lds_bytes_per_cw(kind, Z) : decoder_lds_bytes_per_cw(kind, Z));
        mc.cpb_max[L] = std::max(1, std::min(threads / Z, slots / mc.lds_cw[L]));
        mc.cpb_default[L] = mc.cpb_max[L];  // the most lanes in use (the autotuner may choose fewer)
    }
    if (mc.lds_cw[0] > slots || Z > kMixedThreads) {
        set_last_error("lifting size does not fit one workgroup");
        return LDPC_EINVAL;
    }
    if (p->device >= 0) {
        std::vector<EdgeParam> et;
        build_edge_params(mc.code, &et);
        NRLDPC_HIP(hipMalloc(reinterpret_cast<void**>(&mc.d_etab), et.size() * sizeof(EdgeParam)));
        NRLDPC_HIP(hipMemcpy(mc.d_etab, et.data(), et.size() * sizeof(EdgeParam), hipMemcpyHostToDevice));
    }
    auto res = p->codes.emplace(code_key(kind, Z), mc);
    *out = &res.first->second;
    return LDPC_OK;
}

static size_t llr_elem_bytes(int fmt) { return fmt == kLlrI8 ? 1 : fmt == kLlrF16 ? 2 : 4; }

static int lay_index(const ldpc_mixed* p) { return p->layout == kLayoutSlim ? 1 : 0; }

static int mixed_cpb(const ldpc_mixed* p, const MixedCode& mc) {
    const int L = lay_index(p);
    return mc.cpb[L] > 0 ? std::min(mc.cpb[L], mc.cpb_max[L]) : mc.cpb_default[L];
}

static int wait_upload(ldpc_mixed* p) {
    if (p->upload_pending) {
        NRLDPC_HIP(hipEventSynchronize(p->uploaded));
        p->upload_pending = false;
    }
    return LDPC_OK;
}

// Launches of one object share its plan copy and its work counter, so a launch on stream s
// waits (on the GPU) for the previous launch of the object, whatever stream that used.  A
// stream under capture cannot wait for an event recorded outside it: such a caller orders the
// object's launches itself.
static int order_after_previous(ldpc_mixed* p, hipStream_t s) {
    if (!p->upload_pending) return LDPC_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    NRLDPC_HIP(hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) NRLDPC_HIP(hipStreamWaitEvent(s, p->uploaded, 0));
    return LDPC_OK;
}

static int mixed_plan_locked(ldpc_mixed* p, const ldpc_mixed_cw* cws, int n) {
    if (n < 0 || n > p->max_cw || (n > 0 && !cws)) {
        set_last_error("mixed plan: codeword count outside 0..max_cw");
        return LDPC_EINVAL;
    }
    int st = wait_upload(p);
    if (st) return st;
    p->planned = false;
    std::map<int, std::vector<int>> groups;
    for (int i = 0; i < n; i++) {
        const ldpc_mixed_cw& d = cws[i];
        MixedCode* mc = nullptr;
        if ((st = mixed_get_code(p, d.kind, d.Z, &mc))) {
            set_last_error("mixed plan: descriptor " + std::to_string(i) + ": " + ldpc_last_error());
            return st;
        }
        const int rows_full = mc->code.rows_full;
        if (d.m_rows < 4 || d.m_rows > rows_full || d.llr_offset < 0 || d.hd_offset < 0 || d.reserved != 0) {
            set_last_error("mixed plan: descriptor " + std::to_string(i) + " has rows outside 4.." +
                           std::to_string(rows_full) + " or a negative offset");
            return LDPC_EINVAL;
        }
        groups[code_key(d.kind, d.Z)].push_back(i);
    }
    struct HostTile { MixedTile t; long long cost; int seq; };
    std::vector<HostTile> tiles;
    std::vector<int> slot_of;  // descriptor per slot, in group order
    slot_of.reserve(n);
    int trained_T = 0;
    for (auto& g : groups) {
        MixedCode& mc = p->codes[g.first];
        std::vector<int>& v = g.second;
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return cws[a].m_rows > cws[b].m_rows; });
        const int cpb = mixed_cpb(p, mc);
        if (mc.d_weights) {
            if (trained_T == 0) trained_T = mc.weights_iters;
            else if (trained_T != mc.weights_iters) trained_T = -1;
        }
        for (size_t s = 0; s < v.size(); s += cpb) {
            HostTile h{};
            h.t.etab = mc.d_etab;
            h.t.weights = mc.d_weights;
            h.t.trained = mc.d_weights ? 1 : 0;
            h.t.kind = mc.code.kind;
            h.t.Z = mc.code.Z;
            h.t.m_rows = cws[v[s]].m_rows;
            h.t.count = (int)std::min<size_t>(cpb, v.size() - s);
            h.t.cpb = h.t.count;  // a short last tile uses only the lanes it needs
            h.t.first = (int)(slot_of.size());
            h.t.lds_cw_bytes = mc.lds_cw[lay_index(p)];
            for (int k = 0; k < h.t.count; k++) slot_of.push_back(v[s + k]);
            const int waves = (h.t.cpb * mc.code.Z + 63) / 64;
            h.cost = (long long)((waves + 3) / 4) * edges_in_rows(h.t.kind, h.t.m_rows);
            h.seq = (int)tiles.size();
            tiles.push_back(h);
        }
    }
    std::stable_sort(tiles.begin(), tiles.end(), [](const HostTile& a, const HostTile& b) { return a.cost > b.cost; });
    // staging: tiles in queue order, codewords in tile order
    MixedTile* ht = reinterpret_cast<MixedTile*>(p->h_plan + p->tiles_off);
    MixedCw* hc = reinterpret_cast<MixedCw*>(p->h_plan + p->cws_off);
    p->order.assign(n, 0);
    p->tiles6.clear();
    int slot = 0;
    for (size_t t = 0; t < tiles.size(); t++) {
        MixedTile mt = tiles[t].t;
        const int first_old = mt.first;
        mt.first = slot;
        for (int k = 0; k < mt.count; k++) {
            const int i = slot_of[first_old + k];
            const ldpc_mixed_cw& d = cws[i];
            const int kb = d.kind == kBG1 ? 22 : 10;
            MixedCw c{};
            c.llr_off = d.llr_offset;
            c.hd_off = d.hd_offset;
            c.id = i;
            c.m_rows = d.m_rows;
            c.n_vn = (kb + d.m_rows) * d.Z;
            c.hd_bytes = (c.n_vn + 7) / 8;
            hc[slot] = c;
            p->order[slot] = i;
            slot++;
        }
        ht[t] = mt;
        const int six[6] = {mt.kind, mt.Z, mt.m_rows, mt.cpb, mt.first, mt.count};
        p->tiles6.insert(p->tiles6.end(), six, six + 6);
    }
    long long le = 0, he = 0;
    for (int i = 0; i < n; i++) {
        const int kb = cws[i].kind == kBG1 ? 22 : 10;
        const long long nv = (long long)(kb + cws[i].m_rows) * cws[i].Z;
        le = std::max(le, cws[i].llr_offset + nv);
        he = std::max(he, cws[i].hd_offset + (nv + 7) / 8);
    }
    p->plan_llr_extent = le;
    p->plan_hd_extent = he;
    // header: tile count (read by the decode) and LLR extent (read by the narrow-input conversion)
    *reinterpret_cast<int*>(p->h_plan) = (int)tiles.size();
    *reinterpret_cast<long long*>(p->h_plan + 8) = le;
    p->n_tiles = (int)tiles.size();
    p->n_cw = n;
    p->plan_trained_T = trained_T;
    p->planned = true;
    return LDPC_OK;
}

// The plan's LLR (and APP) floats and hard-decision bytes against the buffer sizes a caller
// gave (0: size not given, not checked).
static int check_plan_fits(const ldpc_mixed* p, size_t llr_floats, size_t hd_bytes) {
    if (llr_floats && (size_t)p->plan_llr_extent > llr_floats) {
        set_last_error("mixed: the plan addresses " + std::to_string(p->plan_llr_extent) +
                       " LLR floats, beyond the " + std::to_string(llr_floats) + " given at capture");
        return LDPC_EINVAL;
    }
    if (hd_bytes && (size_t)p->plan_hd_extent > hd_bytes) {
        set_last_error("mixed: the plan addresses " + std::to_string(p->plan_hd_extent) +
                       " hard-decision bytes, beyond the " + std::to_string(hd_bytes) + " given at capture");
        return LDPC_EINVAL;
    }
    return LDPC_OK;
}

static int mixed_args(ldpc_mixed* p, const float* llr, int n_iter, unsigned flags, uint8_t* hd, float* app_out,
                      int* iters_out, MixedArgs* out) {
    const bool learned = (flags & LDPC_FLAG_LEARNED) != 0;
    if (n_iter < 1) {
        set_last_error("mixed decode: n_iter must be positive");
        return LDPC_EINVAL;
    }
    if (learned && p->plan_trained_T && p->plan_trained_T != n_iter) {
        set_last_error(p->plan_trained_T < 0
                           ? "mixed decode: the plan holds learned weights for different iteration counts"
                           : "mixed decode: learned weights in the plan are for another iteration count");
        return LDPC_EINVAL;
    }
    MixedArgs& a = *out;
    a = MixedArgs{};
    a.tiles = reinterpret_cast<const MixedTile*>(p->d_plan + p->tiles_off);
    a.cws = reinterpret_cast<const MixedCw*>(p->d_plan + p->cws_off);
    a.queue = p->d_queue;
    a.llr = llr;
    a.hd = hd;
    a.app_out = app_out;
    a.iters_out = iters_out;
    a.n_tiles = reinterpret_cast<const int*>(p->d_plan);
    a.n_iter = n_iter;
    a.early_exit = (flags & LDPC_FLAG_EARLY_EXIT) ? 1 : 0;
    a.fallback = (learned && (flags & LDPC_FLAG_FALLBACK_PLAIN)) ? 1 : 0;
    a.slot_off = p->lds - 16;
    a.d1_scratch = p->d_d1;
    a.d1_stride = kMixedSlimD1Stride;
    return LDPC_OK;
}

// Bytes of the plan's two staging regions the current plan uses.
static size_t plan_head_bytes(const ldpc_mixed* p) { return p->tiles_off + (size_t)p->n_tiles * sizeof(MixedTile); }
static size_t plan_body_bytes(const ldpc_mixed* p) { return (size_t)std::max(p->n_cw, 1) * sizeof(MixedCw); }

// Float buffer for narrow inputs of at least n floats (grow-only; a launch still reading the old
// one is waited for before it is freed).
static int conv_reserve(ldpc_mixed* p, size_t n) {
    if (p->conv_cap >= n) return LDPC_OK;
    int st = wait_upload(p);
    if (st) return st;
    if (p->d_conv) NRLDPC_HIP(hipFree(p->d_conv));
    p->d_conv = nullptr;
    p->conv_cap = 0;
    NRLDPC_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_conv), std::max<size_t>(n, 1) * sizeof(float)));
    p->conv_cap = n;
    return LDPC_OK;
}

static LlrConvertArgs convert_args(const ldpc_mixed* p, const void* narrow) {
    LlrConvertArgs c{};
    c.in = narrow;
    c.out = p->d_conv;
    c.count = reinterpret_cast<const long long*>(p->d_plan + 8);
    c.fmt = p->llr_fmt;
    c.scale = p->llr_scale;
    return c;
}

// Enqueue the plan upload, the narrow-input conversion (if any) and the kernel on stream s.
static int mixed_enqueue(ldpc_mixed* p, hipStream_t s, const float* llr, int n_iter, unsigned flags, uint8_t* hd,
                         float* app_out, int* iters_out) {
    const bool narrow = p->llr_fmt != kLlrF32;
    int st;
    if (narrow && (st = conv_reserve(p, (size_t)p->plan_llr_extent))) return st;
    MixedArgs a;
    if ((st = mixed_args(p, narrow ? p->d_conv : llr, n_iter, flags, hd, app_out, iters_out, &a))) return st;
    NRLDPC_HIP(hipMemcpyAsync(p->d_plan, p->h_plan, plan_head_bytes(p), hipMemcpyHostToDevice, s));
    NRLDPC_HIP(hipMemcpyAsync(p->d_plan + p->cws_off, p->h_plan + p->cws_off, plan_body_bytes(p),
                              hipMemcpyHostToDevice, s));
    if (narrow && (st = llr_convert_launch(convert_args(p, llr), s))) return st;
    return mixed_launch((flags & LDPC_FLAG_LEARNED) != 0, p->layout == kLayoutSlim, a, p->grid, p->threads,
                        (size_t)p->lds, s);
}

}  // namespace nrldpc

extern "C" {

int ldpc_mixed_create(ldpc_mixed** out, int device, int max_cw) {
    if (!out || max_cw < 1) return LDPC_EINVAL;
    *out = nullptr;
    ldpc_mixed* p = new ldpc_mixed();
    p->device = device;
    p->max_cw = max_cw;
    p->tiles_off = kHdrBytes;
    p->cws_off = (p->tiles_off + (size_t)max_cw * sizeof(MixedTile) + 255) & ~(size_t)255;
    p->plan_bytes = p->cws_off + (size_t)max_cw * sizeof(MixedCw);
    int st = LDPC_OK;
    auto fail = [&](int code) { ldpc_mixed_destroy(p); return code; };
    if (device < 0) {
        p->lds = (kMaxLdsBytes - 1024) & ~15;
        p->h_plan = static_cast<unsigned char*>(calloc(1, p->plan_bytes));
        if (!p->h_plan) return fail(LDPC_ENOMEM);
        *out = p;
        return LDPC_OK;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device >= n) {
        set_last_error("no such HIP device");
        return fail(LDPC_ENODEV);
    }
    if (hipSetDevice(device) != hipSuccess) return fail(LDPC_EHIP);
    p->lds = max_lds_bytes() & ~15;
    if (hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return fail(LDPC_EHIP);
    const int occ = std::max(1, mixed_occupancy(false, false, p->threads, (size_t)p->lds));
    p->grid = p->grid_full = p->n_cu * occ;
    if (hipHostMalloc(reinterpret_cast<void**>(&p->h_plan), p->plan_bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&p->d_plan), p->plan_bytes) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&p->d_queue), 4 * sizeof(unsigned)) != hipSuccess) {
        set_last_error("mixed: allocation failed");
        return fail(LDPC_ENOMEM);
    }
    memset(p->h_plan, 0, p->plan_bytes);
    if (hipMemset(p->d_queue, 0, 4 * sizeof(unsigned)) != hipSuccess ||
        hipEventCreateWithFlags(&p->uploaded, hipEventDisableTiming) != hipSuccess)
        return fail(LDPC_EHIP);
    (void)st;
    *out = p;
    return LDPC_OK;
}

int ldpc_mixed_destroy(ldpc_mixed* p) {
    if (!p) return LDPC_EINVAL;
    if (p->device >= 0) {
        (void)hipSetDevice(p->device);
        if (p->upload_pending) (void)hipEventSynchronize(p->uploaded);
        if (p->exec) (void)hipGraphExecDestroy(p->exec);
        if (p->graph) (void)hipGraphDestroy(p->graph);
        if (p->uploaded) (void)hipEventDestroy(p->uploaded);
        for (auto& c : p->codes) {
            if (c.second.d_etab) (void)hipFree(c.second.d_etab);
            if (c.second.d_weights) (void)hipFree(c.second.d_weights);
        }
        if (p->d_plan) (void)hipFree(p->d_plan);
        if (p->d_conv) (void)hipFree(p->d_conv);
        if (p->d_queue) (void)hipFree(p->d_queue);
        if (p->d_d1) (void)hipFree(p->d_d1);
        if (p->h_plan) (void)hipHostFree(p->h_plan);
    } else {
        free(p->h_plan);
    }
    delete p;
    return LDPC_OK;
}

int ldpc_mixed_plan(ldpc_mixed* p, const ldpc_mixed_cw* cws, int n_cw) {
    if (!p) return LDPC_EINVAL;
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->device >= 0) NRLDPC_HIP(hipSetDevice(p->device));
    return mixed_plan_locked(p, cws, n_cw);
}

int ldpc_mixed_plan_info(const ldpc_mixed* p, int* n_tiles, int* threads, int* grid, int* lds_bytes) {
    if (!p) return LDPC_EINVAL;
    if (n_tiles) *n_tiles = p->planned ? p->n_tiles : 0;
    if (threads) *threads = p->threads;
    if (grid) *grid = p->grid;
    if (lds_bytes) *lds_bytes = p->lds;
    return LDPC_OK;
}

int ldpc_mixed_tiles(const ldpc_mixed* p, int* out, int max_tiles, int* order) {
    if (!p || (!out && max_tiles > 0)) return LDPC_EINVAL;
    if (!p->planned) { set_last_error("mixed: no plan"); return LDPC_EINVAL; }
    const int n = std::min(max_tiles, p->n_tiles);
    if (n > 0) std::copy(p->tiles6.begin(), p->tiles6.begin() + 6 * n, out);
    if (order) std::copy(p->order.begin(), p->order.end(), order);
    return p->n_tiles;
}

int ldpc_mixed_decode_async(ldpc_mixed* p, void* stream, const float* llr, int n_iter, unsigned flags, uint8_t* hd,
                            float* app_out, int* iters_out) {
    if (!p || !llr || !hd) return LDPC_EINVAL;
    if (p->device < 0) { set_last_error("host-only mixed object"); return LDPC_ENODEV; }
    std::lock_guard<std::mutex> lk(p->mu);
    if (!p->planned) { set_last_error("mixed: no plan"); return LDPC_EINVAL; }
    NRLDPC_HIP(hipSetDevice(p->device));
    if (p->n_tiles == 0) return LDPC_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int st = order_after_previous(p, s);
    if (st) return st;
    if ((st = mixed_enqueue(p, s, llr, n_iter, flags, hd, app_out, iters_out))) return st;
    NRLDPC_HIP(hipEventRecord(p->uploaded, s));
    p->upload_pending = true;
    return LDPC_OK;
}

int ldpc_mixed_capture(ldpc_mixed* p, void* stream, const float* llr, int n_iter, unsigned flags, uint8_t* hd,
                       float* app_out, int* iters_out, const float* llr_host, size_t llr_floats, uint8_t* hd_host,
                       size_t hd_bytes) {
    (void)stream;  // the graph is built node by node; it runs on the stream given to the replay
    if (!p || !llr || !hd) return LDPC_EINVAL;
    if (p->device < 0) { set_last_error("host-only mixed object"); return LDPC_ENODEV; }
    std::lock_guard<std::mutex> lk(p->mu);
    if (!p->planned) { set_last_error("mixed: plan a batch before capturing"); return LDPC_EINVAL; }
    if ((llr_host && !llr_floats) || (hd_host && !hd_bytes)) {
        set_last_error("mixed capture: a host buffer needs its size");
        return LDPC_EINVAL;
    }
    const bool narrow = p->llr_fmt != kLlrF32;
    if (narrow && !llr_floats) {
        set_last_error("mixed capture: narrow LLR inputs need llr_floats (the capacity of llr in elements)");
        return LDPC_EINVAL;
    }
    NRLDPC_HIP(hipSetDevice(p->device));
    int st;
    if ((st = check_plan_fits(p, llr_floats, hd_bytes))) return st;
    if (narrow && (st = conv_reserve(p, llr_floats))) return st;
    MixedArgs a;
    if ((st = mixed_args(p, narrow ? p->d_conv : llr, n_iter, flags, hd, app_out, iters_out, &a))) return st;
    if (p->exec) { NRLDPC_HIP(hipGraphExecDestroy(p->exec)); p->exec = nullptr; }
    if (p->graph) { NRLDPC_HIP(hipGraphDestroy(p->graph)); p->graph = nullptr; }
    p->n_head = p->n_body = nullptr;
    hipGraph_t g = nullptr;
    NRLDPC_HIP(hipGraphCreate(&g, 0));
    p->graph = g;
    // graph: plan header + tiles -> codeword table -> decode; every copy is created at its
    // largest size and cut to the current plan at each replay.  The LLR upload and the
    // hard-decision download are stream copies around the graph launch (replay): graph copy
    // nodes of batch size do not overlap kernels of other streams, stream copies go to the DMA
    // engines and do.
    hipGraphNode_t kern = nullptr;
    NRLDPC_HIP(hipGraphAddMemcpyNode1D(&p->n_head, g, nullptr, 0, p->d_plan, p->h_plan,
                                       p->cws_off, hipMemcpyHostToDevice));
    NRLDPC_HIP(hipGraphAddMemcpyNode1D(&p->n_body, g, &p->n_head, 1, p->d_plan + p->cws_off, p->h_plan + p->cws_off,
                                       (size_t)p->max_cw * sizeof(MixedCw), hipMemcpyHostToDevice));
    hipGraphNode_t before = p->n_body;
    if (narrow) {  // the float copy of this replay's LLRs, as many as the current plan addresses
        hipGraphNode_t conv = nullptr;
        if ((st = llr_convert_graph_node(convert_args(p, llr), g, &p->n_body, 1, &conv))) return st;
        before = conv;
    }
    if ((st = mixed_graph_node((flags & LDPC_FLAG_LEARNED) != 0, p->layout == kLayoutSlim, a, p->grid, p->threads,
                               (size_t)p->lds, g, &before, 1, &kern)))
        return st;
    (void)kern;
    p->g_llr_host = llr_host;
    p->g_llr = const_cast<float*>(llr);
    p->g_llr_floats = llr_floats;
    p->g_hd = hd;
    p->g_hd_host = hd_host;
    p->g_hd_bytes = hd_bytes;
    p->g_n_iter = n_iter;
    p->g_flags = flags;
    p->g_elem = llr_elem_bytes(p->llr_fmt);
    NRLDPC_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return LDPC_OK;
}

int ldpc_mixed_replay(ldpc_mixed* p, void* stream) {
    if (!p) return LDPC_EINVAL;
    std::lock_guard<std::mutex> lk(p->mu);
    if (!p->exec) { set_last_error("mixed: nothing captured"); return LDPC_EINVAL; }
    if (!p->planned) { set_last_error("mixed: no plan"); return LDPC_EINVAL; }
    // The current plan must fit what was captured: the buffers' sizes and, for learned tiles,
    // the weights' iteration count (the graph's kernel reads weights of n_iter iterations).
    int st = check_plan_fits(p, p->g_llr_floats, p->g_hd_bytes);
    if (st) return st;
    if ((p->g_flags & LDPC_FLAG_LEARNED) && p->plan_trained_T && p->plan_trained_T != p->g_n_iter) {
        set_last_error("mixed replay: the plan's learned weights are not for the captured iteration count");
        return LDPC_EINVAL;
    }
    NRLDPC_HIP(hipSetDevice(p->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((st = order_after_previous(p, s))) return st;
    // The previous launch of this object has finished using the plan (planning waited for it),
    // so the copies can be resized to what the current plan reads and writes.
    const size_t lf = p->g_llr_host ? (size_t)p->plan_llr_extent : 0;
    const size_t hb = p->g_hd_host ? (size_t)p->plan_hd_extent : 0;
    NRLDPC_HIP(hipGraphExecMemcpyNodeSetParams1D(p->exec, p->n_head, p->d_plan, p->h_plan, plan_head_bytes(p),
                                                 hipMemcpyHostToDevice));
    NRLDPC_HIP(hipGraphExecMemcpyNodeSetParams1D(p->exec, p->n_body, p->d_plan + p->cws_off, p->h_plan + p->cws_off,
                                                 plan_body_bytes(p), hipMemcpyHostToDevice));
    if (lf) NRLDPC_HIP(hipMemcpyAsync(p->g_llr, p->g_llr_host, lf * p->g_elem, hipMemcpyHostToDevice, s));
    NRLDPC_HIP(hipGraphLaunch(p->exec, s));
    if (hb) NRLDPC_HIP(hipMemcpyAsync(p->g_hd_host, p->g_hd, hb, hipMemcpyDeviceToHost, s));
    NRLDPC_HIP(hipEventRecord(p->uploaded, s));
    p->upload_pending = true;
    return LDPC_OK;
}

int ldpc_mixed_set_weights(ldpc_mixed* p, int kind, int Z, int n_iter, const float* w) {
    if (!p || !w || n_iter < 1) return LDPC_EINVAL;
    if (p->device < 0) { set_last_error("host-only mixed object"); return LDPC_ENODEV; }
    std::lock_guard<std::mutex> lk(p->mu);
    NRLDPC_HIP(hipSetDevice(p->device));
    MixedCode* mc = nullptr;
    int st = mixed_get_code(p, kind, Z, &mc);
    if (st) return st;
    if ((st = wait_upload(p))) return st;
    NRLDPC_HIP(hipDeviceSynchronize());  // no decode may still read the old table
    const int nnz = nnz_full(kind);
    std::vector<float> h((size_t)(n_iter + 1) * nnz * 2, 0.f);
    std::copy(w, w + (size_t)n_iter * nnz * 2, h.begin() + (size_t)nnz * 2);
    if (mc->d_weights && mc->weights_iters != n_iter) {
        NRLDPC_HIP(hipFree(mc->d_weights));
        mc->d_weights = nullptr;
    }
    if (!mc->d_weights) NRLDPC_HIP(hipMalloc(reinterpret_cast<void**>(&mc->d_weights), h.size() * sizeof(float)));
    NRLDPC_HIP(hipMemcpy(mc->d_weights, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    mc->weights_iters = n_iter;
    p->planned = false;  // tiles hold weight pointers: plan again
    return LDPC_OK;
}

int ldpc_mixed_set_input(ldpc_mixed* p, int format, float scale) {
    if (!p || (format != LDPC_LLR_F32 && format != LDPC_LLR_I8 && format != LDPC_LLR_F16) ||
        (format == LDPC_LLR_I8 && !(scale > 0.0f && scale < 3.0e38f))) {
        set_last_error("mixed input: unknown element type or a scale that is not a positive finite number");
        return LDPC_EINVAL;
    }
    std::lock_guard<std::mutex> lk(p->mu);
    p->llr_fmt = format;
    p->llr_scale = format == LDPC_LLR_I8 ? scale : 1.0f;
    return LDPC_OK;
}

int ldpc_mixed_set_grid(ldpc_mixed* p, int workgroups) {
    if (!p || workgroups < 0) return LDPC_EINVAL;
    std::lock_guard<std::mutex> lk(p->mu);
    p->grid = workgroups == 0 ? p->grid_full : std::min(workgroups, p->grid_full);
    return LDPC_OK;
}

int ldpc_mixed_set_layout(ldpc_mixed* p, int layout) {
    if (!p || (layout != LDPC_LAYOUT_AUTO && layout != LDPC_LAYOUT_DUP && layout != LDPC_LAYOUT_SLIM)) {
        set_last_error("mixed layout: LDPC_LAYOUT_AUTO, LDPC_LAYOUT_DUP or LDPC_LAYOUT_SLIM");
        return LDPC_EINVAL;
    }
    std::lock_guard<std::mutex> lk(p->mu);
    const int lay = layout == LDPC_LAYOUT_AUTO ? kMixedDefaultLayout : layout;
    const bool slim = lay == kLayoutSlim;
    const int threads = slim ? kSlimMaxBlock : kMixedThreads;
    if (p->device >= 0) {
        NRLDPC_HIP(hipSetDevice(p->device));
        int st = wait_upload(p);
        if (st) return st;
        const int occ = std::max(1, mixed_occupancy(false, slim, threads, (size_t)p->lds));
        const int grid_full = p->n_cu * occ;
        if (slim && p->d1_slots < grid_full) {
            if (p->d_d1) NRLDPC_HIP(hipFree(p->d_d1));
            p->d_d1 = nullptr;
            p->d1_slots = 0;
            NRLDPC_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_d1), (size_t)grid_full * kMixedSlimD1Stride * sizeof(float)));
            p->d1_slots = grid_full;
        }
        // a captured graph holds the other layout's kernel: capture again
        if (p->exec) { NRLDPC_HIP(hipGraphExecDestroy(p->exec)); p->exec = nullptr; }
        if (p->graph) { NRLDPC_HIP(hipGraphDestroy(p->graph)); p->graph = nullptr; }
        p->grid = p->grid_full = grid_full;
    }
    p->threads = threads;
    p->layout = lay;
    p->planned = false;  // tiles hold the layout's codewords per workgroup
    return LDPC_OK;
}

int ldpc_mixed_autotune_layout(ldpc_mixed* p, const ldpc_mixed_cw* cws, int n_cw, const float* llr, int n_iter,
                               unsigned flags, uint8_t* hd, float* us_dup, float* us_slim, int* best_layout) {
    if (!p || !cws || n_cw < 1 || !llr || !hd || n_iter < 1) return LDPC_EINVAL;
    if (p->device < 0) { set_last_error("host-only mixed object"); return LDPC_ENODEV; }
    float us[2] = {0.f, 0.f};
    const int lays[2] = {LDPC_LAYOUT_DUP, LDPC_LAYOUT_SLIM};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = LDPC_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = LDPC_EHIP;
    for (int L = 0; L < 2 && st == LDPC_OK; L++) {
        if ((st = ldpc_mixed_set_layout(p, lays[L]))) break;
        std::lock_guard<std::mutex> lk(p->mu);
        if ((st = mixed_plan_locked(p, cws, n_cw))) break;
        for (int rep = 0; rep < 5 && st == LDPC_OK; rep++) {  // two warm-up launches, three timed
            if (rep == 2 && hipEventRecord(e0, nullptr) != hipSuccess) st = LDPC_EHIP;
            if (st == LDPC_OK) st = mixed_enqueue(p, nullptr, llr, n_iter, flags, hd, nullptr, nullptr);
        }
        float ms = 0.f;
        if (st == LDPC_OK && (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            st = LDPC_EHIP;
        us[L] = ms * 1000.f / 3.f;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    const int best = us[1] < us[0] ? LDPC_LAYOUT_SLIM : LDPC_LAYOUT_DUP;
    const int st2 = ldpc_mixed_set_layout(p, st == LDPC_OK ? best : LDPC_LAYOUT_AUTO);
    if (st) return st;
    if (us_dup) *us_dup = us[0];
    if (us_slim) *us_slim = us[1];
    if (best_layout) *best_layout = best;
    return st2;
}

int ldpc_mixed_get_layout(const ldpc_mixed* p, int* layout) {
    if (!p || !layout) return LDPC_EINVAL;
    *layout = p->layout;
    return LDPC_OK;
}

static int set_cpb_locked(ldpc_mixed* p, int kind, int Z, int cpb, int L) {
    MixedCode* mc = nullptr;
    int st = mixed_get_code(p, kind, Z, &mc);
    if (st) return st;
    if (cpb > mc->cpb_max[L]) {
        set_last_error("codewords per workgroup exceed the workgroup size or LDS");
        return LDPC_EINVAL;
    }
    mc->cpb[L] = cpb;
    return LDPC_OK;
}

int ldpc_mixed_set_cpb(ldpc_mixed* p, int kind, int Z, int cpb) {
    if (!p || cpb < 0) return LDPC_EINVAL;
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->device >= 0) NRLDPC_HIP(hipSetDevice(p->device));
    return set_cpb_locked(p, kind, Z, cpb, lay_index(p));
}

int ldpc_mixed_get_cpb(ldpc_mixed* p, int kind, int Z, int* cpb) {
    if (!p || !cpb) return LDPC_EINVAL;
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->device >= 0) NRLDPC_HIP(hipSetDevice(p->device));
    MixedCode* mc = nullptr;
    int st = mixed_get_code(p, kind, Z, &mc);
    if (st) return st;
    *cpb = mixed_cpb(p, *mc);
    return LDPC_OK;
}

int ldpc_mixed_autotune(ldpc_mixed* p, int kind, int Z, int n_iter, int n_cw, int* cand_cpb, float* us_per_cw,
                        int max_candidates, int* n_candidates, int* best_cpb) {
    if (!p || n_iter < 1 || n_cw < 1 || n_cw > p->max_cw) return LDPC_EINVAL;
    if (p->device < 0) { set_last_error("host-only mixed object"); return LDPC_ENODEV; }
    std::lock_guard<std::mutex> lk(p->mu);
    NRLDPC_HIP(hipSetDevice(p->device));
    MixedCode* mc = nullptr;
    int st = mixed_get_code(p, kind, Z, &mc);
    if (st) return st;
    const int L = lay_index(p);
    const int nvn = mc->code.n_vn(), hb = mc->code.hd_bytes();
    std::vector<ldpc_mixed_cw> d(n_cw);
    for (int i = 0; i < n_cw; i++)
        d[i] = ldpc_mixed_cw{kind, Z, mc->code.rows_full, 0, (long long)i * nvn, (long long)i * hb};
    // Candidates: every count that fills one more wavefront, plus 1 and the maximum.  Without
    // early exit the work is data independent, so all-zero LLRs time the real decoder.
    std::vector<int> cand{1};
    for (int k = 1; k <= p->threads / 64; k++) cand.push_back(std::max(1, std::min(mc->cpb_max[L], 64 * k / Z)));
    cand.push_back(mc->cpb_max[L]);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    float* llr = nullptr;
    uint8_t* hd = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (llr) (void)hipFree(llr);
        if (hd) (void)hipFree(hd);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    if (hipMalloc(reinterpret_cast<void**>(&llr), (size_t)n_cw * nvn * sizeof(float)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&hd), (size_t)n_cw * hb) != hipSuccess ||
        hipMemset(llr, 0, (size_t)n_cw * nvn * sizeof(float)) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        set_last_error("autotune: allocation failed");
        return LDPC_ENOMEM;
    }
    const int saved = mc->cpb[L];
    const int saved_fmt = p->llr_fmt;  // the candidates are timed on float LLRs
    p->llr_fmt = kLlrF32;
    int best = 1;
    float best_us = 3.4e38f;
    for (size_t c = 0; c < cand.size() && st == LDPC_OK; c++) {
        mc->cpb[L] = cand[c];
        if ((st = mixed_plan_locked(p, d.data(), n_cw))) break;
        float ms = 0.f;
        if ((st = order_after_previous(p, nullptr))) break;
        for (int rep = 0; rep < 4 && st == LDPC_OK; rep++) {
            if (rep == 1) (void)hipEventRecord(e0, nullptr);
            st = mixed_enqueue(p, nullptr, llr, n_iter, 0, hd, nullptr, nullptr);
        }
        if (st) break;
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            st = LDPC_EHIP;
            break;
        }
        const float us = ms * 1000.f / 3.f / (float)n_cw;
        if ((int)c < max_candidates) {
            if (us_per_cw) us_per_cw[c] = us;
            if (cand_cpb) cand_cpb[c] = cand[c];
        }
        if (us < best_us) { best_us = us; best = cand[c]; }
    }
    (void)hipDeviceSynchronize();
    cleanup();
    p->llr_fmt = saved_fmt;
    p->planned = false;
    if (st) { mc->cpb[L] = saved; return st; }
    mc->cpb[L] = best;
    if (n_candidates) *n_candidates = (int)cand.size();
    if (best_cpb) *best_cpb = best;
    return LDPC_OK;
}

int ldpc_mixed_save_tuning(const ldpc_mixed* p, const char* path) {
    if (!p || !path) return LDPC_EINVAL;
    FILE* f = fopen(path, "w");
    if (!f) { set_last_error(std::string("cannot write ") + path); return LDPC_EINVAL; }
    fprintf(f, "# kind Z cpb [slim] (codewords per workgroup of the mixed-batch decoder; \"slim\": the slim layout)\n");
    for (int L = 0; L < 2; L++)
        for (const auto& c : p->codes)
            if (c.second.cpb[L] > 0)
                fprintf(f, "%d %d %d%s\n", c.second.code.kind, c.second.code.Z, c.second.cpb[L], L ? " slim" : "");
    fclose(f);
    return LDPC_OK;
}

int ldpc_mixed_load_tuning(ldpc_mixed* p, const char* path) {
    if (!p || !path) return LDPC_EINVAL;
    FILE* f = fopen(path, "r");
    if (!f) { set_last_error(std::string("cannot read ") + path); return LDPC_EINVAL; }
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->device >= 0 && hipSetDevice(p->device) != hipSuccess) { fclose(f); return LDPC_EHIP; }
    char line[256];
    int st = LDPC_OK;
    while (st == LDPC_OK && fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        int kind = 0, Z = 0, cpb = 0;
        char lay[16] = {0};
        const int got = sscanf(line, "%d %d %d %15s", &kind, &Z, &cpb, lay);
        if (got < 3 || cpb < 0 || (got == 4 && strcmp(lay, "slim") != 0)) {
            set_last_error(std::string("malformed tuning line: ") + line);
            st = LDPC_EINVAL;
            break;
        }
        st = set_cpb_locked(p, kind, Z, cpb, got == 4 ? 1 : 0);
    }
    fclose(f);
    return st;
}

}  // extern "C"
