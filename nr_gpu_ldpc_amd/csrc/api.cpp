// api.cpp -- C ABI (include/ldpc_api.h) over the code descriptions and HIP kernels.
#include "../../include/ldpc_api.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "channel.h"
#include "code.h"
#include "decoder.h"
#include "encoder.h"
#include "status.h"
#include "context.h"

namespace nrldpc {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

int nnz_full(int kind) {
    if (kind == kBG1) return Structure<kBG1>::NNZ;
    if (kind == kBG2) return Structure<kBG2>::NNZ;
    return Structure<kWifi>::NNZ;
}
int n_nd(int kind) {
    if (kind == kBG1) return Structure<kBG1>::NND;
    if (kind == kBG2) return Structure<kBG2>::NND;
    return Structure<kWifi>::NND;
}

// Edge order of the compile-time structure = order of the table entries (row-major).
static void edge_list(const CodeDesc& c, std::vector<int>* row, std::vector<int>* col) {
    row->clear();
    col->clear();
    if (c.kind == kBG1)
        for (const auto& e : kBG1Entries) { row->push_back(e.row); col->push_back(e.col); }
    else if (c.kind == kBG2)
        for (const auto& e : kBG2Entries) { row->push_back(e.row); col->push_back(e.col); }
    else
        for (const auto& e : kWifiEntries) { row->push_back(e.row); col->push_back(e.col); }
}

int default_cpb(const CodeDesc& c) {
    // Fill whole wavefronts when possible; otherwise the most lanes per wavefront.
    const size_t lds = decoder_lds_bytes_per_cw(c.kind, c.Z);
    int best = 1;
    double best_eff = 0.0;
    const size_t lim = (size_t)max_lds_bytes();
    for (int cpb = 1; cpb * c.Z <= kMaxBlock && cpb * lds + 16 <= lim; cpb++) {
        const int th = cpb * c.Z;
        const double eff = (double)th / (double)(((th + 63) / 64) * 64);
        if (eff > best_eff + 1e-9) { best_eff = eff; best = cpb; }
    }
    return best;
}

template <class T>
static int upload(std::vector<T> const& h, T** d) {
    const size_t n = std::max<size_t>(h.size(), 1) * sizeof(T);
    NRLDPC_HIP(hipMalloc(reinterpret_cast<void**>(d), n));
    if (!h.empty()) NRLDPC_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return LDPC_OK;
}

void build_edge_params(const CodeDesc& c, std::vector<EdgeParam>* out) {
    const int Z = c.Z;
    std::vector<int> er, ec;
    edge_list(c, &er, &ec);
    const int NND = n_nd(c.kind);
    std::vector<EdgeParam>& et = *out;
    et.assign(er.size(), EdgeParam{0u, 0u, 0u, 0u});
    for (size_t n = 0; n < er.size(); n++) {
        const int v = c.shift(er[n], ec[n]);
        const int j = ec[n];
        if (j < NND) {
            et[n].off_rd = (uint32_t)((j * 2 * Z + v) * 4);
            et[n].acc = ((uint32_t)(Z - v) << 16) | (uint32_t)((j * Z + v) * 4);
            et[n].acc_off = (uint32_t)((j * Z + v) * 4);
        } else {
            et[n].off_rd = (uint32_t)v;
            et[n].acc = 0;
            et[n].acc_off = 0;
        }
    }
}

int ctx_build_device(ldpc_ctx* x) {
    const CodeDesc& c = x->code;
    std::vector<int> er, ec;
    edge_list(c, &er, &ec);
    std::vector<EdgeParam> et;
    build_edge_params(c, &et);
    int st = upload(et, &x->d_etab);
    if (st) return st;
    // syndrome tables over the rows in use
    std::vector<int> rs(c.rows + 1, 0), scol, svm;
    for (size_t n = 0; n < er.size(); n++)
        if (er[n] < c.rows) { rs[er[n] + 1]++; scol.push_back(ec[n]); svm.push_back(c.shift(er[n], ec[n])); }
    for (int i = 0; i < c.rows; i++) rs[i + 1] += rs[i];
    if ((st = upload(rs, &x->d_row_start))) return st;
    if ((st = upload(scol, &x->d_edge_col))) return st;
    if ((st = upload(svm, &x->d_edge_vm))) return st;
    // encoding program
    std::string err;
    st = build_encode_program(c, &x->prog, &err);
    if (st) { set_last_error(err); return st; }
    if ((st = upload(x->prog.step_target, &x->d_step_target))) return st;
    if ((st = upload(x->prog.step_start, &x->d_step_start))) return st;
    if ((st = upload(x->prog.term_col, &x->d_term_col))) return st;
    if ((st = upload(x->prog.term_shift, &x->d_term_shift))) return st;
    // launch shape: the built-in tuning table of (kind, Z), else the dup layout at the default
    // codewords per workgroup; the slim layout's degree-1 slots when the table uses it
    NRLDPC_HIP(hipDeviceGetAttribute(&x->cus, hipDeviceAttributeMultiprocessorCount, x->device));
    x->m_dup = LayoutModel{};
    x->m_slim = LayoutModel{};
    x->m_dup.cpb = x->cpb_default;
    if (const LayoutTuning* t = builtin_layout_tuning(c.kind, c.Z)) {
        if (t->dup.cpb > 0 && t->dup.cpb <= decoder_dup_cpb_max(c.kind, c.Z)) x->m_dup = t->dup;
        if (t->slim.cpb > 0 && t->slim.cpb <= decoder_slim_cpb_max(c.kind, c.Z)) x->m_slim = t->slim;
    }
    x->cpb_default = x->m_dup.cpb;
    x->cpb = x->cpb_default;
    if (x->m_slim.cpb > 0 && (st = ensure_slim(x, x->m_slim.cpb))) return st;
    return LDPC_OK;
}

}  // namespace nrldpc

using namespace nrldpc;

extern "C" {

const char* ldpc_strerror(int status) {
    switch (status) {
        case LDPC_OK: return "ok";
        case LDPC_EINVAL: return "invalid argument";
        case LDPC_ENOMEM: return "out of memory or batch too large";
        case LDPC_EHIP: return "HIP runtime error";
        case LDPC_ENODEV: return "no usable GPU";
        case LDPC_EUNSUPPORTED: return "unsupported";
        default: return "unknown status";
    }
}

const char* ldpc_last_error(void) { return g_last_error.c_str(); }

int ldpc_device_count(int* count) {
    if (!count) return LDPC_EINVAL;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return LDPC_OK;
}

int ldpc_ctx_create(ldpc_ctx** out, int kind, int Z, int m_rows, int device) {
    if (!out) return LDPC_EINVAL;
    *out = nullptr;
    auto* x = new ldpc_ctx();
    std::string err;
    int st = make_code(kind, Z, m_rows, &x->code, &err);
    if (st) { set_last_error(err); delete x; return st; }
    x->cpb_default = default_cpb(x->code);
    x->cpb = x->cpb_default;
    x->m_dup.cpb = x->cpb_default;
    if (device < 0) {
        // host-only context: code description, encoder, generator tables; no device work
        x->device = -1;
        std::string e2;
        st = build_encode_program(x->code, &x->prog, &e2);
        if (st) { set_last_error(e2); delete x; return st; }
        *out = x;
        return LDPC_OK;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_last_error("no HIP device visible");
        delete x;
        return LDPC_ENODEV;
    }
    if (device < 0 || device >= ndev) { set_last_error("device index out of range"); delete x; return LDPC_EINVAL; }
    x->device = device;
    if (hipSetDevice(device) != hipSuccess) { set_last_error("hipSetDevice failed"); delete x; return LDPC_EHIP; }
    x->cpb_default = default_cpb(x->code);
    x->cpb = x->cpb_default;
    st = ctx_build_device(x);
    if (st) { ldpc_ctx_destroy(x); return st; }
    *out = x;
    return LDPC_OK;
}

int ldpc_ctx_destroy(ldpc_ctx* x) {
    if (!x) return LDPC_EINVAL;
    if (x->device < 0) { delete x; return LDPC_OK; }
    (void)hipSetDevice(x->device);
    if (x->scratch_done) {
        (void)hipEventSynchronize(x->scratch_done);
        (void)hipEventDestroy(x->scratch_done);
    }
    if (x->d1_done) {
        (void)hipEventSynchronize(x->d1_done);
        (void)hipEventDestroy(x->d1_done);
    }
    void* ptrs[] = {x->d_etab, x->d_row_start, x->d_edge_col, x->d_edge_vm, x->d_step_target,
                    x->d_step_start, x->d_term_col, x->d_term_shift, x->d_weights, x->d_scratch, x->d_d1};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete x;
    return LDPC_OK;
}

int ldpc_ctx_info(const ldpc_ctx* x, ldpc_code_info* info) {
    if (!x || !info) return LDPC_EINVAL;
    const CodeDesc& c = x->code;
    info->kind = c.kind;
    info->Z = c.Z;
    info->ils = c.ils;
    info->rows = c.rows;
    info->cols = c.cols;
    info->kb = c.kb;
    info->punct = c.punct;
    info->n_vn = c.n_vn();
    info->K = c.K();
    info->N = c.N();
    info->hd_bytes = (c.n_vn() + 7) / 8;
    info->nnz = c.nnz();
    info->cw_per_block = x->cpb;
    info->nnz_full = nnz_full(c.kind);
    return LDPC_OK;
}

int ldpc_set_launch_layout(ldpc_ctx* x, int layout, int cpb) {
    if (!x || cpb < 0) return LDPC_EINVAL;
    const int kind = x->code.kind, Z = x->code.Z;
    if (layout == LDPC_LAYOUT_AUTO) {
        if (cpb != 0) { set_last_error("the automatic layout takes cw_per_block 0"); return LDPC_EINVAL; }
        x->cpb = x->cpb_default;
        if (x->m_slim.cpb > 0 && x->device >= 0) {
            const int st = ensure_slim(x, x->m_slim.cpb);
            if (st) return st;
        }
        x->layout_set = 0;
        return LDPC_OK;
    }
    if (layout == LDPC_LAYOUT_DUP) {
        const int c = cpb ? cpb : x->cpb_default;
        if (c > decoder_dup_cpb_max(kind, Z)) {
            set_last_error("cw_per_block exceeds the workgroup size or LDS limit of the dup layout");
            return LDPC_EINVAL;
        }
        x->cpb = c;
        x->layout_set = kLayoutDup;
        return LDPC_OK;
    }
    if (layout == LDPC_LAYOUT_SLIM) {
        const int c = cpb ? cpb : (x->m_slim.cpb > 0 ? x->m_slim.cpb : decoder_slim_cpb_max(kind, Z));
        const int st = ensure_slim(x, c);
        if (st) return st;
        x->layout_set = kLayoutSlim;
        return LDPC_OK;
    }
    set_last_error("unknown layout");
    return LDPC_EINVAL;
}

int ldpc_set_launch(ldpc_ctx* x, int cpb) {
    if (!x || cpb < 0) return LDPC_EINVAL;
    if (cpb == 0) return ldpc_set_launch_layout(x, LDPC_LAYOUT_AUTO, 0);
    if (cpb <= decoder_dup_cpb_max(x->code.kind, x->code.Z)) return ldpc_set_launch_layout(x, LDPC_LAYOUT_DUP, cpb);
    if (x->device >= 0 && cpb <= decoder_slim_cpb_max(x->code.kind, x->code.Z))
        return ldpc_set_launch_layout(x, LDPC_LAYOUT_SLIM, cpb);
    set_last_error("cw_per_block exceeds the workgroup size or LDS limit");
    return LDPC_EINVAL;
}

int ldpc_launch_layout(const ldpc_ctx* x, int n_cw, int* layout, int* cw_per_block, int* workgroups) {
    if (!x || n_cw < 1) return LDPC_EINVAL;
    int lay = LDPC_LAYOUT_DUP, c = x->cpb, wg = (n_cw + x->cpb - 1) / x->cpb;
    if (x->device >= 0 && slim_chosen(x, n_cw)) {
        lay = LDPC_LAYOUT_SLIM;
        c = x->slim_cpb;
        wg = std::min((n_cw + c - 1) / c, x->cus);
    }
    if (layout) *layout = lay;
    if (cw_per_block) *cw_per_block = c;
    if (workgroups) *workgroups = wg;
    return LDPC_OK;
}

int ldpc_launch_shape(const ldpc_ctx* x, int n_cw, int* cw_per_block, int* workgroups) {
    if (!x || n_cw < 1 || !cw_per_block || !workgroups) return LDPC_EINVAL;
    return ldpc_launch_layout(x, n_cw, nullptr, cw_per_block, workgroups);
}

int ldpc_autotune(ldpc_ctx* x, int n_iter, int max_candidates, int* cand_layout, int* cand_cpb, int* cand_cw_per_cu,
                  float* cand_round_us, int* n_candidates) {
    if (!x || n_iter < 1 || max_candidates < 0) return LDPC_EINVAL;
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    return autotune_layouts(x, n_iter, max_candidates, cand_layout, cand_cpb, cand_cw_per_cu, cand_round_us,
                            n_candidates);
}

int ldpc_get_layout_model(const ldpc_ctx* x, int layout, int* cpb, int* cw_per_cu, float* round_us) {
    if (!x || (layout != LDPC_LAYOUT_DUP && layout != LDPC_LAYOUT_SLIM)) return LDPC_EINVAL;
    const LayoutModel& m = layout == LDPC_LAYOUT_DUP ? x->m_dup : x->m_slim;
    if (cpb) *cpb = m.cpb;
    if (cw_per_cu) *cw_per_cu = m.cw_per_cu;
    if (round_us) *round_us = m.round_us;
    return LDPC_OK;
}

int ldpc_set_layout_model(ldpc_ctx* x, int layout, int cpb, int cw_per_cu, float round_us) {
    if (!x) return LDPC_EINVAL;
    const int kind = x->code.kind, Z = x->code.Z;
    if (layout == LDPC_LAYOUT_SLIM && cpb == 0) {  // drop the slim layout from the model
        x->m_slim = LayoutModel{};
        x->layout_set = 0;
        x->cpb = x->cpb_default;
        return LDPC_OK;
    }
    if (layout != LDPC_LAYOUT_DUP && layout != LDPC_LAYOUT_SLIM) {
        set_last_error("unknown layout");
        return LDPC_EINVAL;
    }
    const int lim = layout == LDPC_LAYOUT_DUP ? decoder_dup_cpb_max(kind, Z) : decoder_slim_cpb_max(kind, Z);
    if (cpb < 1 || cpb > lim) {
        set_last_error("cw_per_block outside 1 .. the layout's workgroup or LDS limit");
        return LDPC_EINVAL;
    }
    if (cw_per_cu < 1 || !(round_us > 0.f) || !std::isfinite(round_us)) {
        set_last_error("cw_per_cu must be >= 1 and round_us positive and finite");
        return LDPC_EINVAL;
    }
    const LayoutModel m{cpb, cw_per_cu, round_us};
    if (layout == LDPC_LAYOUT_DUP) {
        x->m_dup = m;
        x->cpb_default = cpb;
    } else {
        if (x->device >= 0) {
            const int st = ensure_slim(x, cpb);
            if (st) return st;
        }
        x->m_slim = m;
    }
    x->cpb = x->cpb_default;
    x->layout_set = 0;
    return LDPC_OK;
}

int ldpc_decode_async(ldpc_ctx* x, void* stream, const float* llr, int n_cw, int n_iter, unsigned flags,
                      uint8_t* hd_packed, float* app_out, int* iters_out) {
    if (!x || !llr || !hd_packed || n_cw < 0 || n_iter < 1) return LDPC_EINVAL;
    if (n_cw == 0) return LDPC_OK;
    return decode_with(x, stream, llr, n_cw, n_iter, flags, hd_packed, app_out, iters_out, nullptr, nullptr,
                       nullptr);
}

int ldpc_syndrome_async(ldpc_ctx* x, void* stream, const uint8_t* hd, int n_cw, int* weight) {
    if (!x || !hd || !weight || n_cw < 0) return LDPC_EINVAL;
    if (n_cw == 0) return LDPC_OK;
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(x->device));
    SyndromeArgs a{};
    a.hd = hd;
    a.hd_bytes = (x->code.n_vn() + 7) / 8;
    a.Z = x->code.Z;
    a.rows = x->code.rows;
    a.row_start = x->d_row_start;
    a.edge_col = x->d_edge_col;
    a.edge_vm = x->d_edge_vm;
    a.weight = weight;
    a.n_edges = x->code.nnz();
    a.cols = x->code.cols;
    return syndrome_launch(a, n_cw, static_cast<hipStream_t>(stream));
}

int ldpc_base_matrix(int kind, int ils, int* out, int out_len) {
    int rows, cols;
    if (kind == kBG1) { rows = 46; cols = 68; }
    else if (kind == kBG2) { rows = 42; cols = 52; }
    else if (kind == kWifi) { rows = 12; cols = 24; }
    else return LDPC_EINVAL;
    if (!out || out_len < rows * cols || ils < 0 || ils > 7) return LDPC_EINVAL;
    for (int i = 0; i < rows; i++)
        for (int j = 0; j < cols; j++) out[i * cols + j] = base_value(kind, ils, i, j);
    return LDPC_OK;
}

int ldpc_setup_params(int A, float rate, int* out) {
    if (!out || A <= 0) return LDPC_EINVAL;
    // TS 38.212 5.1 / 5.2.2 / 5.3.2 as restated in setup_ldpc, src/nr_ldpc.cpp:37-112.
    const int l = (A > 3824) ? 24 : 16;
    const int B = A + l;
    int bg, K_cb, K_b;
    if ((A <= 292) || (A < 3824 && (double)rate <= 0.67) || ((double)rate <= 0.25)) {
        bg = 2;
        K_cb = 3840;
        K_b = (B > 640) ? 10 : (B > 560) ? 9 : (B > 192) ? 8 : 6;
    } else {
        bg = 1;
        K_cb = 8448;
        K_b = 22;
    }
    if (B > K_cb) {
        set_last_error("transport block needs segmentation (C > 1), which this path does not implement");
        return LDPC_EUNSUPPORTED;
    }
    const int K_prime = B;
    const int Z_min = K_prime / K_b;  // floor, as the reference computes it
    int Z_c = 1000, ils = -1;
    static const int base[8] = {2, 3, 5, 7, 9, 11, 13, 15};
    for (int i = 0; i < 8; i++)
        for (int z = base[i]; z <= 384; z *= 2)
            if (z >= Z_min && z < Z_c) { Z_c = z; ils = i; }
    if (ils < 0) return LDPC_EINVAL;
    const int rows = bg == 1 ? 46 : 42, cols = bg == 1 ? 68 : 52;
    int mr = 0, mc = 0;
    for (int i = 0; i < rows; i++) {
        int n = 0;
        for (int j = 0; j < cols; j++) n += base_value(bg, ils, i, j) >= 0;
        mr = std::max(mr, n);
    }
    for (int j = 0; j < cols; j++) {
        int n = 0;
        for (int i = 0; i < rows; i++) n += base_value(bg, ils, i, j) >= 0;
        mc = std::max(mc, n);
    }
    const int K = (bg == 1 ? 22 : 10) * Z_c, N = (bg == 1 ? 66 : 50) * Z_c;
    const int vals[10] = {bg, Z_c, K, N, rows, cols, ils, N + 2 * Z_c, mr, mc};
    std::memcpy(out, vals, sizeof(vals));
    return LDPC_OK;
}

int ldpc_encode_host(const ldpc_ctx* x, const uint8_t* info, int n_cw, uint8_t* cw) {
    if (!x || !info || !cw || n_cw < 0) return LDPC_EINVAL;
    const CodeDesc& c = x->code;
    for (int w = 0; w < n_cw; w++)
        encode_host(c, x->prog, info + (size_t)w * c.K(), cw + (size_t)w * c.n_vn());
    return LDPC_OK;
}

int ldpc_generator_shifts(int kind, int Z, short* no_shift, short* pointer_shift, short* gen_shift, int gen_len) {
    CodeDesc c;
    std::string err;
    int st = make_code(kind, Z, 0, &c, &err);
    if (st) { set_last_error(err); return st; }
    if (kind == kWifi) return LDPC_EUNSUPPORTED;
    std::vector<std::vector<int>> lists;
    st = generator_shift_lists(c, &lists, &err);
    if (st) { set_last_error(err); return st; }
    const int prow = c.cols - c.kb;
    int pos = 0;
    for (int i = 0; i < prow; i++)
        for (int j = 0; j < c.kb; j++) {
            const auto& L = lists[(size_t)i * c.kb + j];
            if (no_shift) no_shift[i * c.kb + j] = (short)L.size();
            if (pointer_shift) pointer_shift[i * c.kb + j] = (short)pos;
            for (int v : L) {
                if (gen_shift && pos < gen_len) gen_shift[pos] = (short)v;
                pos++;
            }
        }
    return pos;
}

float ldpc_noise_stddev(float ebn0_db, float rate) {
    // compute_noise_stddev, src/encoding_tools.cpp:27-38, with the double-precision libm
    // bindings of pow and sqrt that a plain C++ host compiler selects for that source.
    const float energy_per_bit = 1 / rate;
    const double eb_n0 = std::pow(10.0, (double)(ebn0_db / 10));
    const float n0 = (float)(energy_per_bit / eb_n0);
    const float var = n0 / 2;
    return (float)std::sqrt((double)var);
}

int ldpc_channel_async(ldpc_ctx* x, void* stream, uint64_t seed, uint64_t stream_base, int n_cw,
                       const float* sigma_dev, float sigma_all, uint8_t* cw_bits, float* llr) {
    if (!x || !llr || n_cw < 0) return LDPC_EINVAL;
    if (n_cw == 0) return LDPC_OK;
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(x->device));
    const CodeDesc& c = x->code;
    ChannelArgs a{};
    a.seed = seed;
    a.stream_base = stream_base;
    a.Z = c.Z;
    a.n_vn = c.n_vn();
    a.K = c.K();
    a.punct = c.punct;
    a.n_steps = (int)x->prog.step_target.size();
    a.step_target = x->d_step_target;
    a.step_start = x->d_step_start;
    a.term_col = x->d_term_col;
    a.term_shift = x->d_term_shift;
    a.sigma_dev = sigma_dev;
    a.sigma_all = sigma_all;
    a.cw_bits = cw_bits;
    a.llr = llr;
    return channel_launch(a, n_cw, static_cast<hipStream_t>(stream));
}

int ldpc_encode_async(ldpc_ctx* x, void* stream, const uint8_t* info, int n_cw, uint8_t* cw) {
    if (!x || !info || !cw || n_cw < 0) return LDPC_EINVAL;
    if (n_cw == 0) return LDPC_OK;
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(x->device));
    const CodeDesc& c = x->code;
    ChannelArgs a{};
    a.Z = c.Z;
    a.n_vn = c.n_vn();
    a.K = c.K();
    a.punct = c.punct;
    a.n_steps = (int)x->prog.step_target.size();
    a.step_target = x->d_step_target;
    a.step_start = x->d_step_start;
    a.term_col = x->d_term_col;
    a.term_shift = x->d_term_shift;
    a.info_in = info;
    a.cw_bits = cw;
    a.llr = nullptr;
    return channel_launch(a, n_cw, static_cast<hipStream_t>(stream));
}

int ldpc_count_errors_async(ldpc_ctx* x, void* stream, const uint8_t* hd, const uint8_t* cw_bits, int n_cw,
                            long long* counts) {
    if (!x || !hd || !cw_bits || !counts || n_cw < 0) return LDPC_EINVAL;
    if (n_cw == 0) return LDPC_OK;
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(x->device));
    ErrorArgs a{};
    a.hd = hd;
    a.cw_bits = cw_bits;
    a.hd_bytes = (x->code.n_vn() + 7) / 8;
    a.n_vn = x->code.n_vn();
    a.K = x->code.K();
    a.counts = counts;
    return count_errors_launch(a, n_cw, static_cast<hipStream_t>(stream));
}

int ldpc_decode_half_async(ldpc_ctx* x, void* stream, const uint16_t* llr, int n_cw, int n_iter, unsigned flags,
                           uint8_t* hd, float* app_out, int* iters_out) {
    if (!x || !llr || !hd || n_cw < 0 || n_iter < 1) return LDPC_EINVAL;
    if (n_cw == 0) return LDPC_OK;
    return decode_half(x, stream, llr, n_cw, n_iter, flags, hd, app_out, iters_out);
}

int ldpc_llr_from_int8_async(void* stream, const int8_t* in, size_t n, float scale, float* out) {
    if ((!in || !out) && n) return LDPC_EINVAL;
    return from_int8_launch(in, n, scale, out, static_cast<hipStream_t>(stream));
}

int ldpc_decode_char_async(ldpc_ctx* x, void* stream, const int8_t* llr, int n_cw, int n_iter, unsigned flags,
                           uint8_t* hd, float* app_out, int* iters_out) {
    if (!x || !llr || !hd || n_cw < 0 || n_iter < 1) return LDPC_EINVAL;
    if (n_cw == 0) return LDPC_OK;
    return decode_char(x, stream, llr, n_cw, n_iter, flags, hd, app_out, iters_out);
}

int ldpc_llr_to_half_async(void* stream, const float* llr, size_t n, uint16_t* out) {
    if ((!llr || !out) && n) return LDPC_EINVAL;
    return to_half_launch(llr, n, out, static_cast<hipStream_t>(stream));
}

int ldpc_hd_to_stream_async(ldpc_ctx* x, void* stream, const uint8_t* hd, int n_cw, uint8_t* out) {
    if (!x || !hd || !out || n_cw < 0) return LDPC_EINVAL;
    if (n_cw == 0) return LDPC_OK;
    if (x->device < 0) { set_last_error("host-only context"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(x->device));
    StreamPackArgs a{};
    a.hd = hd;
    a.out = out;
    a.hd_bytes = x->code.hd_bytes();
    a.n_vn = x->code.n_vn();
    a.n_bits = (long long)n_cw * x->code.n_vn();
    return stream_pack_launch(a, static_cast<hipStream_t>(stream));
}

}  // extern "C"
