// tb_host.cpp -- transport blocks around the decoder (TS 38.212 5.1 CRC, 5.2.2 code-block
// segmentation, 5.3.2 filler bits, 5.4.2 rate matching), host side and C ABI.
//
// The reference stops at one code block (setup_ldpc rejects C > 1, src/nr_ldpc.cpp:78-81) and
// has no rate matching; this file adds both so that a transport block of any size goes
// through the decoder.  The device part (de-rate-matching gather, code-block CRC check) is in
// tb.hip; the host computes the gather map once per configuration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ldpc_api.h"
#include "code.h"
#include "context.h"
#include "status.h"
#include "tb.h"

namespace nrldpc {

uint32_t crc_poly(int type, int* len) {
    switch (type) {
        case LDPC_CRC16: *len = 16; return 0x1021u;     // D^16 + D^12 + D^5 + 1
        case LDPC_CRC24A: *len = 24; return 0x864CFBu;  // gCRC24A, TS 38.212 5.1
        case LDPC_CRC24B: *len = 24; return 0x800063u;  // gCRC24B
        default: *len = 0; return 0u;
    }
}

// Remainder of bits[0..n) * D^L modulo g (bit 0 first = highest power), register initialised to 0.
uint32_t crc_bits(int type, const uint8_t* bits, long long n) {
    int L = 0;
    const uint32_t g = crc_poly(type, &L);
    const uint32_t top = 1u << (L - 1), mask = (L == 32) ? 0xffffffffu : ((1u << L) - 1u);
    uint32_t r = 0;
    for (long long k = 0; k < n; k++) {
        const uint32_t in = (bits[k] & 1u) ? top : 0u;
        r ^= in;
        r = (r & top) ? ((r << 1) ^ g) : (r << 1);
        r &= mask;
    }
    return r;
}

static const int kLift[51] = {2,  3,  4,  5,  6,  7,  8,  9,  10,  11,  12,  13,  14,  15,  16,  18,  20,
                              22, 24, 26, 28, 30, 32, 36, 40, 44,  48,  52,  56,  60,  64,  72,  80,  88,
                              96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384};

int tb_setup(int A, float rate, int bg, TbConfig* t, std::string* err) {
    if (A < 1) { *err = "transport block size must be positive"; return LDPC_EINVAL; }
    TbConfig c{};
    c.A = A;
    c.L_tb = A > 3824 ? 24 : 16;
    c.B = A + c.L_tb;
    // base graph (TS 38.212 7.2.2) unless given
    if (bg == 0) bg = (A <= 292 || (A <= 3824 && (double)rate <= 0.67) || (double)rate <= 0.25) ? 2 : 1;
    if (bg != 1 && bg != 2) { *err = "base graph must be 1 or 2"; return LDPC_EINVAL; }
    c.bg = bg;
    const int Kcb = bg == 1 ? 8448 : 3840;
    int Bp;
    if (c.B <= Kcb) {
        c.L_cb = 0;
        c.C = 1;
        Bp = c.B;
    } else {
        c.L_cb = 24;
        c.C = (c.B + (Kcb - c.L_cb) - 1) / (Kcb - c.L_cb);
        Bp = c.B + c.C * c.L_cb;
    }
    if (Bp % c.C != 0) {
        *err = "B' = " + std::to_string(Bp) + " is not a multiple of C = " + std::to_string(c.C) +
               " (not a TS 38.214 transport block size)";
        return LDPC_EINVAL;
    }
    c.K_prime = Bp / c.C;
    int Kb = 22;
    if (bg == 2) Kb = c.B > 640 ? 10 : c.B > 560 ? 9 : c.B > 192 ? 8 : 6;
    c.Z = 0;
    for (int z : kLift)
        if (Kb * z >= c.K_prime) { c.Z = z; break; }
    if (c.Z == 0) { *err = "no lifting size holds the code block"; return LDPC_EINVAL; }
    c.kb = bg == 1 ? 22 : 10;
    c.cols = bg == 1 ? 68 : 52;
    c.K = c.kb * c.Z;
    c.F = c.K - c.K_prime;
    c.N = (c.cols - 2) * c.Z;
    c.n_vn = c.cols * c.Z;
    *t = c;
    return LDPC_OK;
}

int tb_rate_match_setup(TbConfig* t, int G, int Qm, int NL, int rv, std::string* err) {
    if (G < 1 || Qm < 1 || NL < 1 || rv < 0 || rv > 3 || (Qm != 1 && Qm != 2 && Qm != 4 && Qm != 6 && Qm != 8) ||
        G % (NL * Qm) != 0) {
        *err = "rate matching: G must be a positive multiple of NL * Qm, Qm in {1,2,4,6,8}, rv in 0..3";
        return LDPC_EINVAL;
    }
    t->G = G;
    t->Qm = Qm;
    t->NL = NL;
    t->rv = rv;
    const int Ncb = t->N;  // full circular buffer (I_LBRM = 0)
    const long long Zc = t->Z;
    static const int num1[4] = {0, 17, 33, 56}, num2[4] = {0, 13, 25, 43};
    const int den = t->bg == 1 ? 66 : 50;
    const int num = t->bg == 1 ? num1[rv] : num2[rv];
    t->k0 = (int)(((long long)num * Ncb / (den * Zc)) * Zc);
    // E_r, TS 38.212 5.4.2.1 (all code blocks scheduled)
    const int C = t->C, per = G / (NL * Qm);
    t->E.assign(C, 0);
    for (int r = 0; r < C; r++) {
        if (r <= C - (per % C) - 1) t->E[r] = NL * Qm * (per / C);
        else t->E[r] = NL * Qm * ((per + C - 1) / C);
    }
    t->E_off.assign(C + 1, 0);
    for (int r = 0; r < C; r++) t->E_off[r + 1] = t->E_off[r] + t->E[r];
    return LDPC_OK;
}

// Circular-buffer position of every bit e_k of code block r's bit selection (filler skipped).
static void selection_positions(const TbConfig& t, int E, std::vector<int>* pos) {
    const int Ncb = t.N;
    const int f0 = t.K_prime - 2 * t.Z, f1 = t.K - 2 * t.Z;  // filler positions in d
    pos->resize(E);
    long long j = 0;
    for (int k = 0; k < E;) {
        const int p = (int)((t.k0 + j) % Ncb);
        j++;
        if (p >= f0 && p < f1) continue;
        (*pos)[k++] = p;
    }
}

int tb_build_gather(const TbConfig& t, std::vector<int>* start, std::vector<int>* index) {
    // per code block r and decoder column c: the received positions (ascending k) feeding it
    const int C = t.C, nvn = t.n_vn;
    std::vector<std::vector<int>> lists((size_t)C * nvn);
    std::vector<int> pos;
    for (int r = 0; r < C; r++) {
        const int E = t.E[r], Q = t.Qm;
        selection_positions(t, E, &pos);
        std::vector<int> col_of_e(E);
        for (int k = 0; k < E; k++) col_of_e[k] = 2 * t.Z + pos[k];
        // received bit f_m (m = i + jj*Q) carries e_{i*E/Q + jj} (TS 38.212 5.4.2.2)
        std::vector<int> e_of_f(E);
        for (int jj = 0; jj < E / Q; jj++)
            for (int i = 0; i < Q; i++) e_of_f[i + jj * Q] = i * (E / Q) + jj;
        // gather in ascending e order: invert the interleaver
        std::vector<int> f_of_e(E);
        for (int m = 0; m < E; m++) f_of_e[e_of_f[m]] = m;
        for (int k = 0; k < E; k++) lists[(size_t)r * nvn + col_of_e[k]].push_back(t.E_off[r] + f_of_e[k]);
    }
    start->assign((size_t)C * nvn + 1, 0);
    index->clear();
    for (size_t n = 0; n < lists.size(); n++) {
        (*start)[n + 1] = (*start)[n] + (int)lists[n].size();
        index->insert(index->end(), lists[n].begin(), lists[n].end());
    }
    return LDPC_OK;
}

}  // namespace nrldpc

using namespace nrldpc;

struct ldpc_tb {
    TbConfig cfg;
    int device = -1;
    ldpc_ctx* code = nullptr;  // host-only encoder context of (bg, Z)
    int* d_start = nullptr;
    int* d_index = nullptr;
    std::vector<int> h_start, h_index;
};

extern "C" {

int ldpc_crc(int type, const uint8_t* bits, long long n, uint32_t* crc) {
    int L = 0;
    crc_poly(type, &L);
    if (!crc || (n > 0 && !bits) || n < 0 || L == 0) return LDPC_EINVAL;
    *crc = crc_bits(type, bits, n);
    return LDPC_OK;
}

int ldpc_tb_create(ldpc_tb** out, int A, float rate, int bg, int G, int Qm, int rv, int device) {
    if (!out) return LDPC_EINVAL;
    *out = nullptr;
    ldpc_tb* t = new ldpc_tb();
    std::string err;
    int st = tb_setup(A, rate, bg, &t->cfg, &err);
    if (!st) st = tb_rate_match_setup(&t->cfg, G, Qm, 1, rv, &err);
    if (st) { set_last_error(err); delete t; return st; }
    if ((st = ldpc_ctx_create(&t->code, t->cfg.bg, t->cfg.Z, 0, -1))) { delete t; return st; }
    tb_build_gather(t->cfg, &t->h_start, &t->h_index);
    t->device = device;
    if (device >= 0) {
        auto fail = [&](int s) { ldpc_tb_destroy(t); return s; };
        if (hipSetDevice(device) != hipSuccess) return fail(LDPC_ENODEV);
        if (hipMalloc(reinterpret_cast<void**>(&t->d_start), t->h_start.size() * sizeof(int)) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&t->d_index), std::max<size_t>(1, t->h_index.size()) * sizeof(int)) !=
                hipSuccess)
            return fail(LDPC_ENOMEM);
        if (hipMemcpy(t->d_start, t->h_start.data(), t->h_start.size() * sizeof(int), hipMemcpyHostToDevice) !=
                hipSuccess ||
            (!t->h_index.empty() && hipMemcpy(t->d_index, t->h_index.data(), t->h_index.size() * sizeof(int),
                                              hipMemcpyHostToDevice) != hipSuccess))
            return fail(LDPC_EHIP);
    }
    *out = t;
    return LDPC_OK;
}

int ldpc_tb_destroy(ldpc_tb* t) {
    if (!t) return LDPC_EINVAL;
    if (t->code) ldpc_ctx_destroy(t->code);
    if (t->device >= 0) {
        (void)hipSetDevice(t->device);
        if (t->d_start) (void)hipFree(t->d_start);
        if (t->d_index) (void)hipFree(t->d_index);
    }
    delete t;
    return LDPC_OK;
}

int ldpc_tb_get_info(const ldpc_tb* t, ldpc_tb_info* info, int* E, int max_e) {
    if (!t || !info) return LDPC_EINVAL;
    const TbConfig& c = t->cfg;
    ldpc_tb_info i{};
    i.A = c.A; i.L_tb = c.L_tb; i.B = c.B; i.bg = c.bg; i.C = c.C; i.L_cb = c.L_cb; i.K_prime = c.K_prime;
    i.K = c.K; i.Z = c.Z; i.F = c.F; i.N = c.N; i.n_vn = c.n_vn; i.G = c.G; i.Qm = c.Qm; i.rv = c.rv; i.k0 = c.k0;
    *info = i;
    if (E)
        for (int r = 0; r < std::min(max_e, c.C); r++) E[r] = c.E[r];
    return LDPC_OK;
}

int ldpc_tb_encode_host(const ldpc_tb* t, const uint8_t* a, uint8_t* f, uint8_t* cw) {
    if (!t || !a || !f) return LDPC_EINVAL;
    const TbConfig& c = t->cfg;
    // TB CRC (TS 38.212 7.2.1: CRC24A above 3824 bits, CRC16 otherwise)
    std::vector<uint8_t> b(c.B);
    for (int k = 0; k < c.A; k++) b[k] = a[k] & 1u;
    const uint32_t p = crc_bits(c.L_tb == 24 ? LDPC_CRC24A : LDPC_CRC16, b.data(), c.A);
    for (int k = 0; k < c.L_tb; k++) b[c.A + k] = (p >> (c.L_tb - 1 - k)) & 1u;
    std::vector<uint8_t> blk(c.K), code(c.n_vn);
    std::vector<int> pos;
    int s = 0;
    for (int r = 0; r < c.C; r++) {
        // code block: K' - L_cb data bits, its CRC24B, then F filler bits (encoded as 0)
        const int data = c.K_prime - c.L_cb;
        std::fill(blk.begin(), blk.end(), 0);
        for (int k = 0; k < data; k++) blk[k] = b[s++];
        if (c.L_cb) {
            const uint32_t q = crc_bits(LDPC_CRC24B, blk.data(), data);
            for (int k = 0; k < 24; k++) blk[data + k] = (q >> (23 - k)) & 1u;
        }
        int st = ldpc_encode_host(t->code, blk.data(), 1, code.data());
        if (st) return st;
        if (cw) std::copy(code.begin(), code.end(), cw + (size_t)r * c.n_vn);
        // bit selection from d = columns 2Z .. (circular), then the Qm interleaver
        const int E = c.E[r], Q = c.Qm;
        selection_positions(c, E, &pos);
        uint8_t* fr = f + c.E_off[r];
        for (int jj = 0; jj < E / Q; jj++)
            for (int i = 0; i < Q; i++) fr[i + jj * Q] = code[2 * c.Z + pos[i * (E / Q) + jj]];
    }
    return LDPC_OK;
}

int ldpc_tb_derate_async(const ldpc_tb* t, void* stream, const float* llr_in, float* llr_cb) {
    if (!t || !llr_in || !llr_cb) return LDPC_EINVAL;
    if (t->device < 0) { set_last_error("host-only transport block"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(t->device));
    DerateArgs a{};
    a.in = llr_in;
    a.out = llr_cb;
    a.start = t->d_start;
    a.index = t->d_index;
    a.n = t->cfg.C * t->cfg.n_vn;
    a.n_vn = t->cfg.n_vn;
    a.f0 = t->cfg.K_prime;  // filler columns K' .. K-1 of every code block
    a.f1 = t->cfg.K;
    return derate_launch(a, static_cast<hipStream_t>(stream));
}

int ldpc_tb_check_async(const ldpc_tb* t, void* stream, const uint8_t* hd, int* cb_ok, uint8_t* tb_bits) {
    if (!t || !hd || !cb_ok) return LDPC_EINVAL;
    if (t->device < 0) { set_last_error("host-only transport block"); return LDPC_ENODEV; }
    NRLDPC_HIP(hipSetDevice(t->device));
    TbCheckArgs a{};
    a.hd = hd;
    a.cb_ok = cb_ok;
    a.tb_bits = tb_bits;
    a.C = t->cfg.C;
    a.hd_bytes = (t->cfg.n_vn + 7) / 8;
    a.K_prime = t->cfg.K_prime;
    a.L_cb = t->cfg.L_cb;
    a.B = t->cfg.B;
    return tb_check_launch(a, static_cast<hipStream_t>(stream));
}

int ldpc_tb_decoder_context(const ldpc_tb* t, int device, ldpc_ctx** out) {
    if (!t || !out) return LDPC_EINVAL;
    return ldpc_ctx_create(out, t->cfg.bg, t->cfg.Z, 0, device);
}

}  // extern "C"
