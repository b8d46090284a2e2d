// compat.cpp -- the reference's C++ host API (include/nr_gpu_ldpc.hpp) over the C ABI.
//
// Host-side channel functions restate src/encoding_tools.cpp:13-155 and
// src/host_helpers.cpp:46-64 (the same random-number calls, so a program seeded like the
// reference draws the same bits and noise); memory, grid and decode functions follow
// src/setup_exp.cu:48-312 and src/gpu_decoder_top.cu:16-52, with the decode itself one fused
// launch of ldpc_decode_async.
#include "../../include/nr_gpu_ldpc.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

namespace {

ldpc_ctx* g_ctx = nullptr;  // the process-wide code, as the reference's __constant__ memory
int g_n_cw = 0, g_n_iter = 0, g_n_vn = 0;  // batch geometry of that code (ldpc_gpu_decoder)

// Device buffers of ldpc_gpu_decoder, sized for the process-wide geometry.
struct CEntryBuffers {
    int8_t* in = nullptr;
    uint8_t* hd = nullptr;
    uint8_t* packed = nullptr;
    hipStream_t stream = nullptr;
    size_t n = 0;
};
CEntryBuffers g_cbuf;

void free_centry_buffers() {
    if (g_cbuf.in) (void)hipFree(g_cbuf.in);
    if (g_cbuf.hd) (void)hipFree(g_cbuf.hd);
    if (g_cbuf.packed) (void)hipFree(g_cbuf.packed);
    if (g_cbuf.stream) (void)hipStreamDestroy(g_cbuf.stream);
    g_cbuf = CEntryBuffers{};
}

[[noreturn]] void die_ldpc(int st, const char* what) {
    fprintf(stderr, "nr_gpu_ldpc: %s failed: %s (%s)\n", what, ldpc_strerror(st), ldpc_last_error());
    exit(EXIT_FAILURE);
}

[[noreturn]] void die_hip(hipError_t e, const char* what) {
    fprintf(stderr, "nr_gpu_ldpc: %s failed: %s\n", what, hipGetErrorString(e));
    exit(EXIT_FAILURE);
}

[[noreturn]] void config_error(const char* what) {
    fprintf(stderr, "nr_gpu_ldpc: %s\n", what);
    abort();
}

#define NRLDPC_CHECK(call)                      \
    do {                                        \
        const int st_ = (call);                 \
        if (st_ != LDPC_OK) die_ldpc(st_, #call); \
    } while (0)
#define NRLDPC_CHECK_HIP(call)                   \
    do {                                         \
        const hipError_t e_ = (call);            \
        if (e_ != hipSuccess) die_hip(e_, #call); \
    } while (0)

// Host-only contexts (code tables and encoding programs, no device) per (kind, Z).
ldpc_ctx* host_code(int kind, int Z) {
    static std::map<int, ldpc_ctx*> codes;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    auto it = codes.find(kind * 1024 + Z);
    if (it != codes.end()) return it->second;
    ldpc_ctx* c = nullptr;
    if (ldpc_ctx_create(&c, kind, Z, 0, -1) != LDPC_OK) return nullptr;
    codes[kind * 1024 + Z] = c;
    return c;
}

const short kLiftSizes[51] = {2,  3,  4,  5,  6,  7,  8,  9,  10,  11,  12,  13,  14,  15,  16,  18,  20,
                              22, 24, 26, 28, 30, 32, 36, 40, 44,  48,  52,  56,  60,  64,  72,  80,  88,
                              96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384};

void fill_degrees(ldpc_params* l) {
    l->h_element_count_per_row = static_cast<char*>(calloc(l->H_rows, 1));
    l->h_element_count_per_col = static_cast<char*>(calloc(l->H_cols, 1));
    int mr = 0, mc = 0;
    for (int i = 0; i < l->H_rows; i++)
        for (int j = 0; j < l->H_cols; j++)
            if (l->H[i * l->H_cols + j] >= 0) {
                l->h_element_count_per_row[i]++;
                l->h_element_count_per_col[j]++;
            }
    for (int i = 0; i < l->H_rows; i++) mr = std::max<int>(mr, l->h_element_count_per_row[i]);
    for (int j = 0; j < l->H_cols; j++) mc = std::max<int>(mc, l->h_element_count_per_col[j]);
    l->n_nz_in_row = mr;
    l->n_nz_in_col = mc;
}

// Marsaglia polar AWGN of src/encoding_tools.cpp:40-65 (drand48 draws, float operands, and the
// double-precision log / sqrt that the unqualified calls bind to in a plain C++ translation unit).
void awgn(const float* channel_in, float* channel_out, int n_symbols, float sigma) {
    for (int i = 0; i < n_symbols; i++) {
        float u1, u2, s;
        do {
            const float r1 = (float)drand48();
            const float r2 = (float)drand48();
            u1 = r1 * 2.0f - 1.0f;
            u2 = r2 * 2.0f - 1.0f;
            s = u1 * u1 + u2 * u2;
        } while (s >= 1);
        const float noise = (float)(u1 * std::sqrt((-2.0f * std::log((double)s)) / s));
        channel_out[i] = channel_in[i] + noise * sigma;
    }
}

}  // namespace

ldpc_ctx* nr_gpu_ldpc_context() { return g_ctx; }

ldpc_params setup_ldpc(int n_bits, float code_rate, int n_iterations) {
    int out[10];
    const int st = ldpc_setup_params(n_bits, code_rate, out);
    if (st != LDPC_OK) config_error(ldpc_last_error());
    ldpc_params l{};
    l.A = n_bits;
    l.rate = code_rate;
    l.bg_index = out[0];
    l.Z = out[1];
    l.K = out[2];
    l.N = out[3];
    l.H_rows = out[4];
    l.H_cols = out[5];
    l.N_before_puncture = out[7];
    l.n_iterations = n_iterations;
    l.H = static_cast<int*>(malloc(sizeof(int) * l.H_rows * l.H_cols));
    if (ldpc_base_matrix(l.bg_index, out[6], l.H, l.H_rows * l.H_cols) != LDPC_OK) config_error(ldpc_last_error());
    fill_degrees(&l);
    return l;
}

ldpc_params setup_ldpc_qc12x24(int Z, int n_iterations) {
    ldpc_ctx* c = host_code(LDPC_CODE_QC12X24, Z);
    if (!c) config_error("12 x 24 code: unsupported lifting size");
    ldpc_params l{};
    l.bg_index = LDPC_CODE_QC12X24;
    l.Z = Z;
    l.K = 12 * Z;
    l.A = l.K;
    l.N = 24 * Z;
    l.rate = 0.5f;
    l.H_rows = 12;
    l.H_cols = 24;
    l.N_before_puncture = 24 * Z;
    l.n_iterations = n_iterations;
    l.H = static_cast<int*>(malloc(sizeof(int) * 12 * 24));
    if (ldpc_base_matrix(LDPC_CODE_QC12X24, 0, l.H, 12 * 24) != LDPC_OK) config_error(ldpc_last_error());
    fill_degrees(&l);
    return l;
}

int ldpc_encoder_orig(int* test_input, int* channel_input, short block_length, short BG, unsigned char gen_code) {
    if (gen_code) {
        fprintf(stderr, "ldpc_encoder_orig: code generation (gen_code = 1) is not provided\n");
        return -1;
    }
    if (BG == LDPC_CODE_QC12X24) {  // extension: K = 12 Z, no puncturing, all 24 Z bits out
        const int Zc = block_length / 12;
        ldpc_ctx* c = (block_length % 12 == 0) ? host_code(BG, Zc) : nullptr;
        if (!c) {
            fprintf(stderr, "ldpc_encoder_orig: no 12 x 24 code for %d bits\n", block_length);
            return -1;
        }
        std::vector<uint8_t> in(block_length), cw(24 * Zc);
        for (int i = 0; i < block_length; i++) in[i] = (uint8_t)(test_input[i] & 1);
        if (ldpc_encode_host(c, in.data(), 1, cw.data()) != LDPC_OK) return -1;
        for (int i = 0; i < 24 * Zc; i++) channel_input[i] = cw[i];
        return 0;
    }
    if (BG != 1 && BG != 2) return -1;
    // lifting size and base-graph shape as src/nr_ldpc.cpp:681-716
    const int nrows = BG == 1 ? 46 : 42;
    const int ncols = BG == 1 ? 22 : 10;
    int Kb = 22;
    if (BG == 2) Kb = block_length > 640 ? 10 : block_length > 560 ? 9 : block_length > 192 ? 8 : 6;
    int Zc = 0;
    for (short z : kLiftSizes)
        if (z >= (double)block_length / Kb) {
            Zc = z;
            break;
        }
    ldpc_ctx* c = Zc ? host_code(BG, Zc) : nullptr;
    if (!c || block_length > ncols * Zc) {
        printf("ldpc_encoder_orig: could not determine lifting size\n");
        return -1;
    }
    std::vector<uint8_t> in(ncols * Zc, 0), cw((ncols + nrows) * Zc);
    for (int i = 0; i < block_length; i++) in[i] = (uint8_t)(test_input[i] & 1);
    if (ldpc_encode_host(c, in.data(), 1, cw.data()) != LDPC_OK) return -1;
    // output: information bits from column 2 on, then the parity bits, with the punctured
    // parity columns of src/nr_ldpc.cpp:736-737,905-910 left out
    const int no_punctured_columns = (int)((nrows - 2) * Zc + block_length - block_length * 3) / Zc;
    const int removed_bit = (nrows - no_punctured_columns - 2) * Zc + block_length - (block_length * 3);
    for (int i = 0; i < block_length - 2 * Zc; i++) channel_input[i] = in[2 * Zc + i];
    const int n_par = (nrows - no_punctured_columns) * Zc - removed_bit;
    for (int i = 0; i < n_par; i++) channel_input[block_length - 2 * Zc + i] = cw[ncols * Zc + i];
    return 0;
}

short* choose_generator_matrix(short BG, short Zc) {
    static std::map<int, std::vector<short>> tables;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const int key = BG * 1024 + Zc;
    auto it = tables.find(key);
    if (it != tables.end()) return it->second.data();
    const int kb = BG == 1 ? 22 : 10, rows = BG == 1 ? 46 : 42;
    std::vector<short> no(rows * kb), ptr(rows * kb), gen((size_t)rows * kb * Zc);
    const int n = ldpc_generator_shifts(BG, Zc, no.data(), ptr.data(), gen.data(), (int)gen.size());
    if (n < 0) return nullptr;
    gen.resize(n);
    return tables.emplace(key, std::move(gen)).first->second.data();
}

float compute_noise_stddev(float snr, float code_rate, int /*modulation*/) {
    return ldpc_noise_stddev(snr, code_rate);
}

void transmit(int* info_bin, int* codeword, float* modulate_cwds, unsigned long mem_size_codeword, float* recv,
              int n_info_bits, int n_coded_bits, float sigma, int bg_index) {
    for (int i = 0; i < n_info_bits; i++) info_bin[i] = rand() % 2;
    memset(codeword, 0, mem_size_codeword);
    ldpc_encoder_orig(info_bin, codeword, (short)n_info_bits, (short)bg_index, 0);
    for (int i = 0; i < n_coded_bits; i++) modulate_cwds[i] = codeword[i] == 0 ? 1.0f : -1.0f;
    awgn(modulate_cwds, recv, n_coded_bits, sigma);
}

void calculate_llr(float llr[], float recv[], int n_symbols, float sigma, int z) {
    const int two_cols = 2 * z;
    for (int i = 0; i < n_symbols; i++) llr[i] = i < two_cols ? 0.0f : (recv[i - two_cols] * 2) / (sigma * sigma);
}

int error_check(float trans[], float recv[], int n_symbols) {
    int cnt = 0;
    for (int i = 0; i < n_symbols; i++) cnt += recv[i] * trans[i] < 0;
    return cnt;
}

error_result cuda_error_check(int info[], char hard_decision[], int n_total_cws, int codeword_len, int info_length) {
    error_result e{0, 0, 0, 0};
    for (int i = 0; i < n_total_cws; i++) {
        const char* hd = hard_decision + (size_t)i * codeword_len;
        const int* in = info + (size_t)i * info_length;
        int be = 0;
        for (int j = 0; j < info_length; j++) be += in[j] != hd[j];
        e.bit_error += be;
        e.n_bits += info_length;
        e.frame_error += be != 0;
        e.n_frames++;
    }
    return e;
}

signed char quantize(double D, double x, unsigned char B) {
    double q = std::floor(x / D);
    const short maxlev = (short)(1 << (B - 1));
    if (q <= -maxlev) q = -maxlev;
    else if (q >= maxlev) q = maxlev - 1;
    return (signed char)q;
}

void unpack_char(char* packed_input, char* unpacked_output, int size) {
    for (int i = 0; i < size; i++) unpacked_output[i] = (char)((packed_input[i / 8] >> (i & 7)) & 1);
}

template <typename T>
host_mem<T> setup_host(const params* p, unsigned int n_info_bits, unsigned int n_cw_bits) {
    static_assert(std::is_same<T, float>::value || std::is_same<T, __half>::value || std::is_same<T, char>::value,
                  "the FP32, HALF_PRC and CHAR_PRC decode paths are instantiated");
    host_mem<T> h{};
    h.mem_size_infobits = n_info_bits * sizeof(int);
    h.mem_size_cw = n_cw_bits * sizeof(int);
    h.mem_size_llr = n_cw_bits * sizeof(float);
    h.info_bin = static_cast<int*>(malloc(h.mem_size_infobits));
    h.cwds = static_cast<int*>(malloc(h.mem_size_cw));
    h.modulated_cwds = static_cast<float*>(malloc(h.mem_size_llr));
    h.llr = static_cast<float*>(malloc(h.mem_size_llr));
    h.recv = static_cast<float*>(malloc(h.mem_size_llr));
    h.llr_quan = static_cast<T*>(malloc(n_cw_bits * sizeof(T)));
    const long long n_cw = (long long)p->n_mcw * p->n_cw_per_mcw;
    // The per-stream byte counts are int fields (as in the reference's host_mem); a batch whose
    // LLR buffer would not fit one is refused instead of being allocated with a wrapped size.
    if (n_cw < 1 || n_cw * (long long)n_cw_bits * (long long)sizeof(T) > 0x7fffffffLL)
        config_error("setup_host: n_mcw * n_cw_per_mcw codewords must be >= 1 and their LLR bytes fit an int");
    h.mem_size_llr_cuda = (int)(n_cw * n_cw_bits * sizeof(T));
    h.mem_size_hd_cuda = (int)(n_cw * n_cw_bits);
    h.mem_size_hd_cuda_packed = (int)((n_cw * n_cw_bits + 7) / 8);
    const size_t info_cuda = (size_t)n_cw * n_info_bits * sizeof(int);
    h.info_bin_cuda = new int*[p->n_streams];
    h.llr_cuda = new T*[p->n_streams];
    h.hd_cuda = new char*[p->n_streams];
    h.hd_packed_cuda = new char*[p->n_streams];
    for (int s = 0; s < p->n_streams; s++) {
        NRLDPC_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&h.info_bin_cuda[s]), info_cuda, hipHostMallocDefault));
        NRLDPC_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&h.llr_cuda[s]), h.mem_size_llr_cuda, hipHostMallocDefault));
        NRLDPC_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&h.hd_cuda[s]), h.mem_size_hd_cuda, hipHostMallocDefault));
        NRLDPC_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&h.hd_packed_cuda[s]), h.mem_size_hd_cuda_packed,
                                       hipHostMallocDefault));
    }
    return h;
}

template <typename T>
dev_mem<T> setup_dev_memory(const params* p, const ldpc_params* ldpc, const host_mem<T>* host) {
    dev_mem<T> d{};
    const int n_cw = p->n_mcw * p->n_cw_per_mcw;
    const int hd_bytes = (ldpc->N_before_puncture + 7) / 8;
    const bool repack = ldpc->N_before_puncture % 8 != 0;
    d.streams = new hipStream_t[p->n_streams];
    d.llr = new T*[p->n_streams];
    d.hard_decision = repack ? new char*[p->n_streams] : nullptr;
    d.hard_decision_packed = new char*[p->n_streams];
    for (int s = 0; s < p->n_streams; s++) {
        NRLDPC_CHECK_HIP(hipStreamCreate(&d.streams[s]));
        NRLDPC_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&d.llr[s]), host->mem_size_llr_cuda));
        NRLDPC_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&d.hard_decision_packed[s]), host->mem_size_hd_cuda_packed));
        if (repack) NRLDPC_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&d.hard_decision[s]), (size_t)n_cw * hd_bytes));
    }
    return d;
}

cuda_grid setup_gpu_grid(const ldpc_params* ldpc, int n_cw_per_mcw, int n_mcw) {
    const int bx = (ldpc->Z + 31) / 32 * 32;
    cuda_grid g{};
    g.cnp_blocks = dim3(ldpc->H_rows, n_mcw, 1);
    g.cnp_threads = dim3(bx, n_cw_per_mcw, 1);
    g.vnp_blocks = dim3(ldpc->H_cols, n_mcw, 1);
    g.vnp_threads = dim3(bx, n_cw_per_mcw, 1);
    g.pack_blocks = dim3(ldpc->H_cols, n_mcw, 1);
    g.pack_threads = dim3(bx / 8, n_cw_per_mcw, 1);
    g.early_blocks = dim3(n_mcw, n_cw_per_mcw, 1);
    g.early_threads = dim3(ldpc->Z, ldpc->H_rows, 1);
    ldpc_ctx* c = host_code(ldpc->bg_index, ldpc->Z);
    ldpc_code_info info{};
    if (c && ldpc_ctx_info(c, &info) == LDPC_OK) g.cnp_threads_per_block = (unsigned)(info.cw_per_block * ldpc->Z);
    g.shared_r_cache_size = 0;
    return g;
}

template <typename T>
void generate_new_codewords(const params* p, host_mem<T>* h, const ldpc_params* l, int codewords_in_iteration,
                            float sigma) {
    const int nb = l->N_before_puncture;
    const int z_punct = l->bg_index == LDPC_CODE_QC12X24 ? 0 : l->Z;
    for (int i = 0; i < codewords_in_iteration; i++) {
        transmit(h->info_bin, h->cwds, h->modulated_cwds, h->mem_size_cw, h->recv, l->K, nb, sigma, l->bg_index);
        (void)error_check(h->modulated_cwds, h->recv, nb);
        calculate_llr(h->llr, h->recv, nb, sigma, z_punct);
        // reduce_precision (src/setup_exp.cu:26-46): chars as quantize(0.5, x, 8) makes them,
        // halves rounded to nearest even
        for (int j = 0; j < nb; j++) {
            if constexpr (std::is_same<T, char>::value) h->llr_quan[j] = (char)quantize(0.5, h->llr[j], 8);
            else if constexpr (std::is_same<T, __half>::value) h->llr_quan[j] = __float2half_rn(h->llr[j]);
            else h->llr_quan[j] = (T)h->llr[j];
        }
        // the same codeword goes to every stream (src/setup_exp.cu:187-193)
        for (int s = 0; s < p->n_streams; s++) {
            memcpy(h->info_bin_cuda[s] + (size_t)i * l->K, h->info_bin, h->mem_size_infobits);
            memcpy(h->llr_cuda[s] + (size_t)i * nb, h->llr_quan, nb * sizeof(T));
        }
    }
}

// Row list (h1[k][i]: k-th nonzero block of block row i, ascending column) and column list
// (h2[k][j]: k-th nonzero block of block column j, ascending row) of ldpc->H; unused slots are
// {0, 0, -1, 0}.  h1 is [h1_deg][h1_rows], h2 is [h2_deg][h2_cols], row-major.
void nr_gpu_ldpc_fill_compact(h_element* h1, int h1_deg, int h1_rows, h_element* h2, int h2_deg, int h2_cols,
                              const ldpc_params* ldpc) {
    const int R = ldpc->H_rows, C = ldpc->H_cols;
    if (h1_rows < R || h2_cols < C) config_error("compact matrices smaller than the base matrix");
    for (int k = 0; k < h1_deg; k++)
        for (int i = 0; i < h1_rows; i++) h1[k * h1_rows + i] = {0, 0, -1, 0};
    for (int k = 0; k < h2_deg; k++)
        for (int j = 0; j < h2_cols; j++) h2[k * h2_cols + j] = {0, 0, -1, 0};
    for (int i = 0; i < R; i++) {
        int k = 0;
        for (int j = 0; j < C; j++) {
            const int v = ldpc->H[i * C + j];
            if (v < 0) continue;
            if (k >= h1_deg) config_error("row degree exceeds the compact row list");
            h1[k++ * h1_rows + i] = {(char)i, (char)j, (short)v, 1};
        }
    }
    for (int j = 0; j < C; j++) {
        int k = 0;
        for (int i = 0; i < R; i++) {
            const int v = ldpc->H[i * C + j];
            if (v < 0) continue;
            if (k >= h2_deg) config_error("column degree exceeds the compact column list");
            h2[k++ * h2_cols + j] = {(char)i, (char)j, (short)v, 1};
        }
    }
}

// The process-wide code from its compact lists, which must describe ldpc->H exactly.
void nr_gpu_ldpc_load_compact(const h_element* h1, int h1_deg, int h1_rows, const h_element* h2, int h2_deg,
                              int h2_cols, const ldpc_params* ldpc) {
    const int R = ldpc->H_rows, C = ldpc->H_cols;
    if (h1_rows < R || h2_cols < C) config_error("compact matrices smaller than the base matrix");
    auto same = [](const h_element& e, int i, int j, int v) {
        return e.valid == 1 && e.x == (char)i && e.y == (char)j && e.value == (short)v;
    };
    for (int i = 0; i < R; i++) {
        int k = 0;
        for (int j = 0; j < C; j++) {
            const int v = ldpc->H[i * C + j];
            if (v < 0) continue;
            if (k >= h1_deg || !same(h1[k * h1_rows + i], i, j, v))
                config_error("h_compact1 does not describe ldpc->H");
            k++;
        }
        if (k < h1_deg && h1[k * h1_rows + i].valid) config_error("h_compact1 has blocks ldpc->H does not have");
    }
    for (int j = 0; j < C; j++) {
        int k = 0;
        for (int i = 0; i < R; i++) {
            const int v = ldpc->H[i * C + j];
            if (v < 0) continue;
            if (k >= h2_deg || !same(h2[k * h2_cols + j], i, j, v))
                config_error("h_compact2 does not describe ldpc->H");
            k++;
        }
        if (k < h2_deg && h2[k * h2_cols + j].valid) config_error("h_compact2 has blocks ldpc->H does not have");
    }
    if (g_ctx) {
        ldpc_ctx_destroy(g_ctx);
        g_ctx = nullptr;
    }
    free_centry_buffers();
    g_n_cw = 0;
    g_n_iter = ldpc->n_iterations;
    g_n_vn = ldpc->N_before_puncture;
    int dev = 0;
    NRLDPC_CHECK_HIP(hipGetDevice(&dev));
    NRLDPC_CHECK(ldpc_ctx_create(&g_ctx, ldpc->bg_index, ldpc->Z, ldpc->H_rows, dev));
}

void put_params_in_constant_mem(unsigned int z, unsigned int n_cw_per_mcw, unsigned int n_total_vn,
                                unsigned int n_total_cn, unsigned int n_blk_col, unsigned int threads_per_block) {
    if (!g_ctx) config_error("put_params_in_constant_mem before allocate_constant_memory");
    ldpc_code_info info;
    NRLDPC_CHECK(ldpc_ctx_info(g_ctx, &info));
    if ((int)z != info.Z || (int)n_total_vn != info.n_vn || (int)n_total_cn != info.rows * info.Z ||
        (int)n_blk_col != info.cols || n_cw_per_mcw < 1)
        config_error("kernel parameters do not match the code in constant memory");
    (void)threads_per_block;  // the fused launch chooses its own workgroup shape
}

void put_h_compact_and_params_in_constant_memory(const ldpc_params* ldpc, const params* p, const cuda_grid* g) {
    // the reference's sequence: compact lists of H, upload, then the kernel parameters
    const int d1 = ldpc->n_nz_in_row, d2 = ldpc->n_nz_in_col;
    std::vector<h_element> h1((size_t)d1 * ldpc->H_rows), h2((size_t)d2 * ldpc->H_cols);
    nr_gpu_ldpc_fill_compact(h1.data(), d1, ldpc->H_rows, h2.data(), d2, ldpc->H_cols, ldpc);
    nr_gpu_ldpc_load_compact(h1.data(), d1, ldpc->H_rows, h2.data(), d2, ldpc->H_cols, ldpc);
    put_params_in_constant_mem(ldpc->Z, p ? p->n_cw_per_mcw : 1, ldpc->N_before_puncture, ldpc->H_rows * ldpc->Z,
                               ldpc->H_cols, g ? g->cnp_threads_per_block : 0);
    g_n_cw = g ? (int)(g->cnp_blocks.y * g->cnp_threads.y) : 0;
}

template <typename T>
void perform_gpu_ldpc_decoding(dev_mem<T>* d, host_mem<T>* h, const cuda_grid* g, int n_iterations,
                               int stream_index) {
    if (!g_ctx) config_error("perform_gpu_ldpc_decoding before put_h_compact_and_params_in_constant_memory");
    const hipStream_t st = d->streams[stream_index];
    const int n_cw = (int)(g->cnp_blocks.y * g->cnp_threads.y);
    NRLDPC_CHECK_HIP(hipMemcpyAsync(d->llr[stream_index], h->llr_cuda[stream_index], h->mem_size_llr_cuda,
                                    hipMemcpyHostToDevice, st));
    uint8_t* packed = reinterpret_cast<uint8_t*>(d->hard_decision_packed[stream_index]);
    uint8_t* per_cw = d->hard_decision ? reinterpret_cast<uint8_t*>(d->hard_decision[stream_index]) : packed;
    if constexpr (std::is_same<T, char>::value)  // the CHAR_PRC build
        NRLDPC_CHECK(ldpc_decode_char_async(g_ctx, st, reinterpret_cast<const int8_t*>(d->llr[stream_index]), n_cw,
                                            n_iterations, 0, per_cw, nullptr, nullptr));
    else if constexpr (std::is_same<T, __half>::value)  // the HALF_PRC build
        NRLDPC_CHECK(ldpc_decode_half_async(g_ctx, st, reinterpret_cast<const uint16_t*>(d->llr[stream_index]), n_cw,
                                            n_iterations, 0, per_cw, nullptr, nullptr));
    else
        NRLDPC_CHECK(ldpc_decode_async(g_ctx, st, reinterpret_cast<const float*>(d->llr[stream_index]), n_cw,
                                       n_iterations, 0, per_cw, nullptr, nullptr));
    if (per_cw != packed) NRLDPC_CHECK(ldpc_hd_to_stream_async(g_ctx, st, per_cw, n_cw, packed));
    NRLDPC_CHECK_HIP(hipMemcpyAsync(h->hd_packed_cuda[stream_index], packed, h->mem_size_hd_cuda_packed,
                                    hipMemcpyDeviceToHost, st));
}

extern "C" void ldpc_gpu_decoder(char* inputs, char* outputs) {
    if (!g_ctx || g_n_cw < 1) config_error("ldpc_gpu_decoder before put_h_compact_and_params_in_constant_memory");
    const size_t n = (size_t)g_n_cw * g_n_vn;
    const size_t hb = (size_t)(g_n_vn + 7) / 8, packed = (n + 7) / 8;
    if (g_cbuf.n != n) {
        free_centry_buffers();
        NRLDPC_CHECK_HIP(hipStreamCreate(&g_cbuf.stream));
        NRLDPC_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&g_cbuf.in), n));
        NRLDPC_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&g_cbuf.hd), (size_t)g_n_cw * hb));
        NRLDPC_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&g_cbuf.packed), packed));
        g_cbuf.n = n;
    }
    const hipStream_t s = g_cbuf.stream;
    NRLDPC_CHECK_HIP(hipMemcpyAsync(g_cbuf.in, inputs, n, hipMemcpyHostToDevice, s));
    const bool repack = g_n_vn % 8 != 0;
    uint8_t* hd = repack ? g_cbuf.hd : g_cbuf.packed;
    // char LLRs: the CHAR_PRC schedule, the reference's default build for this entry point
    NRLDPC_CHECK(ldpc_decode_char_async(g_ctx, s, g_cbuf.in, g_n_cw, g_n_iter, 0, hd, nullptr, nullptr));
    if (repack) NRLDPC_CHECK(ldpc_hd_to_stream_async(g_ctx, s, hd, g_n_cw, g_cbuf.packed));
    NRLDPC_CHECK_HIP(hipMemcpyAsync(outputs, g_cbuf.packed, packed, hipMemcpyDeviceToHost, s));
    NRLDPC_CHECK_HIP(hipStreamSynchronize(s));
}

template host_mem<float> setup_host<float>(const params*, unsigned int, unsigned int);
template dev_mem<float> setup_dev_memory<float>(const params*, const ldpc_params*, const host_mem<float>*);
template void generate_new_codewords<float>(const params*, host_mem<float>*, const ldpc_params*, int, float);
template void perform_gpu_ldpc_decoding<float>(dev_mem<float>*, host_mem<float>*, const cuda_grid*, int, int);
template host_mem<__half> setup_host<__half>(const params*, unsigned int, unsigned int);
template dev_mem<__half> setup_dev_memory<__half>(const params*, const ldpc_params*, const host_mem<__half>*);
template void generate_new_codewords<__half>(const params*, host_mem<__half>*, const ldpc_params*, int, float);
template void perform_gpu_ldpc_decoding<__half>(dev_mem<__half>*, host_mem<__half>*, const cuda_grid*, int, int);
template host_mem<char> setup_host<char>(const params*, unsigned int, unsigned int);
template dev_mem<char> setup_dev_memory<char>(const params*, const ldpc_params*, const host_mem<char>*);
template void generate_new_codewords<char>(const params*, host_mem<char>*, const ldpc_params*, int, float);
template void perform_gpu_ldpc_decoding<char>(dev_mem<char>*, host_mem<char>*, const cuda_grid*, int, int);
