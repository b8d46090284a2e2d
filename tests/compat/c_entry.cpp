// ldpc_gpu_decoder (the reference's proposed C entry point) against the C ABI's CHAR_PRC decode
// of the same int8 LLRs; prints one JSON object.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "nr_gpu_ldpc.hpp"

int main() {
    params p = {1.0f, 1.0f, 1.0f, 8, 8424, 1.0f / 3.0f, 2, 4, 1};
    const ldpc_params l = setup_ldpc(p.n_info_bits, p.code_rate, 10);
    const cuda_grid g = setup_gpu_grid(&l, p.n_cw_per_mcw, p.n_mcw);
    put_h_compact_and_params_in_constant_memory(&l, &p, &g);
    const int n_cw = 8, nvn = l.N_before_puncture;
    srand(5);
    srand48(5);
    std::vector<int> info(l.K), cw(nvn);
    std::vector<float> mod(nvn), recv(nvn), llr(nvn);
    std::vector<char> q((size_t)n_cw * nvn);
    const float sigma = compute_noise_stddev(1.5f, l.rate);
    for (int c = 0; c < n_cw; c++) {
        transmit(info.data(), cw.data(), mod.data(), nvn * sizeof(int), recv.data(), l.K, nvn, sigma, l.bg_index);
        calculate_llr(llr.data(), recv.data(), nvn, sigma, l.Z);
        for (int i = 0; i < nvn; i++) q[(size_t)c * nvn + i] = (char)quantize(0.5, llr[i], 8);
    }
    const size_t packed = ((size_t)n_cw * nvn + 7) / 8;
    std::vector<char> out(packed);
    ldpc_gpu_decoder(q.data(), out.data());
    // the same through the C ABI
    ldpc_ctx* c = nr_gpu_ldpc_context();
    int8_t* d_llr = nullptr;
    uint8_t* d_hd = nullptr;
    (void)hipMalloc(reinterpret_cast<void**>(&d_llr), q.size());
    (void)hipMalloc(reinterpret_cast<void**>(&d_hd), packed);
    (void)hipMemcpy(d_llr, q.data(), q.size(), hipMemcpyHostToDevice);
    const int st = ldpc_decode_char_async(c, nullptr, d_llr, n_cw, 10, 0, d_hd, nullptr, nullptr);
    std::vector<char> want(packed);
    (void)hipMemcpy(want.data(), d_hd, packed, hipMemcpyDeviceToHost);
    int diff = 0, ones = 0;
    for (size_t i = 0; i < packed; i++) {
        diff += out[i] != want[i];
        ones += __builtin_popcount((unsigned char)out[i]);
    }
    printf("{\"status\": %d, \"bytes\": %zu, \"differing_bytes\": %d, \"ones\": %d}\n", st, packed, diff, ones);
    return 0;
}
