// ldpc_gpu_decoder (the C entry point proposed by the reference, Linking_With_Other_Projects.md:19-25)
// on int8 LLRs made by the reference-named host chain (transmit, calculate_llr, quantize).
//
//   c_entry A rate n_cw_per_mcw n_mcw n_iter ebn0_db prefix
//
// Writes prefix.llr (the int8 inputs, [n_cw][N_before_puncture]) and prefix.out (the packed
// output stream, ceil(n_cw * N_before_puncture / 8) bytes) for the oracle comparison, and prints
// one JSON object with the status and the comparison against the C ABI's CHAR_PRC decode.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "nr_gpu_ldpc.hpp"

int main(int argc, char** argv) {
    if (argc != 8) {
        fprintf(stderr, "usage: %s A rate n_cw_per_mcw n_mcw n_iter ebn0_db prefix\n", argv[0]);
        return 2;
    }
    const int A = atoi(argv[1]);
    const float rate = (float)atof(argv[2]);
    const int cpm = atoi(argv[3]), n_mcw = atoi(argv[4]), n_iter = atoi(argv[5]);
    const float ebn0 = (float)atof(argv[6]);
    const std::string prefix = argv[7];
    params p = {1.0f, 1.0f, 1.0f, cpm * n_mcw, A, rate, cpm, n_mcw, 1};
    const ldpc_params l = setup_ldpc(p.n_info_bits, p.code_rate, n_iter);
    const cuda_grid g = setup_gpu_grid(&l, p.n_cw_per_mcw, p.n_mcw);
    put_h_compact_and_params_in_constant_memory(&l, &p, &g);
    const int n_cw = cpm * n_mcw, nvn = l.N_before_puncture;
    srand(5);
    srand48(5);
    std::vector<int> info(l.K), cw(nvn);
    std::vector<float> mod(nvn), recv(nvn), llr(nvn);
    std::vector<char> q((size_t)n_cw * nvn);
    const float sigma = compute_noise_stddev(ebn0, l.rate);
    for (int c = 0; c < n_cw; c++) {
        transmit(info.data(), cw.data(), mod.data(), nvn * sizeof(int), recv.data(), l.K, nvn, sigma, l.bg_index);
        calculate_llr(llr.data(), recv.data(), nvn, sigma, l.Z);
        for (int i = 0; i < nvn; i++) q[(size_t)c * nvn + i] = (char)quantize(0.5, llr[i], 8);
    }
    const size_t packed = ((size_t)n_cw * nvn + 7) / 8;
    std::vector<char> out(packed);
    ldpc_gpu_decoder(q.data(), out.data());
    FILE* f = fopen((prefix + ".llr").c_str(), "wb");
    if (!f || fwrite(q.data(), 1, q.size(), f) != q.size()) return 3;
    fclose(f);
    f = fopen((prefix + ".out").c_str(), "wb");
    if (!f || fwrite(out.data(), 1, out.size(), f) != out.size()) return 3;
    fclose(f);
    // the same through the C ABI (per-codeword packing, then the reference's stream layout)
    ldpc_ctx* c = nr_gpu_ldpc_context();
    ldpc_code_info ci;
    int st = ldpc_ctx_info(c, &ci);
    int8_t* d_llr = nullptr;
    uint8_t* d_hd = nullptr;
    uint8_t* d_stream = nullptr;
    (void)hipMalloc(reinterpret_cast<void**>(&d_llr), q.size());
    (void)hipMalloc(reinterpret_cast<void**>(&d_hd), (size_t)n_cw * ci.hd_bytes);
    (void)hipMalloc(reinterpret_cast<void**>(&d_stream), packed);
    (void)hipMemcpy(d_llr, q.data(), q.size(), hipMemcpyHostToDevice);
    if (!st) st = ldpc_decode_char_async(c, nullptr, d_llr, n_cw, n_iter, 0, d_hd, nullptr, nullptr);
    if (!st) st = ldpc_hd_to_stream_async(c, nullptr, d_hd, n_cw, d_stream);
    std::vector<char> want(packed);
    (void)hipMemcpy(want.data(), d_stream, packed, hipMemcpyDeviceToHost);
    int diff = 0, ones = 0;
    for (size_t i = 0; i < packed; i++) {
        diff += out[i] != want[i];
        ones += __builtin_popcount((unsigned char)out[i]);
    }
    printf("{\"status\": %d, \"bg\": %d, \"Z\": %d, \"n_vn\": %d, \"n_cw\": %d, \"bytes\": %zu, "
           "\"differing_bytes\": %d, \"ones\": %d}\n",
           st, l.bg_index, l.Z, nvn, n_cw, packed, diff, ones);
    return 0;
}
