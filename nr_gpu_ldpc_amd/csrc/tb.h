// tb.h -- transport-block configuration and the device launches of tb.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace nrldpc {

struct TbConfig {
    int A = 0, L_tb = 0, B = 0, bg = 0, C = 0, L_cb = 0, K_prime = 0, K = 0, Z = 0, F = 0, N = 0, n_vn = 0;
    int kb = 0, cols = 0;
    int G = 0, Qm = 0, NL = 1, rv = 0, k0 = 0;
    std::vector<int> E, E_off;  // rate-matched bits per code block, prefix offsets
};

uint32_t crc_poly(int type, int* len);
uint32_t crc_bits(int type, const uint8_t* bits, long long n);
int tb_setup(int A, float rate, int bg, TbConfig* t, std::string* err);
int tb_rate_match_setup(TbConfig* t, int G, int Qm, int NL, int rv, std::string* err);
int tb_build_gather(const TbConfig& t, std::vector<int>* start, std::vector<int>* index);

struct DerateArgs {
    const float* in;     // received LLRs, code block r's E_r values from offset E_off[r]
    float* out;          // [C][n_vn] decoder LLRs
    const int* start;    // [C*n_vn + 1] CSR row starts
    const int* index;    // received positions per decoder column, ascending bit-selection order
    int n, n_vn;         // n = C * n_vn
    int f0, f1;          // filler columns f0 .. f1-1 of every code block (known zeros)
};

struct TbCheckArgs {
    const uint8_t* hd;   // [C][hd_bytes] packed decisions
    int* cb_ok;          // [C]: 1 when the code block's CRC24B holds (always 1 when C == 1)
    uint8_t* tb_bits;    // [B] reassembled transport block + its CRC, one byte per bit (nullable)
    int C, hd_bytes, K_prime, L_cb, B;
};

int derate_launch(const DerateArgs& a, hipStream_t s);
int tb_check_launch(const TbCheckArgs& a, hipStream_t s);

}  // namespace nrldpc
