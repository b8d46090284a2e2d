// Host-side checks of the reference-named API (no GPU): prints one JSON object.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "nr_gpu_ldpc.hpp"

static void print_ints(const char* name, const int* v, int n, bool last = false) {
    printf("\"%s\": [", name);
    for (int i = 0; i < n; i++) printf("%s%d", i ? "," : "", v[i]);
    printf("]%s\n", last ? "" : ",");
}

static h_element g_h1[24][46];  // wider than the largest row degree (19)
static h_element g_h2[32][68];  // wider than the largest column degree (30)

int main(int argc, char** argv) {
    const ldpc_params a = setup_ldpc(8424, 1.0f / 3.0f, 5);
    fill_in_compact_matrices(g_h1, g_h2, &a);
    if (argc > 1 && argv[1][0] == 'h') {
        // 100,000 BG1 Z=384 codewords per stream: 10.4 GB of LLRs, beyond the int size fields
        params p{};
        p.n_cw_per_mcw = 1000;
        p.n_mcw = 100;
        p.n_streams = 1;
        setup_host<float>(&p, a.K, a.N_before_puncture);
        return 0;
    }
    if (argc > 1) {
        // a row list that no longer describes H is refused before any device work
        g_h1[3][5].value = (short)(g_h1[3][5].value + 1);
        allocate_constant_memory(g_h1, g_h2, &a, sizeof g_h1, sizeof g_h2);
        return 0;
    }
    printf("{\n");
    const int fa[] = {a.bg_index, a.Z, a.K, a.N, a.H_rows, a.H_cols, a.N_before_puncture, a.n_nz_in_row,
                      a.n_nz_in_col, a.n_iterations, a.h_element_count_per_row[0], a.h_element_count_per_col[0]};
    print_ints("setup_bg1", fa, 12);
    const ldpc_params b = setup_ldpc(2064, 0.2f, 10);
    const int fb[] = {b.bg_index, b.Z, b.K, b.N, b.H_rows, b.H_cols, b.N_before_puncture, b.n_nz_in_row, b.n_nz_in_col};
    print_ints("setup_bg2", fb, 9);
    const ldpc_params q = setup_ldpc_qc12x24(96, 5);
    const int fq[] = {q.bg_index, q.Z, q.K, q.N, q.H_rows, q.H_cols, q.N_before_puncture};
    print_ints("setup_qc", fq, 7);
    print_ints("H_bg1_row0", a.H, 68);

    // encoder output layout: BG1 Z=2 (44 info bits -> 40 + 92 outputs), BG2 40 info bits (Kb = 6,
    // Z = 7 -> 26 + 94 outputs)
    std::vector<int> info(44), out(132, -7);
    for (int i = 0; i < 44; i++) info[i] = (i * 7 + 3) % 5 == 0;
    printf("\"enc_bg1_rc\": %d,\n", ldpc_encoder_orig(info.data(), out.data(), 44, 1, 0));
    print_ints("enc_bg1_info", info.data(), 44);
    print_ints("enc_bg1_out", out.data(), 132);
    std::vector<int> info2(40), out2(120, -7);
    for (int i = 0; i < 40; i++) info2[i] = (i * 3 + 1) % 4 == 0;
    printf("\"enc_bg2_rc\": %d,\n", ldpc_encoder_orig(info2.data(), out2.data(), 40, 2, 0));
    print_ints("enc_bg2_info", info2.data(), 40);
    print_ints("enc_bg2_out", out2.data(), 120);
    printf("\"enc_gen_code_rc\": %d,\n", ldpc_encoder_orig(info.data(), out.data(), 44, 1, 1));

    printf("\"sigma\": [%.9g, %.9g],\n", compute_noise_stddev(1.5f, 1.0f / 3.0f), compute_noise_stddev(-2.0f, 0.5f));
    float recv[6] = {0.5f, -1.25f, 2.0f, 0.0f, -0.75f, 1.0f}, llr[8];
    calculate_llr(llr, recv, 8, 0.8f, 1);
    printf("\"llr\": [");
    for (int i = 0; i < 8; i++) printf("%s%.9g", i ? "," : "", llr[i]);
    printf("],\n");
    float tr[4] = {1, -1, 1, -1}, rc[4] = {0.5f, 0.2f, -0.1f, -2.0f};
    printf("\"error_check\": %d,\n", error_check(tr, rc, 4));
    int inf[6] = {0, 1, 1, 0, 0, 1};
    char hd[8] = {0, 1, 1, 9, 1, 1, 0, 9};  // two codewords of length 4, info length 3
    const error_result e = cuda_error_check(inf, hd, 2, 4, 3);
    const int fe[] = {e.bit_error, e.n_bits, e.frame_error, e.n_frames};
    print_ints("cuda_error_check", fe, 4);
    const int qv[] = {quantize(0.5, 1.3, 8), quantize(0.5, -100.0, 8), quantize(0.5, 100.0, 8), quantize(0.5, -0.1, 8)};
    print_ints("quantize", qv, 4);
    char packed[2] = {(char)0xA5, (char)0x03}, unp[12];
    unpack_char(packed, unp, 12);
    int ui[12];
    for (int i = 0; i < 12; i++) ui[i] = unp[i];
    print_ints("unpack", ui, 12);

    // transmit: rand()%2 info bits, drand48 polar noise; seeded here, replayed by the test
    srand(4242);
    srand48(99);
    const int K = 44, nb = 68 * 2;
    std::vector<int> ib(K), cw(nb);
    std::vector<float> mod(nb), rv(nb);
    transmit(ib.data(), cw.data(), mod.data(), nb * sizeof(int), rv.data(), K, nb, 0.7f, 1);
    print_ints("tx_info", ib.data(), K);
    print_ints("tx_cw", cw.data(), nb);
    printf("\"tx_recv\": [");
    for (int i = 0; i < nb; i++) printf("%s%.9g", i ? "," : "", rv[i]);
    printf("],\n");
    // compact lists: valid entries, then the first row's list and the first column's list
    int n1 = 0, n2 = 0;
    for (auto& row : g_h1)
        for (auto& e : row) n1 += e.valid;
    for (auto& row : g_h2)
        for (auto& e : row) n2 += e.valid;
    int r0[3 * 19], c0[3 * 30];
    for (int k = 0; k < 19; k++) {
        r0[3 * k] = g_h1[k][0].x; r0[3 * k + 1] = g_h1[k][0].y; r0[3 * k + 2] = g_h1[k][0].value;
    }
    for (int k = 0; k < 30; k++) {
        c0[3 * k] = g_h2[k][0].x; c0[3 * k + 1] = g_h2[k][0].y; c0[3 * k + 2] = g_h2[k][0].value;
    }
    const int counts[] = {n1, n2, g_h1[19][0].value, g_h1[19][0].valid, g_h2[30][0].value};
    print_ints("compact_counts", counts, 5);
    print_ints("compact_row0", r0, 3 * 19);
    print_ints("compact_col0", c0, 3 * 30);
    const short* g = choose_generator_matrix(1, 384);
    printf("\"gen_first\": [%d, %d, %d]\n", g[0], g[1], g[2]);
    printf("}\n");
    return 0;
}
