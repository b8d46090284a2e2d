// ldpc_ber.cpp -- BER/FER/latency/throughput driver over the reference-named host API.
//
// The experiment loop of the reference's src/main.cpp:92-187 (run_exp, run_test_on_batch,
// test_latency, test_throughput, measure_ber) on top of include/nr_gpu_ldpc.hpp: for every SNR
// point, batches of n_cw_per_mcw * n_mcw codewords are generated on the host (rand / drand48
// seeded with srand), copied to every stream, decoded, and counted over the first K bits.
// One JSON line per SNR point goes to stdout.
//
//   ldpc_ber [--code nr|qc] [--info-bits A] [--rate R] [--Z Z] [--iters T] [--snr-min S0]
//            [--snr-max S1] [--snr-step dS] [--codewords N] [--cw-per-mcw C] [--mcw M]
//            [--streams S] [--seed X] [--dump PREFIX] [--prec float|half|char]
//
// --prec char runs the reference's default CHAR_PRC build (LLRs quantised by quantize(0.5, x,
// 8), the signed-char schedule), --prec half its HALF_PRC build (binary16 LLRs and storage);
// float (the default here) the FP32 one.
// --dump PREFIX writes the first batch's LLRs (float32, binary16 bit patterns with --prec half,
// or int8 with --prec char,
// [n_cw][N_before_puncture]) and packed hard decisions of stream 0 to PREFIX.llr / PREFIX.hd,
// for checking against the oracle.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/nr_gpu_ldpc.hpp"

namespace {

struct Options {
    std::string code = "nr";
    int info_bits = 8424;
    float rate = 1.0f / 3.0f;
    int Z = 96;
    int iters = 5;
    float snr_min = -5, snr_max = 20, snr_step = 5;
    int codewords = 5000, cw_per_mcw = 2, mcw = 20, streams = 6;
    long seed = 69012;
    std::string dump;
    std::string prec = "float";
};

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
float test_latency(dev_mem<T>* d, host_mem<T>* h, const cuda_grid* g, int n_iterations) {
    const double t0 = now_ms();
    perform_gpu_ldpc_decoding<T>(d, h, g, n_iterations, 0);
    (void)hipDeviceSynchronize();
    return (float)(now_ms() - t0);
}

template <typename T>
float test_throughput(dev_mem<T>* d, host_mem<T>* h, const cuda_grid* g, int n_iterations, int n_streams) {
    const double t0 = now_ms();
    for (int s = 0; s < n_streams; s++) perform_gpu_ldpc_decoding<T>(d, h, g, n_iterations, s);
    (void)hipDeviceSynchronize();
    return (float)(now_ms() - t0);
}

template <typename T>
error_result measure_ber(const host_mem<T>* h, int n_streams, int n_cw_per_mcw, int n_mcw, int codeword_len,
                         int info_len) {
    error_result e{0, 0, 0, 0};
    for (int s = 0; s < n_streams; s++) {
        unpack_char(h->hd_packed_cuda[s], h->hd_cuda[s], h->mem_size_hd_cuda);
        e = e + cuda_error_check(h->info_bin_cuda[s], h->hd_cuda[s], n_cw_per_mcw * n_mcw, codeword_len, info_len);
    }
    return e;
}

template <typename T>
gpu_result run_test_on_batch(int cws_in_iteration, float sigma, host_mem<T>* h, dev_mem<T>* d,
                             const cuda_grid* g, const ldpc_params* ldpc, const params* p, const Options& o,
                             bool first) {
    generate_new_codewords<T>(p, h, ldpc, cws_in_iteration, sigma);
    const float latency = test_latency<T>(d, h, g, ldpc->n_iterations);
    const float run_time = test_throughput<T>(d, h, g, ldpc->n_iterations, p->n_streams);
    const error_result e = measure_ber<T>(h, p->n_streams, p->n_cw_per_mcw, p->n_mcw, ldpc->N_before_puncture,
                                          ldpc->K);
    if (first && !o.dump.empty()) {
        FILE* f = fopen((o.dump + ".llr").c_str(), "wb");
        if (f) { fwrite(h->llr_cuda[0], 1, h->mem_size_llr_cuda, f); fclose(f); }
        f = fopen((o.dump + ".hd").c_str(), "wb");
        if (f) { fwrite(h->hd_packed_cuda[0], 1, h->mem_size_hd_cuda_packed, f); fclose(f); }
    }
    return {latency, run_time, e.n_bits, e.bit_error, e.n_frames, e.frame_error};
}

template <typename T>
void run_exp(const ldpc_params* ldpc, host_mem<T>* h, dev_mem<T>* d, const cuda_grid* g, const params* p,
             const Options& o) {
    bool first = true;
    for (float snr = p->snr_min; snr <= p->snr_max + 1e-6f; snr += p->snr_step) {
        const float sigma = compute_noise_stddev(snr, ldpc->rate);
        gpu_result r = {1e30f, 0, 0, 0, 0, 0};
        const int cws_in_iteration = p->n_cw_per_mcw * p->n_mcw;
        for (int done = 0; done < p->n_codewords; done += cws_in_iteration) {
            r = r + run_test_on_batch(cws_in_iteration, sigma, h, d, g, ldpc, p, o, first);
            first = false;
        }
        const double ber = r.n_bits_total ? (double)r.n_bit_errors / r.n_bits_total : 0.0;
        const double fer = r.n_frames ? (double)r.n_frame_errors / r.n_frames : 0.0;
        const double code_mbps = r.n_bits_total / (double)r.cpu_run_time / 1000.0;  // bits per ms / 1000
        printf("{\"snr_db\": %g, \"sigma\": %.9g, \"ber\": %.9g, \"fer\": %.9g, \"bit_errors\": %d, \"frames\": %d, "
               "\"frame_errors\": %d, \"best_latency_us\": %.3f, \"user_throughput_mbps\": %.3f, "
               "\"codeword_throughput_mbps\": %.3f}\n",
               snr, sigma, ber, fer, r.n_bit_errors, r.n_frames, r.n_frame_errors, r.best_latency * 1000.0,
               code_mbps * ldpc->rate, code_mbps);
        fflush(stdout);
    }
}

bool parse(int argc, char** argv, Options* o) {
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (i + 1 >= argc) return false;
        const char* v = argv[++i];
        if (a == "--code") o->code = v;
        else if (a == "--info-bits") o->info_bits = atoi(v);
        else if (a == "--rate") o->rate = (float)atof(v);
        else if (a == "--Z") o->Z = atoi(v);
        else if (a == "--iters") o->iters = atoi(v);
        else if (a == "--snr-min") o->snr_min = (float)atof(v);
        else if (a == "--snr-max") o->snr_max = (float)atof(v);
        else if (a == "--snr-step") o->snr_step = (float)atof(v);
        else if (a == "--codewords") o->codewords = atoi(v);
        else if (a == "--cw-per-mcw") o->cw_per_mcw = atoi(v);
        else if (a == "--mcw") o->mcw = atoi(v);
        else if (a == "--streams") o->streams = atoi(v);
        else if (a == "--seed") o->seed = atol(v);
        else if (a == "--dump") o->dump = v;
        else if (a == "--prec") o->prec = v;
        else return false;
    }
    return o->snr_step > 0 && o->cw_per_mcw > 0 && o->mcw > 0 && o->streams > 0 && o->iters > 0 &&
           (o->prec == "float" || o->prec == "half" || o->prec == "char");
}

// T = float: the FP32 build; T = char: the reference's default CHAR_PRC build.
template <typename T>
void run(const ldpc_params& ldpc, params& p, const Options& o) {
    host_mem<T> host = setup_host<T>(&p, ldpc.K, ldpc.N_before_puncture);
    dev_mem<T> dev = setup_dev_memory<T>(&p, &ldpc, &host);
    const cuda_grid grid = setup_gpu_grid(&ldpc, p.n_cw_per_mcw, p.n_mcw);
    put_h_compact_and_params_in_constant_memory(&ldpc, &p, &grid);
    // one untimed decode per stream first: the first launch of the process loads the code
    // object and the first use of a stream sets it up, which would otherwise land in the first
    // SNR point's latency and throughput
    for (int s = 0; s < p.n_streams; s++) {
        memset(host.llr_cuda[s], 0, host.mem_size_llr_cuda);
        perform_gpu_ldpc_decoding<T>(&dev, &host, &grid, ldpc.n_iterations, s);
    }
    (void)hipDeviceSynchronize();
    run_exp<T>(&ldpc, &host, &dev, &grid, &p, o);
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse(argc, argv, &o)) {
        fprintf(stderr, "usage: ldpc_ber [--code nr|qc] [--info-bits A] [--rate R] [--Z Z] [--iters T] "
                        "[--snr-min S] [--snr-max S] [--snr-step S] [--codewords N] [--cw-per-mcw C] [--mcw M] "
                        "[--streams S] [--seed X] [--dump PREFIX] [--prec float|half|char]\n");
        return 2;
    }
    params p = {o.snr_min, o.snr_max, o.snr_step, o.codewords, o.info_bits, o.rate, o.cw_per_mcw, o.mcw, o.streams};
    srand((unsigned)o.seed);
    const ldpc_params ldpc = o.code == "qc" ? setup_ldpc_qc12x24(o.Z, o.iters) : setup_ldpc(o.info_bits, o.rate, o.iters);
    p.code_rate = ldpc.rate;
    fprintf(stderr, "code bg=%d Z=%d K=%d N=%d rows=%d cols=%d iters=%d streams=%d batch=%dx%d\n", ldpc.bg_index,
            ldpc.Z, ldpc.K, ldpc.N, ldpc.H_rows, ldpc.H_cols, ldpc.n_iterations, p.n_streams, p.n_mcw, p.n_cw_per_mcw);
    if (o.prec == "char") run<char>(ldpc, p, o);
    else if (o.prec == "half") run<__half>(ldpc, p, o);
    else run<float>(ldpc, p, o);
    return 0;
}
