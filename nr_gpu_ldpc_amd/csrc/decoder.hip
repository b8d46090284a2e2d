// decoder.hip -- host-side launch of the fused decoder kernels (decoder_kernel.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "code.h"
#include "decoder.h"
#include "status.h"

namespace nrldpc {

// Instantiated in decoder_bg1.hip / decoder_bg2.hip / decoder_qc.hip.
template <int KIND, int ZC, bool LEARNED>
int launch_decode(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s);

template <int KIND>
int launch_decode_train(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s);

template <int KIND, int ZC>
int launch_decode_half(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s);
template <int KIND, int ZC, bool LEARNED, bool HALF>
int launch_decode_slim(const DecodeArgs& a, int grid, int threads, size_t lds, hipStream_t s);
template <int KIND, int ZC>
int launch_decode_char_pair(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s);
template <int KIND, int ZC>
int launch_decode_char_split(const DecodeArgs& a, int blocks, int threads, size_t lds, hipStream_t s);

// Lifting sizes with a dedicated kernel (all offsets compile-time): the hot configurations.
constexpr int kBG1Special = 384, kBG2Special = 208, kQCSpecial = 96;

// Dynamic LDS available to one workgroup on the current device (queried once per process).
int max_lds_bytes() {
    static int v = -1;
    if (v < 0) {
        int dev = 0, bytes = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            bytes = 64 * 1024;
        v = bytes - 1024;  // headroom for the kernels' static LDS
    }
    return v;
}

size_t decoder_lds_bytes_per_cw(int kind, int Z) {
    int nnd = 0;
    if (kind == kBG1) nnd = Structure<kBG1>::NND;
    else if (kind == kBG2) nnd = Structure<kBG2>::NND;
    else nnd = Structure<kWifi>::NND;
    return ((size_t)3 * nnd * 4 * Z + 16 + 15) & ~(size_t)15;
}

static int n_nd1(int kind) {
    if (kind == kBG1) return Structure<kBG1>::ND1;
    if (kind == kBG2) return Structure<kBG2>::ND1;
    return Structure<kWifi>::ND1;
}

size_t decoder_slim_lds_bytes_per_cw(int kind, int Z) {
    int nnd = 0;
    if (kind == kBG1) nnd = Structure<kBG1>::NND;
    else if (kind == kBG2) nnd = Structure<kBG2>::NND;
    else nnd = Structure<kWifi>::NND;
    return ((size_t)2 * nnd * 4 * Z + 16 + 15) & ~(size_t)15;
}

// Codewords per workgroup that fit `threads` and the LDS (+16 bytes: the workgroup flag word).
static int cpb_max_for(int Z, int threads, size_t lds_cw) {
    const size_t lim = (size_t)max_lds_bytes();
    if (Z < 1 || Z > threads || lds_cw + 16 > lim) return 0;
    return (int)std::min<size_t>((size_t)(threads / Z), (lim - 16) / lds_cw);
}

int decoder_dup_cpb_max(int kind, int Z) { return cpb_max_for(Z, kMaxBlock, decoder_lds_bytes_per_cw(kind, Z)); }
int decoder_slim_cpb_max(int kind, int Z) { return cpb_max_for(Z, kSlimMaxBlock, decoder_slim_lds_bytes_per_cw(kind, Z)); }

size_t decoder_slim_slot_bytes(int kind, int Z, int cpb) {
    return (size_t)cpb * n_nd1(kind) * Z * sizeof(float);
}

int layout_for_batch(const LayoutModel& dup, const LayoutModel& slim, int n_cw, int cus) {
    if (slim.cpb <= 0 || slim.cw_per_cu <= 0 || cus <= 0 || n_cw <= 0) return kLayoutDup;
    if (dup.cpb <= 0 || dup.cw_per_cu <= 0) return kLayoutSlim;
    const long long per_d = (long long)cus * dup.cw_per_cu, per_s = (long long)cus * slim.cw_per_cu;
    const double t_d = (double)((n_cw + per_d - 1) / per_d) * dup.round_us;
    const double t_s = (double)((n_cw + per_s - 1) / per_s) * slim.round_us;
    return t_s < t_d ? kLayoutSlim : kLayoutDup;
}

// Instantiated in decoder_slim_{bg1,bg2,qc}.hip and decoder_{bg1,bg2,qc}.hip.
template <int KIND, int ZC, bool LEARNED>
int occupancy_decode(int threads, size_t lds);

int decoder_dup_occupancy(int kind, int Z, bool learned, int threads, size_t lds) {
    if (kind == kBG1) {
        if (Z == kBG1Special) return learned ? occupancy_decode<kBG1, kBG1Special, true>(threads, lds)
                                             : occupancy_decode<kBG1, kBG1Special, false>(threads, lds);
        return learned ? occupancy_decode<kBG1, 0, true>(threads, lds) : occupancy_decode<kBG1, 0, false>(threads, lds);
    }
    if (kind == kBG2) {
        if (Z == kBG2Special) return learned ? occupancy_decode<kBG2, kBG2Special, true>(threads, lds)
                                             : occupancy_decode<kBG2, kBG2Special, false>(threads, lds);
        return learned ? occupancy_decode<kBG2, 0, true>(threads, lds) : occupancy_decode<kBG2, 0, false>(threads, lds);
    }
    if (Z == kQCSpecial) return learned ? occupancy_decode<kWifi, kQCSpecial, true>(threads, lds)
                                        : occupancy_decode<kWifi, kQCSpecial, false>(threads, lds);
    return learned ? occupancy_decode<kWifi, 0, true>(threads, lds) : occupancy_decode<kWifi, 0, false>(threads, lds);
}

// One code kind's slim kernels: compiled-in offsets at ZS, the per-edge table elsewhere.
template <int KIND, int ZS>
static int slim_for_kind(bool learned, bool half, const DecodeArgs& a, int grid, int threads, size_t lds,
                         hipStream_t s) {
    if (a.Z == ZS) {
        if (half) return launch_decode_slim<KIND, ZS, false, true>(a, grid, threads, lds, s);
        return learned ? launch_decode_slim<KIND, ZS, true, false>(a, grid, threads, lds, s)
                       : launch_decode_slim<KIND, ZS, false, false>(a, grid, threads, lds, s);
    }
    if (half) return launch_decode_slim<KIND, 0, false, true>(a, grid, threads, lds, s);
    return learned ? launch_decode_slim<KIND, 0, true, false>(a, grid, threads, lds, s)
                   : launch_decode_slim<KIND, 0, false, false>(a, grid, threads, lds, s);
}

int decoder_launch_slim(int kind, bool learned, const DecodeArgs& a, int grid, hipStream_t s, bool half) {
    const int threads = a.cpb * a.Z;
    const size_t lds = (size_t)a.cpb * a.lds_cw_bytes + 16;  // + workgroup flag slot
    if (a.cpb < 1 || a.cpb > decoder_slim_cpb_max(kind, a.Z) || threads > kSlimMaxBlock ||
        lds > (size_t)max_lds_bytes() || grid < 1 || (n_nd1(kind) > 0 && !a.d1_scratch) || (half && learned) ||
        (half && !a.llr_half)) {
        set_last_error("slim decoder launch: unsupported shape");
        return LDPC_EINVAL;
    }
    if (kind == kBG1) return slim_for_kind<kBG1, kBG1Special>(learned, half, a, grid, threads, lds, s);
    if (kind == kBG2) return slim_for_kind<kBG2, kBG2Special>(learned, half, a, grid, threads, lds, s);
    return slim_for_kind<kWifi, kQCSpecial>(learned, half, a, grid, threads, lds, s);
}

int decoder_launch(int kind, bool learned, const DecodeArgs& a, hipStream_t s) {
    const int blocks = (a.n_cw + a.cpb - 1) / a.cpb;
    const int threads = a.cpb * a.Z;
    const size_t lds = (size_t)a.cpb * a.lds_cw_bytes + 16;  // + workgroup flag slot
    if (threads > kMaxBlock || lds > (size_t)max_lds_bytes() || blocks < 1) {
        set_last_error("launch shape exceeds the workgroup size or LDS limit");
        return LDPC_EINVAL;
    }
    const bool generic = a.force_generic != 0;
    if (kind == kBG1) {
        if (a.Z == kBG1Special && !generic)
            return learned ? launch_decode<kBG1, kBG1Special, true>(a, blocks, threads, lds, s)
                           : launch_decode<kBG1, kBG1Special, false>(a, blocks, threads, lds, s);
        return learned ? launch_decode<kBG1, 0, true>(a, blocks, threads, lds, s)
                       : launch_decode<kBG1, 0, false>(a, blocks, threads, lds, s);
    }
    if (kind == kBG2) {
        if (a.Z == kBG2Special && !generic)
            return learned ? launch_decode<kBG2, kBG2Special, true>(a, blocks, threads, lds, s)
                           : launch_decode<kBG2, kBG2Special, false>(a, blocks, threads, lds, s);
        return learned ? launch_decode<kBG2, 0, true>(a, blocks, threads, lds, s)
                       : launch_decode<kBG2, 0, false>(a, blocks, threads, lds, s);
    }
    if (a.Z == kQCSpecial && !generic)
        return learned ? launch_decode<kWifi, kQCSpecial, true>(a, blocks, threads, lds, s)
                       : launch_decode<kWifi, kQCSpecial, false>(a, blocks, threads, lds, s);
    return learned ? launch_decode<kWifi, 0, true>(a, blocks, threads, lds, s)
                   : launch_decode<kWifi, 0, false>(a, blocks, threads, lds, s);
}

}  // namespace nrldpc

namespace nrldpc {

int decoder_launch_half(int kind, const DecodeArgs& a, hipStream_t s) {
    const int blocks = (a.n_cw + a.cpb - 1) / a.cpb;
    const int threads = a.cpb * a.Z;
    const size_t lds = (size_t)a.cpb * a.lds_cw_bytes + 16;
    if (threads > kMaxBlock || lds > (size_t)max_lds_bytes() || blocks < 1 || !a.llr_half) {
        set_last_error("launch shape exceeds the workgroup size or LDS limit");
        return LDPC_EINVAL;
    }
    if (kind == kBG1)
        return a.Z == kBG1Special ? launch_decode_half<kBG1, kBG1Special>(a, blocks, threads, lds, s)
                                  : launch_decode_half<kBG1, 0>(a, blocks, threads, lds, s);
    if (kind == kBG2)
        return a.Z == kBG2Special ? launch_decode_half<kBG2, kBG2Special>(a, blocks, threads, lds, s)
                                  : launch_decode_half<kBG2, 0>(a, blocks, threads, lds, s);
    return a.Z == kQCSpecial ? launch_decode_half<kWifi, kQCSpecial>(a, blocks, threads, lds, s)
                             : launch_decode_half<kWifi, 0>(a, blocks, threads, lds, s);
}

size_t decoder_pair_lds_bytes(int kind, int Z) {
    int nnd = 0;
    if (kind == kBG1) nnd = Structure<kBG1>::NND;
    else if (kind == kBG2) nnd = Structure<kBG2>::NND;
    else nnd = Structure<kWifi>::NND;
    // read copy and accumulator, both [NND][2Z] words, and the pair's vote word
    return ((size_t)4 * nnd * 4 * Z + 16 + 15) & ~(size_t)15;
}

int decoder_pair_slots(int kind, int Z) {
    // fill whole wavefronts when possible, else the most lanes per wavefront, within the LDS
    const size_t lds = decoder_pair_lds_bytes(kind, Z);
    const size_t lim = (size_t)max_lds_bytes();
    int best = 1;
    double best_eff = 0.0;
    for (int cpb = 1; cpb * Z <= kMaxBlock && cpb * lds + 16 <= lim; cpb++) {
        const int th = cpb * Z;
        const double eff = (double)th / (double)(((th + 63) / 64) * 64);
        if (eff > best_eff + 1e-9) { best_eff = eff; best = cpb; }
    }
    return best;
}

#ifndef NRLDPC_CHAR_MAX_GROUPS
#define NRLDPC_CHAR_MAX_GROUPS 2
#endif
// Row groups of the CHAR_PRC pair kernel: two (the block rows dealt to two sets of wavefronts,
// twice the threads) for the base graphs where cpb pairs fill whole wavefronts and twice that
// fits a workgroup, else one.  Measured at T = 10 on gfx950: two groups decode 1.08-1.49x faster
// at 1, 2, 4, 5 and 6 wavefronts per group (BG1 Z = 384: 2.75 -> 2.01 ms per 4096 codewords),
// but slower with three (0.86x at BG1 Z = 192, 0.89x at BG1 Z = 96 with two pairs, 1.03x at
// BG2 Z = 192), so groups of three wavefronts stay single.
int decoder_pair_groups(int kind, int Z, int cpb) {
    const int th = cpb * Z;
    if (NRLDPC_CHAR_MAX_GROUPS < 2 || (kind != kBG1 && kind != kBG2)) return 1;
    return th % 64 == 0 && th != 192 && 2 * th <= kSlimMaxBlock ? 2 : 1;
}

// The CHAR_PRC schedule, two codewords per lane: a.cpb codeword pairs per workgroup, a.lds_cw_bytes
// LDS bytes per pair.
int decoder_launch_char(int kind, const DecodeArgs& a, hipStream_t s) {
    const int n_pairs = (a.n_cw + 1) / 2;
    const int blocks = (n_pairs + a.cpb - 1) / a.cpb;
    const int threads = a.cpb * a.Z;
    const size_t lds = (size_t)a.cpb * a.lds_cw_bytes + 16;
    if (threads > kMaxBlock || lds > (size_t)max_lds_bytes() || blocks < 1 || !a.llr_i8) {
        set_last_error("launch shape exceeds the workgroup size or LDS limit");
        return LDPC_EINVAL;
    }
    if (decoder_pair_groups(kind, a.Z, a.cpb) == 2) {
        if (kind == kBG1)
            return a.Z == kBG1Special ? launch_decode_char_split<kBG1, kBG1Special>(a, blocks, 2 * threads, lds, s)
                                      : launch_decode_char_split<kBG1, 0>(a, blocks, 2 * threads, lds, s);
        return launch_decode_char_split<kBG2, 0>(a, blocks, 2 * threads, lds, s);
    }
    if (kind == kBG1)
        return a.Z == kBG1Special ? launch_decode_char_pair<kBG1, kBG1Special>(a, blocks, threads, lds, s)
                                  : launch_decode_char_pair<kBG1, 0>(a, blocks, threads, lds, s);
    if (kind == kBG2)
        return a.Z == kBG2Special ? launch_decode_char_pair<kBG2, kBG2Special>(a, blocks, threads, lds, s)
                                  : launch_decode_char_pair<kBG2, 0>(a, blocks, threads, lds, s);
    return a.Z == kQCSpecial ? launch_decode_char_pair<kWifi, kQCSpecial>(a, blocks, threads, lds, s)
                             : launch_decode_char_pair<kWifi, 0>(a, blocks, threads, lds, s);
}

int decoder_launch_train(int kind, const DecodeArgs& a, hipStream_t s) {
    const int blocks = (a.n_cw + a.cpb - 1) / a.cpb;
    const int threads = a.cpb * a.Z;
    const size_t lds = (size_t)a.cpb * a.lds_cw_bytes + 16;
    if (threads > kMaxBlock || lds > (size_t)max_lds_bytes() || blocks < 1) {
        set_last_error("launch shape exceeds the workgroup size or LDS limit");
        return LDPC_EINVAL;
    }
    if (kind == kBG1) return launch_decode_train<kBG1>(a, blocks, threads, lds, s);
    if (kind == kBG2) return launch_decode_train<kBG2>(a, blocks, threads, lds, s);
    return launch_decode_train<kWifi>(a, blocks, threads, lds, s);
}

}  // namespace nrldpc
