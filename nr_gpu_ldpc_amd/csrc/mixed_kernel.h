// mixed_kernel.h -- device code of the persistent mixed-batch decoder (included by the four
// instantiation units mixed_k_{dup,slim}_{plain,learned}.hip, compiled in parallel).
//
// The host planner (mixed_plan.cpp) groups the batch into tiles of one (base graph, Z) each, at
// most cpb codewords per tile, ordered longest first.  A grid sized to the resident
// workgroups pulls tiles from a device work counter until the list is exhausted; every tile
// runs the fused decoder body with per-codeword row counts, early exit and outputs (compiled-in
// offsets for BG1 Z=384 and BG2 Z=208, the per-edge table for every other size).
// The last workgroup to finish clears the counters, so consecutive launches (and graph
// replays) need no reset.
#pragma once
#include "status.h"
#include "decoder_kernel.h"

namespace nrldpc {

using CTile = const MixedTile __attribute__((address_space(4)));

// One tile of (KIND, ZC): ZC > 0 is a lifting size with every circulant offset compiled in.
// SLIM: the tile's codewords use the slim layout (single LDS read copy, degree-1 APPs in the
// workgroup's slot of m.d1_scratch).
template <int KIND, int ZC, bool LEARNED, bool SLIM>
__device__ __forceinline__ void mixed_tile(const DecodeArgs& a, const MixedCw* cws, CTile* t, bool fb,
                                           uint32_t slot) {
    if (LEARNED && t->trained) {
        // learned weights; with fallback, the syndrome of the final hard decision decides which
        // codewords plain min-sum decodes again (its output replaces the learned one only where
        // it passes the check)
        const bool ok = decode_block<KIND, ZC, true, false, true, false, SLIM>(a, 0, cws, true, fb, false);
        const bool need = fb && (int)threadIdx.x < t->count * a.Z && !ok;
        if (fb && block_any(need, slot))
            decode_block<KIND, ZC, false, false, true, false, SLIM>(a, 0, cws, need, true, true);
    } else {
        decode_block<KIND, ZC, false, false, true, false, SLIM>(a, 0, cws);
    }
}

template <bool LEARNED, bool SLIM>
__device__ __forceinline__ void mixed_body(const MixedArgs& m) {
    const uint32_t slot = lds_base() + (uint32_t)m.slot_off;
    const int n_tiles = *m.n_tiles;
    for (;;) {
        __syncthreads();  // the previous tile is done with the broadcast word
        if (threadIdx.x == 0) lds_stu(slot, atomicAdd(&m.queue[0], 1u));
        __syncthreads();
        const int ti = __builtin_amdgcn_readfirstlane((int)lds_ldu(slot));
        if (ti >= n_tiles) break;
        // Tile fields are wave-uniform: read them through the constant address space so that
        // they land in SGPRs (scalar loads).
        CTile* t = (CTile*)(m.tiles) + ti;
        const int kind = t->kind;
        DecodeArgs a{};
        a.llr = m.llr;
        a.hd = m.hd;
        a.app_out = m.app_out;
        a.iters_out = m.iters_out;
        a.etab = t->etab;
        a.weights = t->weights;
        a.n_cw = t->count;
        a.Z = t->Z;
        a.m_rows = t->m_rows;
        a.n_iter = m.n_iter;
        a.cpb = t->cpb;
        a.early_exit = m.early_exit;
        a.lds_cw_bytes = t->lds_cw_bytes;
        a.nnz = kind == kBG1 ? Structure<kBG1>::NNZ : Structure<kBG2>::NNZ;
        a.d1_scratch = m.d1_scratch;
        a.d1_stride = m.d1_stride;
        const MixedCw* cws = m.cws + t->first;
        const bool fb = m.fallback != 0;
        // the two lifting sizes of the NR configurations that carry most traffic (BG1 Z=384,
        // BG2 Z=208) run with compiled-in offsets, every other one with the per-edge table
        if (kind == kBG1) {
            if (a.Z == 384) mixed_tile<kBG1, 384, LEARNED, SLIM>(a, cws, t, fb, slot);
            else mixed_tile<kBG1, 0, LEARNED, SLIM>(a, cws, t, fb, slot);
        } else {
            if (a.Z == 208) mixed_tile<kBG2, 208, LEARNED, SLIM>(a, cws, t, fb, slot);
            else mixed_tile<kBG2, 0, LEARNED, SLIM>(a, cws, t, fb, slot);
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&m.queue[1], 1u) == gridDim.x - 1) {  // every other workgroup has left
            atomicExch(&m.queue[0], 0u);
            atomicExch(&m.queue[1], 0u);
        }
    }
}

template <bool LEARNED>
__global__ void __launch_bounds__(kMaxBlock) k_decode_mixed(MixedArgs m) {
    mixed_body<LEARNED, false>(m);
}

// The slim layout: up to 12 wavefronts per workgroup at three per SIMD (one workgroup per CU).
template <bool LEARNED>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, kSlimMaxBlock), amdgpu_waves_per_eu(3)))
k_decode_mixed_slim(MixedArgs m) {
    mixed_body<LEARNED, true>(m);
}


}  // namespace nrldpc
