// decoder.h -- internal launch interface of the fused decoder kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace nrldpc {

constexpr int kMaxBlock = 512;          // threads per workgroup (2 waves per SIMD)
constexpr int kSlimMaxBlock = 768;      // slim-layout kernel: up to 12 wavefronts (3 per SIMD)
constexpr int kMaxLdsBytes = 160 * 1024; // gfx950 LDS per CU (upper bound; see max_lds_bytes)

// Per base-graph edge, in row-major edge order (four dwords, read through the scalar cache).
// For an edge into an LDS-resident column (nd = block column j < NND):
//   off_rd  = (j * 2Z + V) * 4                 byte offset into the read copy
//   acc     = (Z - V) << 16 | (j * Z + V) * 4  wrap threshold | byte offset into the accumulator
//   acc_off = (j * Z + V) * 4                  the offset alone, so that the kernel adds it
//                                              without a scalar mask per edge
// For an edge into a degree-1 column: off_rd = V (mod Z), acc = acc_off = 0.
struct alignas(16) EdgeParam {
    uint32_t off_rd;
    uint32_t acc;
    uint32_t acc_off;
    uint32_t reserved;
};

// Element types of mixed-batch channel LLRs (LDPC_LLR_* of the C ABI).
constexpr int kLlrF32 = 0, kLlrI8 = 1, kLlrF16 = 2;

struct DecodeArgs {
    const float* llr;      // [n_cw][n_vn]
    uint8_t* hd;           // [n_cw][hd_bytes]
    float* app_out;        // [n_cw][n_vn] or null
    int* iters_out;        // [n_cw] or null
    const EdgeParam* etab; // [nnz_full]
    const float2* weights; // learned: [(n_iter+1)][nnz] (alpha, beta); row 0 all zero
    int n_cw, Z, m_rows, n_vn, hd_bytes, n_iter, cpb, early_exit;
    int lds_cw_bytes;
    int nnz;
    int force_generic;     // 1: use the any-Z kernel even where a specialised one exists
    // training records (learned forward with records): [n_cw][n_iter][rows_full][Z]
    float* rec_m1;
    float* rec_m2;
    uint32_t* rec_w;       // Q sign bits (edge k at bit d-1-k) | arg-min << 20 | 2nd arg-min << 25
    int* fail_out;         // [n_cw] or null: 0 where the output hard decision satisfies every
                           // parity check, else 1 (checked after the last iteration)
    const int* list;       // or null: decode only codewords list[0 .. *list_count - 1]
    const int* list_count; // device count for `list`
    const uint16_t* llr_half;  // binary16 LLRs [n_cw][n_vn] for the HALF_PRC schedule (llr unused)
    const int8_t* llr_i8;      // signed-char LLRs [n_cw][n_vn] for the CHAR_PRC schedule (llr unused)
    float* d1_scratch;         // slim-layout kernel: APP of the degree-1 columns, one slot of
                               // d1_stride floats per resident workgroup, [cpb][ND1][Z] used
    int d1_stride;
};

// One codeword of a mixed batch (device table, built by the planner).
struct MixedCw {
    long long llr_off;  // float offset of the codeword's LLR row (also its APP row)
    long long hd_off;   // byte offset of its packed hard decisions
    int id;             // position in the caller's descriptor array (iteration output)
    int m_rows;         // parity rows kept
    int n_vn;           // (kb + m_rows) * Z
    int hd_bytes;       // ceil(n_vn / 8)
};

// One work item of the persistent mixed-batch kernel: up to cpb codewords of one (kind, Z).
struct MixedTile {
    const EdgeParam* etab;  // per-edge table of (kind, Z)
    const float2* weights;  // learned weights of (kind, Z) or null
    int kind, Z;
    int m_rows;             // largest parity-row count among the tile's codewords
    int cpb;                // LDS codeword slots (lanes used: cpb * Z)
    int first, count;       // codewords first .. first + count - 1 of the MixedCw list
    int lds_cw_bytes;
    int trained;            // 1: learned weights of this (kind, Z) are set (weights != null)
};

struct MixedArgs {
    const MixedTile* tiles;
    const MixedCw* cws;
    unsigned* queue;        // [0] next tile, [1] finished workgroups (both zero between launches)
    const float* llr;
    uint8_t* hd;
    float* app_out;         // or null
    int* iters_out;         // [n descriptors] or null
    const int* n_tiles;     // device word: tiles in the list (read by the kernel, so a captured
                            // launch serves any plan that fits the buffers)
    int n_iter, early_exit;
    int fallback;           // learned tiles: re-decode syndrome failures with plain min-sum
    int slot_off;           // LDS byte offset of the work-queue broadcast word
    float* d1_scratch;      // slim layout: degree-1 APP slots, d1_stride floats per workgroup
    int d1_stride;
};

// Narrow mixed-batch LLRs (int8 q -> q * scale, or binary16) -> float, for the element count in
// the device word *count (the plan's LLR extent, so that one captured launch serves every plan).
struct LlrConvertArgs {
    const void* in;
    float* out;
    const long long* count;
    int fmt;
    float scale;
};
int llr_convert_launch(const LlrConvertArgs& a, hipStream_t s);
int llr_convert_graph_node(const LlrConvertArgs& a, hipGraph_t g, const hipGraphNode_t* deps, size_t n_deps,
                           hipGraphNode_t* node);

// The persistent mixed kernel of one layout (slim: kLayoutSlim tiles, up to kSlimMaxBlock threads
// at three wavefronts per SIMD, a.d1_scratch slots; else kLayoutDup tiles, up to kMaxBlock).
int mixed_launch(bool learned, bool slim, const MixedArgs& a, int grid, int threads, size_t lds, hipStream_t s);
// The same launch as a kernel node of graph g after the given dependencies.
int mixed_graph_node(bool learned, bool slim, const MixedArgs& a, int grid, int threads, size_t lds, hipGraph_t g,
                     const hipGraphNode_t* deps, size_t n_deps, hipGraphNode_t* node);
// Resident workgroups per CU of the mixed kernel at this shape (0 on error).
int mixed_occupancy(bool learned, bool slim, int threads, size_t lds);
// Floats of one workgroup's degree-1 slot in the slim mixed kernel (any tile of either base graph).
constexpr int kMixedSlimD1Stride = 42 * kSlimMaxBlock;

// LDS tilings of the fused FP32 decoder (plain and learned), the autotuner's layout axis:
//   kLayoutDup   read copy [NND][2Z] (sub-columns stored twice: the circulant read never wraps)
//                and accumulator [NND][Z] per codeword, degree-1 APPs in VGPRs; workgroups of
//                up to kMaxBlock threads, one per cpb codewords (k_decode)
//   kLayoutSlim  read copy [NND][Z] (the read selects between two bases like the accumulator
//                update) and accumulator [NND][Z], degree-1 APPs in the workgroup's slot of a
//                device buffer; up to kSlimMaxBlock threads at three wavefronts per SIMD, a
//                persistent grid of one workgroup per CU (k_decode_slim)
// Both run the same arithmetic in the same order: their outputs are identical.
constexpr int kLayoutDup = 1, kLayoutSlim = 2;

size_t decoder_lds_bytes_per_cw(int kind, int Z);
size_t decoder_slim_lds_bytes_per_cw(int kind, int Z);
// Largest codewords per workgroup of each layout (threads and LDS); 0 if none fits.
int decoder_dup_cpb_max(int kind, int Z);
int decoder_slim_cpb_max(int kind, int Z);
size_t decoder_slim_slot_bytes(int kind, int Z, int cpb);  // a.d1_scratch bytes per workgroup slot
// Persistent slim launch of `grid` workgroups of a.cpb codewords (a.d1_scratch holds grid slots);
// half: the HALF_PRC schedule (binary16 LLRs in a.llr_half, plain min-sum).
int decoder_launch_slim(int kind, bool learned, const DecodeArgs& a, int grid, hipStream_t s, bool half = false);
// Resident workgroups per CU of the dup-layout kernel at this shape (0 on error).
int decoder_dup_occupancy(int kind, int Z, bool learned, int threads, size_t lds);

// Cost model of one layout for one (kind, Z): a round is every CU decoding cw_per_cu codewords
// (cpb per workgroup) and takes round_us; n codewords take ceil(n / (cw_per_cu * CUs)) rounds.
struct LayoutModel {
    int cpb = 0;          // 0: layout not in use
    int cw_per_cu = 0;
    float round_us = 0.f;
};
struct LayoutTuning {
    int kind, Z;
    LayoutModel dup, slim;
};
// The measured table compiled into the library (decode_tuning.inc), or null for an untuned code.
const LayoutTuning* builtin_layout_tuning(int kind, int Z);
// The layout with the shorter modelled time for n_cw codewords on `cus` CUs (kLayoutDup on ties
// or when slim is not in use).
int layout_for_batch(const LayoutModel& dup, const LayoutModel& slim, int n_cw, int cus);
int max_lds_bytes();  // per-workgroup dynamic LDS limit of the current device
int decoder_launch(int kind, bool learned, const DecodeArgs& a, hipStream_t s);
// The reference's HALF_PRC schedule (binary16 LLRs in a.llr_half).
int decoder_launch_half(int kind, const DecodeArgs& a, hipStream_t s);
// The reference's CHAR_PRC schedule (signed-char LLRs in a.llr_i8), two codewords per lane:
// a.cpb = codeword pairs per workgroup (decoder_pair_slots), a.lds_cw_bytes = LDS bytes per pair
// (decoder_pair_lds_bytes).
int decoder_launch_char(int kind, const DecodeArgs& a, hipStream_t s);
size_t decoder_pair_lds_bytes(int kind, int Z);
int decoder_pair_slots(int kind, int Z);
// row groups of the pair kernel for cpb pairs per workgroup (threads = groups * cpb * Z)
int decoder_pair_groups(int kind, int Z, int cpb);
// Learned forward that also writes the training records (generic-Z kernel).
int decoder_launch_train(int kind, const DecodeArgs& a, hipStream_t s);

}  // namespace nrldpc
