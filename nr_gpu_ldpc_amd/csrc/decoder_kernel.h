// decoder_kernel.h -- the fused decoder kernel template (included by the per-code
// instantiation units decoder_bg1.hip, decoder_bg2.hip, decoder_qc.hip).
#pragma once
// decoder.hip -- fused flooding (learned) scaled min-sum decoder for gfx950.
//
// One workgroup decodes `cpb` codewords; lane (cw, r) owns expanded row r of every block row
// and sub-column r of every block column of codeword cw.  All iterations run inside one
// launch with the whole decoder state on chip:
//
//   LDS, per codeword   APP of the block columns of degree > 1, twice:
//                       * read copy  [NND][2Z]  (sub-columns 0..Z-1 stored twice so the
//                         circulant read at (r + V) never wraps: one VALU add per edge)
//                       * accumulator [NND][Z]  (APP_new, updated in place)
//   VGPRs, per lane     check-node state of every block row: the two scaled minima
//                       (|R| values), the sign bits of R and the arg-min index
//                       (12 bytes/row instead of the reference's R per edge), and the APP
//                       of every degree-1 (extension parity) column the lane's row feeds.
//
// Arithmetic follows the reference kernels exactly (cnp_kernel_1st_iter / cnp_kernel /
// vnp_kernel_normal / vnp_kernel_last_iter, src/ldpc_kernels.cu:72-468):
//   Q = APP_old - R_old;  two-min with the 1000.0f start value;  R = +-fl(0.75 m);
//   dt = fl(R_new - R_old);  APP_new = fl(...fl(APP_old + dt_i0) + dt_i1 ...) over incident
//   block rows in ascending order;  HD = APP > 0 ? 0 : 1.
// The arg-min index is replaced by the equality test |Q_k| == m1, which selects the same R
// (a tie at the minimum forces m2 == m1), and (m1, m2) are updated with min / med3, which
// reproduce the `<` / `else if <` rule exactly.  Per-column update order is kept by
// processing block rows in ascending order and placing a workgroup barrier before the VN
// update of every block row that shares an LDS column with the rows since the previous
// barrier (compile-time "epochs").
#include "code.h"
#include "decoder.h"

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace nrldpc {

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

__device__ __forceinline__ float as_f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t as_u(float f) { return __float_as_uint(f); }

// Rows of at most this many edges keep each edge's wrap-selected base from the pass-1 read to the
// VN update in the slim layout; larger rows recompute it.  Measured at BG1 Z=384, T=10, 4096
// codewords: 10 (198 of the 266 wrapping edges) 3.27 ms per launch, 19 (every row) 3.52 ms (VGPR
// spills 56 -> 94), none 3.29 ms.
constexpr int kSlimShareDeg = 10;

extern __shared__ __attribute__((aligned(16))) unsigned char g_smem[];

// LDS accesses through address-space-3 pointers made from 32-bit byte addresses: the
// address goes straight into the ds_read/ds_write VGPR (no per-access base add).  Every
// address below already includes lds_base(), the offset of the dynamic LDS block.
using lds_f32 = __attribute__((address_space(3))) float;
using lds_u32 = __attribute__((address_space(3))) uint32_t;
using lds_u8 = __attribute__((address_space(3))) uint8_t;
__device__ __forceinline__ uint32_t lds_base() {
    return (uint32_t)(uintptr_t)(__attribute__((address_space(3))) unsigned char*)g_smem;
}
__device__ __forceinline__ float lds_ld(uint32_t a) { return *(lds_f32*)(uintptr_t)a; }
__device__ __forceinline__ void lds_st(uint32_t a, float v) { *(lds_f32*)(uintptr_t)a = v; }
__device__ __forceinline__ uint32_t lds_ldu(uint32_t a) { return *(lds_u32*)(uintptr_t)a; }
__device__ __forceinline__ void lds_stu(uint32_t a, uint32_t v) { *(lds_u32*)(uintptr_t)a = v; }
__device__ __forceinline__ void lds_stb(uint32_t a, uint8_t v) { *(lds_u8*)(uintptr_t)a = v; }

// VN update of one LDS-resident column: APP_new += dt.  Within one epoch every column is
// updated by at most one lane, so a plain read-modify-write is race free.
__device__ __forceinline__ void lds_acc(uint32_t a, float dt) { lds_st(a, lds_ld(a) + dt); }

// Workgroup-wide "any lane true" through one LDS word (no static LDS, so the dynamic LDS
// region starts at address 0 and every offset folds into the instruction).
__device__ __forceinline__ bool block_any(bool v, uint32_t slot) {
    __syncthreads();
    if (threadIdx.x == 0) lds_stu(slot, 0u);
    __syncthreads();
    // one store per wavefront (same-address LDS stores of many lanes are serialised)
    if (__ballot(v) != 0u && (threadIdx.x & 63u) == 0u) lds_stu(slot, 1u);
    __syncthreads();
    const bool any = lds_ldu(slot) != 0u;
    __syncthreads();
    return any;
}

// Two workgroup-wide "any lane true" votes at once (bit 0: a, bit 1: b).
__device__ __forceinline__ uint32_t block_any2(bool a, bool b, uint32_t slot) {
    __syncthreads();
    if (threadIdx.x == 0) lds_stu(slot, 0u);
    __syncthreads();
    // reduced over the wavefront first: one atomic per wavefront instead of one per lane
    const uint32_t v = (__ballot(a) != 0u ? 1u : 0u) | (__ballot(b) != 0u ? 2u : 0u);
    if (v != 0u && (threadIdx.x & 63u) == 0u)
        __hip_atomic_fetch_or((lds_u32*)(uintptr_t)slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    const uint32_t m = lds_ldu(slot);
    __syncthreads();
    return m;
}

// binary16 storage of the HALF_PRC schedule: round to nearest even and back.
__device__ __forceinline__ float round_half(float x) { return (float)(_Float16)x; }
__device__ __forceinline__ float half_bits_to_float(uint16_t u) { return (float)__builtin_bit_cast(_Float16, u); }

// Dword 3 of a raw buffer resource on gfx9 (no format conversion, bounds by byte count).
constexpr int kBufferDword3 = 0x00020000;

// Per-edge table entries are read through the constant address space so that the compiler
// issues scalar (s_load) reads of these wave-uniform values.
using CEdge = const EdgeParam __attribute__((address_space(4)));
using CW2 = const float2 __attribute__((address_space(4)));

template <int KIND>
constexpr bool nd_columns_leading() {
    using S = Structure<KIND>;
    for (int j = 0; j < S::NND; j++)
        if (S::T.nd_index[j] != j) return false;
    return true;
}


constexpr int lifting_set_constexpr(int Z) {
    const int base[8] = {2, 3, 5, 7, 9, 11, 13, 15};
    for (int i = 0; i < 8; i++)
        for (int z = base[i]; z <= 384; z *= 2)
            if (z == Z) return i;
    return -1;
}

// V mod Z of edge e of base graph KIND at lifting size ZC (compile time; ZC > 0).
template <int KIND, int ZC>
constexpr int edge_vm(int e) {
    if constexpr (KIND == kBG1) return kBG1Entries[e].v[lifting_set_constexpr(ZC)] % ZC;
    else if constexpr (KIND == kBG2) return kBG2Entries[e].v[lifting_set_constexpr(ZC)] % ZC;
    else return kWifiEntries[e].v % ZC;
}


// The degree-1 slot that block row i feeds, or -1 (none, or more than one).
template <int KIND>
constexpr int row_d1_slot(int i) {
    using S = Structure<KIND>;
    int slot = -1, n = 0;
    for (int e = S::T.row_start[i]; e < S::T.row_start[i + 1]; e++)
        if (S::T.d1_slot[S::T.ecol[e]] >= 0) { slot = S::T.d1_slot[S::T.ecol[e]]; n++; }
    return n == 1 ? slot : -1;
}

// Read copy of block column J (< NND) at sub-column (r + V) for edge E of this lane.
template <int ZC, int KIND, int E, int J, class ET>
__device__ __forceinline__ float read_copy(uint32_t rRd, uint32_t rRdHi, ET et) {
    if constexpr (ZC > 0) {
        constexpr uint32_t off = (uint32_t)(J * 2 * ZC + edge_vm<KIND, ZC>(E)) * 4u;
        if constexpr (off < 0x10000u) return lds_ld(rRd + off);
        else return lds_ld(rRdHi + (off - 0x10000u));
    } else {
        return lds_ld(rRd + et[E].off_rd);
    }
}

// Accumulator address of block column J (< NND), sub-column (r + V) mod Z, for edge E.
template <int ZC, int KIND, int E, int J, class ET>
__device__ __forceinline__ uint32_t acc_addr(int r, uint32_t rHi, uint32_t rAcc, uint32_t rAccW, ET et) {
    if constexpr (ZC > 0) {
        constexpr int vm = edge_vm<KIND, ZC>(E);
        constexpr uint32_t off = (uint32_t)(J * ZC + vm) * 4u;
        static_assert(off < 0x10000u, "accumulator offset exceeds the LDS immediate");
        if constexpr (vm == 0) {
            return rAcc + off;
        } else if constexpr (ZC - vm > 64) {
            // The threshold is no inline constant: compare against it as a literal operand
            // (the compiler would otherwise load it into an SGPR with a scalar move, which costs
            // the wavefront an issue slot per edge).  Same two VALU instructions, VCC clobbered.
            uint32_t res;
            asm("v_cmp_gt_i32_e32 vcc, %1, %2\n\tv_cndmask_b32_e32 %0, %3, %4, vcc"
                : "=v"(res)
                : "i"(ZC - vm), "v"(r), "v"(rAccW), "v"(rAcc)
                : "vcc");
            return res + off;
        } else {
            return (r >= ZC - vm ? rAccW : rAcc) + off;
        }
    } else {
        const uint32_t acc = et[E].acc;
        return (rHi >= acc ? rAccW : rAcc) + et[E].acc_off;
    }
}

// Slim layout with compiled-in offsets: the accumulator lies NND*Z*4 bytes above the read copy,
// so one circulant wrap select per edge serves both of its accesses (pass-1 read at sel + rd(),
// VN update at sel + ac()).  Edges whose accumulator offset would exceed the 16-bit LDS
// immediate take bases raised by `shift`.
template <int KIND, int ZC>
struct SlimImm {
    static constexpr uint32_t accd = (uint32_t)Structure<KIND>::NND * ZC * 4u;
    static constexpr uint32_t max_off = (uint32_t)(Structure<KIND>::NND * ZC - 1) * 4u;
    static constexpr uint32_t shift = max_off + accd > 0xFFFCu ? max_off + accd - 0xFFFCu : 0u;
    template <int E, int J>
    static constexpr uint32_t off() { return (uint32_t)(J * ZC + edge_vm<KIND, ZC>(E)) * 4u; }
    template <int E, int J>
    static constexpr bool hi() { return off<E, J>() + accd > 0xFFFCu; }
    template <int E, int J>
    static constexpr uint32_t rd() {
        static_assert(!hi<E, J>() || off<E, J>() >= shift, "shifted read offset below zero");
        return hi<E, J>() ? off<E, J>() - shift : off<E, J>();
    }
    template <int E, int J>
    static constexpr uint32_t ac() { return rd<E, J>() + accd; }
};

// The base of edge E's circulant access: b, or bw for lanes whose (r + V) wraps (ZC > 0).
template <int ZC, int KIND, int E>
__device__ __forceinline__ uint32_t wrap_sel(int r, uint32_t b, uint32_t bw) {
    constexpr int vm = edge_vm<KIND, ZC>(E);
    if constexpr (vm == 0) {
        return b;
    } else if constexpr (ZC - vm > 64) {
        uint32_t res;  // literal-operand compare, as in acc_addr
        asm("v_cmp_gt_i32_e32 vcc, %1, %2\n\tv_cndmask_b32_e32 %0, %3, %4, vcc"
            : "=v"(res)
            : "i"(ZC - vm), "v"(r), "v"(bw), "v"(b)
            : "vcc");
        return res;
    } else {
        return r >= ZC - vm ? bw : b;
    }
}

// One workgroup's share of a decode: codewords blk*cpb .. blk*cpb+cpb-1.
// ZC > 0: specialised for one lifting size, every circulant offset is an immediate.
// ZC == 0: any lifting size, offsets come from the per-edge table (scalar loads).
// TRAIN (learned only): also store, per iteration and expanded row, the two minima, the
// first-strict-minimum and second-minimum edge indices and the Q sign bits, for the
// backward pass (layout [cw][iteration][row][Z]).
// MIXED (persistent mixed-batch kernel, any ZC): the workgroup decodes one tile of a mixed
// batch.  `mcw` lists the tile's codewords; each has its own parity-row count (at most
// a.m_rows, the tile's maximum), length and output location.  Rows beyond a codeword's own
// count are computed with zero channel LLRs on their extension columns and their VN updates
// are forced to +0.0, which leaves every APP bit pattern unchanged (APP is never -0.0).
// Lanes at or above cpb*Z shadow lane (cpb-1, tid mod Z): they compute the same values and
// skip the one non-idempotent store (the accumulator read-modify-write) and the outputs.
// MIXED only: `sel` marks the lane's codeword for decoding; `final_check` also evaluates the
// syndrome after the last iteration; `emit_only_ok` then writes outputs only for codewords
// whose hard decision passes it.  Returns, per lane, whether its codeword converged (early
// exit) or passed the final check.
// HALF: the reference's HALF_PRC schedule -- binary16 channel LLRs (a.llr_half), and R, dt and
// the APP between iterations rounded to binary16 where the reference stores them, all
// arithmetic in single precision.
// SLIM: the state of one codeword cut to fit two workgroups per CU -- a single read copy
// [NND][Z] (the circulant read selects between two bases, like the accumulator update) and the
// APP of the degree-1 columns in the workgroup's slot of a.d1_scratch instead of VGPRs.  Same
// arithmetic in the same order, so the results are bit-identical to the default layout.
template <int KIND, int ZC, bool LEARNED, bool TRAIN, bool MIXED, bool HALF = false, bool SLIM = false>
__device__ __forceinline__ bool decode_block(const DecodeArgs& a, const int blk,
                                             const MixedCw* __restrict__ mcw, const bool sel = true,
                                             const bool final_check = false,
                                             const bool emit_only_ok = false) {
    using S = Structure<KIND>;
    static_assert(nd_columns_leading<KIND>(), "LDS-resident columns must be the leading ones");
    static_assert(!(MIXED && TRAIN), "mixed tiles use the plain/learned forward path");
    static_assert(!(HALF && (LEARNED || TRAIN || MIXED)), "the binary16 schedule is plain min-sum");
    static_assert(!(SLIM && TRAIN), "the slim layout serves the plain, learned and binary16 decodes");
    constexpr int ROWS = S::ROWS;
    constexpr int NND = S::NND;
    constexpr int ND1 = S::ND1 > 0 ? S::ND1 : 1;

    const int Z = ZC > 0 ? ZC : a.Z;
    const int tid = threadIdx.x;
    int cwl = tid / Z;
    const int r = tid - cwl * Z;
    bool dead = false;  // MIXED: lane beyond the tile's cpb*Z lanes
    if constexpr (MIXED) {
        if (cwl >= a.cpb) { dead = true; cwl = a.cpb - 1; }
    }
    int cw = blk * a.cpb + cwl;
    bool live;
    int cwc;
    if constexpr (MIXED) {
        live = !dead && cwl < a.n_cw && sel;
        cwc = cwl < a.n_cw ? cwl : a.n_cw - 1;  // index into mcw
    } else if (a.list != nullptr) {
        // decode only the listed codewords: slot blk * cpb + cwl takes list[slot]
        const int n_list = *a.list_count;
        if (blk * a.cpb >= n_list) return false;  // uniform over the workgroup
        live = cw < n_list;
        cw = live ? a.list[cw] : 0;
        cwc = live ? cw : a.n_cw - 1;
    } else {
        live = cw < a.n_cw;
        cwc = live ? cw : a.n_cw - 1;
    }
    // One codeword per workgroup (and no mixed tile): every per-codeword flag is uniform over
    // the workgroup, so no workgroup vote is needed.
    const bool uniform_cw = !MIXED && a.cpb == 1;
    // All lanes of this wavefront decode the same codeword (dead lanes count as the codeword
    // they shadow).
    const bool wave_one_cw = __ballot(cwl != __builtin_amdgcn_readfirstlane(cwl)) == 0u;
    int m_lane = a.m_rows;  // parity rows of this lane's codeword
    if constexpr (MIXED) m_lane = mcw[cwc].m_rows;
    // MIXED: a wavefront whose lanes are all beyond cpb*Z only takes part in the barriers.
    const bool idle = MIXED && (int)(__builtin_amdgcn_readfirstlane(tid) & ~63) >= a.cpb * Z;
    const uint32_t Z4 = 4u * Z;
    const uint32_t base = lds_base() + cwl * a.lds_cw_bytes;
    constexpr uint32_t RDN = SLIM ? 1u : 2u;       // copies of each sub-column in the read copy
    uint32_t rRd = base + 4u * r;                  // read copy, block column 0
    uint32_t rAcc = base + RDN * NND * Z4 + 4u * r; // accumulator, block column 0
    uint32_t rAccW = rAcc - Z4;                    // same, for lanes whose (r+V) wraps
    uint32_t rRdHi = SLIM ? rRd - Z4 : rRd + 0x10000u;  // SLIM: read copy for wrapping lanes;
                                                   // else a second base (16-bit immediates)
    // SLIM with compiled-in offsets: the same two bases raised for the high columns (SlimImm)
    constexpr uint32_t kShift = (SLIM && ZC > 0) ? SlimImm<KIND, (ZC > 0 ? ZC : 1)>::shift : 0u;
    uint32_t rRdS = rRd + kShift, rRdHiS = rRdHi + kShift;
    const uint32_t rHi = ((uint32_t)r << 16) | 0xFFFFu;  // compared with EdgeParam::acc
    const uint32_t flagAddr = base + (RDN + 1u) * NND * Z4;
    // SLIM: this lane's degree-1 APPs, slot s at byte d1v + s * 4Z of the workgroup's scratch
    // slot, addressed through a buffer resource (wave-uniform base, one lane offset for all
    // slots: no per-slot 64-bit address is formed or kept)
    const __amdgpu_buffer_rsrc_t d1rs = __builtin_amdgcn_make_buffer_rsrc(
        SLIM ? a.d1_scratch + (size_t)blockIdx.x * a.d1_stride : nullptr, 0, SLIM ? a.cpb * ND1 * Z * 4 : 0,
        kBufferDword3);
    const uint32_t d1v = SLIM ? (uint32_t)(cwl * ND1 * Z + r) * 4u : 0u;
    auto d1_ld = [&](int slot) -> float {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(d1rs, d1v, slot * Z * 4, 0));
    };
    auto d1_st = [&](int slot, float v) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), d1rs, d1v, slot * Z * 4, 0);
    };
    const uint32_t wgFlag = lds_base() + (uint32_t)a.cpb * a.lds_cw_bytes;  // workgroup flag slot
    CEdge* et = (CEdge*)(a.etab);
    const int m_rows = a.m_rows;

    constexpr int NW = S::NW;
    float A1[ROWS], A2[ROWS];  // plain: fl(0.75 m1), fl(0.75 m2); learned: m1, m2
    uint32_t W[NW];            // per row: sign of R per edge + arg-min index (see Structure)
    float app1[SLIM ? 1 : ND1];  // APP of the degree-1 columns fed by this lane's rows
#pragma unroll
    for (int i = 0; i < ROWS; i++) { A1[i] = 0.f; A2[i] = 0.f; }
#pragma unroll
    for (int i = 0; i < NW; i++) W[i] = 0u;

    // ---- load channel LLRs
    const float* __restrict__ L = MIXED ? a.llr + mcw[cwc].llr_off : a.llr + (size_t)cwc * a.n_vn;
    const uint16_t* __restrict__ Lh = HALF ? a.llr_half + (size_t)cwc * a.n_vn : nullptr;
    auto llr_at = [&](int n) -> float {
        if constexpr (HALF) return half_bits_to_float(Lh[n]);
        else return L[n];
    };
    if (!idle) {
#pragma unroll 2
        for (int j = 0; j < NND; j++) {
            const float v = llr_at(j * Z + r) + 0.0f;  // -0.0 -> +0.0 (see the sign rule below)
            lds_st(rRd + j * RDN * Z4, v);
            if constexpr (!SLIM) lds_st(rRd + j * 2 * Z4 + Z4, v);
            lds_st(rAcc + j * Z4, v);
        }
    }
    if (!idle) static_for<NND, S::COLS>([&](auto JC) {
        constexpr int j = JC;
        constexpr int e = S::T.d1_edge[j];
        constexpr int i = S::T.d1_row[j];
        if (i < m_rows) {
            int c;
            if constexpr (ZC > 0) c = r + edge_vm<KIND, ZC>(e);
            else c = r + (int)et[e].off_rd;
            if (c >= Z) c -= Z;
            // MIXED: rows beyond the codeword's own count start from 0 (their columns lie past the
            // codeword's LLRs); lanes that shadow another lane leave its degree-1 slots alone
            const float v = (!MIXED || i < m_lane) ? llr_at(j * Z + c) + 0.0f : 0.0f;
            if constexpr (SLIM) {
                if (!MIXED || !dead) d1_st(S::T.d1_slot[j], v);
            } else {
                app1[S::T.d1_slot[j]] = v;
            }
        }
    });
    __syncthreads();

    bool emitted = false;
    bool final_ok = false;
    const bool early = a.early_exit != 0;
    // learned: [(n_iter+1)][nnz] (alpha, beta), row 0 zero; read-only during the launch, so read
    // through the constant address space (scalar loads into SGPRs, as for the edge table)
    CW2* wt = (CW2*)(a.weights);

    // Writes hard decisions (and APP) of every lane's codeword for which `mine` holds.
    auto emit = [&](bool mine, int iters) {
        if (mine) {
            // Loop-variant copies: no address of this path is precomputed outside it.
            uint32_t eb = base;
            int er = r;
            float* ap_out = a.app_out;
            CEdge* ee = et;
            asm volatile("" : "+v"(eb), "+v"(er), "+s"(ap_out), "+s"(ee));
            int ml = m_rows, nv = a.n_vn;  // this codeword's parity rows and length
            if constexpr (MIXED) {
                if (ap_out) ap_out += mcw[cwc].llr_off;
                ml = mcw[cwc].m_rows;
                nv = mcw[cwc].n_vn;
            } else if (ap_out) {
                ap_out += (size_t)cw * a.n_vn;
            }
#pragma unroll 1
            for (int j = 0; j < NND; j++) {
                const float x = lds_ld(eb + RDN * NND * Z4 + 4u * er + j * Z4);
                lds_stb(eb + j * Z + er, x > 0.f ? 0 : 1);
                if (ap_out) ap_out[j * Z + er] = x;
            }
            static_for<NND, S::COLS>([&](auto JC) {
                constexpr int j = JC;
                constexpr int e = S::T.d1_edge[j];
                constexpr int i = S::T.d1_row[j];
                if (i < ml) {
                    int c;
                    if constexpr (ZC > 0) c = er + edge_vm<KIND, ZC>(e);
                    else c = er + (int)ee[e].off_rd;
                    if (c >= Z) c -= Z;
                    float x;
                    if constexpr (SLIM) x = d1_ld(S::T.d1_slot[j]);
                    else x = app1[S::T.d1_slot[j]];
                    lds_stb(eb + j * Z + c, x > 0.f ? 0 : 1);
                    if (ap_out) ap_out[j * Z + c] = x;
                }
            });
            // zero the staging bytes up to the next multiple of 8 columns (the packed padding
            // bits), stepping by Z: fewer than 8 lanes per codeword when Z < 8
            for (int p = er; p < 8; p += Z) lds_stb(eb + nv + p, 0);
        }
        __syncthreads();
        if (mine) {
            int hb = a.hd_bytes;
            uint8_t* hrow = a.hd + (size_t)cw * a.hd_bytes;
            int* itp = a.iters_out ? a.iters_out + cw : nullptr;
            if constexpr (MIXED) {
                hb = mcw[cwc].hd_bytes;
                hrow = a.hd + mcw[cwc].hd_off;
                itp = a.iters_out ? a.iters_out + mcw[cwc].id : nullptr;
            }
            for (int b = r; b < hb; b += Z) {
                const uint32_t v0 = lds_ldu(base + 8u * b), v1 = lds_ldu(base + 8u * b + 4u);
                const uint32_t lo = (v0 | (v0 >> 7) | (v0 >> 14) | (v0 >> 21)) & 0xFu;
                const uint32_t hi = (v1 | (v1 >> 7) | (v1 >> 14) | (v1 >> 21)) & 0xFu;
                hrow[b] = (uint8_t)(lo | (hi << 4));
            }
            if (r == 0 && itp) *itp = iters;
        }
        __syncthreads();
    };

    // A wavefront none of whose lanes decodes a codeword still running (each one emitted at
    // convergence, not live, or beyond the tile) leaves the iteration loop for the barrier-only
    // loop below, which takes the workgroup's barriers and votes in the same order: its issue slots
    // go to the wavefronts of the codewords still decoding.  Wave-uniform.
    bool shadow = __builtin_amdgcn_readfirstlane((int)(idle || __ballot(live) == 0u)) != 0;
    int t = 0;
    for (; t < a.n_iter && !shadow; ++t) {
        // Loop-variant copies of the lane bases keep address arithmetic from being hoisted
        // out of the iteration loop into long-lived VGPRs.
        asm volatile("" : "+v"(rRd), "+v"(rAcc), "+v"(rAccW), "+v"(rRdHi));
        if constexpr (SLIM && ZC > 0) asm volatile("" : "+v"(rRdS), "+v"(rRdHiS));
        CW2* w_old = LEARNED ? wt + (size_t)t * a.nnz : nullptr;
        CW2* w_new = LEARNED ? wt + (size_t)(t + 1) * a.nnz : nullptr;
        static_for<0, ROWS>([&](auto IC) {
            constexpr int i = IC;
            constexpr int D = S::deg(i);
            constexpr int E0 = S::T.row_start[i];
            constexpr int WW = S::T.w_word[i];
            constexpr int WO = S::T.w_off[i];                 // sign bits start here
            constexpr int IO = S::half(i) ? WO + 11 : 24;     // arg-min index starts here
            int mr = m_rows;
            asm volatile("" : "+s"(mr));
            if (i >= mr) return;
            // Keep this row's table reads and wrap tests inside the row (bounded SGPR live
            // ranges): both are loop invariant and would otherwise be hoisted and spilled.
            asm volatile("" : "+s"(et));
            int rr = r;
            uint32_t rrHi = rHi;
            asm volatile("" : "+v"(rr), "+v"(rrHi));
            // SLIM: the row's degree-1 APP is loaded first, ahead of the LDS reads
            float d1x = 0.0f;
            if constexpr (SLIM && row_d1_slot<KIND>(i) >= 0) d1x = d1_ld(row_d1_slot<KIND>(i));
            float dtv[D];
            float xd1[D];  // SLIM: APP_old of the row's degree-1 edges
            // SLIM, compiled-in offsets: the wrap-selected base of each edge, from the pass-1 read
            // to the VN update
            constexpr bool SHARE = SLIM && ZC > 0 && D <= kSlimShareDeg;
            uint32_t selv[D];
            if (!idle) {
                float q[D], ro[D];
                float m1 = 1000.0f, m2 = 1000.0f;
                uint32_t neg = 0u;
                const float a1 = A1[i], a2 = A2[i];
                const uint32_t wo = W[WW];
                uint32_t a1i = 31u, a2i = 31u;  // TRAIN: arg-min / second arg-min (31 = none)
                uint32_t idxo = (wo >> IO) & 31u;
                asm volatile("" : "+v"(idxo));
                // pass 1: Q = APP_old - R_old, two-min, signs
                static_for<0, D>([&](auto KC) {
                    constexpr int k = KC;
                    constexpr int e = E0 + k;
                    constexpr int j = S::T.ecol[e];
                    float x;
                    if constexpr (j < NND) {
                        if constexpr (SHARE) {
                            using SI = SlimImm<KIND, (ZC > 0 ? ZC : 1)>;
                            constexpr bool hi = SI::template hi<e, j>();
                            const uint32_t sb = wrap_sel<ZC, KIND, e>(rr, hi ? rRdS : rRd, hi ? rRdHiS : rRdHi);
                            selv[k] = sb;
                            x = lds_ld(sb + SI::template rd<e, j>());
                        } else if constexpr (SLIM) {
                            x = lds_ld(acc_addr<ZC, KIND, e, j>(rr, rrHi, rRd, rRdHi, et));
                        } else {
                            x = read_copy<ZC, KIND, e, j>(rRd, rRdHi, et);
                        }
                    } else if constexpr (SLIM) {
                        if constexpr (row_d1_slot<KIND>(i) == S::T.d1_slot[j]) x = d1x;
                        else x = d1_ld(S::T.d1_slot[j]);
                        xd1[k] = x;
                    } else {
                        x = app1[S::T.d1_slot[j]];
                    }
                    const float mo = (idxo == (uint32_t)k) ? a2 : a1;
                    float rv;
                    if constexpr (LEARNED) {
                        const float wx = w_old[e].x, wy = w_old[e].y;
                        const float mag = wx * fmaxf(mo - wy, 0.0f);
                        rv = as_f(as_u(mag) ^ ((wo << (32 - WO - D + k)) & 0x80000000u));
                    } else {
                        rv = as_f((as_u(mo) & 0x7fffffffu) | ((wo << (32 - WO - D + k)) & 0x80000000u));
                    }
                    const float qv = x - rv;
                    q[k] = qv;
                    ro[k] = rv;
                    const float aq = fabsf(qv);
                    if constexpr (TRAIN) {
                        const bool lt1 = aq < m1, lt2 = aq < m2;
                        a2i = lt1 ? a1i : (lt2 ? (uint32_t)k : a2i);
                        a1i = lt1 ? (uint32_t)k : a1i;
                    }
                    m2 = __builtin_amdgcn_fmed3f(m1, m2, aq);
                    m1 = fminf(m1, aq);
                    // sign bit of Q; edge k lands at bit D-1-k.  APP is never -0.0 (the channel
                    // LLRs are canonicalised on load and x + y = -0.0 needs x = y = -0.0), so Q is
                    // never -0.0 and the raw sign bit equals the reference's (Q < 0).
                    neg = __builtin_amdgcn_alignbit(neg, as_u(qv), 31);
                });
                const uint32_t sgn = __popc(neg) & 1u;
                const uint32_t sbit = sgn << 31;
                if constexpr (TRAIN) {
                    if (live) {
                        const size_t ri = (((size_t)cw * a.n_iter + t) * ROWS + i) * (size_t)Z + r;
                        a.rec_m1[ri] = m1;
                        a.rec_m2[ri] = m2;
                        a.rec_w[ri] = neg | (a1i << 20) | (a2i << 25);
                    }
                }
                float n1, n2;
                if constexpr (LEARNED) { n1 = m1; n2 = m2; }
                else { n1 = 0.75f * m1; n2 = 0.75f * m2; }
                const float n1s = as_f(as_u(n1) ^ sbit), n2s = as_f(as_u(n2) ^ sbit);
                const bool row_on = !MIXED || i < m_lane;  // MIXED: row kept by this codeword
                // pass 2: R_new = (product of the other signs) * scaled min, dt = R_new - R_old
                uint32_t idxn = 0u;
                static_for<0, D>([&](auto KC) {
                    constexpr int k = KC;
                    constexpr int e = E0 + k;
                    const bool eq = fabsf(q[k]) == m1;
                    idxn = eq ? (uint32_t)k : idxn;
                    float ms;
                    if constexpr (LEARNED) {
                        const float wx = w_new[e].x, wy = w_new[e].y;
                        const float mag = wx * fmaxf((eq ? n2 : n1) - wy, 0.0f);
                        ms = as_f(as_u(mag) ^ sbit);
                    } else {
                        ms = eq ? n2s : n1s;
                    }
                    const float rn = as_f(as_u(ms) ^ (as_u(q[k]) & 0x80000000u));
                    if constexpr (MIXED) dtv[k] = row_on ? rn - ro[k] : 0.0f;
                    else if constexpr (HALF) dtv[k] = round_half(rn - ro[k]);
                    else dtv[k] = rn - ro[k];
                });
                if constexpr (HALF) {  // R is stored in binary16; dt used the unrounded R_new
                    A1[i] = round_half(n1);
                    A2[i] = round_half(n2);
                } else {
                    A1[i] = n1;
                    A2[i] = n2;
                }
                constexpr uint32_t dmask = (1u << D) - 1u;
                const uint32_t fresh_w = ((sgn ? ~neg : neg) & dmask) | (idxn << (IO - WO));
                if constexpr (!S::half(i)) W[WW] = fresh_w;
                else if constexpr (WO == 0) W[WW] = (W[WW] & 0xFFFF0000u) | fresh_w;
                else W[WW] = (W[WW] & 0x0000FFFFu) | (fresh_w << 16);
            }  // !idle
            if constexpr (S::T.epoch_start[i] != 0 && i != 0) __syncthreads();
            if (idle) return;
            // VN update, ascending block row per column
            // read-modify-write in chunks of 8 edges: loads first, then stores
            static_for<0, (D + 7) / 8>([&](auto CC) {
                constexpr int C0 = CC * 8;
                constexpr int CN = (D - C0) < 8 ? (D - C0) : 8;
                uint32_t addr[CN];
                float cur[CN];
                static_for<0, CN>([&](auto KC) {
                    constexpr int k = C0 + KC;
                    constexpr int e = E0 + k;
                    constexpr int j = S::T.ecol[e];
                    if constexpr (j < NND) {
                        if constexpr (SHARE) {
                            using SI = SlimImm<KIND, (ZC > 0 ? ZC : 1)>;
                            addr[KC] = selv[k] + SI::template ac<e, j>();
                        } else {
                            addr[KC] = acc_addr<ZC, KIND, e, j>(rr, rrHi, rAcc, rAccW, et);
                        }
                        cur[KC] = lds_ld(addr[KC]);
                    }
                });
                static_for<0, CN>([&](auto KC) {
                    constexpr int k = C0 + KC;
                    constexpr int e = E0 + k;
                    constexpr int j = S::T.ecol[e];
                    if constexpr (j < NND) {
                        if (!MIXED || !dead) lds_st(addr[KC], cur[KC] + dtv[k]);
                    } else if constexpr (SLIM) {
                        float nv = xd1[k] + dtv[k];
                        // HALF: stored in binary16 between iterations (the column's only update)
                        if constexpr (HALF) nv = t + 1 < a.n_iter ? round_half(nv) : nv;
                        if (!MIXED || !dead) d1_st(S::T.d1_slot[j], nv);
                    } else {
                        float& ap = app1[S::T.d1_slot[j]];
                        ap = ap + dtv[k];
                    }
                });
            });
        });
        __syncthreads();
        const bool last = t + 1 == a.n_iter;
        bool fin = false;  // MIXED: syndrome of the final hard decision
        if constexpr (MIXED) fin = final_check && last;
        else fin = a.fail_out != nullptr && last;
        if (last && !fin) break;
        // APP_new becomes the read copy of the next iteration.
        if (!idle) {
#pragma unroll 2
            for (int j = 0; j < NND; j++) {
                float v = lds_ld(rAcc + j * Z4);
                if constexpr (HALF) {  // the APP between iterations is stored in binary16
                    if (!last) {
                        v = round_half(v);
                        lds_st(rAcc + j * Z4, v);
                    }
                }
                lds_st(rRd + j * RDN * Z4, v);
                if constexpr (!SLIM) lds_st(rRd + j * 2 * Z4 + Z4, v);
            }
            if constexpr (HALF) {
                if (!last) {
#pragma unroll
                    for (int n = 0; n < ND1; n++) app1[n] = round_half(app1[n]);
                }
            }
        }
        const bool check = early || fin;
        if (check && r == 0 && !idle) lds_stu(flagAddr, 0u);
        // the workgroup vote word of this check is cleared here too (the previous check's read
        // of it is separated from this store by the barrier after the row loop)
        if (check && !uniform_cw && tid == 0) lds_stu(wgFlag, 0u);
        __syncthreads();
        if (check) {
            // Syndrome of the current hard decision; any unsatisfied check keeps decoding.  The
            // parity of a check row is taken over wavefront ballots of "APP > 0" (hard decision
            // 0): one compare per edge, the XOR of the 64-lane masks on the scalar unit.
            // A wavefront stops (a uniform branch) once every codeword among its lanes has an
            // unsatisfied check: before convergence that is almost always one of the first rows.
            uint64_t badm = 0u;
            // lanes of this wavefront that belong to this lane's codeword: [lo, hi) (lanes of
            // one codeword are contiguous; dead lanes lie beyond the codeword they shadow)
            int lane = tid & 63;
            asm volatile("" : "+v"(lane));  // computed here, not kept live across the loop
            const int wb = tid - lane;
            const int lo = max(cwl * Z - wb, 0);
            const int hi = min((cwl + 1) * Z - wb, 64);
            const uint64_t segm = (hi >= 64 ? ~0ull : ((1ull << hi) - 1ull)) & ~((1ull << lo) - 1ull);
            bool stop = false;
            if (!idle) static_for<0, ROWS>([&](auto IC) {
                constexpr int i = IC;
                constexpr int D = S::deg(i);
                constexpr int E0 = S::T.row_start[i];
                int mr = m_rows;
                asm volatile("" : "+s"(mr));
                if (i >= mr || stop) return;
                asm volatile("" : "+s"(et));
                uint64_t zeros = 0u;  // lanes with an odd number of hard-decision zeros
                float xs[D];
                constexpr int NLD = S::row_lds_edges(i);
                static_for<0, D>([&](auto KC) {
                    constexpr int e = E0 + KC;
                    constexpr int j = S::T.ecol[e];
                    if constexpr (j < NND) {
                        if constexpr (SLIM) xs[KC] = lds_ld(acc_addr<ZC, KIND, e, j>(r, rHi, rRd, rRdHi, et));
                        else xs[KC] = read_copy<ZC, KIND, e, j>(rRd, rRdHi, et);
                    } else if constexpr (SLIM) {
                        xs[KC] = d1_ld(S::T.d1_slot[j]);
                    } else {
                        xs[KC] = app1[S::T.d1_slot[j]];
                    }
                });
                // All of the row's LDS reads are issued before the first compare: under the
                // kernel's register pressure the scheduler would otherwise wait for each read
                // before issuing the next.
                if constexpr (NLD > 0) __builtin_amdgcn_sched_group_barrier(0x100, NLD, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, 2 * D, 0);
                static_for<0, D>([&](auto KC) { zeros ^= __ballot(xs[KC] > 0.f); });
                // ones = D - zeros: odd parity of the ones where that of the zeros differs from D's
                // (complemented masks are cut to the active lanes of a partial last wavefront)
                uint64_t odd = (D & 1) ? (~zeros & __ballot(true)) : zeros;
                if constexpr (MIXED) odd &= __ballot(i < m_lane);
                badm |= odd;
                if (badm != 0u) stop = wave_one_cw || __ballot((badm & segm) == 0u) == 0u;
            });
            // One store per codeword and wavefront, by the codeword's first lane in it, if any of
            // the codeword's lanes here has an unsatisfied check.
            if (lane == lo && (badm & segm) != 0u) lds_stu(flagAddr, 1u);
            __syncthreads();
            const bool conv = lds_ldu(flagAddr) == 0u;
            if (fin) {
                final_ok = conv;
                break;
            }
            const bool fresh = live && conv && !emitted;
            const bool rest = live && !emitted && !fresh;
            bool any_fresh, any_rest;
            if (uniform_cw) {  // one codeword per workgroup: both are workgroup-uniform already
                any_fresh = __builtin_amdgcn_readfirstlane((int)fresh) != 0;
                any_rest = __builtin_amdgcn_readfirstlane((int)rest) != 0;
            } else {
                // one atomic per wavefront into the vote word cleared before the check
                const uint32_t v = (__ballot(fresh) != 0u ? 1u : 0u) | (__ballot(rest) != 0u ? 2u : 0u);
                if (v != 0u && (tid & 63) == 0)
                    __hip_atomic_fetch_or((lds_u32*)(uintptr_t)wgFlag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __syncthreads();
                const uint32_t m = lds_ldu(wgFlag);
                any_fresh = (m & 1u) != 0u;
                any_rest = (m & 2u) != 0u;
            }
            if (any_fresh) {
                // staging reuses the read copy of the converged codewords only
                emit(fresh, t + 1);
                if (fresh) emitted = true;
                shadow = __builtin_amdgcn_readfirstlane((int)(__ballot(live && !emitted) == 0u)) != 0;
            }
            if (!any_rest) {
                shadow = false;  // the whole workgroup is done
                break;
            }
        }
    }
    if (shadow) {
        // The rest of the workgroup's iterations, barriers and votes only (the loop above with
        // every lane's work dropped: no codeword here is running, so it stores no syndrome flag,
        // casts no vote and emits nothing).
        for (; t < a.n_iter; ++t) {
            static_for<0, ROWS>([&](auto IC) {
                constexpr int i = IC;
                if constexpr (S::T.epoch_start[i] != 0 && i != 0) {
                    if (i < m_rows) __syncthreads();
                }
            });
            __syncthreads();
            const bool last = t + 1 == a.n_iter;
            bool fin = false;
            if constexpr (MIXED) fin = final_check && last;
            else fin = a.fail_out != nullptr && last;
            if (last && !fin) break;
            const bool check = early || fin;
            if (check && !uniform_cw && tid == 0) lds_stu(wgFlag, 0u);
            __syncthreads();
            if (check) {
                __syncthreads();
                if (fin) {
                    final_ok = lds_ldu(flagAddr) == 0u;
                    break;
                }
                bool any_fresh = false, any_rest = false;
                if (!uniform_cw) {
                    __syncthreads();
                    const uint32_t m = lds_ldu(wgFlag);
                    any_fresh = (m & 1u) != 0u;
                    any_rest = (m & 2u) != 0u;
                }
                if (any_fresh) emit(false, t + 1);
                if (!any_rest) break;
            }
        }
    }
    bool out = live && !emitted;
    if constexpr (MIXED) out = out && (!emit_only_ok || final_ok);
    if (uniform_cw ? __builtin_amdgcn_readfirstlane((int)out) != 0 : block_any(out, wgFlag)) emit(out, t < a.n_iter ? t + 1 : a.n_iter);
    if constexpr (!MIXED) {
        if (a.fail_out != nullptr && live && r == 0) a.fail_out[cw] = (emitted || final_ok) ? 0 : 1;
    }
    return emitted || final_ok;
}

template <int KIND, int ZC, bool LEARNED, bool TRAIN = false, bool HALF = false>
__global__ void __launch_bounds__(kMaxBlock) k_decode(DecodeArgs a) {
    decode_block<KIND, ZC, LEARNED, TRAIN, false, HALF>(a, blockIdx.x, nullptr);
}

// The slim layout with a persistent grid: every workgroup owns one slot of a.d1_scratch and
// decodes blocks blockIdx.x, blockIdx.x + gridDim.x, ...  Register budget for three wavefronts
// per SIMD (two 6-wavefront workgroups per CU, or one of 12).
template <int KIND, int ZC, bool LEARNED, bool HALF = false>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, kSlimMaxBlock), amdgpu_waves_per_eu(3)))
k_decode_slim(DecodeArgs a) {
    const int nblk = (a.n_cw + a.cpb - 1) / a.cpb;
    for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        decode_block<KIND, ZC, LEARNED, false, false, HALF, true>(a, blk, nullptr);
        __syncthreads();  // the next block's loads overwrite this block's LDS
    }
}

}  // namespace nrldpc
