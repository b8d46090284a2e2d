// decoder_pair.h -- the reference's CHAR_PRC schedule (its default build, src/globals.h:12-13)
// with two codewords per lane in packed 16-bit integer arithmetic (included by
// decoder_pair.hip).
//
// Every value of this schedule is an integer: the channel LLRs are signed chars, R and dt are
// truncated to signed char (src/ldpc_kernels.cu:173-175, 304-306), the APP is clamped to
// [-128, 127] between iterations (:387-390) and a saturated APP (127 or -128) enters the check
// node without R_old (:258-264).  So the whole decode runs exactly in 16-bit integers, and one
// lane carries codeword 2g in the low half and codeword 2g+1 in the high half of each register
// (v_pk_* instructions): every instruction advances two codewords.
//
// One workgroup holds `cpb` codeword pairs; lane (pair slot p, r) owns expanded row r of every
// block row and sub-column r of every block column of both codewords of pair p.  With two row
// groups (k_decode_char_split) the workgroup holds two such lane sets, each running half of the
// block rows (PairRegs deals them by edge count) and holding only those rows' registers.
//
//   VGPRs, per lane   R of every edge into a column of degree > 1, for both codewords, as signed
//                     bytes (two edges per register: bytes 0/2 = even edge of codeword 2g / 2g+1,
//                     bytes 1/3 = odd edge); and per degree-1 column the lane's rows feed one
//                     register with that column's APP and its single edge's R: bytes [R, A, R, A]
//                     (the APP is clamped at its one update per iteration; after the last
//                     iteration the register holds the unclamped APP as an int16 pair)
//   LDS, per pair     read copy [NND][2Z] of (x, nf) words: bytes 1/3 = APP of the two codewords,
//                     bytes 0/2 = 0xFF where that APP is not saturated (else 0), sub-columns
//                     stored twice so that the read at (r + V) never wraps;
//                     accumulator [NND][2Z]: the new APP as a 32-bit "linear" pair lo + 65536 hi;
//                     each edge adds its dt at (r + V) with one LDS atomic, and the two halves
//                     of a sub-column (c and c + Z) are summed once per iteration (integer sums
//                     are exact in any order, so neither the ascending-row order nor a circulant
//                     wrap is needed).
//
// Per edge:  Q = x - R_old (x when saturated) = v_pk_mad(R_old, nf, x);  |Q|;  the two minima
// (min / max-min);  the sign parity by XOR;  then R = +-min(trunc(0.75 m), 128), clamped to 127,
// with m = m2 on the edge with |Q| == m1 (the same R as the reference's first-arg-min rule, a tie
// forcing m2 == m1);  dt = clamp(R - R_old, -128, 127).
// Early exit: a codeword whose hard decision has zero syndrome after an iteration is frozen
// (its dt masked to 0, so its APP keeps the value the reference outputs) and the workgroup stops
// once every codeword has converged or the iterations are spent.
#pragma once

#include "code.h"
#include "decoder.h"

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace nrldpc {
namespace pairk {

typedef short s2 __attribute__((ext_vector_type(2)));
using u32 = uint32_t;

template <int B, int E, class F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

__device__ __forceinline__ s2 S(u32 u) { return __builtin_bit_cast(s2, u); }
__device__ __forceinline__ u32 U(s2 v) { return __builtin_bit_cast(u32, v); }
__device__ __forceinline__ s2 smin(s2 a, s2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ s2 smax(s2 a, s2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ s2 sp(short v) { return (s2){v, v}; }
typedef unsigned short u2 __attribute__((ext_vector_type(2)));
// a - b per half, saturated at 0 (unsigned): one v_pk_sub_u16 with clamp
__device__ __forceinline__ s2 usub_sat(s2 a, s2 b) {
    return __builtin_bit_cast(s2, __builtin_elementwise_sub_sat(__builtin_bit_cast(u2, a), __builtin_bit_cast(u2, b)));
}
// a - b per half, saturated to [-32768, 32767]: one v_pk_sub_i16 with clamp
__device__ __forceinline__ s2 ssub_sat(s2 a, s2 b) { return __builtin_elementwise_sub_sat(a, b); }
// x * 256 per half as unsigned 16-bit (x in [0, 255])
__device__ __forceinline__ s2 shl8(s2 x) { return __builtin_bit_cast(s2, __builtin_bit_cast(u2, x) << (unsigned short)8); }
// min per half, unsigned
__device__ __forceinline__ s2 umin(s2 a, s2 b) {
    return __builtin_bit_cast(s2, __builtin_elementwise_min(__builtin_bit_cast(u2, a), __builtin_bit_cast(u2, b)));
}

extern __shared__ __attribute__((aligned(16))) unsigned char g_psmem[];
using lds_u32 = __attribute__((address_space(3))) uint32_t;
using lds_u8 = __attribute__((address_space(3))) uint8_t;
__device__ __forceinline__ u32 lds_base() {
    return (u32)(uintptr_t)(__attribute__((address_space(3))) unsigned char*)g_psmem;
}
__device__ __forceinline__ u32 ld(u32 a) { return *(lds_u32*)(uintptr_t)a; }
__device__ __forceinline__ void st(u32 a, u32 v) { *(lds_u32*)(uintptr_t)a = v; }
__device__ __forceinline__ void stb(u32 a, uint8_t v) { *(lds_u8*)(uintptr_t)a = v; }
__device__ __forceinline__ void lds_add(u32 a, u32 v) {
    __hip_atomic_fetch_add((lds_u32*)(uintptr_t)a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Pair (lo, hi) -> lo + 65536 * hi as a 32-bit two's-complement integer: (v ^ 0x8000) - 0x8000.
// Sums of such values are the linear encodings of the per-half sums while each half stays in
// [-32768, 32767].
__device__ __forceinline__ u32 to_lin(s2 v, u32 k8000, u32 kneg8000) {
    u32 r;
    asm("v_xad_u32 %0, %1, %2, %3" : "=v"(r) : "v"(U(v)), "v"(k8000), "v"(kneg8000));
    return r;
}
// Inverse of to_lin: (x + 0x8000) ^ 0x8000.
__device__ __forceinline__ s2 from_lin(u32 x) { return S((x + 0x8000u) ^ 0x8000u); }

// Read-copy word of an APP pair x (both halves in [-128, 127]): bytes 1/3 = x, bytes 0/2 =
// 0xFF unless x is saturated.  x(x + 1) = 16256 exactly at x = 127 and x = -128, and is at most
// 16002 elsewhere, so (x(x + 1) - 16256) >> 15 is -1 for an unsaturated x and 0 otherwise.
__device__ __forceinline__ u32 rd_word(s2 x) {
    const s2 nf = (x * x + x - sp(16256)) >> 15;
    return (U(nf) & 0x00FF00FFu) | U(x << 8);
}
// Q = x - R_old, or x where x is saturated (nf = 0 there, -1 elsewhere): one v_pk_mad.
__device__ __forceinline__ s2 q_of(u32 word, s2 rold) {
    const s2 x = S(word) >> 8;
    const s2 nf = S(__builtin_amdgcn_perm(word, word, 0x02020000u));
    return rold * nf + x;
}
__device__ __forceinline__ s2 nf_of(s2 x) { return (x * x + x - sp(16256)) >> 15; }

// R of the N-th edge into a degree > 1 column (two codewords) from the packed edge registers.
template <int N>
__device__ __forceinline__ s2 r_get(const u32* RW) {
    if constexpr (N & 1) return S(RW[N >> 1]) >> 8;
    else return (S(RW[N >> 1]) << 8) >> 8;
}
// The same R times 256 (its byte in the high byte of each half, low byte 0).
template <int N>
__device__ __forceinline__ s2 r256_get(const u32* RW) {
    if constexpr (N & 1) return S(RW[N >> 1] & 0xFF00FF00u);
    else return S(RW[N >> 1]) << 8;
}

template <int KIND, int ZC>
constexpr int edge_vm_p(int e) {
    constexpr int base[8] = {2, 3, 5, 7, 9, 11, 13, 15};
    int ils = -1;
    for (int i = 0; i < 8; i++)
        for (int z = base[i]; z <= 384; z *= 2)
            if (z == ZC) ils = i;
    if constexpr (KIND == kBG1) return kBG1Entries[e].v[ils] % ZC;
    else if constexpr (KIND == kBG2) return kBG2Entries[e].v[ils] % ZC;
    else return kWifiEntries[e].v % ZC;
}

using CEdge = const EdgeParam __attribute__((address_space(4)));

// Row groups and register words.  With NG row groups the block rows are dealt to NG sets of
// wavefronts (each set covers every lane position r of every pair), greedily by edge count, so
// that each set runs about 1/NG of the check-node work and holds only its own rows' state.
// Register word of each edge into a column of degree > 1: the group's edges numbered in
// row-major order skipping the degree-1 columns' edges, two per register; degree-1 columns get
// register slots among their row's group.
template <int KIND, int NG = 1>
struct PairRegs {
    using Sx = Structure<KIND>;
    struct Tab {
        int grp[Sx::ROWS];                   // row group of each block row
        int idx[Sx::NNZ > 0 ? Sx::NNZ : 1];  // index among the group's edges into degree > 1 columns, or -1
        int d1[Sx::COLS];                    // register slot of a degree-1 column in its group, else -1
        int n[NG];                           // edges into degree > 1 columns, per group
        int nd1[NG];                         // degree-1 columns, per group
    };
    static constexpr Tab make() {
        Tab t{};
        int load[NG] = {};
        for (int g = 0; g < NG; g++) t.n[g] = t.nd1[g] = 0;
        for (int i = 0; i < Sx::ROWS; i++) {
            int g = 0;
            for (int h = 1; h < NG; h++)
                if (load[h] < load[g]) g = h;
            t.grp[i] = g;
            load[g] += Sx::T.row_start[i + 1] - Sx::T.row_start[i];
            for (int e = Sx::T.row_start[i]; e < Sx::T.row_start[i + 1]; e++)
                t.idx[e] = Sx::T.nd_index[Sx::T.ecol[e]] >= 0 ? t.n[g]++ : -1;
        }
        for (int j = 0; j < Sx::COLS; j++) t.d1[j] = Sx::T.d1_slot[j] >= 0 ? t.nd1[t.grp[Sx::T.d1_row[j]]]++ : -1;
        return t;
    }
    static constexpr Tab T = make();
    static constexpr int max_of(const int* v) {
        int m = 0;
        for (int g = 0; g < NG; g++) m = v[g] > m ? v[g] : m;
        return m;
    }
    static constexpr int NRW = (max_of(T.n) + 1) / 2;
    static constexpr int ND1 = max_of(T.nd1);
    // Whether the other half of edge e's register word belongs to the neighbouring edge of the
    // same row (k: e's position in a row of degree d): then both halves are written together.
    static constexpr bool partner_in_row(int e, int k, int d) {
        const int n = T.idx[e];
        if (n & 1) return k > 0 && T.idx[e - 1] == n - 1;
        return k + 1 < d && T.idx[e + 1] == n + 1;
    }
};

// Byte offset (relative to a region base) of edge E's sub-column in a doubled [NND][2Z] array.
template <int KIND, int ZC, int E, int J>
__device__ __forceinline__ u32 off2(CEdge* et) {
    if constexpr (ZC > 0) return (u32)(J * 2 * ZC + edge_vm_p<KIND, ZC>(E)) * 4u;
    else return et[E].off_rd;
}
// LDS word at region offset `off` through two bases (instruction offsets are 16-bit).
__device__ __forceinline__ u32 ld_at(u32 lo, u32 hi, u32 off) {
    return off < 0x10000u ? ld(lo + off) : ld(hi + (off - 0x10000u));
}

// The decoder body.  NG row groups: the workgroup holds NG * cpb * Z lanes, group g being lanes
// [g cpb Z, (g + 1) cpb Z), whole wavefronts each (the launch guarantees it); a group runs the
// check nodes of its rows for every pair of the workgroup, and the groups meet only in the LDS
// accumulator (integer sums, order-free) and at the per-iteration barrier.
template <int KIND, int ZC, bool EARLY, int NG>
__device__ __forceinline__ void decode_pair(const DecodeArgs& a) {
    using Sx = Structure<KIND>;
    constexpr int ROWS = Sx::ROWS;
    constexpr int NND = Sx::NND;
    using PR = PairRegs<KIND, NG>;
    constexpr int ND1 = PR::ND1 > 0 ? PR::ND1 : 1;
    constexpr int NRW = PR::NRW > 0 ? PR::NRW : 1;

    const int Z = ZC > 0 ? ZC : a.Z;
    const int tid = threadIdx.x;
    const int gsz = a.cpb * Z;  // lanes of one row group
    // row group of this wavefront (uniform: groups are whole wavefronts)
    const int grp = NG == 1 ? 0 : __builtin_amdgcn_readfirstlane(tid / gsz);
    const int tg = tid - grp * gsz;
    const int slot = tg / Z;
    const int r = tg - slot * Z;
    const int n_pairs = (a.n_cw + 1) >> 1;
    const int g = blockIdx.x * a.cpb + slot;
    const bool slot_live = g < n_pairs;
    const int gc = slot_live ? g : n_pairs - 1;
    const int c0 = 2 * gc;
    const int c1 = 2 * gc + 1 < a.n_cw ? 2 * gc + 1 : c0;  // a lone last codeword pairs with itself
    const bool hi_real = slot_live && 2 * g + 1 < a.n_cw;
    // halves that hold a codeword still being decoded (0xFFFF per half)
    u32 live = slot_live ? (hi_real ? 0xFFFFFFFFu : 0x0000FFFFu) : 0u;
    u32 its = (u32)a.n_iter | ((u32)a.n_iter << 16);  // iterations run, per half
    const int m_rows = a.m_rows;
    const u32 Z4 = 4u * (u32)Z;
    const u32 base = lds_base() + (u32)slot * (u32)a.lds_cw_bytes;
    const u32 RD = base;                       // read copy [NND][2Z]
    const u32 AC = base + 2u * NND * Z4;       // accumulator [NND][2Z]
    const u32 FLAG = base + 4u * NND * Z4;     // per-pair vote word
    const u32 WGF = lds_base() + (u32)a.cpb * (u32)a.lds_cw_bytes;  // workgroup vote word
    u32 rRd = RD + 4u * r, rRdHi = rRd + 0x10000u;
    u32 rAc = AC + 4u * r, rAcHi = rAc + 0x10000u;
    CEdge* et = (CEdge*)(a.etab);
    u32 k8000 = 0x8000u, kneg = 0xFFFF8000u;
    asm volatile("" : "+v"(k8000), "+v"(kneg));

    u32 RW[NRW];
#pragma unroll
    for (int i = 0; i < NRW; i++) RW[i] = 0u;
    u32 d1w[ND1];  // degree-1 columns: bytes [R, A, R, A] (int16 APP pair after the last iteration)

    // ---- channel LLRs (signed chars, used as the LLR values)
    const int8_t* __restrict__ L0 = a.llr_i8 + (size_t)c0 * a.n_vn;
    const int8_t* __restrict__ L1 = a.llr_i8 + (size_t)c1 * a.n_vn;
    auto pair_at = [&](int n) -> s2 { return (s2){(short)L0[n], (short)L1[n]}; };
#pragma unroll 2
    for (int j = grp; j < NND; j += NG) {
        const s2 x = pair_at(j * Z + r);
        const u32 w = rd_word(x);
        st(rRd + j * 2 * Z4, w);
        st(rRd + j * 2 * Z4 + Z4, w);
        st(rAc + j * 2 * Z4, to_lin(x, k8000, kneg));
        st(rAc + j * 2 * Z4 + Z4, 0u);
    }
    sfor<NND, Sx::COLS>([&](auto JC) {
        constexpr int j = JC;
        constexpr int e = Sx::T.d1_edge[j];
        constexpr int i = Sx::T.d1_row[j];
        if (PR::T.grp[i] != grp) return;
        if (i < m_rows) {
            int c;
            if constexpr (ZC > 0) c = r + edge_vm_p<KIND, ZC>(e);
            else c = r + (int)et[e].off_rd;
            if (c >= Z) c -= Z;
            d1w[PR::T.d1[j]] = U(pair_at(j * Z + c) << 8);  // R = 0
        } else {
            d1w[PR::T.d1[j]] = 0u;
        }
    });
    __syncthreads();

    bool ran_last = false;  // the last iteration ran (degree-1 registers hold int16 APP pairs)
    int t = 0;
    for (; t < a.n_iter; ++t) {
        asm volatile("" : "+v"(rRd), "+v"(rRdHi), "+v"(rAc), "+v"(rAcHi));
        const bool last = t + 1 == a.n_iter;
        sfor<0, ROWS>([&](auto IC) {
            constexpr int i = IC;
            constexpr int D = Sx::deg(i);
            constexpr int E0 = Sx::T.row_start[i];
            if constexpr (NG > 1)
                if (PR::T.grp[i] != grp) return;
            int mr = m_rows;
            asm volatile("" : "+s"(mr));
            if (i >= mr) return;
            asm volatile("" : "+s"(et));
            // rows of at most 12 edges hold |Q| and R_old from pass 1 to pass 2; wider ones (the
            // register peak) recompute them there
            constexpr bool KEEP_A = D <= 12;
            s2 q[D], aq[KEEP_A ? D : 1];
            s2 rod1 = sp(0), rod1_256 = sp(0);  // R_old of the row's degree-1 edge (and times 256), if any
            s2 m1 = sp(1000), m2 = sp(1000);
            u32 par = 0u;
            // pass 1: Q, |Q|, the two minima, the sign parity
            sfor<0, D>([&](auto KC) {
                constexpr int k = KC;
                constexpr int e = E0 + k;
                constexpr int j = Sx::T.ecol[e];
                s2 qv;
                if constexpr (j < NND) {
                    const s2 ro = r_get<PR::T.idx[e]>(RW);
                    const u32 off = off2<KIND, ZC, e, j>(et);
                    qv = q_of(ld_at(rRd, rRdHi, off), ro);
                } else {
                    const u32 w = d1w[PR::T.d1[j]];
                    const s2 x = S(w) >> 8;
                    rod1_256 = S(w) << 8;
                    rod1 = rod1_256 >> 8;
                    qv = rod1 * nf_of(x) + x;
                }
                const s2 av = smax(qv, sp(0) - qv);
                q[k] = qv;
                if constexpr (KEEP_A) aq[k] = av;
                m2 = smax(m1, smin(m2, av));
                m1 = smin(m1, av);
                par ^= U(qv);
            });
            // scaled magnitudes with the product sign, times 256: +-min(256 trunc(0.75 m), 32767).
            // Pass 2 keeps R times 256 (R in the high byte): 32767 and -32767 stand for R = 127
            // and -128 (the clamp of +128 to 127 comes for free), and dt = clamp(R - R_old, -128,
            // 127) is the high byte of the saturating difference of the two scaled values.
            const s2 pm = S(par) >> 15;
            const s2 n1 = umin(shl8(m1 * sp(3) >> 2), sp(32767));
            const s2 n2 = umin(shl8(m2 * sp(3) >> 2), sp(32767));
            const s2 n1s = (n1 ^ pm) - pm, n2s = (n2 ^ pm) - pm;
            const s2 d21 = n2s - n1s;
            const s2 m1p = m1 + sp(1);
            // pass 2, stage by stage over chunks of CH edges (dependent packed instructions issued
            // back to back cost a wait state each): R, dt, the VN update; R stored back two edges
            // per register
            constexpr int CH = 4;
            s2 pend = sp(0);  // R of an even-numbered edge whose odd partner follows in this row
            sfor<0, (D + CH - 1) / CH>([&](auto CC) {
                constexpr int K0 = CC * CH;
                constexpr int CN = (D - K0) < CH ? (D - K0) : CH;
                s2 v[CN], sq[CN];
                if constexpr (!KEEP_A)
                    sfor<0, CN>([&](auto KC) { v[KC] = smax(q[K0 + KC], sp(0) - q[K0 + KC]); });
                sfor<0, CN>([&](auto KC) {  // (m1 + 1) - |Q| saturated at 0: 1 on the edge(s) with |Q| == m1
                    if constexpr (KEEP_A) v[KC] = usub_sat(m1p, aq[K0 + KC]);
                    else v[KC] = usub_sat(m1p, v[KC]);
                    asm("" : "+v"(v[KC]));
                    sq[KC] = q[K0 + KC] >> 15;
                });
                sfor<0, CN>([&](auto KC) { v[KC] = v[KC] * d21 + n1s; });  // +-scaled magnitude
                sfor<0, CN>([&](auto KC) { v[KC] = v[KC] ^ sq[KC]; });
                sfor<0, CN>([&](auto KC) { v[KC] = v[KC] - sq[KC]; });    // own sign: R_new times 256
                s2 dt[CN];
                sfor<0, CN>([&](auto KC) {
                    constexpr int e = E0 + K0 + KC;
                    constexpr int j = Sx::T.ecol[e];
                    if constexpr (j >= NND) dt[KC] = ssub_sat(v[KC], rod1_256);
                    else dt[KC] = ssub_sat(v[KC], r256_get<PR::T.idx[e]>(RW));
                });
                sfor<0, CN>([&](auto KC) {
                    dt[KC] = dt[KC] >> 8;
                    if constexpr (EARLY) dt[KC] = S(U(dt[KC]) & live);
                });
                sfor<0, CN>([&](auto KC) {
                    constexpr int k = K0 + KC;
                    constexpr int e = E0 + k;
                    constexpr int j = Sx::T.ecol[e];
                    const s2 rn = v[KC];
                    if constexpr (j < NND) {
                        // store R (R_old of this edge has been read)
                        constexpr int n = PR::T.idx[e];
                        constexpr bool partner_here = PR::partner_in_row(e, k, D);
                        // rn holds R in the high byte of each half
                        if constexpr (n & 1) {
                            if constexpr (partner_here) RW[n >> 1] = __builtin_amdgcn_perm(U(rn), U(pend), 0x07030501u);
                            else RW[n >> 1] = __builtin_amdgcn_perm(U(rn), RW[n >> 1], 0x07020500u);
                        } else {
                            if constexpr (partner_here) pend = rn;
                            else RW[n >> 1] = __builtin_amdgcn_perm(RW[n >> 1], U(rn), 0x07030501u);
                        }
                        const u32 off = off2<KIND, ZC, e, j>(et);
                        const u32 lv = to_lin(dt[KC], k8000, kneg);
                        if (off < 0x10000u) lds_add(rAc + off, lv);
                        else lds_add(rAcHi + (off - 0x10000u), lv);
                    } else {
                        // the degree-1 column's only update: clamped between iterations
                        u32& w = d1w[PR::T.d1[j]];
                        const s2 an = (S(w) >> 8) + dt[KC];
                        if (last) {
                            w = U(an);
                        } else {
                            const s2 ac = smin(smax(an, sp(-128)), sp(127));
                            w = __builtin_amdgcn_perm(U(ac << 8), U(rn), 0x07030501u);  // [R, A, R, A]
                        }
                    }
                });
            });
        });
        __syncthreads();
        if (last) {
            ran_last = true;
            break;
        }
        // fold the two halves of every sub-column into the APP, clamp, rebuild the read copy
#pragma unroll 2
        for (int j = grp; j < NND; j += NG) {
            const u32 sum = ld(rAc + j * 2 * Z4) + ld(rAc + j * 2 * Z4 + Z4);
            const s2 x = smin(smax(from_lin(sum), sp(-128)), sp(127));
            const u32 w = rd_word(x);
            st(rRd + j * 2 * Z4, w);
            st(rRd + j * 2 * Z4 + Z4, w);
            st(rAc + j * 2 * Z4, to_lin(x, k8000, kneg));
            st(rAc + j * 2 * Z4 + Z4, 0u);
        }
        if constexpr (EARLY) {
            if (r == 0) st(FLAG, 0u);
            if (tid == 0) st(WGF, 0u);
            __syncthreads();
            // syndrome of the hard decisions (1 where APP <= 0): the parity of a check is the sign
            // of the XOR of (x - 1) over its edges, per half.  A wavefront stops walking rows once
            // every live codeword among its lanes has an unsatisfied check.
            const int lane = tid & 63;
            const int wb = tid - lane;
            const int pb = grp * gsz + slot * Z;  // first lane of this lane's pair in its group
            const int lo = max(pb - wb, 0), hi = min(pb + Z - wb, 64);
            // lanes of this wavefront that belong to this lane's pair (contiguous)
            const uint64_t segm = (hi >= 64 ? ~0ull : ((1ull << hi) - 1ull)) & ~((1ull << lo) - 1ull);
            u32 bad = 0u;
            bool stop = false;
            sfor<0, ROWS>([&](auto IC) {
                constexpr int i = IC;
                constexpr int D = Sx::deg(i);
                constexpr int E0 = Sx::T.row_start[i];
                if constexpr (NG > 1)
                    if (PR::T.grp[i] != grp) return;
                int mr = m_rows;
                asm volatile("" : "+s"(mr));
                if (i >= mr || stop) return;
                asm volatile("" : "+s"(et));
                u32 p = 0u;
                sfor<0, D>([&](auto KC) {
                    constexpr int e = E0 + KC;
                    constexpr int j = Sx::T.ecol[e];
                    s2 x;
                    if constexpr (j < NND) x = S(ld_at(rRd, rRdHi, off2<KIND, ZC, e, j>(et))) >> 8;
                    else x = S(d1w[PR::T.d1[j]]) >> 8;
                    p ^= U(x - sp(1));
                });
                bad |= p;
                const uint64_t fl0 = __ballot((bad & 0x8000u) != 0u) & segm;
                const uint64_t fl1 = __ballot((bad & 0x80000000u) != 0u) & segm;
                const bool ok = ((live & 0xFFFFu) == 0u || fl0 != 0u) && ((live & 0xFFFF0000u) == 0u || fl1 != 0u);
                stop = __ballot(!ok) == 0u;
            });
            // per pair: bit 0 / 1 = an unsatisfied check of codeword 2g / 2g+1 among its lanes;
            // one atomic per wavefront and pair, by the pair's first lane in the wavefront
            const uint64_t b0 = __ballot((bad & 0x8000u) != 0u) & segm;
            const uint64_t b1 = __ballot((bad & 0x80000000u) != 0u) & segm;
            const u32 f = (b0 ? 1u : 0u) | (b1 ? 2u : 0u);
            if (lane == lo && f != 0u)
                __hip_atomic_fetch_or((lds_u32*)(uintptr_t)FLAG, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __syncthreads();
            const u32 fl = ld(FLAG);
            u32 conv = 0u;
            if ((live & 0xFFFFu) && !(fl & 1u)) conv |= 0xFFFFu;
            if ((live & 0xFFFF0000u) && !(fl & 2u)) conv |= 0xFFFF0000u;
            its = (its & ~conv) | (conv & ((u32)(t + 1) | ((u32)(t + 1) << 16)));
            live &= ~conv;
            if (__ballot(live != 0u) != 0u && (tid & 63) == 0)
                __hip_atomic_fetch_or((lds_u32*)(uintptr_t)WGF, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __syncthreads();
            if (ld(WGF) == 0u) { ++t; break; }
        } else {
            __syncthreads();
        }
    }
    __syncthreads();  // every lane is done with the read copy: it becomes the staging area

    // ---- outputs: the APP is the sum of the two accumulator halves (the second is 0 after an
    // iteration whose fold ran); hard decision 1 where APP <= 0, staged as bytes in the read copy
    // (codeword 2g at STG, 2g+1 at STG + nvp) and packed 8 -> 1 (bit k of byte b = column 8b + k)
    const int nv = a.n_vn;
    const u32 nvp = ((u32)nv + 7u) & ~7u;
    const u32 STG = RD;
    float* ap0 = a.app_out ? a.app_out + (size_t)c0 * nv : nullptr;
    float* ap1 = a.app_out ? a.app_out + (size_t)c1 * nv : nullptr;
    const bool w0 = slot_live && ap0 != nullptr, w1 = hi_real && ap1 != nullptr;
    for (int j = grp; j < NND; j += NG) {
        const s2 x = from_lin(ld(rAc + j * 2 * Z4) + ld(rAc + j * 2 * Z4 + Z4));
        const int col = j * Z + r;
        stb(STG + col, x.x > 0 ? 0 : 1);
        stb(STG + nvp + col, x.y > 0 ? 0 : 1);
        if (w0) ap0[col] = (float)x.x;
        if (w1) ap1[col] = (float)x.y;
    }
    sfor<NND, Sx::COLS>([&](auto JC) {
        constexpr int j = JC;
        constexpr int e = Sx::T.d1_edge[j];
        constexpr int i = Sx::T.d1_row[j];
        if (i < m_rows && PR::T.grp[i] == grp) {
            int c;
            if constexpr (ZC > 0) c = r + edge_vm_p<KIND, ZC>(e);
            else c = r + (int)et[e].off_rd;
            if (c >= Z) c -= Z;
            const u32 w = d1w[PR::T.d1[j]];
            const s2 x = ran_last ? S(w) : S(w) >> 8;
            const int col = j * Z + c;
            stb(STG + col, x.x > 0 ? 0 : 1);
            stb(STG + nvp + col, x.y > 0 ? 0 : 1);
            if (w0) ap0[col] = (float)x.x;
            if (w1) ap1[col] = (float)x.y;
        }
    });
    // zero padding up to the next multiple of 8 columns (stepping by Z: fewer than 8 lanes per
    // codeword pair when Z < 8)
    for (u32 p = (u32)(r + grp * Z); (u32)nv + p < nvp; p += (u32)(NG * Z)) {
        stb(STG + nv + p, 0);
        stb(STG + nvp + nv + p, 0);
    }
    __syncthreads();
    const int hb = a.hd_bytes;
    for (int h = 0; h < 2; h++) {
        if (!(h == 0 ? slot_live : hi_real)) continue;
        uint8_t* hrow = a.hd + (size_t)(h == 0 ? c0 : c1) * hb;
        const u32 src = STG + (h == 0 ? 0u : nvp);
        for (int b = r + grp * Z; b < hb; b += NG * Z) {
            const u32 v0 = ld(src + 8u * b), v1 = ld(src + 8u * b + 4u);
            const u32 lo4 = (v0 | (v0 >> 7) | (v0 >> 14) | (v0 >> 21)) & 0xFu;
            const u32 hi4 = (v1 | (v1 >> 7) | (v1 >> 14) | (v1 >> 21)) & 0xFu;
            hrow[b] = (uint8_t)(lo4 | (hi4 << 4));
        }
    }
    if (r == 0 && grp == 0 && a.iters_out) {
        if (slot_live) a.iters_out[c0] = (int)(its & 0xFFFFu);
        if (hi_real) a.iters_out[c1] = (int)(its >> 16);
    }
}

template <int KIND, int ZC, bool EARLY>
__global__ void __launch_bounds__(kMaxBlock) k_decode_char_pair(DecodeArgs a) {
    decode_pair<KIND, ZC, EARLY, 1>(a);
}

// Two row groups (2 cpb Z lanes, whole wavefronts per group): at BG1 Z = 384 twelve wavefronts
// per workgroup, three on each SIMD, in place of six placed 2-2-1-1.
template <int KIND, int ZC, bool EARLY>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, kSlimMaxBlock), amdgpu_waves_per_eu(3)))
k_decode_char_split(DecodeArgs a) {
    decode_pair<KIND, ZC, EARLY, 2>(a);
}

}  // namespace pairk
}  // namespace nrldpc
