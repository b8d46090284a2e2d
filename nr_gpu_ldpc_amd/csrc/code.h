// code.h -- quasi-cyclic LDPC code descriptions (5G NR BG1/BG2 and the 12x24 example code).
//
// Two views of the same data:
//  * Structure<Kind>: compile-time sparsity pattern (which base-graph positions are nonzero),
//    identical for every lifting set of a base graph.  The decoder kernels unroll over it.
//  * CodeDesc: run-time description for one (kind, Z, kept parity rows): shift values
//    V mod Z (TS 38.212 5.3.2) and the derived sizes.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace nrldpc {

struct nrldpc_bg_entry { int row, col, v[8]; };
struct nrldpc_qc_entry { int row, col, v; };
#define NRLDPC_TABLE_QUAL static constexpr
#include "bg_tables.inc"
#undef NRLDPC_TABLE_QUAL

enum CodeKind : int { kBG1 = 1, kBG2 = 2, kWifi = 3 };

// ------------------------------------------------------------------ compile-time structure

template <int KIND> struct Dims;
template <> struct Dims<kBG1> { static constexpr int ROWS = 46, COLS = 68, KB = 22, NNZ = NRLDPC_BG1_NNZ; };
template <> struct Dims<kBG2> { static constexpr int ROWS = 42, COLS = 52, KB = 10, NNZ = NRLDPC_BG2_NNZ; };
template <> struct Dims<kWifi> { static constexpr int ROWS = 12, COLS = 24, KB = 12, NNZ = NRLDPC_WIFI_NNZ; };

template <int KIND> constexpr int entry_row(int n) {
    if constexpr (KIND == kBG1) return kBG1Entries[n].row;
    else if constexpr (KIND == kBG2) return kBG2Entries[n].row;
    else return kWifiEntries[n].row;
}
template <int KIND> constexpr int entry_col(int n) {
    if constexpr (KIND == kBG1) return kBG1Entries[n].col;
    else if constexpr (KIND == kBG2) return kBG2Entries[n].col;
    else return kWifiEntries[n].col;
}

// Sparsity-derived tables.  Edges are numbered in row-major order (rows ascending, columns
// ascending inside a row), which is the order of the reference's compact row list
// h_compact1 (src/setup_exp.cu:241-250).
template <int KIND> struct Structure {
    using D = Dims<KIND>;
    static constexpr int ROWS = D::ROWS, COLS = D::COLS, KB = D::KB, NNZ = D::NNZ;
    struct Tables {
        int row_start[ROWS + 1];
        int ecol[NNZ];
        int coldeg[COLS];
        int nd_index[COLS];    // index among columns kept in LDS (degree > 1), else -1
        int d1_slot[COLS];     // register slot of a degree-1 column, else -1
        int d1_edge[COLS];     // the single edge of a degree-1 column, else -1
        int d1_row[COLS];      // its block row, else -1
        int epoch_start[ROWS]; // 1 when row i must wait for all previous rows' VN updates
        int w_word[ROWS];      // register holding row i's sign bits and arg-min index
        int w_off[ROWS];       // bit offset of row i inside that register (0 or 16)
        int n_w;               // registers used for all rows
        int max_deg;
        int n_nd, n_d1;
    };
    static constexpr Tables make() {
        Tables t{};
        for (int i = 0; i <= ROWS; i++) t.row_start[i] = 0;
        for (int n = 0; n < NNZ; n++) t.row_start[entry_row<KIND>(n) + 1]++;
        for (int i = 0; i < ROWS; i++) t.row_start[i + 1] += t.row_start[i];
        for (int n = 0; n < NNZ; n++) t.ecol[n] = entry_col<KIND>(n);
        for (int j = 0; j < COLS; j++) t.coldeg[j] = 0;
        for (int n = 0; n < NNZ; n++) t.coldeg[t.ecol[n]]++;
        t.n_nd = 0;
        t.n_d1 = 0;
        for (int j = 0; j < COLS; j++) {
            t.nd_index[j] = -1;
            t.d1_slot[j] = -1;
            t.d1_edge[j] = -1;
            t.d1_row[j] = -1;
            if (t.coldeg[j] > 1) {
                t.nd_index[j] = t.n_nd++;
            } else {
                t.d1_slot[j] = t.n_d1++;
                for (int n = 0; n < NNZ; n++)
                    if (t.ecol[n] == j) { t.d1_edge[j] = n; t.d1_row[j] = entry_row<KIND>(n); }
            }
        }
        t.max_deg = 0;
        for (int i = 0; i < ROWS; i++) {
            const int d = t.row_start[i + 1] - t.row_start[i];
            if (d > t.max_deg) t.max_deg = d;
        }
        // Sign bits (one per edge) and the 5-bit arg-min index of a row of degree <= 11 fit in
        // 16 bits (signs at bits 0..10, index at 11..15): two such rows share one register.
        // Wider rows use a full register (signs at 0..d-1, index at 24..28).
        t.n_w = 0;
        int half_open = -1;
        for (int i = 0; i < ROWS; i++) {
            const int d = t.row_start[i + 1] - t.row_start[i];
            if (d <= 11) {
                if (half_open >= 0) { t.w_word[i] = half_open; t.w_off[i] = 16; half_open = -1; }
                else { t.w_word[i] = t.n_w++; t.w_off[i] = 0; half_open = t.w_word[i]; }
            } else {
                t.w_word[i] = t.n_w++;
                t.w_off[i] = 0;
            }
        }
        // Rows that share no LDS-resident column with the rows since the last epoch start can
        // update variable nodes without waiting: per-column update order stays ascending.
        bool used[COLS] = {};
        for (int i = 0; i < ROWS; i++) {
            bool clash = false;
            for (int n = t.row_start[i]; n < t.row_start[i + 1]; n++)
                if (t.nd_index[t.ecol[n]] >= 0 && used[t.ecol[n]]) clash = true;
            t.epoch_start[i] = (i == 0 || clash) ? 1 : 0;
            if (clash)
                for (int j = 0; j < COLS; j++) used[j] = false;
            for (int n = t.row_start[i]; n < t.row_start[i + 1]; n++)
                if (t.nd_index[t.ecol[n]] >= 0) used[t.ecol[n]] = true;
        }
        return t;
    }
    static constexpr Tables T = make();
    static constexpr int deg(int i) { return T.row_start[i + 1] - T.row_start[i]; }
    static constexpr int NND = T.n_nd;
    static constexpr int ND1 = T.n_d1;
    static constexpr int MAXDEG = T.max_deg;
    static constexpr int NW = T.n_w;
    static constexpr bool half(int i) { return deg(i) <= 11; }
    // edges of row i whose column is LDS-resident (degree > 1)
    static constexpr int row_lds_edges(int i) {
        int n = 0;
        for (int e = T.row_start[i]; e < T.row_start[i + 1]; e++) n += T.nd_index[T.ecol[e]] >= 0;
        return n;
    }
};

// Every LDS-resident column lies in the information + core part, which parity-row truncation
// never removes; degree-1 columns are the extension parity columns.
template <int KIND> constexpr bool nd_columns_in_core() {
    using S = Structure<KIND>;
    for (int j = 0; j < S::COLS; j++)
        if (S::T.nd_index[j] >= 0 && KIND != kWifi && j >= S::KB + 4) return false;
    return true;
}
static_assert(nd_columns_in_core<kBG1>() && nd_columns_in_core<kBG2>(), "unexpected base graph");
static_assert(Structure<kBG1>::NND == 26 && Structure<kBG1>::ND1 == 42, "BG1 structure");
static_assert(Structure<kBG2>::NND == 14 && Structure<kBG2>::ND1 == 38, "BG2 structure");
static_assert(Structure<kWifi>::NND == 24 && Structure<kWifi>::ND1 == 0, "12x24 structure");
static_assert(Structure<kBG1>::NW == 25 && Structure<kBG2>::NW == 21 && Structure<kWifi>::NW == 6,
              "packed sign words");

// ------------------------------------------------------------------ run-time description

int lifting_set_index(int Z);  // TS 38.212 Table 5.3.2-1; -1 when Z is not a lifting size

struct CodeDesc {
    int kind = 0;
    int Z = 0;
    int ils = 0;        // lifting-set index (0 for the 12x24 code)
    int rows = 0;       // block rows in use (parity-row truncation keeps the first rows)
    int cols = 0;       // block columns in use
    int rows_full = 0, cols_full = 0;
    int kb = 0;         // information block columns
    int punct = 0;      // punctured leading block columns (2 for NR, 0 for the 12x24 code)
    std::vector<int> V; // rows_full x cols_full, V mod Z or -1

    int n_vn() const { return cols * Z; }            // decoder columns incl. punctured ones
    int hd_bytes() const { return (n_vn() + 7) / 8; } // packed hard-decision bytes
    int K() const { return kb * Z; }                 // information bits incl. punctured ones
    int N() const { return (cols - punct) * Z; }     // transmitted bits
    int M() const { return rows * Z; }
    int shift(int i, int j) const { return V[i * cols_full + j]; }
    int nnz() const;                                 // block edges in use
};

// Returns 0 on success, a negative status on an invalid (kind, Z, rows) combination.
int make_code(int kind, int Z, int m_rows, CodeDesc* out, std::string* err);
// Raw table value (not reduced mod Z) for (kind, ils, row, col), -1 for a zero block.
int base_value(int kind, int ils, int row, int col);

}  // namespace nrldpc
