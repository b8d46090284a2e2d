// code.cpp -- run-time code descriptions.
#include "code.h"

#include "status.h"

namespace nrldpc {

int lifting_set_index(int Z) {
    // Table 5.3.2-1: the eight sets {a * 2^j <= 384} for a in {2,3,5,7,9,11,13,15}.
    static const int base[8] = {2, 3, 5, 7, 9, 11, 13, 15};
    for (int i = 0; i < 8; i++)
        for (int z = base[i]; z <= 384; z *= 2)
            if (z == Z) return i;
    return -1;
}

int base_value(int kind, int ils, int row, int col) {
    if (kind == kBG1) {
        for (const auto& e : kBG1Entries)
            if (e.row == row && e.col == col) return e.v[ils];
    } else if (kind == kBG2) {
        for (const auto& e : kBG2Entries)
            if (e.row == row && e.col == col) return e.v[ils];
    } else if (kind == kWifi) {
        for (const auto& e : kWifiEntries)
            if (e.row == row && e.col == col) return e.v;
    }
    return -1;
}

int CodeDesc::nnz() const {
    int n = 0;
    for (int i = 0; i < rows; i++)
        for (int j = 0; j < cols; j++) n += shift(i, j) >= 0;
    return n;
}

int make_code(int kind, int Z, int m_rows, CodeDesc* c, std::string* err) {
    auto fail = [&](const std::string& m) {
        if (err) *err = m;
        return LDPC_EINVAL;
    };
    *c = CodeDesc{};
    c->kind = kind;
    c->Z = Z;
    if (kind == kBG1) {
        c->rows_full = 46; c->cols_full = 68; c->kb = 22; c->punct = 2;
    } else if (kind == kBG2) {
        c->rows_full = 42; c->cols_full = 52; c->kb = 10; c->punct = 2;
    } else if (kind == kWifi) {
        c->rows_full = 12; c->cols_full = 24; c->kb = 12; c->punct = 0;
    } else {
        return fail("unknown code kind " + std::to_string(kind));
    }
    if (Z < 1 || Z > 384) return fail("lifting size Z must be in 1..384");
    if (kind == kWifi) {
        c->ils = 0;
        c->rows = 12;
        if (m_rows > 0 && m_rows != 12) return fail("the 12x24 code has no parity-row truncation");
        c->cols = 24;
    } else {
        c->ils = lifting_set_index(Z);
        if (c->ils < 0) return fail("Z=" + std::to_string(Z) + " is not a TS 38.212 lifting size");
        c->rows = m_rows <= 0 ? c->rows_full : m_rows;
        if (c->rows < 4 || c->rows > c->rows_full)
            return fail("kept parity rows must be in 4.." + std::to_string(c->rows_full));
        c->cols = c->kb + c->rows;
    }
    c->V.assign(c->rows_full * c->cols_full, -1);
    auto put = [&](int r, int col, int v) { c->V[r * c->cols_full + col] = v % Z; };
    if (kind == kBG1)
        for (const auto& e : kBG1Entries) put(e.row, e.col, e.v[c->ils]);
    else if (kind == kBG2)
        for (const auto& e : kBG2Entries) put(e.row, e.col, e.v[c->ils]);
    else
        for (const auto& e : kWifiEntries) put(e.row, e.col, e.v);
    return LDPC_OK;
}

}  // namespace nrldpc
