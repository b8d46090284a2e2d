// encoder.cpp -- encoding programs, host encoder and generator shift lists.
#include "encoder.h"

#include <algorithm>
#include <map>

#include "status.h"

namespace nrldpc {

int build_encode_program(const CodeDesc& c, EncodeProgram* p, std::string* err) {
    const int Z = c.Z;
    *p = EncodeProgram{};
    std::vector<int> coldeg(c.cols, 0);
    for (int i = 0; i < c.rows; i++)
        for (int j = 0; j < c.cols; j++) coldeg[j] += c.shift(i, j) >= 0;
    // Core rows: rows without a degree-1 (extension) parity column.
    std::vector<int> core;
    for (int i = 0; i < c.rows; i++) {
        bool ext = false;
        for (int j = c.kb; j < c.cols; j++)
            if (c.shift(i, j) >= 0 && coldeg[j] == 1) ext = true;
        if (!ext) core.push_back(i);
    }
    const int first = c.kb;
    // In the sum of the core rows every core parity column other than `first` must cancel
    // (shifts pair up) and `first` must reduce to a single circulant P^b.
    int b = -1;
    for (int j = c.kb; j < c.cols; j++) {
        std::map<int, int> mult;
        for (int i : core)
            if (c.shift(i, j) >= 0) mult[c.shift(i, j)] ^= 1;
        std::vector<int> odd;
        for (auto& kv : mult)
            if (kv.second) odd.push_back(kv.first);
        if (j == first) {
            if (odd.size() != 1) {
                if (err) *err = "core parity structure is not double-diagonal";
                return LDPC_EUNSUPPORTED;
            }
            b = odd[0];
        } else if (!odd.empty()) {
            bool in_core = false;
            for (int i : core) in_core |= c.shift(i, j) >= 0;
            if (in_core) {
                if (err) *err = "core parity columns do not cancel";
                return LDPC_EUNSUPPORTED;
            }
        }
    }
    auto add_step = [&](int target, const std::map<std::pair<int, int>, int>& terms) {
        p->step_target.push_back(target);
        p->step_start.push_back((int)p->term_col.size());
        for (auto& kv : terms)
            if (kv.second) {
                p->term_col.push_back(kv.first.first);
                p->term_shift.push_back(kv.first.second);
            }
    };
    {
        // c[first][x] = sum_{core i} sum_{info j} c[j][(x - b + V_ij) mod Z]
        std::map<std::pair<int, int>, int> terms;
        for (int i : core)
            for (int j = 0; j < c.kb; j++)
                if (c.shift(i, j) >= 0) terms[{j, ((c.shift(i, j) - b) % Z + Z) % Z}] ^= 1;
        add_step(first, terms);
    }
    std::vector<char> known(c.cols, 0);
    for (int j = 0; j <= first; j++) known[j] = 1;
    int remaining = c.cols - first - 1;
    while (remaining > 0) {
        bool progress = false;
        for (int i = 0; i < c.rows && remaining > 0; i++) {
            int unknown = -1, n_unknown = 0;
            for (int j = 0; j < c.cols; j++)
                if (c.shift(i, j) >= 0 && !known[j]) { unknown = j; n_unknown++; }
            if (n_unknown != 1) continue;
            const int vt = c.shift(i, unknown);
            std::map<std::pair<int, int>, int> terms;
            for (int j = 0; j < c.cols; j++)
                if (j != unknown && c.shift(i, j) >= 0)
                    terms[{j, ((c.shift(i, j) - vt) % Z + Z) % Z}] ^= 1;
            add_step(unknown, terms);
            known[unknown] = 1;
            remaining--;
            progress = true;
        }
        if (!progress) {
            if (err) *err = "parity columns cannot be solved by peeling";
            return LDPC_EUNSUPPORTED;
        }
    }
    p->step_start.push_back((int)p->term_col.size());
    return LDPC_OK;
}

void encode_host(const CodeDesc& c, const EncodeProgram& p, const uint8_t* info, uint8_t* cw) {
    const int Z = c.Z;
    std::fill(cw, cw + (size_t)c.n_vn(), 0);
    for (int n = 0; n < c.K(); n++) cw[n] = info[n] & 1;
    for (size_t s = 0; s + 1 < p.step_start.size(); s++) {
        uint8_t* dst = cw + (size_t)p.step_target[s] * Z;
        for (int t = p.step_start[s]; t < p.step_start[s + 1]; t++) {
            const uint8_t* src = cw + (size_t)p.term_col[t] * Z;
            const int sh = p.term_shift[t];
            for (int x = 0; x < Z; x++) {
                int y = x + sh;
                if (y >= Z) y -= Z;
                dst[x] ^= src[y];
            }
        }
    }
}

int generator_shift_lists(const CodeDesc& c, std::vector<std::vector<int>>* lists, std::string* err) {
    EncodeProgram prog;
    int st = build_encode_program(c, &prog, err);
    if (st != LDPC_OK) return st;
    const int Z = c.Z, prow = c.cols - c.kb;
    lists->assign((size_t)prow * c.kb, {});
    std::vector<uint8_t> info(c.K()), cw(c.n_vn());
    for (int j = 0; j < c.kb; j++) {
        std::fill(info.begin(), info.end(), 0);
        info[(size_t)j * Z] = 1;  // unit vector at (j, 0)
        encode_host(c, prog, info.data(), cw.data());
        for (int i = 0; i < prow; i++) {
            auto& L = (*lists)[(size_t)i * c.kb + j];
            // parity bit (i, x) depends on info bit (j, (x + s + 1) mod Z): here (x + s + 1) = 0
            for (int x = 0; x < Z; x++)
                if (cw[(size_t)(c.kb + i) * Z + x]) L.push_back(((-x - 1) % Z + Z) % Z);
            std::sort(L.begin(), L.end());
        }
    }
    return LDPC_OK;
}

}  // namespace nrldpc
