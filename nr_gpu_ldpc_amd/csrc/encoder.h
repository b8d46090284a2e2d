// encoder.h -- systematic QC-LDPC encoding from the parity-check matrix.
//
// Encoding is compiled once per code into an "encoding program": an ordered list of steps,
// each producing one parity block column as an XOR of circularly shifted known block columns:
//     c[target][x] = XOR over terms (j, s) of c[j][(x + s) mod Z].
// The first step solves the core (double-diagonal) part by summing the core rows, in which
// every core parity column except the first cancels; every later step peels a row that has
// exactly one unknown column.  The CPU encoder and the device encoder run the same program.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "code.h"

namespace nrldpc {

struct EncodeProgram {
    std::vector<int> step_target;  // block column produced by step s
    std::vector<int> step_start;   // terms of step s: [step_start[s], step_start[s+1])
    std::vector<int> term_col;     // source block column
    std::vector<int> term_shift;   // circular shift (mod Z)
};

int build_encode_program(const CodeDesc& c, EncodeProgram* p, std::string* err);
// info: K bits (one byte each); cw: n_vn bits (one byte each), info bits first.
void encode_host(const CodeDesc& c, const EncodeProgram& p, const uint8_t* info, uint8_t* cw);

// Generator shift lists in the layout of the OAI tables used by the reference encoder
// (src/nr_ldpc.cpp:864-891): for parity block row i and information block column j, the
// shifts s such that parity bit (i, x) includes information bit (j, (x + s + 1) mod Z) after
// the reference's per-step rotation.  Returns the lists per (i, j), rows * kb entries.
int generator_shift_lists(const CodeDesc& c, std::vector<std::vector<int>>* lists, std::string* err);

}  // namespace nrldpc
