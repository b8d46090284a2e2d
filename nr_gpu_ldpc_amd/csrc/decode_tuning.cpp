// decode_tuning.cpp -- the launch-shape table compiled into the library (decode_tuning.inc,
// written by the autotuner: nr_gpu_ldpc_amd/tune.py --decode).
#include "decoder.h"

namespace nrldpc {

static const LayoutTuning kLayoutTable[] = {
#include "decode_tuning.inc"
};

const LayoutTuning* builtin_layout_tuning(int kind, int Z) {
    for (const LayoutTuning& t : kLayoutTable)
        if (t.kind == kind && t.Z == Z) return &t;
    return nullptr;
}

}  // namespace nrldpc
