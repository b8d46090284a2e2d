// mixed.hip -- host side of the persistent mixed-batch decoder (launch, graph node, occupancy;
// the kernels are instantiated in mixed_k_{dup,slim}_{plain,learned}.hip from mixed_kernel.h)
// and the narrow-input conversion kernel.
#include "status.h"
#include "decoder_kernel.h"

namespace nrldpc {

const void* mixed_kernel_dup_plain();
const void* mixed_kernel_dup_learned();
const void* mixed_kernel_slim_plain();
const void* mixed_kernel_slim_learned();

static const void* mixed_kernel(bool learned, bool slim) {
    if (slim) return learned ? mixed_kernel_slim_learned() : mixed_kernel_slim_plain();
    return learned ? mixed_kernel_dup_learned() : mixed_kernel_dup_plain();
}

static bool mixed_shape_ok(bool slim, const MixedArgs& a, int grid, int threads, size_t lds) {
    if (threads > (slim ? kSlimMaxBlock : kMaxBlock) || threads < 64 || lds > (size_t)max_lds_bytes() || grid < 1 ||
        (slim && (!a.d1_scratch || a.d1_stride < kMixedSlimD1Stride))) {
        set_last_error("mixed launch exceeds the workgroup size or LDS limit");
        return false;
    }
    return true;
}

int mixed_launch(bool learned, bool slim, const MixedArgs& a, int grid, int threads, size_t lds, hipStream_t s) {
    if (!mixed_shape_ok(slim, a, grid, threads, lds)) return LDPC_EINVAL;
    MixedArgs args = a;
    void* params[] = {&args};
    NRLDPC_HIP(hipLaunchKernel(mixed_kernel(learned, slim), dim3(grid), dim3(threads), params, lds, s));
    return LDPC_OK;
}

int mixed_graph_node(bool learned, bool slim, const MixedArgs& a, int grid, int threads, size_t lds, hipGraph_t g,
                     const hipGraphNode_t* deps, size_t n_deps, hipGraphNode_t* node) {
    if (!mixed_shape_ok(slim, a, grid, threads, lds)) return LDPC_EINVAL;
    MixedArgs args = a;  // the node keeps its own copy of the argument values
    void* params[] = {&args};
    hipKernelNodeParams kp{};
    kp.func = const_cast<void*>(mixed_kernel(learned, slim));
    kp.gridDim = dim3(grid);
    kp.blockDim = dim3(threads);
    kp.sharedMemBytes = (unsigned)lds;
    kp.kernelParams = params;
    kp.extra = nullptr;
    NRLDPC_HIP(hipGraphAddKernelNode(node, g, deps, n_deps, &kp));
    return LDPC_OK;
}

__global__ void __launch_bounds__(256) k_llr_convert(LlrConvertArgs a) {
    const long long n = *a.count;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float v;
        if (a.fmt == kLlrI8) v = (float)static_cast<const int8_t*>(a.in)[i] * a.scale;
        else v = half_bits_to_float(static_cast<const uint16_t*>(a.in)[i]);
        a.out[i] = v;
    }
}

constexpr int kConvertBlocks = 2048;

int llr_convert_launch(const LlrConvertArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_llr_convert, dim3(kConvertBlocks), dim3(256), 0, s, a);
    NRLDPC_HIP(hipGetLastError());
    return LDPC_OK;
}

int llr_convert_graph_node(const LlrConvertArgs& a, hipGraph_t g, const hipGraphNode_t* deps, size_t n_deps,
                           hipGraphNode_t* node) {
    LlrConvertArgs args = a;
    void* params[] = {&args};
    hipKernelNodeParams kp{};
    kp.func = reinterpret_cast<void*>(k_llr_convert);
    kp.gridDim = dim3(kConvertBlocks);
    kp.blockDim = dim3(256);
    kp.sharedMemBytes = 0;
    kp.kernelParams = params;
    kp.extra = nullptr;
    NRLDPC_HIP(hipGraphAddKernelNode(node, g, deps, n_deps, &kp));
    return LDPC_OK;
}

int mixed_occupancy(bool learned, bool slim, int threads, size_t lds) {
    int n = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mixed_kernel(learned, slim), threads, lds) == hipSuccess
               ? n : 0;
}

}  // namespace nrldpc
