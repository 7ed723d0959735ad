// The context behind the opaque darcy_ctx handle of the C ABI: one device,
// one stream, timing events, the last error and the growable workspace.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/darcy_hip.h"

namespace darcy_hip {

// A growable device allocation owned by the context.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        const hipError_t e = hipMalloc(&ptr, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

}  // namespace darcy_hip

struct darcy_ctx {
    int device = 0;
    int cus = 0;
    std::string arch;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    std::string error;
    std::mutex mu;
    darcy_hip::DevBuf slots, probs, stats, order, counter, staging, grid, fields;
    // Single solves with n > grid_min_n run on the grid path, with
    // levels of side <= grid_tail_side in one workgroup (tuned values);
    // grid_graph != 0 replays the cycle from a captured HIP graph.
    int grid_min_n = 64;
    int grid_tail_side = 64;
    int grid_graph = 1;
    // Workgroups per CU of the streaming CG kernels (rz, update) on the grid path.
    int grid_cg_per_cu = 4;
    // Cells per row-pipeline step of the persistent batch kernel's global
    // levels (tuned value).
    int batch_step_cells = 1024;       // FP64 cycle
    int batch_step_cells_fp32 = 4096;  // FP32 cycle
    // Grid path, FP64 cycle: 1 (default) applies the levels of side <= 8 as
    // one dense FP64 matrix per visit (the recursive cycle's linear map,
    // formed in FP64 from unit vectors); 0 runs the recursion itself, which
    // reproduces the reference's cycle bit for bit.
    int grid_fp64_dense_tail = 1;
    // Largest side the grid path's tail applies as one dense matrix: 8 (the
    // levels of side <= 8, kept in LDS) or 16 (the levels from the first of
    // side <= 16, read from L2).  The FP64 cycle uses either only with
    // grid_fp64_dense_tail.
    int grid_dense_side = 16;
    // Persistent batch kernel: a batch whose largest side is >= this runs on
    // 1024-thread workgroups, one per CU (0: always 256-thread workgroups,
    // four per CU).  Same bits either way.
    int batch_wide_side = 384;
    // Device address and size of the library's static scratch array (the
    // allocation-free entry point), looked up when the context is created.
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    hipError_t scratch_status = hipSuccess;
    // The last captured cycle (+ the rz kernel) of the grid path and what it
    // was captured for: the workspace address, n, precision, options and
    // tail level.  Reused while all of them match.
    hipGraphExec_t cycle_exec = nullptr;
    hipGraph_t cycle_graph = nullptr;
    const void* cycle_base = nullptr;
    int cycle_key[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool graph_checked = false;  // launch limits of the captured kernels verified
};
