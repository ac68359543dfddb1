// Host-side plumbing shared by the libnfx translation units: error state, context, launch checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/nfx.h"

namespace nfx {

void set_error(const char* fmt, ...);

#define NFX_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            ::nfx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                     \
            return NFX_ERR_HIP;                                                             \
        }                                                                                   \
    } while (0)

#define NFX_CHECK_LAUNCH()                                                                   \
    do {                                                                                     \
        hipError_t e_ = hipGetLastError();                                                   \
        if (e_ != hipSuccess) {                                                              \
            ::nfx::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                      \
            return NFX_ERR_HIP;                                                              \
        }                                                                                    \
    } while (0)

#define NFX_REQUIRE(cond, ...)            \
    do {                                  \
        if (!(cond)) {                    \
            ::nfx::set_error(__VA_ARGS__); \
            return NFX_ERR_ARG;           \
        }                                 \
    } while (0)

// reduction scratch: per-block FP64 partials and a completion counter, allocated once per context
constexpr int kMaxReduceBlocks = 2048;
constexpr int kGraphBlock = 100;  // steps per captured graph
constexpr int kMaxGraphs = 32;    // captured graphs a context keeps

}  // namespace nfx

struct nfx_ctx {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double* partials = nullptr;       // kMaxReduceBlocks * 2 doubles
    unsigned int* counter = nullptr;  // 1 counter (reset by the last block of each reduction)
    int64_t* step_dev = nullptr;      // device step counter for graph replays
    double* hist_dev = nullptr;       // kGraphBlock doubles, tip force of the replayed block
    double* tip_rows = nullptr;       // node-fused steps: per tile and step of a block, the indenter force
    int64_t tip_rows_cap = 0;         // doubles allocated at tip_rows
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // batched stepping: the batch being enqueued (NULL otherwise)
    const nfx_batch* batch = nullptr;
    // ... and, while its force pass is enqueued with the step's curve sample riding in it, that step
    bool curve_fused = false;
    int64_t curve_step = 0;
    const int64_t* curve_step_dev = nullptr;
    int32_t curve_off = 0;
    // captured blocks of kGraphBlock steps, one per (plan, batch, history mode), least recently
    // used evicted past kMaxGraphs
    struct GraphEntry {
        nfx_plan plan;
        nfx_batch batch;
        bool is_batch, hist;
        hipGraph_t graph;
        hipGraphExec_t exec;
        uint64_t last_use;
    };
    std::vector<GraphEntry> graphs;
    uint64_t graph_clock = 0;
    int64_t captures = 0;  // graphs captured so far (nfx_graph_captures)
};
