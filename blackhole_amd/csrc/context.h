// context.h - internal: the context object behind the C ABI and the kernel
// launchers, shared by api.hip (single frames) and batch.hip (batches).
#pragma once

#include "../../include/bhx.h"
#include "host_setup.h"
#include "params.h"

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <tuple>

namespace bhx {
hipError_t launch_raygen(const RaygenArgs& args, hipStream_t stream);
// Ray generation / shading of many scenes in one launch: block b works for
// scene block_scene[b], as that scene's block b - block_first[scene].
hipError_t launch_raygen_batch(const RaygenArgs* table, const int32_t* block_scene, const int32_t* block_first,
                               int blocks, hipStream_t stream);
hipError_t launch_shade_batch(const ShadeArgs* table, const int32_t* block_scene, const int32_t* block_first,
                              int blocks, hipStream_t s);
hipError_t launch_trace(const TraceArgs& args, int grid, int block, hipStream_t stream);
// Tracer blocks resident per CU, for the lean kernel (scenes without jets or
// the Sun, three waves per SIMD) or the full one (two).
hipError_t trace_occupancy(int block, bool lean, int* blocks_per_cu);
bool trace_is_lean(const TraceArgs& args);
hipError_t launch_pilot(const RaygenArgs& gen, const TraceArgs& args, int block, hipStream_t stream);
hipError_t launch_trace_batch(const TraceArgs& args, int grid, int block, hipStream_t stream);
hipError_t trace_batch_occupancy(int block, int n_scenes, bool lean, int* blocks_per_cu);
hipError_t launch_shade(const ShadeArgs& args, int blocks, hipStream_t s);
hipError_t launch_develop(const double* xyz, int64_t n, double exposure, int tm, double log_k,
                          uint8_t* rgb, hipStream_t s);
hipError_t launch_develop_auto(const double* xyz, int64_t n, const ExposureState* state, double key,
                               double* e_out, int tm, double log_k, uint8_t* rgb, hipStream_t s);
hipError_t launch_exposure(const double* xyz, int64_t n, ExposureState* state, unsigned int* hist,
                           hipStream_t s);
hipError_t launch_bb(const double* tab, const double* T, int64_t n, double* out, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s);
hipError_t launch_zero_u32(unsigned int* p, int n, hipStream_t s);
}  // namespace bhx

// Small per-context device block: counters of one trace + one shade.
struct Counters {
    unsigned long long totals[4];       // steps, attempts, max norm bits, max Carter bits
    unsigned long long classes[8];      // 6 classes + jet rays
    unsigned int tile_counter;
    unsigned int pilot_counter;          // queue of the cost-order pilot
    unsigned long long pilot_totals[4];  // the pilot's totals (not reported)
};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct bhx_context {
    int device = 0;
    int n_cu = 0;
    std::string err;
    bhx_launch launch{256, 0, 32, 1, BHX_ORDER_COST};   // chunk 32, refill 1: fastest of a sweep on the 1024² frame
    // tables
    double* d_bb = nullptr;
    // catalogue (cell CSR + star rows)
    bool custom_cat = false;
    bhx::host::Catalogue custom;
    bool cat_valid = false;
    int cat_count = -1;
    uint64_t cat_seed = 0;
    bool cat_is_custom = false;
    DevBuf cell_start, star_rows;
    // queue order cache, one entry per frame shape (W, H, S): tiles
    // centre-first and the first queue slot of each tile.  Entries are never
    // rewritten while they may be read by a launch in flight.
    struct TileSet {
        DevBuf order, start;
        int tiles_x = 0, n_tiles = 0;
    };
    std::map<std::tuple<int, int, int, bool, double>, TileSet> tiles;   // (W, H, S, axis band, roll)
    const TileSet* cur_tiles = nullptr;
    // cost-ordered queue of the last frame: pilot rays (init rows, slot map,
    // attempts, classes), sort keys/values, tile sizes, the sorted tile order
    // and its queue offsets, and the sort/scan workspace
    DevBuf pilot_init, pilot_q2ray, pilot_att, pilot_cls, order_keys, order_vals, order_sizes;
    DevBuf cost_order, cost_start, order_tmp;
    // PNG / APNG writer: host images uploaded, filtered scanlines, hash keys
    // and positions (in and out of the sort), chain links, matches (lengths,
    // distances), sort workspace, changed-pixel box, cropped sub-frame
    DevBuf png_in, png_raw, png_key, png_pos, png_prev, png_match, png_tmp, png_box, png_sub;
    // initial ray states in queue order (kInitRows x n) and slot -> ray map
    DevBuf init, q2ray;
    // jet / Sun parameters of the tracer's event branches (last uploaded copy)
    DevBuf evp;
    bhx::EventParams ev_host{};
    bool ev_valid = false;
    // counters
    Counters* d_cnt = nullptr;
    // The cost-ordered queue is built on a stream of its own (non-blocking,
    // default priority): a frame's pilot and sort only wait until the last
    // frame's ray generation has read the previous order (order_free), so they
    // run while that frame still traces and shades.
    hipStream_t order_stream = nullptr;
    hipEvent_t order_free = nullptr;
    DevBuf flux_partial, exp_hist, exp_state;
    // frame buffers of bhx_render
    DevBuf r_cls, r_y, r_jet, r_xyz;
    // veiling glare: row pass, the two blurred images, the taps
    DevBuf glare_tmp, glare_near, glare_far, glare_taps;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

// message of the calling thread's last failed bhx_trace_noalloc_dev (noalloc.hip)
const char* bhx_noalloc_last_error();

// cost-ordered queue of the current frame into ctx->cost_order / cost_start (order.hip)
int cost_order(bhx_context* ctx, const bhx::RaygenArgs& G, const bhx::TraceArgs& P, hipStream_t st);

inline int bhx_fail(bhx_context* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIPCHK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return bhx_fail(ctx, BHX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
