// order.hip - cost-ordered work queue of a single-scene frame.
//
// The persistent tracer runs lanes until the queue is empty; the frame then
// ends when its last-dispatched rays finish.  With the centre-first tile order
// those are the image's periphery, sky rays of 200+ DP5 attempts, so the last
// ray lifetime of the frame runs on a draining machine.  Longest-first order
// ends the frame on the shortest rays instead.  The cost of a tile is
// measured, not modelled: a pilot traces one ray per 8 x 8 tile (its centre
// pixel) with the same tracer at loose tolerances, hmax x16 and a 32-step cap
// (~0.3 ms at 1024², 4% of the frame: the pilot's duration is its cap times a
// lone wave's attempt latency); the tile's key is the largest pilot
// attempt count over its 3 x 3 tile neighbourhood (a tile that straddles a
// disc edge or the shadow holds longer rays than its centre ray shows).  A
// stable radix sort, descending, of the centre-first order by that key, and a
// scan of the tile sizes give the queue: same rays, same per-ray results
// (every ray is integrated independently of the schedule), shorter tail.
#include "context.h"

#include <hipcub/hipcub.hpp>

namespace bhx {

namespace {

// key of the i-th tile of the centre-first order: max pilot attempts over its
// 3 x 3 tile neighbourhood (clamped at the image edge), at most 16 bits
__global__ void __launch_bounds__(256) k_pilot_keys(const int32_t* __restrict__ att, const int32_t* __restrict__ cf,
                                                    int tiles_x, int tiles_y, uint32_t* __restrict__ keys,
                                                    int32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = tiles_x * tiles_y;
    if (i >= n) return;
    const int t = cf[i];
    const int tx = t % tiles_x, ty = t / tiles_x;
    int k = 0;
    for (int dy = -1; dy <= 1; ++dy) {
        const int yy = ty + dy;
        if (yy < 0 || yy >= tiles_y) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int xx = tx + dx;
            if (xx < 0 || xx >= tiles_x) continue;
            k = max(k, att[yy * tiles_x + xx]);
        }
    }
    keys[i] = static_cast<uint32_t>(min(k, 0xffff));
    vals[i] = t;
}

// samples in each tile of the sorted order; slot 0 of the start array is 0
__global__ void __launch_bounds__(256) k_tile_sizes(const int32_t* __restrict__ order, int n, int tiles_x, int W,
                                                    int H, int S2, int32_t* __restrict__ sizes,
                                                    int32_t* __restrict__ start) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) start[0] = 0;
    if (i >= n) return;
    const int t = order[i];
    const int tx = t % tiles_x, ty = t / tiles_x;
    sizes[i] = min(8, W - tx * 8) * min(8, H - ty * 8) * S2;
}

}  // namespace

hipError_t launch_pilot(const RaygenArgs& gen, const TraceArgs& args, int block, hipStream_t stream);

}  // namespace bhx

using namespace bhx;

// Fills ctx->cost_order / cost_start with the cost-ordered queue of the
// frame (tile set ctx->cur_tiles, centre-first) on stream st.  G and P are the
// frame's ray-generation and tracer arguments.
int cost_order(bhx_context* ctx, const RaygenArgs& G, const TraceArgs& P, hipStream_t st) {
    const bhx_context::TileSet& ts = *ctx->cur_tiles;
    const int n = ts.n_tiles;
    const int tiles_y = n / ts.tiles_x;
    HIPCHK(ctx->pilot_init.ensure(sizeof(double) * kInitRows * n));
    HIPCHK(ctx->pilot_q2ray.ensure(sizeof(int32_t) * n));
    HIPCHK(ctx->pilot_att.ensure(sizeof(int32_t) * n));
    HIPCHK(ctx->pilot_cls.ensure(n));
    HIPCHK(ctx->order_keys.ensure(2 * sizeof(uint32_t) * n));
    HIPCHK(ctx->order_vals.ensure(sizeof(int32_t) * n));
    HIPCHK(ctx->order_sizes.ensure(sizeof(int32_t) * n));
    HIPCHK(ctx->cost_order.ensure(sizeof(int32_t) * n));
    HIPCHK(ctx->cost_start.ensure(sizeof(int32_t) * (n + 1)));
    uint32_t* keys_in = static_cast<uint32_t*>(ctx->order_keys.p);
    uint32_t* keys_out = keys_in + n;
    int32_t* vals_in = static_cast<int32_t*>(ctx->order_vals.p);
    int32_t* order = static_cast<int32_t*>(ctx->cost_order.p);
    int32_t* start = static_cast<int32_t*>(ctx->cost_start.p);
    int32_t* sizes = static_cast<int32_t*>(ctx->order_sizes.p);
    size_t sort_bytes = 0, scan_bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, sort_bytes, keys_in, keys_out, vals_in, order, n,
                                                         0, 16, st));
    HIPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, sizes, start + 1, n, st));
    HIPCHK(ctx->order_tmp.ensure(std::max(sort_bytes, scan_bytes)));

    // pilot rays and their trace
    RaygenArgs g = G;
    g.init = static_cast<double*>(ctx->pilot_init.p);
    g.q2ray = static_cast<int32_t*>(ctx->pilot_q2ray.p);
    g.q2scene = nullptr;
    TraceArgs p = P;
    p.n_rays = n;
    p.init = g.init;
    p.q2ray = g.q2ray;
    p.counter = &ctx->d_cnt->pilot_counter;
    p.totals = ctx->d_cnt->pilot_totals;
    p.rtol = std::max(P.rtol, kPilotRtol);
    p.atol = std::max(P.atol, kPilotAtol);
    p.max_steps = std::min(P.max_steps, kPilotSteps);
    p.has_jet = 0;   // emission only: the jet changes no ray's path
    p.chunk = 64;
    p.refill = 64;
    p.cls = static_cast<uint8_t*>(ctx->pilot_cls.p);
    p.attempts = static_cast<int32_t*>(ctx->pilot_att.p);
    p.y = nullptr;
    p.jet = nullptr;
    p.steps = nullptr;
    p.norm_err = nullptr;
    p.carter = nullptr;
    // one-wave blocks: the pilot's few waves (n_tiles / 64) spread over as many
    // CUs as there are waves; measured 0.244 ms vs 0.258 ms with 4-wave blocks
    HIPCHK(launch_pilot(g, p, 64, st));

    // keys in centre-first order, stable sort descending, queue offsets
    const int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_pilot_keys, dim3(blocks), dim3(256), 0, st, static_cast<const int32_t*>(p.attempts),
                       static_cast<const int32_t*>(ts.order.p), ts.tiles_x, tiles_y, keys_in, vals_in);
    HIPCHK(hipGetLastError());
    size_t bytes = sort_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairsDescending(ctx->order_tmp.p, bytes, keys_in, keys_out, vals_in, order,
                                                         n, 0, 16, st));
    const int S = std::max(1, G.S);
    hipLaunchKernelGGL(k_tile_sizes, dim3(blocks), dim3(256), 0, st, static_cast<const int32_t*>(order), n,
                       ts.tiles_x, G.W, G.H, S * S, sizes, start);
    HIPCHK(hipGetLastError());
    bytes = scan_bytes;
    HIPCHK(hipcub::DeviceScan::InclusiveSum(ctx->order_tmp.p, bytes, sizes, start + 1, n, st));
    return BHX_OK;
}
