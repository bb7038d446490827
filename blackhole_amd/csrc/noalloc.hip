// noalloc.hip - bhx_trace_noalloc_dev: the tracer stage with no device
// allocation, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// Everything bhx_trace_dev keeps in its context - the initial ray states, the
// queue-slot -> ray map, the tile order and queue offsets, the sort
// workspace, the counters and the event parameters - lives here in one
// file-scope device array, `scratch`, of kScratchBytes (about 140 MiB, held on
// every GPU the library runs on until the process ends).  Frames are limited
// to kMaxRays samples and kMaxTiles 8 x 8 tiles.  The kernels are the
// tracer's own (k_raygen, k_trace); the queue is the centre-first tile order,
// built on the device (tile keys, a stable radix sort, a scan of the tile
// sizes) rather than on the host.  Every ray is integrated independently of
// the schedule, so the per-ray results are those of bhx_trace_dev, bit for bit.
//
// In host code the name `scratch` is only a host-side stand-in: its device
// address on each GPU is looked up once per device (hipGetSymbolAddress, on the
// first call with that device current) and only that address is ever handed to
// kernels.  Calls that overlap on one GPU share the array and race.
#include "context.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>

using namespace bhx;

namespace {

constexpr int64_t kMaxRays = int64_t(1) << 20;   // one 1024 x 1024 frame at one sample per pixel
constexpr int kMaxTiles = 1 << 15;               // 8 x 8-pixel tiles per frame
constexpr size_t kSortTempBytes = size_t(4) << 20;
constexpr int kBlock = 256, kChunk = 32, kRefill = 1;   // bhx_trace_dev's default launch shape
constexpr int kMaxDevices = 64;

constexpr size_t align256(size_t n) { return (n + 255) & ~size_t(255); }
// layout of the scratch array (byte offsets)
constexpr size_t kOffInit = 0;                                                          // kInitRows x n doubles
constexpr size_t kOffQ2ray = kOffInit + align256(sizeof(double) * kInitRows * kMaxRays);  // n int32
constexpr size_t kOffKeys = kOffQ2ray + align256(sizeof(int32_t) * kMaxRays);           // 2 x tiles u64
constexpr size_t kOffVals = kOffKeys + align256(2 * sizeof(uint64_t) * kMaxTiles);      // 2 x tiles int32
constexpr size_t kOffSizes = kOffVals + align256(2 * sizeof(int32_t) * kMaxTiles);      // tiles int32
constexpr size_t kOffStart = kOffSizes + align256(sizeof(int32_t) * kMaxTiles);         // tiles + 1 int32
constexpr size_t kOffTemp = kOffStart + align256(sizeof(int32_t) * (kMaxTiles + 1));    // sort / scan workspace
constexpr size_t kOffCnt = kOffTemp + align256(kSortTempBytes);                         // Counters
constexpr size_t kOffEv = kOffCnt + align256(sizeof(Counters));                         // EventParams
constexpr size_t kScratchBytes = kOffEv + align256(sizeof(EventParams));

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Per-device set-up: the device address of `scratch` and the launch grid.
struct DeviceState {
    std::once_flag once;
    int rc = BHX_E_HIP;
    unsigned char* base = nullptr;   // device address of scratch on this GPU
    int grid = 0;        // persistent grid of the full tracer kernel
    int grid_lean = 0;   // and of the lean one (scenes without jets or the Sun)
};
DeviceState g_device[kMaxDevices];

thread_local char g_err[256];

int fail(int code, const char* msg) {
    std::snprintf(g_err, sizeof g_err, "bhx_trace_noalloc_dev: %s", msg);
    return code;
}

void set_up(int device, DeviceState& d) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return;
    int n_cu = 0, per_cu = 0, per_cu_lean = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return;
    if (trace_occupancy(kBlock, false, &per_cu) != hipSuccess || per_cu <= 0 || n_cu <= 0) return;
    if (trace_occupancy(kBlock, true, &per_cu_lean) != hipSuccess || per_cu_lean <= 0) return;
    d.base = static_cast<unsigned char*>(p);
    d.grid = per_cu * n_cu;
    d.grid_lean = per_cu_lean * n_cu;
    d.rc = BHX_OK;
}

// Centre-first key of every 8 x 8 tile: squared distance of its centre from
// the image centre, or from the spin axis' image when that is nearer (frames
// with the conical jet); the non-negative double's bits order like the value.
__global__ void __launch_bounds__(256) k_centre_keys(int W, int H, int tiles_x, int n, int band, double cr, double sr,
                                                     uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double cx = (i % tiles_x) * 8 + 4 - 0.5 * W, cy = (i / tiles_x) * 8 + 4 - 0.5 * H;
    double key = cx * cx + cy * cy;
    if (band) {
        const double perp = cr * cx + sr * cy;
        key = fmin(key, perp * perp);
    }
    keys[i] = static_cast<uint64_t>(__double_as_longlong(key));
    vals[i] = i;
}

// Samples in each tile of the queue order; start[0] = 0 (the scan fills the rest).
__global__ void __launch_bounds__(256) k_queue_sizes(const int32_t* __restrict__ order, int n, int tiles_x, int W, int H,
                                                     int S2, int32_t* __restrict__ sizes, int32_t* __restrict__ start) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) start[0] = 0;
    if (i >= n) return;
    const int t = order[i];
    const int tx = t % tiles_x, ty = t / tiles_x;
    sizes[i] = min(8, W - tx * 8) * min(8, H - ty * 8) * S2;
}

// The scene's jet / Sun parameters, passed by value, into scratch.
__global__ void k_store_events(const EventParams ev, EventParams* dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = ev;
}

}  // namespace

// The message of the calling thread's last failed bhx_trace_noalloc_dev ("" if none).
const char* bhx_noalloc_last_error() { return g_err; }

extern "C" int bhx_trace_noalloc_dev(int32_t device, const bhx_scene* s, const bhx_rays* rays, bhx_stats* stats,
                                     void* stream) {
    g_err[0] = '\0';
    if (!s) return fail(BHX_E_INVALID, "scene is NULL");
    if (!rays) return fail(BHX_E_INVALID, "rays is NULL");
    if (s->width <= 0 || s->height <= 0) return fail(BHX_E_INVALID, "width/height must be > 0");
    if (device < 0 || device >= kMaxDevices) return fail(BHX_E_INVALID, "device index out of range");
    const int S = std::max(1, s->sqrt_spp);
    const double n_samples = static_cast<double>(s->width) * s->height * S * S;
    if (n_samples > static_cast<double>(kMaxRays))
        return fail(BHX_E_TOO_LARGE, "frame exceeds 1048576 samples (width * height * sqrt_spp^2)");
    const int tiles_x = (s->width + 7) / 8, tiles_y = (s->height + 7) / 8;
    if (static_cast<int64_t>(tiles_x) * tiles_y > kMaxTiles)
        return fail(BHX_E_TOO_LARGE, "frame exceeds 32768 tiles of 8 x 8 pixels");
    const int n_tiles = tiles_x * tiles_y;
    const int32_t n_rays = static_cast<int32_t>(n_samples);

    if (hipSetDevice(device) != hipSuccess) return fail(BHX_E_HIP, "hipSetDevice failed");
    DeviceState& d = g_device[device];
    std::call_once(d.once, set_up, device, std::ref(d));
    if (d.rc != BHX_OK) return fail(d.rc, "scratch address or launch grid unavailable on this device");
    hipStream_t st = static_cast<hipStream_t>(stream);

    unsigned char* base = d.base;
    double* init = reinterpret_cast<double*>(base + kOffInit);
    int32_t* q2ray = reinterpret_cast<int32_t*>(base + kOffQ2ray);
    uint64_t* keys_in = reinterpret_cast<uint64_t*>(base + kOffKeys);
    uint64_t* keys_out = keys_in + kMaxTiles;
    int32_t* vals_in = reinterpret_cast<int32_t*>(base + kOffVals);
    int32_t* order = vals_in + kMaxTiles;
    int32_t* sizes = reinterpret_cast<int32_t*>(base + kOffSizes);
    int32_t* start = reinterpret_cast<int32_t*>(base + kOffStart);
    void* temp = base + kOffTemp;
    Counters* cnt = reinterpret_cast<Counters*>(base + kOffCnt);
    EventParams* evp = reinterpret_cast<EventParams*>(base + kOffEv);

    size_t sort_bytes = 0, scan_bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys_in, keys_out, vals_in, order, n_tiles, 0, 64,
                                           st) != hipSuccess ||
        hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, sizes, start + 1, n_tiles, st) != hipSuccess)
        return fail(BHX_E_HIP, "sort / scan workspace query failed");
    if (sort_bytes > kSortTempBytes || scan_bytes > kSortTempBytes)
        return fail(BHX_E_TOO_LARGE, "sort workspace exceeds its scratch region");

    // queue: centre-first tiles and their first slots
    const bool band = host::axis_band(*s);
    const int blocks = (n_tiles + 255) / 256;
    hipLaunchKernelGGL(k_centre_keys, dim3(blocks), dim3(256), 0, st, s->width, s->height, tiles_x, n_tiles,
                       band ? 1 : 0, std::cos(s->cam_roll), std::sin(s->cam_roll), keys_in, vals_in);
    if (hipGetLastError() != hipSuccess) return fail(BHX_E_HIP, "tile key launch failed");
    size_t bytes = sort_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys_in, keys_out, vals_in, order, n_tiles, 0, 64, st) !=
        hipSuccess)
        return fail(BHX_E_HIP, "tile sort failed");
    hipLaunchKernelGGL(k_queue_sizes, dim3(blocks), dim3(256), 0, st, static_cast<const int32_t*>(order), n_tiles,
                       tiles_x, s->width, s->height, S * S, sizes, start);
    if (hipGetLastError() != hipSuccess) return fail(BHX_E_HIP, "tile size launch failed");
    bytes = scan_bytes;
    if (hipcub::DeviceScan::InclusiveSum(temp, bytes, sizes, start + 1, n_tiles, st) != hipSuccess)
        return fail(BHX_E_HIP, "tile scan failed");

    RaygenArgs G = host::raygen_args(*s);
    G.tiles_x = tiles_x;
    G.n_tiles = n_tiles;
    G.tile_order = order;
    G.tile_start = start;
    G.init = init;
    G.q2ray = q2ray;
    TraceArgs P = host::trace_args(*s);
    P.ev = evp;
    if (P.has_jet || P.has_sun) {
        hipLaunchKernelGGL(k_store_events, dim3(1), dim3(64), 0, st, host::event_params(*s), evp);
        if (hipGetLastError() != hipSuccess) return fail(BHX_E_HIP, "event parameter store failed");
    }
    P.refill = kRefill;
    P.chunk = kChunk;
    P.init = init;
    P.q2ray = q2ray;
    P.counter = &cnt->tile_counter;
    P.cls = rays->cls;
    P.y = rays->y;
    P.jet = rays->jet;
    P.steps = rays->steps;
    P.attempts = rays->attempts;
    P.norm_err = rays->norm_err;
    P.carter = rays->carter;
    P.totals = cnt->totals;
    // counters cleared by a fill kernel: the array keeps the last call's contents
    if (launch_zero_u64(reinterpret_cast<unsigned long long*>(cnt), sizeof(Counters) / 8, st) != hipSuccess)
        return fail(BHX_E_HIP, "counter clear failed");
    if (launch_raygen(G, st) != hipSuccess) return fail(BHX_E_HIP, "ray generation launch failed");
    // fewer rays than resident lanes: one-wave blocks, as bhx_trace_dev does
    int grid = trace_is_lean(P) ? d.grid_lean : d.grid, block = kBlock;
    if (n_rays < static_cast<int64_t>(grid) * block) {
        block = 64;
        grid = (n_rays + 63) / 64;
    }
    if (launch_trace(P, grid, block, st) != hipSuccess)
        return fail(BHX_E_HIP, "tracer launch failed");
    if (stats) {
        Counters c;
        if (hipStreamSynchronize(st) != hipSuccess ||
            hipMemcpy(&c, cnt, sizeof c, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(BHX_E_HIP, "reading the counters failed");
        double mn, mc;
        std::memcpy(&mn, &c.totals[2], 8);
        std::memcpy(&mc, &c.totals[3], 8);
        stats->rays += n_rays;
        stats->steps += static_cast<int64_t>(c.totals[0]);
        stats->attempts += static_cast<int64_t>(c.totals[1]);
        stats->rhs_evals += static_cast<int64_t>(c.totals[1]) * 6 + n_rays;
        stats->max_norm_error = std::max(stats->max_norm_error, mn);
        stats->max_carter_drift = std::max(stats->max_carter_drift, mc);
        stats->trace_ms = 0.0;   // no events: creating them is not this call's to do
        stats->order_ms = 0.0;
    }
    return BHX_OK;
}
