// adjoint.hip - the differentiable render on gfx950: a recording forward
// kernel and the discrete-adjoint kernel over the per-ray pieces of
// adjoint_dev.h, behind the bhx_adjoint_* entry points.
//
//   k_adj_forward   one ray per lane; a lane whose ray ends takes the next queue
//                   slot from a global counter in the same iteration, so every
//                   iteration is one DP5 attempt on every lane.  Queue slots map
//                   to rays in the tracer's queue order (8 x 8-pixel tiles,
//                   centre first, the spin axis' image first with the conical
//                   jet), so a wave holds neighbouring rays of similar cost.  Writes the
//                   class, hit state, accepted-step count and intensity
//                   (disc plus jet) of each sample, and a checkpoint of the
//                   integrator state every K accepted steps into a pool of
//                   8-record chunks that each ray takes as it goes (an atomic
//                   counter), each chunk linked back to the ray's previous
//                   one: the pool holds what the rays' real step counts need.
//   k_adj_backward  one lit ray (I > 0) per lane, the rays sorted by step
//                   count and taken 64 at a time by whichever wave is done,
//                   so the lanes of a wave replay and sweep segments of equal
//                   length together and the waves holding the longest rays
//                   take nothing else (a wave's few jet-lit lanes share their
//                   jet samples' node evaluations with the other lanes); each
//                   lane's (K + 1)-step segment buffer is a strided HBM
//                   scratch.  Each ray's contribution is kept and summed per
//                   frame in a fixed order (k_adj_slice_sums, k_adj_frame_sums).
#include "adjoint_frame.h"
#include "context.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

using namespace bhx;
using namespace bhx::adj;

namespace {

// Checkpoint pool: chunks of 2^lg records of kCk doubles (kStep with K = 1;
// lg from K: a ray of typical length fills about one chunk); link[c] is the
// chunk its ray filled before c (-1 for the ray's first).  Per ray, the last
// chunk and the number of records.
constexpr int64_t kMaxPoolBytes = int64_t(128) << 30;   // 128 GiB
inline int chunk_lg(int K) { return K >= 64 ? 2 : (K >= 16 ? 3 : (K > 1 ? 4 : 7)); }

// The forward's store of one ray: record j goes to slot j mod 2^lg of the
// ray's current chunk.  A ray's first chunk is chunk q, q its queue slot
// (the pool's first n_rays chunks: starting a ray takes no atomic); every
// further 2^lg records take a chunk n_rays + c, c from the counter.  Past the
// pool's end the counter still counts (it tells the host the size needed)
// and the records are dropped: the ray is marked lost (current chunk kLost).
// The lane's current chunk and record count live in an LDS slot (x, y; z
// holds the lane's ray index): they are touched once every K steps, and
// registers are what the forward is short of.
constexpr int32_t kLost = -2;
struct PoolWriter {
    double* pool;
    int32_t* link;
    uint32_t* counter;
    uint32_t cap;              // chunks in the pool (>= n_first)
    uint32_t n_first;          // first chunks: one per queue slot
    int32_t lg;
    int32_t rs;                // doubles per record
    int4* slot;
    __device__ void start(int32_t ray, int32_t q) { *slot = make_int4(q, 0, ray, 0); }
    __device__ double* put(int j) {
        int32_t cur = slot->x;
        const int32_t mask = (1 << lg) - 1;
        if (j == 0) {
            link[cur] = -1;
        } else if ((j & mask) == 0) {
            const uint64_t c = static_cast<uint64_t>(n_first) + atomicAdd(counter, 1u);
            if (c < cap && cur != kLost) {
                link[c] = cur;
                cur = static_cast<int32_t>(c);
            } else {
                cur = kLost;
            }
        }
        slot->x = cur;
        slot->y = j + 1;
        return cur == kLost ? nullptr : pool + ((static_cast<int64_t>(cur) << lg) + (j & mask)) * rs;
    }
};

// The backward's view of one ray's records: the chunk of record j is found
// by walking the links back from the ray's last chunk.
struct PoolReader {
    const double* pool;
    const int32_t* link;
    int32_t last, nrec, lg, rs;
    __device__ const double* get(int j) const {
        int32_t c = last;
        for (int k = (nrec - 1) >> lg; k > (j >> lg); --k) c = link[c];
        return pool + ((static_cast<int64_t>(c) << lg) + (j & ((1 << lg) - 1))) * rs;
    }
    // the walk back from the last record (n = nrec): one link per chunk
    __device__ CkPos rpos(int n) const { return {last, n - 1}; }
    __device__ const double* at(const CkPos& p) const {
        return pool + ((static_cast<int64_t>(p.c) << lg) + (p.j & ((1 << lg) - 1))) * rs;
    }
    __device__ void step_back(CkPos& p) const {
        if ((p.j & ((1 << lg) - 1)) == 0) p.c = link[p.c];
        --p.j;
    }
};

struct AdjArgs {
    int64_t n_rays;
    // forward
    uint32_t* counter;
    const int32_t* qmap;      // queue slot -> ray index
    double* pool;             // checkpoint chunks: pool_cap x 2^ck_lg x ck_rs doubles
    int32_t* link;            // pool_cap chunk links
    uint32_t* pool_counter;   // chunks taken beyond the first ones (the pool is short when
                              // n_rays + this exceeds pool_cap)
    uint32_t pool_cap;
    int32_t q_chunk;          // queue slots a wave takes at once
    int32_t refill;           // idle lanes a wave waits for before it refills them
    int32_t ck_lg;
    int32_t ck_rs;            // doubles per record: kCk, or kStep with K = 1
    int32_t* ck_last;         // per ray: its last chunk (-1: records lost)
    int32_t* ck_nrec;         // per ray: records written
    uint8_t* cls;
    double* y;                // 8 x n_rays SoA (layout of bhx_rays.y)
    int32_t* steps;
    double* I;                // n_rays intensity: disc sigma (g T)^4 / pi + jet kappa J
    int32_t* key;             // sort key: accepted steps of a lit ray, else 0
    unsigned int* n_lit;      // lit rays (I > 0) of the forward
    // backward
    const double* w;          // dL/dI per sample
    const int32_t* order;     // rays by key, descending: the lit rays first
    double* seg;              // (K + 1) * kSeg doubles per lane, strided by seg_stride
    int64_t seg_stride;
    unsigned long long* q_next;   // the backward's next slot of order (per wave, 64 at a time)
    unsigned int* mismatch;
    // frames of a batch (kernels instantiated for batches): ray r belongs to
    // frame r / n_per, whose camera is cams[r / n_per]; the backward leaves
    // each processed ray's d(w I)/d(params) in contrib (4 x n_rays SoA) for
    // the per-frame sums (one frame: n_per = n_rays)
    const FrameCam* cams;
    int64_t n_per;
    double* contrib;
};

struct AdjConst {
    Frame F;
    FrameCam cam;             // the camera of a single frame
};

template <bool kBatch>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_adj_forward(const AdjConst* __restrict__ C, const AdjArgs P) {
    // A step's conical-jet sample (32 nodes, ~10 steps of work) is not taken in
    // the step: its dense output waits in the lane's LDS slot, and at the loop
    // head, where every lane of the wave is active, each pass serves two
    // waiting lanes with one node per lane; the owner then sums its nodes in
    // order, as jet_step does.  A ray that ended writes its intensity after
    // its last sample.
    __shared__ double pend[256 * 17];
    __shared__ double nodes[4][64];
    const Frame& F = C->F;
    const int lane = __lane_id();
    double* my_pend = pend + threadIdx.x * 17;
    double* wave_nodes = nodes[threadIdx.x / 64];
    Ray s;
    s.jet_wait = 0;
    s.pt = s.pph = s.jet = 0.0;
    bool live = false, out_wait = false;
    // per lane in LDS: checkpoint chunk, record count, ray index; the disc
    // intensity of a ray that ended, until its last jet sample is in
    __shared__ int4 lane_ray[256];
    __shared__ double lane_idisc[256];
    PoolWriter wr{P.pool, P.link, P.pool_counter, P.pool_cap, static_cast<uint32_t>(P.n_rays), P.ck_lg, P.ck_rs,
                  &lane_ray[threadIdx.x]};
    // the wave's share of the queue: slots [c_next, c_end), taken q_chunk at
    // a time; idle lanes get consecutive slots (ballot + prefix rank), once
    // `refill` of them wait or the whole wave is idle, so that the rays'
    // set-up (camera, first right-hand side, checkpoint 0) runs on many lanes
    // at once rather than once per finished lane
    int64_t c_next = 0, c_end = 0;
    bool q_done = false;
    for (;;) {
        unsigned long long owners = __ballot(s.jet_wait != 0);
        while (owners != 0ull) {
            const int A = __builtin_ctzll(owners);
            owners &= owners - 1;
            int B = -1;
            if (owners != 0ull) {
                B = __builtin_ctzll(owners);
                owners &= owners - 1;
            }
            const int owner = (lane >= 32 && B >= 0) ? B : A;
            const double* op = pend + (threadIdx.x - lane + owner) * 17;
            const double opt = __shfl(s.pt, owner), opph = __shfl(s.pph, owner);
            wave_nodes[lane] = jet_value_node(F, op, op[15], opt, opph, lane & 31);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == A || lane == B) {
                const double* v = wave_nodes + ((lane == A) ? 0 : 32);
                double sum = 0.0;
                for (int j = 0; j < 4 * kConeSub; ++j) sum += gl_w(j & 3) * v[j];
                s.jet += F.j0 * sum * my_pend[16] * my_pend[15] / kConeSub;
                s.jet_wait = 0;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (out_wait) {   // the ray ended on the last step
            const int64_t ray = lane_ray[threadIdx.x].z;
            const double I = lane_idisc[threadIdx.x] + F.kappa * s.jet;
            P.I[ray] = I;
            P.key[ray] = (I > 0.0) ? max(1, min(s.nsteps, 32767)) : 0;   // 15-bit key, >= 1 for a lit ray
            if (I > 0.0) atomicAdd(P.n_lit, 1u);
            out_wait = false;
        }
        unsigned long long idle = __ballot(!live);
        const int n_idle = __popcll(idle);
        if (!q_done && (n_idle >= P.refill || n_idle == 64)) {
            while (idle != 0ull) {
                if (c_next >= c_end) {
                    uint32_t b = 0;
                    if (lane == 0) b = atomicAdd(P.counter, static_cast<uint32_t>(P.q_chunk));
                    b = __builtin_amdgcn_readfirstlane(b);
                    if (static_cast<int64_t>(b) >= P.n_rays) {
                        q_done = true;
                        break;
                    }
                    c_next = b;
                    c_end = min(P.n_rays, static_cast<int64_t>(b) + P.q_chunk);
                }
                const int64_t avail = c_end - c_next;
                const int rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(idle >> 32),
                                                           __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(idle), 0u));
                if (!live && rank < avail) {
                    const int64_t q = c_next + rank;
                    const int32_t ray = P.qmap[q];
                    wr.start(ray, static_cast<int32_t>(q));
                    if (kBatch) {
                        const int64_t frame = ray / P.n_per;
                        fwd_init(F, P.cams[frame], ray - frame * P.n_per, s, wr);
                    } else {
                        fwd_init(F, C->cam, ray, s, wr);
                    }
                    live = true;
                }
                c_next += min(static_cast<int64_t>(__popcll(idle)), avail);
                idle = __ballot(!live);
            }
        }
        // a refill runs whenever all 64 lanes are idle and only stops with
        // every lane live or the queue drained: no live lane, no work left
        if (__ballot(live) == 0ull) break;
        if (!live) continue;
        double yo[6], se = 0.0;
        const int term = fwd_step(F, s, wr, yo, &se, my_pend);
        if (term >= 0) {
            const int64_t N = P.n_rays;
            const int4 ck = *wr.slot;
            const int64_t ray = ck.z;
            P.ck_last[ray] = ck.x == kLost ? -1 : ck.x;
            P.ck_nrec[ray] = ck.y;
            P.cls[ray] = (uint8_t)term;
            if (P.y) {
                P.y[ray] = yo[0]; P.y[N + ray] = yo[1]; P.y[2 * N + ray] = yo[2]; P.y[3 * N + ray] = yo[3];
                P.y[4 * N + ray] = s.pt; P.y[5 * N + ray] = yo[4]; P.y[6 * N + ray] = yo[5]; P.y[7 * N + ray] = s.pph;
            }
            P.steps[ray] = s.nsteps;
            lane_idisc[threadIdx.x] = (term == BHX_DISC) ? disc_intensity(F, yo[1], s.pt, s.pph) : 0.0;
            out_wait = true;
            live = false;
        }
    }
}

__global__ void k_iota(int32_t* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = static_cast<int32_t>(i);
}

template <bool kBatch>
__global__ void __launch_bounds__(256) k_adj_backward(const AdjConst* __restrict__ C, const AdjArgs P) {
    __shared__ double jet_lds[4][kJetLds];   // per wave: the jet nodes shared across its lanes
    const Frame& F = C->F;
    const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned int mism = 0;
    const int64_t n_order = *P.n_lit;
    // A wave takes the next 64 rays of the order (most accepted steps first)
    // whenever all its lanes are done: the waves holding the longest rays
    // take nothing more, the others share out the rest.  Every processed ray
    // leaves its own contribution, summed per frame afterwards in a fixed
    // order, so the gradient does not depend on which lane took which ray.
    for (;;) {
        unsigned long long base = 0;
        if (__lane_id() == 0) base = atomicAdd(P.q_next, 64ull);
        base = __shfl(base, 0);
        if (base >= static_cast<unsigned long long>(n_order)) break;
        const int64_t i = static_cast<int64_t>(base) + __lane_id();
        if (i >= n_order) continue;
        const int64_t ray = P.order[i];
        const double w = P.w[ray];
        if (w == 0.0) continue;
        const PoolReader rd{P.pool, P.link, P.ck_last[ray], P.ck_nrec[ray], P.ck_lg, P.ck_rs};
        double gr[4] = {0.0, 0.0, 0.0, 0.0};
        if (kBatch) {
            const int64_t frame = ray / P.n_per;
            mism += bwd_ray(F, P.cams[frame], ray - frame * P.n_per, P.steps[ray], P.cls[ray], w, rd, P.seg + lane,
                            P.seg_stride, gr, F.has_jet ? jet_lds[threadIdx.x / 64] : nullptr);
        } else {
            mism += bwd_ray(F, C->cam, ray, P.steps[ray], P.cls[ray], w, rd, P.seg + lane, P.seg_stride, gr,
                            F.has_jet ? jet_lds[threadIdx.x / 64] : nullptr);
        }
        for (int c = 0; c < 4; ++c) P.contrib[c * P.n_rays + ray] = gr[c];
    }
    for (int off = 32; off > 0; off >>= 1) mism += __shfl_xor(mism, off);
    if (__lane_id() == 0 && mism) atomicAdd(P.mismatch, mism);
}

// Per frame (of a batch, or the one frame), the sum of its processed rays'
// contributions (lit rays with dL/dI != 0) in a fixed order, in two passes:
// the frame's rays in kSumSlices contiguous slices, one block each (each
// thread a fixed stride of the slice, then a fixed tree), then per frame the
// slices' sums in slice order.  Deterministic, and the same order for a
// frame alone as inside a batch.
constexpr int kSumSlices = 64;
__global__ void __launch_bounds__(256) k_adj_slice_sums(const AdjArgs P, const double* __restrict__ w,
                                                        double* __restrict__ part) {
    __shared__ double red[4][256];
    const int64_t frame = blockIdx.y, slice = blockIdx.x;
    const int64_t base = frame * P.n_per;
    const int64_t lo = P.n_per * slice / kSumSlices, hi = P.n_per * (slice + 1) / kSumSlices;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t r = lo + threadIdx.x; r < hi; r += blockDim.x) {
        const int64_t ray = base + r;
        if (P.key[ray] > 0 && w[ray] != 0.0)
            for (int c = 0; c < 4; ++c) acc[c] += P.contrib[c * P.n_rays + ray];
    }
    for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1) {
        if ((int)threadIdx.x < half)
            for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[(frame * kSumSlices + slice) * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_adj_frame_sums(const double* __restrict__ part, int64_t n_frames, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // frame * 4 + component
    if (i >= n_frames * 4) return;
    const int64_t frame = i / 4, c = i % 4;
    double sum = 0.0;
    for (int sl = 0; sl < kSumSlices; ++sl) sum += part[(frame * kSumSlices + sl) * 4 + c];
    out[i] = sum;
}

}  // namespace

struct bhx_adjoint {
    bhx_context* ctx = nullptr;
    std::vector<bhx_scene> scenes;   // the frame (one), or the frames of a batch
    bool is_batch = false;
    int K = 0;
    int ck_lg = 3;                   // checkpoint records per pool chunk: 2^ck_lg
    int ck_rs = kCk;                 // doubles per record (kStep with K = 1)
    int64_t n_per = 0;               // samples per frame
    int64_t n_rays = 0;              // samples of all frames
    AdjConst host{};
    DevBuf cst, cams, pool, link, ck_last, ck_nrec, cls, y, steps, I, key, key_sorted, iota, order, sort_tmp, seg,
        counters, qmap, contrib, frame_grad, slice_sums;
    uint32_t pool_cap = 0;           // chunks in the checkpoint pool (>= n_rays: a first chunk per ray)
    int32_t q_chunk = 32;            // the forward's queue: slots per wave take
    int32_t refill = 1;              // idle lanes a wave waits for before it refills them
    int grid_f = 0, grid_b = 0;
    int fwd_per_cu = 0;              // recording-forward blocks per CU that fit (one wave per SIMD each)
    int fwd_waves = 0;               // waves per SIMD the recording forward runs (0: as many as fit)
    double fwd_ms = 0.0, bwd_ms = 0.0;
    bool fwd_pending = false;
    bool have_steps = false;         // steps holds a forward's step counts (the next queue order)
    unsigned int mismatches = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t chunk_bytes() const { return sizeof(double) * ck_rs * (size_t(1) << ck_lg); }
};

namespace {

// Frame constants of the adjoint's frames (with its K): one Frame shared by
// all of them, a camera per frame.
int adj_set_scenes(bhx_adjoint* A, const bhx_scene* s, int n) {
    bhx_context* ctx = A->ctx;
    if (!s || n < 1) return bhx_fail(ctx, BHX_E_INVALID, "scene is NULL");
    std::vector<FrameCam> cams(static_cast<size_t>(n));
    AdjConst c;
    for (int f = 0; f < n; ++f) {
        const int S = std::max(1, s[f].sqrt_spp);
        const int64_t m = static_cast<int64_t>(s[f].width) * s[f].height * S * S;
        if (m != A->n_per) return bhx_fail(ctx, BHX_E_INVALID, "scene size differs from the adjoint's");
        if (s[f].max_steps < 1) return bhx_fail(ctx, BHX_E_INVALID, "max_steps must be > 0");
        Frame F;
        if (!make_frame(s[f], A->K, F, cams[f]))
            return bhx_fail(ctx, BHX_E_INVALID,
                            "only disc scenes (optionally with the conical jet) without neutron star, Sun or hot spot "
                            "are differentiable");
        if (!(s[f].rtol > 0.0) || !(s[f].atol > 0.0)) return bhx_fail(ctx, BHX_E_INVALID, "rtol/atol must be > 0");
        if (!(s[f].eddington > 0.0)) return bhx_fail(ctx, BHX_E_INVALID, "eddington must be > 0");
        if (f == 0) c.F = F;
        else if (std::memcmp(&F, &c.F, sizeof F) != 0)
            return bhx_fail(ctx, BHX_E_INVALID, "the frames of a batch must differ in their camera only");
    }
    c.cam = cams[0];
    // the constants are read by kernels of earlier calls: replace them in order
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(A->cst.p, &c, sizeof c, hipMemcpyHostToDevice));
    if (A->is_batch) HIPCHK(hipMemcpy(A->cams.p, cams.data(), sizeof(FrameCam) * n, hipMemcpyHostToDevice));
    A->host = c;
    A->scenes.assign(s, s + n);
    return BHX_OK;
}

// Recording forward on stream st (I: the caller's intensity buffer, or NULL
// for the adjoint's own).
int adj_record(bhx_adjoint* A, double* I, hipStream_t st) {
    bhx_context* ctx = A->ctx;
    AdjArgs P{};
    P.n_rays = A->n_rays;
    P.counter = static_cast<uint32_t*>(A->counters.p);
    P.qmap = static_cast<const int32_t*>(A->qmap.p);
    P.pool = static_cast<double*>(A->pool.p);
    P.link = static_cast<int32_t*>(A->link.p);
    P.pool_counter = reinterpret_cast<uint32_t*>(static_cast<char*>(A->counters.p) + 16);
    P.pool_cap = A->pool_cap;
    P.q_chunk = A->q_chunk;
    P.refill = A->refill;
    P.ck_lg = A->ck_lg;
    P.ck_rs = A->ck_rs;
    P.ck_last = static_cast<int32_t*>(A->ck_last.p);
    P.ck_nrec = static_cast<int32_t*>(A->ck_nrec.p);
    P.cls = static_cast<uint8_t*>(A->cls.p);
    P.y = static_cast<double*>(A->y.p);   // NULL for a batch: hit states are not kept
    P.steps = static_cast<int32_t*>(A->steps.p);
    P.I = I ? I : static_cast<double*>(A->I.p);
    P.key = static_cast<int32_t*>(A->key.p);
    P.n_lit = reinterpret_cast<unsigned int*>(static_cast<char*>(A->counters.p) + 8);
    P.cams = static_cast<const FrameCam*>(A->cams.p);
    P.n_per = A->n_per;
    // the queue: after a first forward in the frames' tile order, the rays by
    // the last forward's accepted steps, most first, so that the longest rays
    // start first and no lane is left with one at the end (the order changes
    // which lane traces a ray, never what the ray computes)
    if (A->have_steps) {
        size_t tq = A->sort_tmp.bytes;
        HIPCHK(hipcub::DeviceRadixSort::SortPairsDescending(A->sort_tmp.p, tq, static_cast<const int32_t*>(A->steps.p),
                                                            static_cast<int32_t*>(A->key_sorted.p),
                                                            static_cast<const int32_t*>(A->iota.p),
                                                            static_cast<int32_t*>(A->qmap.p), A->n_rays, 0, 31, st));
    }
    // counters: [0] queue, [8] lit rays, [16] pool chunks taken, [32] replay mismatches
    HIPCHK(launch_zero_u64(static_cast<unsigned long long*>(A->counters.p), 8, st));
    HIPCHK(hipEventRecord(A->ev[0], st));
    if (A->is_batch)
        hipLaunchKernelGGL(k_adj_forward<true>, dim3(A->grid_f), dim3(256), 0, st,
                           static_cast<const AdjConst*>(A->cst.p), P);
    else
        hipLaunchKernelGGL(k_adj_forward<false>, dim3(A->grid_f), dim3(256), 0, st,
                           static_cast<const AdjConst*>(A->cst.p), P);
    HIPCHK(hipGetLastError());
    // the backward's ray order: lit rays first, most accepted steps first
    // (a stable radix sort on the 15-bit step count: deterministic)
    size_t tmp = A->sort_tmp.bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairsDescending(A->sort_tmp.p, tmp, static_cast<const int32_t*>(A->key.p),
                                                        static_cast<int32_t*>(A->key_sorted.p),
                                                        static_cast<const int32_t*>(A->iota.p),
                                                        static_cast<int32_t*>(A->order.p), A->n_rays, 0, 15, st));
    HIPCHK(hipEventRecord(A->ev[1], st));
    A->fwd_pending = true;
    A->have_steps = true;
    return BHX_OK;
}

// A forward that ran out of pool chunks counted what it needed: grow the pool
// to that and record the same forward again (deterministic: the same rays,
// steps and intensities; the caller's intensity buffer is not written again).
int adj_grow_if_short(bhx_adjoint* A, hipStream_t st) {
    bhx_context* ctx = A->ctx;
    uint32_t more = 0;
    HIPCHK(hipMemcpyAsync(&more, static_cast<char*>(A->counters.p) + 16, sizeof more, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int64_t need = A->n_rays + static_cast<int64_t>(more);
    if (need <= static_cast<int64_t>(A->pool_cap)) return BHX_OK;
    if (static_cast<double>(need) * A->chunk_bytes() > static_cast<double>(kMaxPoolBytes) || need > 0xffffffffLL)
        return bhx_fail(ctx, BHX_E_TOO_LARGE, "checkpoint pool above 128 GiB: raise K");
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(A->pool.ensure(A->chunk_bytes() * static_cast<size_t>(need)));
    HIPCHK(A->link.ensure(sizeof(int32_t) * static_cast<size_t>(need)));
    A->pool_cap = static_cast<uint32_t>(need);
    return adj_record(A, nullptr, st);
}

// The adjoint sweep over the last forward's lit rays (synchronises st):
// grad gets 4 values per frame.
int adj_sweep(bhx_adjoint* A, const double* w, double* grad, hipStream_t st) {
    bhx_context* ctx = A->ctx;
    AdjArgs P{};
    P.n_rays = A->n_rays;
    P.pool = static_cast<double*>(A->pool.p);
    P.link = static_cast<int32_t*>(A->link.p);
    P.ck_lg = A->ck_lg;
    P.ck_rs = A->ck_rs;
    P.ck_last = static_cast<int32_t*>(A->ck_last.p);
    P.ck_nrec = static_cast<int32_t*>(A->ck_nrec.p);
    P.cls = static_cast<uint8_t*>(A->cls.p);
    P.steps = static_cast<int32_t*>(A->steps.p);
    P.key = static_cast<int32_t*>(A->key.p);
    P.w = w;
    P.order = static_cast<const int32_t*>(A->order.p);
    P.seg = static_cast<double*>(A->seg.p);
    P.seg_stride = static_cast<int64_t>(A->grid_b) * 256;
    P.q_next = reinterpret_cast<unsigned long long*>(static_cast<char*>(A->counters.p) + 40);
    P.mismatch = reinterpret_cast<unsigned int*>(static_cast<char*>(A->counters.p) + 32);
    P.n_lit = reinterpret_cast<unsigned int*>(static_cast<char*>(A->counters.p) + 8);
    P.cams = static_cast<const FrameCam*>(A->cams.p);
    P.n_per = A->n_per;
    P.contrib = static_cast<double*>(A->contrib.p);
    const int n_frames = static_cast<int>(A->scenes.size());
    HIPCHK(launch_zero_u32(P.mismatch, 1, st));
    HIPCHK(launch_zero_u64(P.q_next, 1, st));
    HIPCHK(hipEventRecord(A->ev[2], st));
    if (A->is_batch)
        hipLaunchKernelGGL(k_adj_backward<true>, dim3(A->grid_b), dim3(256), 0, st,
                           static_cast<const AdjConst*>(A->cst.p), P);
    else
        hipLaunchKernelGGL(k_adj_backward<false>, dim3(A->grid_b), dim3(256), 0, st,
                           static_cast<const AdjConst*>(A->cst.p), P);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_adj_slice_sums, dim3(kSumSlices, n_frames), dim3(256), 0, st, P, w,
                       static_cast<double*>(A->slice_sums.p));
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_adj_frame_sums, dim3((4 * n_frames + 255) / 256), dim3(256), 0, st,
                       static_cast<const double*>(A->slice_sums.p), static_cast<int64_t>(n_frames),
                       static_cast<double*>(A->frame_grad.p));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(A->ev[3], st));
    unsigned int mism = 0;
    HIPCHK(hipMemcpyAsync(grad, A->frame_grad.p, sizeof(double) * 4 * n_frames, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&mism, P.mismatch, sizeof mism, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    A->mismatches = mism;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, A->ev[2], A->ev[3]) == hipSuccess) A->bwd_ms = ms;
    if (A->fwd_pending && hipEventElapsedTime(&ms, A->ev[0], A->ev[1]) == hipSuccess) A->fwd_ms = ms;
    A->fwd_pending = false;
    return BHX_OK;
}

// The checkpoint interval: the segment buffer holds K + 1 steps per lane and
// the pool's chunks 2^chunk_lg(K) records; the constants carry K.
int adj_set_K(bhx_adjoint* A, int K) {
    bhx_context* ctx = A->ctx;
    const int64_t lanes = static_cast<int64_t>(A->grid_b) * 256;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(A->seg.ensure(sizeof(double) * kSeg * (K + 1) * lanes));
    A->K = K;
    A->ck_lg = chunk_lg(K);
    A->ck_rs = K == 1 ? kStep : kCk;
    // at least a first chunk per ray
    if (static_cast<double>(A->n_rays) * A->chunk_bytes() > static_cast<double>(kMaxPoolBytes))
        return bhx_fail(ctx, BHX_E_TOO_LARGE, "checkpoint pool above 128 GiB: raise K");
    HIPCHK(A->pool.ensure(A->chunk_bytes() * static_cast<size_t>(A->n_rays)));
    HIPCHK(A->link.ensure(sizeof(int32_t) * static_cast<size_t>(A->n_rays)));
    A->pool_cap = static_cast<uint32_t>(
        std::min<size_t>({A->pool.bytes / A->chunk_bytes(), A->link.bytes / 4, size_t(0xffffffffu)}));
    return adj_set_scenes(A, A->scenes.data(), static_cast<int>(A->scenes.size()));
}

int adj_create(bhx_context* ctx, const bhx_scene* s, int32_t n, int32_t K, bool batch, bhx_adjoint** out) {
    if (!ctx || !out) return BHX_E_INVALID;
    *out = nullptr;
    if (!s || n < 1) return bhx_fail(ctx, BHX_E_INVALID, "scene is NULL");
    if (n > 65535) return bhx_fail(ctx, BHX_E_TOO_LARGE, "more than 65535 frames in a batch");
    if (K < 1 || K > 4096) return bhx_fail(ctx, BHX_E_INVALID, "checkpoint interval K must be in [1, 4096]");
    if (s[0].width <= 0 || s[0].height <= 0 || s[0].max_steps < 1)
        return bhx_fail(ctx, BHX_E_INVALID, "width/height/max_steps must be > 0");
    const int S = std::max(1, s[0].sqrt_spp);
    const int64_t n_per = static_cast<int64_t>(s[0].width) * s[0].height * S * S;
    for (int f = 1; f < n; ++f)
        if (s[f].width != s[0].width || s[f].height != s[0].height || std::max(1, s[f].sqrt_spp) != S)
            return bhx_fail(ctx, BHX_E_INVALID, "the frames of a batch must have one size");
    const int64_t n_rays = n_per * n;
    if (n_rays > 2000000000LL) return bhx_fail(ctx, BHX_E_TOO_LARGE, "more than 2e9 samples");
    HIPCHK(hipSetDevice(ctx->device));
    bhx_adjoint* A = new bhx_adjoint();
    A->ctx = ctx;
    A->is_batch = batch;
    // a frame's rays fill the resident lanes a few times over: refill each
    // idle lane at once; a batch's many frames run in steady state, where
    // refilling 4 lanes at a time pays (measured on the 1024^2 sweep frames)
    A->refill = batch ? 4 : 1;
    A->K = K;
    A->ck_lg = chunk_lg(K);
    A->ck_rs = K == 1 ? kStep : kCk;
    A->n_per = n_per;
    A->n_rays = n_rays;
    auto bail = [&](int rc) {
        bhx_adjoint_destroy(A);
        return rc;
    };
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adj_forward<false>, 256, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    A->fwd_per_cu = per_cu;
    A->grid_f = per_cu * ctx->n_cu;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adj_backward<false>, 256, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    A->grid_b = per_cu * ctx->n_cu;
    const int64_t lanes = static_cast<int64_t>(A->grid_b) * 256;
    // first guess of the checkpoint pool: single frames rays of up to 512
    // accepted steps (a frame at rtol 1e-10 averages ~300), batches 256 (with
    // K = 1, a record per step: 320 and 160 steps); a forward that finds it
    // short tells the first backward after it, which grows the pool to the need
    const int64_t rec = K == 1 ? (batch ? 160 : 320) : (batch ? 256 : 512) / K + 1;
    const int64_t per_ray = (rec + (int64_t(1) << A->ck_lg) - 1) >> A->ck_lg;
    const int64_t max_chunks = std::min<int64_t>(kMaxPoolBytes / static_cast<int64_t>(A->chunk_bytes()), 0xffffffffLL);
    if (n_rays > max_chunks) return bail(bhx_fail(ctx, BHX_E_TOO_LARGE, "checkpoint pool above 128 GiB: raise K"));
    const int64_t chunks = std::min<int64_t>(n_rays * per_ray, max_chunks);
    A->pool_cap = static_cast<uint32_t>(chunks);
    if (A->cst.ensure(sizeof(AdjConst)) != hipSuccess ||
        (batch && A->cams.ensure(sizeof(FrameCam) * n) != hipSuccess) ||
        A->pool.ensure(A->chunk_bytes() * static_cast<size_t>(chunks)) != hipSuccess ||
        A->link.ensure(sizeof(int32_t) * static_cast<size_t>(chunks)) != hipSuccess ||
        A->ck_last.ensure(sizeof(int32_t) * n_rays) != hipSuccess ||
        A->ck_nrec.ensure(sizeof(int32_t) * n_rays) != hipSuccess || A->cls.ensure(n_rays) != hipSuccess ||
        (!batch && A->y.ensure(sizeof(double) * 8 * n_rays) != hipSuccess) ||
        A->steps.ensure(sizeof(int32_t) * n_rays) != hipSuccess || A->I.ensure(sizeof(double) * n_rays) != hipSuccess ||
        A->order.ensure(sizeof(int32_t) * n_rays) != hipSuccess || A->key.ensure(sizeof(int32_t) * n_rays) != hipSuccess ||
        A->key_sorted.ensure(sizeof(int32_t) * n_rays) != hipSuccess ||
        A->iota.ensure(sizeof(int32_t) * n_rays) != hipSuccess ||
        A->seg.ensure(sizeof(double) * kSeg * (K + 1) * lanes) != hipSuccess ||
        A->contrib.ensure(sizeof(double) * 4 * n_rays) != hipSuccess ||
        A->frame_grad.ensure(sizeof(double) * 4 * n) != hipSuccess ||
        A->slice_sums.ensure(sizeof(double) * 4 * kSumSlices * n) != hipSuccess || A->counters.ensure(64) != hipSuccess)
        return bail(bhx_fail(ctx, BHX_E_NOMEM, "adjoint buffers"));
    for (auto& e : A->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(bhx_fail(ctx, BHX_E_HIP, "hipEventCreate"));
    // the forward's queue: frame after frame, each in the tracer's tile order
    {
        std::vector<int32_t> map(static_cast<size_t>(n_rays));
        const int W = s[0].width, H = s[0].height, tiles_x = (W + 7) / 8, S2 = S * S;
        for (int f = 0; f < n; ++f) {
            std::vector<int32_t> order, start;
            host::tile_queue(W, H, S, 0, order, start, host::axis_band(s[f]), s[f].cam_roll);
            const int64_t base = n_per * f;
            for (size_t t = 0; t < order.size(); ++t) {
                const int tx = order[t] % tiles_x, ty = order[t] / tiles_x;
                const int x0 = 8 * tx, y0 = 8 * ty;
                const int tw = std::min(8, W - x0), th = std::min(8, H - y0);
                for (int i = 0; i < tw * th * S2; ++i) {
                    const int pin = i / S2, k = i - pin * S2;
                    const int px = x0 + pin % tw, py = y0 + pin / tw;
                    map[static_cast<size_t>(base + start[t] + i)] = static_cast<int32_t>(base + (py * W + px) * S2 + k);
                }
            }
        }
        if (A->qmap.ensure(sizeof(int32_t) * n_rays) != hipSuccess) return bail(bhx_fail(ctx, BHX_E_NOMEM, "queue map"));
        if (hipMemcpy(A->qmap.p, map.data(), sizeof(int32_t) * n_rays, hipMemcpyHostToDevice) != hipSuccess)
            return bail(bhx_fail(ctx, BHX_E_HIP, "queue map upload"));
    }
    // ray ids 0..n-1 (sort values) and the radix sort's scratch
    hipLaunchKernelGGL(k_iota, dim3(static_cast<unsigned>((n_rays + 255) / 256)), dim3(256), 0, 0,
                       static_cast<int32_t*>(A->iota.p), n_rays);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return bail(bhx_fail(ctx, BHX_E_HIP, "k_iota"));
    size_t tmp = 0, tmp31 = 0;
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, static_cast<const int32_t*>(A->key.p),
                                                     static_cast<int32_t*>(A->key_sorted.p),
                                                     static_cast<const int32_t*>(A->iota.p),
                                                     static_cast<int32_t*>(A->order.p), n_rays, 0, 15) != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp31, static_cast<const int32_t*>(A->steps.p),
                                                     static_cast<int32_t*>(A->key_sorted.p),
                                                     static_cast<const int32_t*>(A->iota.p),
                                                     static_cast<int32_t*>(A->qmap.p), n_rays, 0, 31) != hipSuccess ||
        A->sort_tmp.ensure(std::max<size_t>({tmp, tmp31, size_t(16)})) != hipSuccess)
        return bail(bhx_fail(ctx, BHX_E_NOMEM, "radix sort scratch"));
    const int rc = adj_set_scenes(A, s, n);
    if (rc) return bail(rc);
    *out = A;
    return BHX_OK;
}

}  // namespace

extern "C" {

int bhx_adjoint_create(bhx_context* ctx, const bhx_scene* s, int32_t K, bhx_adjoint** out) {
    return adj_create(ctx, s, 1, K, false, out);
}

int bhx_adjoint_create_batch(bhx_context* ctx, const bhx_scene* scenes, int32_t n, int32_t K, bhx_adjoint** out) {
    return adj_create(ctx, scenes, n, K, true, out);
}

int bhx_adjoint_destroy(bhx_adjoint* A) {
    if (!A) return BHX_OK;
    (void)hipSetDevice(A->ctx->device);
    (void)hipDeviceSynchronize();
    for (DevBuf* d : {&A->cst, &A->cams, &A->pool, &A->link, &A->ck_last, &A->ck_nrec, &A->cls, &A->y, &A->steps,
                      &A->I, &A->key, &A->key_sorted, &A->iota, &A->order, &A->sort_tmp, &A->seg,
                      &A->counters, &A->qmap, &A->contrib, &A->frame_grad, &A->slice_sums})
        d->release();
    for (auto& e : A->ev)
        if (e) (void)hipEventDestroy(e);
    delete A;
    return BHX_OK;
}

int bhx_adjoint_set_scene(bhx_adjoint* A, const bhx_scene* s) {
    if (!A) return BHX_E_INVALID;
    bhx_context* ctx = A->ctx;
    if (A->is_batch) return bhx_fail(ctx, BHX_E_INVALID, "a batch keeps the frames it was created with");
    HIPCHK(hipSetDevice(ctx->device));
    return adj_set_scenes(A, s, 1);
}

int bhx_adjoint_forward(bhx_adjoint* A, double* I, void* stream) {
    if (!A) return BHX_E_INVALID;
    bhx_context* ctx = A->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    return adj_record(A, I, static_cast<hipStream_t>(stream));
}

int bhx_adjoint_backward(bhx_adjoint* A, const double* w, double* grad, void* stream) {
    if (!A || !w || !grad) return BHX_E_INVALID;
    bhx_context* ctx = A->ctx;
    if (A->is_batch) return bhx_fail(ctx, BHX_E_INVALID, "a batch's gradients come from bhx_adjoint_backward_batch");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int rc = adj_grow_if_short(A, st);
    if (rc) return rc;
    return adj_sweep(A, w, grad, st);
}

int bhx_adjoint_backward_batch(bhx_adjoint* A, const double* w, double* grad, void* stream) {
    if (!A || !w || !grad) return BHX_E_INVALID;
    bhx_context* ctx = A->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int rc = adj_grow_if_short(A, st);
    if (rc) return rc;
    return adj_sweep(A, w, grad, st);
}

int bhx_adjoint_rays(const bhx_adjoint* A, bhx_rays* rays) {
    if (!A || !rays) return BHX_E_INVALID;
    std::memset(rays, 0, sizeof *rays);
    rays->cls = static_cast<uint8_t*>(A->cls.p);
    rays->y = static_cast<double*>(A->y.p);
    rays->steps = static_cast<int32_t*>(A->steps.p);
    return BHX_OK;
}

int bhx_adjoint_copy_rays(bhx_adjoint* A, uint8_t* cls, double* y, int32_t* steps, void* stream) {
    if (!A) return BHX_E_INVALID;
    bhx_context* ctx = A->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    if (cls) HIPCHK(hipMemcpy(cls, A->cls.p, A->n_rays, hipMemcpyDeviceToHost));
    if (y && A->y.p) HIPCHK(hipMemcpy(y, A->y.p, sizeof(double) * 8 * A->n_rays, hipMemcpyDeviceToHost));
    if (steps) HIPCHK(hipMemcpy(steps, A->steps.p, sizeof(int32_t) * A->n_rays, hipMemcpyDeviceToHost));
    return BHX_OK;
}

int bhx_adjoint_autotune(bhx_adjoint* A, const double* w, const int32_t* Ks, int32_t n_K, int32_t reps,
                         int32_t* best_K, double* best_ms, void* stream) {
    if (!A || !w || !Ks || n_K < 1 || n_K > 16 || reps < 1) return BHX_E_INVALID;
    bhx_context* ctx = A->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    for (int i = 0; i < n_K; ++i)
        if (Ks[i] < 1 || Ks[i] > 4096) return bhx_fail(ctx, BHX_E_INVALID, "checkpoint interval K must be in [1, 4096]");
    int K_best = A->K;
    double ms_best = 0.0;
    bool any = false;
    std::vector<double> grad(4 * A->scenes.size());
    for (int i = 0; i < n_K; ++i) {
        // a K whose records would not fit the pool's cap is no candidate
        int rc = adj_set_K(A, Ks[i]);
        if (rc == BHX_E_TOO_LARGE) continue;
        if (rc) return rc;
        // one untimed pass (grows the checkpoint pool if this K needs more)
        rc = adj_record(A, nullptr, st);
        if (!rc) rc = adj_grow_if_short(A, st);
        if (rc == BHX_E_TOO_LARGE) continue;
        if (!rc) rc = adj_sweep(A, w, grad.data(), st);
        if (rc) return rc;
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps && !rc; ++r) {
            rc = adj_record(A, nullptr, st);
            if (!rc) rc = adj_sweep(A, w, grad.data(), st);
        }
        if (rc) return rc;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
        if (!any || ms < ms_best) {
            ms_best = ms;
            K_best = Ks[i];
            any = true;
        }
    }
    if (!any) {
        (void)adj_set_K(A, K_best);
        return bhx_fail(ctx, BHX_E_TOO_LARGE, "no candidate checkpoint interval fits the 128 GiB pool");
    }
    int rc = adj_set_K(A, K_best);
    if (rc) return rc;
    // then the recording forward's waves per SIMD at that K: a small frame
    // lasts as long as its longest ray, which fewer waves per SIMD finish
    // sooner; a large one is throughput-bound and wants them all
    const int w_start = A->grid_f / ctx->n_cu;
    int w_best = w_start;
    for (int wv = A->fwd_per_cu; wv >= 1; --wv) {
        A->grid_f = wv * ctx->n_cu;
        rc = adj_record(A, nullptr, st);
        if (!rc) rc = adj_grow_if_short(A, st);
        if (!rc) rc = adj_sweep(A, w, grad.data(), st);
        if (rc) return rc;
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps && !rc; ++r) {
            rc = adj_record(A, nullptr, st);
            if (!rc) rc = adj_sweep(A, w, grad.data(), st);
        }
        if (rc) return rc;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
        if (ms < ms_best) {
            ms_best = ms;
            w_best = wv;
        }
    }
    A->grid_f = w_best * ctx->n_cu;
    A->fwd_waves = w_best;
    if (best_K) *best_K = K_best;
    if (best_ms) *best_ms = ms_best;
    return BHX_OK;
}

int bhx_adjoint_get_interval(const bhx_adjoint* A, int32_t* K) {
    if (!A || !K) return BHX_E_INVALID;
    *K = A->K;
    return BHX_OK;
}

int bhx_adjoint_set_launch(bhx_adjoint* A, int32_t q_chunk, int32_t refill) {
    if (!A) return BHX_E_INVALID;
    bhx_context* ctx = A->ctx;
    if (q_chunk < 0 || q_chunk > 4096) return bhx_fail(ctx, BHX_E_INVALID, "queue chunk must be in [0, 4096]");
    if (refill < 0 || refill > 64) return bhx_fail(ctx, BHX_E_INVALID, "refill must be in [0, 64]");
    if (q_chunk) A->q_chunk = q_chunk;
    if (refill) A->refill = refill;
    return BHX_OK;
}

int bhx_adjoint_set_forward_waves(bhx_adjoint* A, int32_t waves) {
    if (!A) return BHX_E_INVALID;
    if (waves < 0 || waves > 16) return bhx_fail(A->ctx, BHX_E_INVALID, "forward waves per SIMD must be in [0, 16]");
    A->fwd_waves = waves;
    const int w = waves > 0 ? std::min(waves, A->fwd_per_cu) : A->fwd_per_cu;
    A->grid_f = w * A->ctx->n_cu;
    return BHX_OK;
}

int bhx_adjoint_get_forward_waves(const bhx_adjoint* A, int32_t* waves) {
    if (!A || !waves) return BHX_E_INVALID;
    *waves = A->grid_f / A->ctx->n_cu;
    return BHX_OK;
}

int bhx_adjoint_info(bhx_adjoint* A, int64_t* ck_bytes, int64_t* lit_rays, int64_t* mismatches,
                     double* fwd_ms, double* bwd_ms) {
    if (!A) return BHX_E_INVALID;
    bhx_context* ctx = A->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipDeviceSynchronize());
    if (A->fwd_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, A->ev[0], A->ev[1]) == hipSuccess) A->fwd_ms = ms;
    }
    unsigned int nd = 0;
    HIPCHK(hipMemcpy(&nd, static_cast<char*>(A->counters.p) + 8, sizeof nd, hipMemcpyDeviceToHost));
    if (ck_bytes) *ck_bytes = static_cast<int64_t>(A->pool.bytes + A->link.bytes);
    if (lit_rays) *lit_rays = nd;
    if (mismatches) *mismatches = A->mismatches;
    if (fwd_ms) *fwd_ms = A->fwd_ms;
    if (bwd_ms) *bwd_ms = A->bwd_ms;
    return BHX_OK;
}

}  // extern "C"
