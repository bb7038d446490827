// png.hip - PNG / APNG writer with the device doing the data-parallel part.
//
// Same bytes as the reference's img::write_png / img::write_apng
// (src/image.cpp:193-778): per-scanline filter of minimum |signed residual|
// sum, one zlib stream holding one dynamic-Huffman DEFLATE block from a greedy
// LZ77 parse over a 3-byte hash chain, stored-block fallback, PNG / APNG 1.0
// chunks.  The split:
//
//   device  k_png_filter   one workgroup per scanline: the five filters'
//                          scores (block reduction), the first minimum, the
//                          filtered row;
//           k_png_hash     the 15-bit hash of every position with 2 bytes
//                          after it (image.cpp:419-423);
//           radix sort     (hash, position) pairs, stable: each hash's
//                          positions in order;
//           k_png_links    prev[p] = the previous position with p's hash;
//           k_png_match    the greedy parse's match search at EVERY position.
//   host    the parse itself, Huffman lengths and canonical codes, the bit
//           stream, Adler-32, CRC-32 and the chunks.
//
// Why the match search can run everywhere at once: the reference inserts
// every position i (with i + 2 < n) into the hash chain exactly once, in
// increasing order, whether i starts a token or lies inside a match
// (image.cpp:438-441, 463-467).  The chain seen from i is therefore "all
// earlier positions with i's hash, newest first", the same for every parse,
// and the best match at i (first longest among at most 192 candidates within
// 32 KiB, image.cpp:426-437) is a function of i alone.  The host's parse then
// reads it at the positions where tokens start.
#include "context.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace bhx {
namespace {

constexpr int kWin = 32768, kMaxMatch = 258, kMaxChain = 192;

__device__ __forceinline__ int paeth_pred(int a, int b, int c) {
    const int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

__device__ __forceinline__ int predictor(int f, int a, int b, int c) {
    return f == 0 ? 0 : f == 1 ? a : f == 2 ? b : f == 3 ? ((a + b) >> 1) : paeth_pred(a, b, c);
}

// one workgroup of 256 per scanline y of a w x h RGB image
__global__ void __launch_bounds__(256) k_png_filter(const uint8_t* __restrict__ rgb, int w,
                                                    uint8_t* __restrict__ raw) {
    __shared__ long long part[4][5];
    __shared__ int pick;
    const int y = blockIdx.x;
    const int64_t stride = 3ll * w;
    const uint8_t* cur = rgb + y * stride;
    const uint8_t* up = y ? cur - stride : nullptr;
    long long s[5] = {0, 0, 0, 0, 0};
    for (int64_t x = threadIdx.x; x < stride; x += blockDim.x) {
        const int v = cur[x], a = x >= 3 ? cur[x - 3] : 0, b = up ? up[x] : 0, c = (up && x >= 3) ? up[x - 3] : 0;
#pragma unroll
        for (int f = 0; f < 5; ++f) s[f] += abs(static_cast<int>(static_cast<int8_t>(static_cast<uint8_t>(v - predictor(f, a, b, c)))));
    }
#pragma unroll
    for (int f = 0; f < 5; ++f)
        for (int off = 32; off > 0; off >>= 1) s[f] += __shfl_down(s[f], off);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int f = 0; f < 5; ++f) part[wv][f] = s[f];
    __syncthreads();
    if (threadIdx.x == 0) {
        int best = 0;
        long long bs = -1;
        for (int f = 0; f < 5; ++f) {
            long long t = 0;
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += part[k][f];
            if (bs < 0 || t < bs) { bs = t; best = f; }   // the first minimum (image.cpp:500-505)
        }
        pick = best;
    }
    __syncthreads();
    const int f = pick;
    uint8_t* o = raw + y * (stride + 1);
    if (threadIdx.x == 0) o[0] = static_cast<uint8_t>(f);
    for (int64_t x = threadIdx.x; x < stride; x += blockDim.x) {
        const int v = cur[x], a = x >= 3 ? cur[x - 3] : 0, b = up ? up[x] : 0, c = (up && x >= 3) ? up[x - 3] : 0;
        o[1 + x] = static_cast<uint8_t>(v - predictor(f, a, b, c));
    }
}

// key = the 15-bit hash of bytes i..i+2, value = i, for i < m = n - 2
__global__ void __launch_bounds__(256) k_png_hash(const uint8_t* __restrict__ d, int64_t m,
                                                  uint16_t* __restrict__ key, int32_t* __restrict__ pos) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    key[i] = static_cast<uint16_t>(((uint32_t(d[i]) << 10) ^ (uint32_t(d[i + 1]) << 5) ^ uint32_t(d[i + 2])) & 0x7FFFu);
    pos[i] = static_cast<int32_t>(i);
}

// prev[p] = the previous position with p's hash (-1 if none), from the
// stably sorted (hash, position) pairs
__global__ void __launch_bounds__(256) k_png_links(const uint16_t* __restrict__ key, const int32_t* __restrict__ pos,
                                                   int64_t m, int32_t* __restrict__ prev) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= m) return;
    prev[pos[k]] = (k > 0 && key[k - 1] == key[k]) ? pos[k - 1] : -1;
}

// the longest match at i among the chain's first 192 candidates within the
// window; the first one found wins ties; a full-length match stops the walk.
// A candidate can only be longer than the best so far if it agrees at the
// best's length, so that one byte is tested before the full comparison (the
// result is the same: a candidate that fails it is at most as long).
__global__ void __launch_bounds__(256) k_png_match(const uint8_t* __restrict__ d, int64_t n,
                                                   const int32_t* __restrict__ prev, uint16_t* __restrict__ mlen,
                                                   uint16_t* __restrict__ mdist) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i + 2 >= n) return;
    const int cap = static_cast<int>(min<int64_t>(kMaxMatch, n - i));
    const uint8_t* q = d + i;
    int best = 0, bdist = 0;
    uint8_t qb = q[0];   // q[best]
    int cand = prev[i];
    for (int chain = 0; cand >= 0 && chain < kMaxChain; ++chain) {
        const int64_t dist = i - cand;
        if (dist > kWin) break;
        const uint8_t* p = d + cand;
        const int next = prev[cand];
        if (p[best] == qb) {
            int l = 0;
            while (l < cap && p[l] == q[l]) ++l;
            if (l > best) {
                best = l;
                bdist = static_cast<int>(dist);
                if (best >= cap) break;
                qb = q[best];
            }
        }
        cand = next;
    }
    mlen[i] = static_cast<uint16_t>(best);
    mdist[i] = static_cast<uint16_t>(bdist);   // <= 32768
}

// bounding box of the pixels that differ between two frames: box = {x0, y0, x1, y1}
__global__ void __launch_bounds__(256) k_apng_box(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b,
                                                  int w, int64_t npix, int* box) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= npix) return;
    if (a[3 * p] != b[3 * p] || a[3 * p + 1] != b[3 * p + 1] || a[3 * p + 2] != b[3 * p + 2]) {
        const int x = static_cast<int>(p % w), y = static_cast<int>(p / w);
        atomicMin(&box[0], x);
        atomicMin(&box[1], y);
        atomicMax(&box[2], x);
        atomicMax(&box[3], y);
    }
}

__global__ void k_box_init(int* box, int w, int h) {
    if (threadIdx.x == 0) { box[0] = w; box[1] = h; box[2] = -1; box[3] = -1; }
}

// the fw x fh rectangle at (fx, fy) of a w-wide frame, packed
__global__ void __launch_bounds__(256) k_apng_crop(const uint8_t* __restrict__ src, int w, int fx, int fy, int fw,
                                                   int64_t nbytes, uint8_t* __restrict__ dst) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= nbytes) return;
    const int64_t row = k / (3ll * fw), col = k - row * 3ll * fw;
    dst[k] = src[((fy + row) * (int64_t)w + fx) * 3 + col];
}

// ---- host: DEFLATE from the device's matches -------------------------------
struct BitSink {
    std::vector<uint8_t> out;
    uint64_t acc = 0;
    int n = 0;
    void put(uint32_t v, int bits) {   // least-significant bit first
        acc |= static_cast<uint64_t>(v & ((1u << bits) - 1u)) << n;
        n += bits;
        while (n >= 8) {
            out.push_back(static_cast<uint8_t>(acc));
            acc >>= 8;
            n -= 8;
        }
    }
    void pad() {
        if (n > 0) put(0, 8 - n);
    }
};

uint32_t reverse_bits(uint32_t v, int len) {
    uint32_t r = 0;
    for (int i = 0; i < len; ++i) r |= ((v >> i) & 1u) << (len - 1 - i);
    return r;
}

// Huffman code lengths with the reference's choice of nodes (image.cpp:270-334):
// the two smallest by (weight, id), leaves id -(symbol + 1), internal nodes
// 0, 1, ... in creation order.  Internal nodes are created with non-decreasing
// weights and increasing ids, so a queue of them in creation order and the
// leaves sorted by (weight, id) give the same sequence as a heap.  Too deep:
// halve (rounding up) the frequencies and rebuild.
void code_lengths(std::vector<uint32_t> freq, int max_bits, std::vector<uint8_t>& len) {
    const int n = static_cast<int>(freq.size());
    len.assign(n, 0);
    for (;;) {
        std::vector<std::pair<uint64_t, int>> leaf;
        for (int s = 0; s < n; ++s)
            if (freq[s]) leaf.push_back({freq[s], -s - 1});
        if (leaf.empty()) return;
        if (leaf.size() == 1) {
            const int s = -leaf[0].second - 1;
            len[s] = 1;
            len[s == 0 ? 1 : 0] = 1;
            return;
        }
        std::sort(leaf.begin(), leaf.end());
        std::vector<std::pair<uint64_t, int>> node;       // internal: (weight, id = index)
        std::vector<std::pair<int, int>> kid;
        size_t li = 0, ni = 0;
        auto take = [&]() {
            const bool from_leaf = li < leaf.size() && (ni >= node.size() || leaf[li] < node[ni]);
            return from_leaf ? leaf[li++] : node[ni++];
        };
        while ((leaf.size() - li) + (node.size() - ni) > 1) {
            const auto a = take(), b = take();
            kid.push_back({a.second, b.second});
            node.push_back({a.first + b.first, static_cast<int>(node.size())});
        }
        // depths: a node's children were created before it
        std::vector<int> depth(node.size(), 0);
        int deepest = 0;
        for (int j = static_cast<int>(node.size()) - 1; j >= 0; --j)
            for (int c : {kid[j].first, kid[j].second}) {
                if (c >= 0) {
                    depth[c] = depth[j] + 1;
                } else {
                    len[-c - 1] = static_cast<uint8_t>(depth[j] + 1);
                    deepest = std::max(deepest, depth[j] + 1);
                }
            }
        if (deepest <= max_bits) return;
        for (uint32_t& f : freq)
            if (f) f = (f + 1) >> 1;
        std::fill(len.begin(), len.end(), 0);
    }
}

// canonical codes (RFC 1951 3.2.2), stored bit-reversed for LSB-first output
void codes_of(const std::vector<uint8_t>& len, std::vector<uint32_t>& rev) {
    int top = 0;
    for (uint8_t l : len) top = std::max<int>(top, l);
    rev.assign(len.size(), 0);
    if (!top) return;
    std::vector<uint32_t> count(top + 1, 0), next(top + 2, 0);
    for (uint8_t l : len)
        if (l) ++count[l];
    uint32_t c = 0;
    for (int b = 1; b <= top; ++b) next[b] = c = (c + count[b - 1]) << 1;
    for (size_t i = 0; i < len.size(); ++i)
        if (len[i]) rev[i] = reverse_bits(next[len[i]]++, len[i]);
}

struct LenSym { uint8_t sym, nbits; uint8_t extra; };

// code lengths run-length coded with 16 / 17 / 18 as the reference does
// (image.cpp:355-384): zero runs take 18 (11..138) then 17 (3..10); other
// values are sent once and repeated with 16 (3..6)
void run_lengths(const std::vector<uint8_t>& len, std::vector<LenSym>& out) {
    size_t i = 0;
    while (i < len.size()) {
        const uint8_t v = len[i];
        size_t run = 1;
        while (i + run < len.size() && len[i + run] == v) ++run;
        i += run;
        if (v != 0) {
            out.push_back({v, 0, 0});
            --run;
        }
        if (v == 0)
            for (; run >= 11; run -= std::min<size_t>(run, 138))
                out.push_back({18, 7, static_cast<uint8_t>(std::min<size_t>(run, 138) - 11)});
        const size_t big = v == 0 ? 10 : 6;
        for (; run >= 3; run -= std::min(run, big))
            out.push_back({static_cast<uint8_t>(v == 0 ? 17 : 16), static_cast<uint8_t>(v == 0 ? 3 : 2),
                           static_cast<uint8_t>(std::min(run, big) - 3)});
        for (; run > 0; --run) out.push_back({v, 0, 0});
    }
}

// length and distance codes (RFC 1951 3.2.5) as lookup tables
struct Tables {
    uint8_t lsym[259], lbits[259];
    uint16_t lextra[259];
    uint8_t dsym[32769];
    Tables() {
        static const int base[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                                     35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
        static const int dbase[31] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385,
                                      513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577, 32769};
        for (int c = 0; c < 29; ++c) {
            const int end = c == 28 ? 259 : base[c + 1];
            for (int l = base[c]; l < end; ++l) {
                lsym[l] = static_cast<uint8_t>(c);
                lextra[l] = static_cast<uint16_t>(l - base[c]);
            }
        }
        for (int c = 0; c < 30; ++c)
            for (int d = dbase[c]; d < dbase[c + 1]; ++d) dsym[d] = static_cast<uint8_t>(c);
    }
};
const int kLenXBits[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const int kDistXBits[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
const int kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385,
                           513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const int kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                          35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};

const Tables& tables() {
    static const Tables t;
    return t;
}

// one dynamic-Huffman block over d (image.cpp:484-543), the parse reading
// the device's match at each token start
std::vector<uint8_t> deflate_block(const std::vector<uint8_t>& d, const std::vector<uint16_t>& mlen,
                                   const std::vector<uint16_t>& mdist) {
    const Tables& T = tables();
    const size_t n = d.size();
    // tokens: literal v < 256, or (length << 16 | distance) tagged by bit 31
    std::vector<uint32_t> tok;
    tok.reserve(n / 4 + 16);
    std::vector<uint32_t> lf(288, 0), df(30, 0);
    for (size_t i = 0; i < n;) {
        if (i + 2 < n && mlen[i] >= 3) {
            const int l = mlen[i], ds = mdist[i];
            tok.push_back(0x80000000u | (uint32_t(l) << 16) | uint32_t(ds - 1));
            ++lf[257 + T.lsym[l]];
            ++df[T.dsym[ds]];
            i += l;
        } else {
            tok.push_back(d[i]);
            ++lf[d[i]];
            ++i;
        }
    }
    ++lf[256];
    std::vector<uint8_t> ll, dl;
    code_lengths(lf, 15, ll);
    code_lengths(df, 15, dl);
    if (std::all_of(dl.begin(), dl.end(), [](uint8_t v) { return v == 0; })) dl[0] = dl[1] = 1;
    std::vector<uint32_t> lc, dc;
    codes_of(ll, lc);
    codes_of(dl, dc);
    int hlit = 286, hdist = 30;
    while (hlit > 257 && ll[hlit - 1] == 0) --hlit;
    while (hdist > 1 && dl[hdist - 1] == 0) --hdist;
    std::vector<uint8_t> lens(ll.begin(), ll.begin() + hlit);
    lens.insert(lens.end(), dl.begin(), dl.begin() + hdist);
    std::vector<LenSym> rl;
    run_lengths(lens, rl);
    std::vector<uint32_t> cf(19, 0);
    for (const LenSym& r : rl) ++cf[r.sym];
    std::vector<uint8_t> cl;
    code_lengths(cf, 7, cl);
    std::vector<uint32_t> cc;
    codes_of(cl, cc);
    static const int kOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
    int hclen = 19;
    while (hclen > 4 && cl[kOrder[hclen - 1]] == 0) --hclen;

    BitSink w;
    w.out.reserve(n / 2 + 64);
    w.put(1, 1);   // BFINAL
    w.put(2, 2);   // dynamic Huffman
    w.put(hlit - 257, 5);
    w.put(hdist - 1, 5);
    w.put(hclen - 4, 4);
    for (int k = 0; k < hclen; ++k) w.put(cl[kOrder[k]], 3);
    for (const LenSym& r : rl) {
        w.put(cc[r.sym], cl[r.sym]);
        if (r.nbits) w.put(r.extra, r.nbits);
    }
    for (uint32_t t : tok) {
        if (!(t & 0x80000000u)) {
            w.put(lc[t], ll[t]);
            continue;
        }
        const int l = (t >> 16) & 0x1FF, ds = int(t & 0xFFFF) + 1;
        const int ls = T.lsym[l], dsy = T.dsym[ds];
        w.put(lc[257 + ls], ll[257 + ls]);
        if (kLenXBits[ls]) w.put(l - kLenBase[ls], kLenXBits[ls]);
        w.put(dc[dsy], dl[dsy]);
        if (kDistXBits[dsy]) w.put(ds - kDistBase[dsy], kDistXBits[dsy]);
    }
    w.put(lc[256], ll[256]);
    w.pad();
    return std::move(w.out);
}

std::vector<uint8_t> stored_blocks(const std::vector<uint8_t>& d) {   // image.cpp:546-569
    std::vector<uint8_t> out;
    size_t off = 0;
    for (;;) {
        const size_t len = std::min<size_t>(65535, d.size() - off);
        const bool last = off + len >= d.size();
        const uint8_t hdr[5] = {static_cast<uint8_t>(last), static_cast<uint8_t>(len), static_cast<uint8_t>(len >> 8),
                                static_cast<uint8_t>(~len), static_cast<uint8_t>(~len >> 8)};
        out.insert(out.end(), hdr, hdr + 5);
        out.insert(out.end(), d.begin() + off, d.begin() + off + len);
        off += len;
        if (last) return out;
    }
}

uint32_t adler32(const uint8_t* p, size_t n) {
    uint32_t a = 1, b = 0;
    while (n) {   // 5552 bytes keep b below 2^32 before the reduction
        const size_t k = std::min<size_t>(n, 5552);
        for (size_t i = 0; i < k; ++i) {
            a += p[i];
            b += a;
        }
        a %= 65521u;
        b %= 65521u;
        p += k;
        n -= k;
    }
    return (b << 16) | a;
}

uint32_t crc32(const uint8_t* p, size_t n) {
    static uint32_t tab[256];
    static const bool ready = [] {
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t c = k;
            for (int j = 0; j < 8; ++j) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
            tab[k] = c;
        }
        return true;
    }();
    (void)ready;
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = tab[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    const uint8_t b[4] = {uint8_t(x >> 24), uint8_t(x >> 16), uint8_t(x >> 8), uint8_t(x)};
    v.insert(v.end(), b, b + 4);
}

void add_chunk(std::vector<uint8_t>& png, const char* tag, const uint8_t* body, size_t n) {
    put_be32(png, static_cast<uint32_t>(n));
    const size_t at = png.size();
    png.insert(png.end(), tag, tag + 4);
    if (n) png.insert(png.end(), body, body + n);
    put_be32(png, crc32(png.data() + at, 4 + n));
}

void add_chunk(std::vector<uint8_t>& png, const char* tag, const std::vector<uint8_t>& body) {
    add_chunk(png, tag, body.data(), body.size());
}

const uint8_t kSignature[8] = {137, 'P', 'N', 'G', 13, 10, 26, 10};

std::vector<uint8_t> header_body(int w, int h) {
    std::vector<uint8_t> b;
    put_be32(b, static_cast<uint32_t>(w));
    put_be32(b, static_cast<uint32_t>(h));
    const uint8_t tail[5] = {8, 2, 0, 0, 0};   // 8-bit truecolour, deflate, adaptive filters, no interlace
    b.insert(b.end(), tail, tail + 5);
    return b;
}

// The zlib stream of a w x h device image (image.cpp:598-650): filter and
// match search on the device, the rest here.  Synchronises the stream.
int image_stream(bhx_context* ctx, const uint8_t* rgb, int w, int h, hipStream_t st, std::vector<uint8_t>& z) {
    const int64_t stride = 3ll * w, n = (stride + 1) * h, m = n >= 3 ? n - 2 : 0;
    HIPCHK(ctx->png_raw.ensure(n));
    HIPCHK(ctx->png_key.ensure(4 * std::max<int64_t>(m, 1)));   // keys in and out (u16)
    HIPCHK(ctx->png_pos.ensure(8 * std::max<int64_t>(m, 1)));   // positions in and out
    HIPCHK(ctx->png_prev.ensure(4 * std::max<int64_t>(n, 1)));
    HIPCHK(ctx->png_match.ensure(4 * std::max<int64_t>(n, 1))); // lengths then distances (u16)
    uint8_t* raw = static_cast<uint8_t*>(ctx->png_raw.p);
    uint16_t* key = static_cast<uint16_t*>(ctx->png_key.p);
    int32_t* pos = static_cast<int32_t*>(ctx->png_pos.p);
    int32_t* prev = static_cast<int32_t*>(ctx->png_prev.p);
    uint16_t* mlen = static_cast<uint16_t*>(ctx->png_match.p);
    uint16_t* mdist = mlen + n;
    hipLaunchKernelGGL(k_png_filter, dim3(h), dim3(256), 0, st, rgb, w, raw);
    HIPCHK(hipGetLastError());
    const unsigned gm = static_cast<unsigned>((m + 255) / 256), gn = static_cast<unsigned>((n + 255) / 256);
    if (m > 0) {
        hipLaunchKernelGGL(k_png_hash, dim3(gm), dim3(256), 0, st, raw, m, key, pos);
        HIPCHK(hipGetLastError());
        size_t tmp = 0;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key, key + m, pos, pos + m, static_cast<int>(m), 0,
                                                  15, st));
        HIPCHK(ctx->png_tmp.ensure(tmp));
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(ctx->png_tmp.p, tmp, key, key + m, pos, pos + m,
                                                  static_cast<int>(m), 0, 15, st));
        hipLaunchKernelGGL(k_png_links, dim3(gm), dim3(256), 0, st, key + m, pos + m, m, prev);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_png_match, dim3(gn), dim3(256), 0, st, raw, n, prev, mlen, mdist);
        HIPCHK(hipGetLastError());
    }
    std::vector<uint8_t> d(n);
    std::vector<uint16_t> ml(n, 0), md(n, 0);
    HIPCHK(hipMemcpyAsync(d.data(), raw, n, hipMemcpyDeviceToHost, st));
    if (m > 0) {
        HIPCHK(hipMemcpyAsync(ml.data(), mlen, 2 * m, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(md.data(), mdist, 2 * m, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    std::vector<uint8_t> body = deflate_block(d, ml, md);
    if (body.size() >= d.size() + 5) body = stored_blocks(d);   // image.cpp:572-578
    z.clear();
    z.reserve(body.size() + 6);
    z.push_back(0x78);
    z.push_back(0x01);
    z.insert(z.end(), body.begin(), body.end());
    put_be32(z, adler32(d.data(), d.size()));
    return BHX_OK;
}

int deliver(bhx_context* ctx, const std::vector<uint8_t>& png, bhx_bytes* out) {
    uint8_t* p = static_cast<uint8_t*>(std::malloc(png.size()));
    if (!p) return bhx_fail(ctx, BHX_E_NOMEM, "out of host memory for the PNG stream");
    std::memcpy(p, png.data(), png.size());
    out->data = p;
    out->size = png.size();
    return BHX_OK;
}

int check_image(bhx_context* ctx, int w, int h) {
    if (w <= 0 || h <= 0) return bhx_fail(ctx, BHX_E_INVALID, "width/height must be > 0");
    if ((3ll * w + 1) * h > (int64_t(1) << 31) - 16)
        return bhx_fail(ctx, BHX_E_TOO_LARGE, "image data above 2 GiB");
    return BHX_OK;
}

int png_dev(bhx_context* ctx, const uint8_t* rgb, int w, int h, hipStream_t st, bhx_bytes* out) {
    std::vector<uint8_t> z;
    int rc = image_stream(ctx, rgb, w, h, st, z);
    if (rc) return rc;
    std::vector<uint8_t> png(kSignature, kSignature + 8);
    add_chunk(png, "IHDR", header_body(w, h));
    add_chunk(png, "IDAT", z);
    add_chunk(png, "IEND", nullptr, 0);
    return deliver(ctx, png, out);
}

// write_apng (image.cpp:661-762) over device frames
int apng_dev(bhx_context* ctx, const uint8_t* const* frames, int n, int w, int h, int fps, int loops, hipStream_t st,
             bhx_bytes* out) {
    const int64_t npix = int64_t(w) * h;
    HIPCHK(ctx->png_box.ensure(4 * sizeof(int)));
    int* box = static_cast<int*>(ctx->png_box.p);
    std::vector<uint8_t> png(kSignature, kSignature + 8);
    add_chunk(png, "IHDR", header_body(w, h));
    std::vector<uint8_t> actl;
    put_be32(actl, static_cast<uint32_t>(n));
    put_be32(actl, static_cast<uint32_t>(std::max(0, loops)));
    add_chunk(png, "acTL", actl);
    uint32_t seq = 0;
    std::vector<uint8_t> z;
    for (int i = 0; i < n; ++i) {
        int fx = 0, fy = 0, fw = w, fh = h;
        if (i > 0) {
            hipLaunchKernelGGL(k_box_init, dim3(1), dim3(64), 0, st, box, w, h);
            HIPCHK(hipGetLastError());
            hipLaunchKernelGGL(k_apng_box, dim3(static_cast<unsigned>((npix + 255) / 256)), dim3(256), 0, st,
                               frames[i - 1], frames[i], w, npix, box);
            HIPCHK(hipGetLastError());
            int b[4];
            HIPCHK(hipMemcpyAsync(b, box, sizeof b, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (b[2] < b[0]) {   // no pixel changed: one pixel is stored
                fw = fh = 1;
            } else {
                fx = b[0]; fy = b[1]; fw = b[2] - b[0] + 1; fh = b[3] - b[1] + 1;
            }
        }
        const uint8_t* src = frames[i];
        if (!(fx == 0 && fy == 0 && fw == w && fh == h)) {
            const int64_t nb = 3ll * fw * fh;
            HIPCHK(ctx->png_sub.ensure(nb));
            hipLaunchKernelGGL(k_apng_crop, dim3(static_cast<unsigned>((nb + 255) / 256)), dim3(256), 0, st, frames[i],
                               w, fx, fy, fw, nb, static_cast<uint8_t*>(ctx->png_sub.p));
            HIPCHK(hipGetLastError());
            src = static_cast<const uint8_t*>(ctx->png_sub.p);
        }
        std::vector<uint8_t> fctl;
        put_be32(fctl, seq++);
        put_be32(fctl, static_cast<uint32_t>(fw));
        put_be32(fctl, static_cast<uint32_t>(fh));
        put_be32(fctl, static_cast<uint32_t>(fx));
        put_be32(fctl, static_cast<uint32_t>(fy));
        const uint8_t tail[6] = {0, 1, uint8_t((fps >> 8) & 0xFF), uint8_t(fps & 0xFF), 0, 0};   // delay 1/fps, dispose NONE, blend SOURCE
        fctl.insert(fctl.end(), tail, tail + 6);
        add_chunk(png, "fcTL", fctl);
        int rc = image_stream(ctx, src, fw, fh, st, z);
        if (rc) return rc;
        if (i == 0) {
            add_chunk(png, "IDAT", z);
        } else {
            std::vector<uint8_t> fd;
            put_be32(fd, seq++);
            fd.insert(fd.end(), z.begin(), z.end());
            add_chunk(png, "fdAT", fd);
        }
    }
    add_chunk(png, "IEND", nullptr, 0);
    return deliver(ctx, png, out);
}

}  // namespace
}  // namespace bhx

using namespace bhx;

extern "C" {

int bhx_png_encode_dev(bhx_context* ctx, const uint8_t* rgb, int32_t width, int32_t height, void* stream,
                       bhx_bytes* out) {
    if (!ctx) return BHX_E_INVALID;
    if (!rgb || !out) return bhx_fail(ctx, BHX_E_INVALID, "rgb / out is NULL");
    int rc = check_image(ctx, width, height);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    return png_dev(ctx, rgb, width, height, static_cast<hipStream_t>(stream), out);
}

int bhx_png_encode(bhx_context* ctx, const uint8_t* rgb, int32_t width, int32_t height, bhx_bytes* out) {
    if (!ctx) return BHX_E_INVALID;
    if (!rgb || !out) return bhx_fail(ctx, BHX_E_INVALID, "rgb / out is NULL");
    int rc = check_image(ctx, width, height);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t nb = size_t(3) * width * height;
    HIPCHK(ctx->png_in.ensure(nb));
    HIPCHK(hipMemcpy(ctx->png_in.p, rgb, nb, hipMemcpyHostToDevice));
    return png_dev(ctx, static_cast<const uint8_t*>(ctx->png_in.p), width, height, nullptr, out);
}

int bhx_apng_encode_dev(bhx_context* ctx, const uint8_t* const* frames, int32_t n_frames, int32_t width,
                        int32_t height, int32_t fps, int32_t loops, void* stream, bhx_bytes* out) {
    if (!ctx) return BHX_E_INVALID;
    if (!frames || !out || n_frames <= 0) return bhx_fail(ctx, BHX_E_INVALID, "no frames / out is NULL");
    for (int32_t i = 0; i < n_frames; ++i)
        if (!frames[i]) return bhx_fail(ctx, BHX_E_INVALID, "a frame pointer is NULL");
    int rc = check_image(ctx, width, height);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    return apng_dev(ctx, frames, n_frames, width, height, fps, loops, static_cast<hipStream_t>(stream), out);
}

int bhx_apng_encode(bhx_context* ctx, const uint8_t* const* frames, int32_t n_frames, int32_t width,
                    int32_t height, int32_t fps, int32_t loops, bhx_bytes* out) {
    if (!ctx) return BHX_E_INVALID;
    if (!frames || !out || n_frames <= 0) return bhx_fail(ctx, BHX_E_INVALID, "no frames / out is NULL");
    for (int32_t i = 0; i < n_frames; ++i)
        if (!frames[i]) return bhx_fail(ctx, BHX_E_INVALID, "a frame pointer is NULL");
    int rc = check_image(ctx, width, height);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t nb = size_t(3) * width * height;
    HIPCHK(ctx->png_in.ensure(nb * n_frames));
    std::vector<const uint8_t*> dev(n_frames);
    for (int32_t i = 0; i < n_frames; ++i) {
        uint8_t* p = static_cast<uint8_t*>(ctx->png_in.p) + nb * i;
        HIPCHK(hipMemcpy(p, frames[i], nb, hipMemcpyHostToDevice));
        dev[i] = p;
    }
    return apng_dev(ctx, dev.data(), n_frames, width, height, fps, loops, nullptr, out);
}

void bhx_bytes_free(bhx_bytes* b) {
    if (!b) return;
    std::free(b->data);
    b->data = nullptr;
    b->size = 0;
}

}  // extern "C"
