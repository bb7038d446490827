// png.cpp - CPU oracle of the reference's PNG / APNG writer.
//
// TEST INFRASTRUCTURE ONLY: the checker of the GPU-assisted encoder in
// blackhole_amd/csrc/png.hip.  Nothing in blackhole_amd/ includes, links or
// loads this file.
//
// A restatement of MattPar-Cell/Blackhole src/image.cpp:193-778 (write_png,
// write_apng) that produces the same bytes: per-scanline filter chosen by the
// minimum sum of |signed residual| (first minimum of types 0..4,
// image.cpp:478-509), a zlib stream (header 78 01, Adler-32,
// image.cpp:511-518) holding one DEFLATE block with dynamic Huffman codes
// (image.cpp:484-543) built from a greedy LZ77 parse whose match finder walks
// a 3-byte hash chain (15-bit hash, 192 candidates, 32 KiB window,
// image.cpp:406-478), with a stored-block fallback (image.cpp:546-569), and
// the chunk layout of PNG and APNG 1.0 (image.cpp:571-762).  Written from
// the algorithm; the code shares no text with the reference.

#include "oracle.h"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

using Bytes = std::vector<uint8_t>;

// ---- checksums ------------------------------------------------------------
uint32_t crc_update(uint32_t crc, const uint8_t* p, size_t n) {   // image.cpp:200-216
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t c = k;
            for (int b = 0; b < 8; ++b) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[k] = c;
        }
        ready = true;
    }
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return crc;
}

uint32_t adler(const uint8_t* p, size_t n) {   // image.cpp:218-225
    uint32_t s1 = 1, s2 = 0;
    for (size_t i = 0; i < n; ++i) {
        s1 = (s1 + p[i]) % 65521u;
        s2 = (s2 + s1) % 65521u;
    }
    return (s2 << 16) | s1;
}

// ---- bit output (LSB-first fields, MSB-first Huffman codes; image.cpp:228-250)
struct Bits {
    Bytes out;
    uint32_t acc = 0;
    int nacc = 0;
    void put(uint32_t v, int n) {
        acc |= (v & ((1u << n) - 1u)) << nacc;
        nacc += n;
        while (nacc >= 8) {
            out.push_back(static_cast<uint8_t>(acc));
            acc >>= 8;
            nacc -= 8;
        }
    }
    void put_code(uint32_t code, int len) {
        for (int b = len - 1; b >= 0; --b) put((code >> b) & 1u, 1);
    }
    void pad() {
        while (nacc > 0) put(0, 1);
    }
};

const int kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                          35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const int kLenBits[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2,
                          3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const int kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193,
                           257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145,
                           8193, 12289, 16385, 24577};
const int kDistBits[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6,
                           7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

// ---- Huffman code lengths (image.cpp:270-334) ------------------------------
// Huffman's algorithm over the used symbols; the next two nodes are the
// smallest by (weight, id) with leaf ids -(symbol + 1) and internal nodes
// numbered 0, 1, ... as created.  If the tree is deeper than max_bits, every
// non-zero frequency becomes (f + 1) / 2 and the tree is rebuilt.
void huffman_lengths(std::vector<uint32_t> freq, int max_bits, std::vector<uint8_t>& len) {
    const size_t n = freq.size();
    len.assign(n, 0);
    for (;;) {
        std::vector<int> sym;
        for (size_t i = 0; i < n; ++i)
            if (freq[i]) sym.push_back(static_cast<int>(i));
        if (sym.empty()) return;
        if (sym.size() == 1) {
            len[sym[0]] = 1;
            len[sym[0] == 0 ? 1 : 0] = 1;
            return;
        }
        struct Item { uint64_t w; int id; };
        auto later = [](const Item& a, const Item& b) { return a.w != b.w ? a.w > b.w : a.id > b.id; };
        std::vector<Item> pq;
        for (int s : sym) pq.push_back({freq[s], -s - 1});
        std::vector<std::pair<int, int>> kids;
        while (pq.size() > 1) {
            Item two[2];
            for (Item& t : two) {   // extract the (weight, id) minimum
                auto it = std::min_element(pq.begin(), pq.end(),
                                           [&](const Item& a, const Item& b) { return later(b, a); });
                t = *it;
                pq.erase(it);
            }
            kids.push_back({two[0].id, two[1].id});
            pq.push_back({two[0].w + two[1].w, static_cast<int>(kids.size()) - 1});
        }
        std::fill(len.begin(), len.end(), 0);
        int deepest = 0;
        std::vector<std::pair<int, int>> todo{{pq[0].id, 0}};
        while (!todo.empty()) {
            const std::pair<int, int> nd = todo.back();
            todo.pop_back();
            if (nd.first < 0) {
                const int s = -nd.first - 1;
                len[s] = static_cast<uint8_t>(std::max(1, nd.second));
                deepest = std::max(deepest, static_cast<int>(len[s]));
            } else {
                todo.push_back({kids[nd.first].first, nd.second + 1});
                todo.push_back({kids[nd.first].second, nd.second + 1});
            }
        }
        if (deepest <= max_bits) return;
        for (uint32_t& f : freq)
            if (f) f = (f + 1) >> 1;
    }
}

// canonical codes (RFC 1951 3.2.2; image.cpp:337-352)
void canonical_codes(const std::vector<uint8_t>& len, std::vector<uint32_t>& code) {
    code.assign(len.size(), 0);
    int top = 0;
    for (uint8_t l : len) top = std::max<int>(top, l);
    if (top == 0) return;
    std::vector<uint32_t> count(top + 1, 0), next(top + 2, 0);
    for (uint8_t l : len)
        if (l) ++count[l];
    uint32_t c = 0;
    for (int b = 1; b <= top; ++b) {
        c = (c + count[b - 1]) << 1;
        next[b] = c;
    }
    for (size_t i = 0; i < len.size(); ++i)
        if (len[i]) code[i] = next[len[i]]++;
}

// run-length code of the code lengths with symbols 16 / 17 / 18 (image.cpp:355-384)
struct Rl { int sym, nbits, extra; };
void rle(const std::vector<uint8_t>& len, std::vector<Rl>& out) {
    size_t i = 0;
    while (i < len.size()) {
        const int v = len[i];
        size_t run = 1;
        while (i + run < len.size() && len[i + run] == v) ++run;
        if (v == 0) {
            for (; run >= 11; ) { const size_t k = std::min<size_t>(run, 138); out.push_back({18, 7, int(k) - 11}); run -= k; i += k; }
            for (; run >= 3; ) { const size_t k = std::min<size_t>(run, 10); out.push_back({17, 3, int(k) - 3}); run -= k; i += k; }
            for (; run > 0; --run, ++i) out.push_back({0, 0, 0});
        } else {
            out.push_back({v, 0, 0});
            ++i;
            --run;
            for (; run >= 3; ) { const size_t k = std::min<size_t>(run, 6); out.push_back({16, 2, int(k) - 3}); run -= k; i += k; }
            for (; run > 0; --run, ++i) out.push_back({v, 0, 0});
        }
    }
}

// ---- DEFLATE -------------------------------------------------------------
struct Tok { int sym, lbits, lextra, dsym, dbits, dextra; bool match; };

// greedy LZ77 with a 3-byte hash chain (image.cpp:406-478)
void lz77(const Bytes& d, std::vector<Tok>& toks, std::vector<uint32_t>& lf, std::vector<uint32_t>& df) {
    const size_t n = d.size();
    lf.assign(288, 0);
    df.assign(30, 0);
    std::vector<int> head(1 << 15, -1), prev(n, -1);
    auto hash = [&](size_t i) {
        return ((uint32_t(d[i]) << 10) ^ (uint32_t(d[i + 1]) << 5) ^ uint32_t(d[i + 2])) & 0x7FFFu;
    };
    auto insert = [&](size_t i) {
        const uint32_t h = hash(i);
        prev[i] = head[h];
        head[h] = static_cast<int>(i);
    };
    size_t i = 0;
    while (i < n) {
        int blen = 0, bdist = 0;
        if (i + 2 < n) {
            const size_t cap = std::min<size_t>(258, n - i);
            int c = head[hash(i)];
            for (int steps = 0; c >= 0 && steps < 192; ++steps) {
                const size_t dist = i - static_cast<size_t>(c);
                if (dist == 0 || dist > 32768) break;
                size_t l = 0;
                while (l < cap && d[c + l] == d[i + l]) ++l;
                if (static_cast<int>(l) > blen) {
                    blen = static_cast<int>(l);
                    bdist = static_cast<int>(dist);
                    if (l >= cap) break;
                }
                c = prev[c];
            }
            insert(i);
        }
        Tok t{};
        if (blen >= 3) {
            int lc = 0;
            while (lc < 28 && kLenBase[lc + 1] <= blen) ++lc;
            int dc = 0;
            while (dc < 29 && kDistBase[dc + 1] <= bdist) ++dc;
            t = {257 + lc, kLenBits[lc], blen - kLenBase[lc], dc, kDistBits[dc], bdist - kDistBase[dc], true};
            ++lf[t.sym];
            ++df[dc];
            for (size_t k = i + 1; k < i + blen && k + 2 < n; ++k) insert(k);
            i += blen;
        } else {
            t.sym = d[i];
            ++lf[t.sym];
            ++i;
        }
        toks.push_back(t);
    }
    ++lf[256];
}

Bytes deflate_dynamic(const Bytes& d) {   // image.cpp:484-543
    std::vector<Tok> toks;
    std::vector<uint32_t> lf, df;
    lz77(d, toks, lf, df);
    std::vector<uint8_t> ll, dl;
    huffman_lengths(lf, 15, ll);
    huffman_lengths(df, 15, dl);
    if (std::all_of(dl.begin(), dl.end(), [](uint8_t v) { return v == 0; })) dl[0] = dl[1] = 1;
    std::vector<uint32_t> lc, dc;
    canonical_codes(ll, lc);
    canonical_codes(dl, dc);
    int hlit = 286, hdist = 30;
    while (hlit > 257 && ll[hlit - 1] == 0) --hlit;
    while (hdist > 1 && dl[hdist - 1] == 0) --hdist;
    std::vector<uint8_t> both(ll.begin(), ll.begin() + hlit);
    both.insert(both.end(), dl.begin(), dl.begin() + hdist);
    std::vector<Rl> rl;
    rle(both, rl);
    std::vector<uint32_t> cf(19, 0);
    for (const Rl& r : rl) ++cf[r.sym];
    std::vector<uint8_t> cl;
    huffman_lengths(cf, 7, cl);
    std::vector<uint32_t> cc;
    canonical_codes(cl, cc);
    static const int kOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
    int hclen = 19;
    while (hclen > 4 && cl[kOrder[hclen - 1]] == 0) --hclen;
    Bits w;
    w.put(1, 1);
    w.put(2, 2);
    w.put(hlit - 257, 5);
    w.put(hdist - 1, 5);
    w.put(hclen - 4, 4);
    for (int k = 0; k < hclen; ++k) w.put(cl[kOrder[k]], 3);
    for (const Rl& r : rl) {
        w.put_code(cc[r.sym], cl[r.sym]);
        if (r.nbits) w.put(r.extra, r.nbits);
    }
    for (const Tok& t : toks) {
        w.put_code(lc[t.sym], ll[t.sym]);
        if (t.match) {
            if (t.lbits) w.put(t.lextra, t.lbits);
            w.put_code(dc[t.dsym], dl[t.dsym]);
            if (t.dbits) w.put(t.dextra, t.dbits);
        }
    }
    w.put_code(lc[256], ll[256]);
    w.pad();
    return w.out;
}

Bytes deflate_stored(const Bytes& d) {   // image.cpp:546-569
    Bytes out;
    size_t off = 0;
    do {
        const size_t len = std::min<size_t>(65535, d.size() - off);
        const bool last = off + len >= d.size();
        out.push_back(last ? 1 : 0);
        out.push_back(uint8_t(len));
        out.push_back(uint8_t(len >> 8));
        out.push_back(uint8_t(~len));
        out.push_back(uint8_t(~len >> 8));
        out.insert(out.end(), d.begin() + off, d.begin() + off + len);
        off += len;
        if (last) break;
    } while (off < d.size());
    return out;
}

Bytes deflate(const Bytes& d) {   // image.cpp:572-578
    if (d.empty()) return deflate_stored(d);
    Bytes dyn = deflate_dynamic(d);
    if (dyn.size() >= d.size() + 5) return deflate_stored(d);
    return dyn;
}

// ---- PNG ------------------------------------------------------------------
void be32(Bytes& v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back(uint8_t(x >> s));
}

void chunk(Bytes& png, const char* tag, const Bytes& body) {   // image.cpp:587-594
    be32(png, static_cast<uint32_t>(body.size()));
    const size_t at = png.size();
    png.insert(png.end(), tag, tag + 4);
    png.insert(png.end(), body.begin(), body.end());
    be32(png, crc_update(0xFFFFFFFFu, png.data() + at, png.size() - at) ^ 0xFFFFFFFFu);
}

int paeth(int a, int b, int c) {   // image.cpp:605-610
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

// filtered scanlines wrapped in a zlib stream (image.cpp:598-650)
Bytes image_zlib(const uint8_t* rgb, int w, int h) {
    const size_t stride = static_cast<size_t>(w) * 3;
    Bytes raw;
    raw.reserve((stride + 1) * h);
    Bytes cand(stride), keep(stride);
    for (int y = 0; y < h; ++y) {
        const uint8_t* cur = rgb + y * stride;
        const uint8_t* up = y ? cur - stride : nullptr;
        int best = 0;
        long best_score = -1;
        for (int f = 0; f < 5; ++f) {
            long score = 0;
            for (size_t x = 0; x < stride; ++x) {
                const int a = x >= 3 ? cur[x - 3] : 0, b = up ? up[x] : 0, c = (up && x >= 3) ? up[x - 3] : 0;
                const int pred = f == 0 ? 0 : f == 1 ? a : f == 2 ? b : f == 3 ? (a + b) >> 1 : paeth(a, b, c);
                cand[x] = static_cast<uint8_t>(cur[x] - pred);
                score += std::abs(static_cast<int8_t>(cand[x]));
            }
            if (best_score < 0 || score < best_score) {
                best_score = score;
                best = f;
                keep = cand;
            }
        }
        raw.push_back(static_cast<uint8_t>(best));
        raw.insert(raw.end(), keep.begin(), keep.end());
    }
    Bytes z{0x78, 0x01};
    const Bytes body = deflate(raw);
    z.insert(z.end(), body.begin(), body.end());
    be32(z, adler(raw.data(), raw.size()));
    return z;
}

Bytes ihdr(int w, int h) {
    Bytes b;
    be32(b, static_cast<uint32_t>(w));
    be32(b, static_cast<uint32_t>(h));
    b.insert(b.end(), {8, 2, 0, 0, 0});
    return b;
}

const uint8_t kSig[8] = {137, 'P', 'N', 'G', 13, 10, 26, 10};

int deliver(const Bytes& b, uint8_t* out, size_t cap, size_t* len) {
    *len = b.size();
    if (!out || cap < b.size()) return 1;
    std::memcpy(out, b.data(), b.size());
    return 0;
}

}  // namespace

extern "C" {

int oracle_png_encode(const uint8_t* rgb, int32_t w, int32_t h, uint8_t* out, size_t cap, size_t* len) {
    if (w <= 0 || h <= 0 || !rgb || !len) return -1;
    Bytes png(kSig, kSig + 8);
    chunk(png, "IHDR", ihdr(w, h));
    chunk(png, "IDAT", image_zlib(rgb, w, h));
    chunk(png, "IEND", {});
    return deliver(png, out, cap, len);
}

// write_apng (image.cpp:661-762): frames concatenated, n x (w h 3) bytes
int oracle_apng_encode(const uint8_t* frames, int32_t n, int32_t w, int32_t h, int32_t fps, int32_t loops,
                       uint8_t* out, size_t cap, size_t* len) {
    if (n <= 0 || w <= 0 || h <= 0 || !frames || !len) return -1;
    const size_t fsz = static_cast<size_t>(w) * h * 3, stride = static_cast<size_t>(w) * 3;
    Bytes png(kSig, kSig + 8);
    chunk(png, "IHDR", ihdr(w, h));
    Bytes actl;
    be32(actl, static_cast<uint32_t>(n));
    be32(actl, static_cast<uint32_t>(std::max(0, loops)));
    chunk(png, "acTL", actl);
    uint32_t seq = 0;
    for (int i = 0; i < n; ++i) {
        const uint8_t* cur = frames + i * fsz;
        int fx = 0, fy = 0, fw = w, fh = h;
        if (i > 0) {   // bounding box of the pixels that changed
            const uint8_t* pre = cur - fsz;
            int x0 = w, y0 = h, x1 = -1, y1 = -1;
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x)
                    if (std::memcmp(pre + y * stride + 3 * x, cur + y * stride + 3 * x, 3) != 0) {
                        x0 = std::min(x0, x); x1 = std::max(x1, x);
                        y0 = std::min(y0, y); y1 = std::max(y1, y);
                    }
            if (x1 < x0) { fw = fh = 1; }
            else { fx = x0; fy = y0; fw = x1 - x0 + 1; fh = y1 - y0 + 1; }
        }
        Bytes sub(static_cast<size_t>(fw) * fh * 3);
        for (int y = 0; y < fh; ++y)
            std::memcpy(sub.data() + static_cast<size_t>(y) * fw * 3, cur + (static_cast<size_t>(fy + y) * w + fx) * 3,
                        static_cast<size_t>(fw) * 3);
        Bytes fctl;
        be32(fctl, seq++);
        be32(fctl, fw); be32(fctl, fh); be32(fctl, fx); be32(fctl, fy);
        fctl.insert(fctl.end(), {0, 1, uint8_t(fps >> 8), uint8_t(fps), 0, 0});
        chunk(png, "fcTL", fctl);
        const Bytes z = image_zlib(sub.data(), fw, fh);
        if (i == 0) {
            chunk(png, "IDAT", z);
        } else {
            Bytes fd;
            be32(fd, seq++);
            fd.insert(fd.end(), z.begin(), z.end());
            chunk(png, "fdAT", fd);
        }
    }
    chunk(png, "IEND", {});
    return deliver(png, out, cap, len);
}

// Huffman code lengths of a frequency table (the writer's builder), for the
// property checks of the builder itself.
void oracle_huffman_lengths(const uint32_t* freq, int32_t n, int32_t max_bits, uint8_t* len) {
    std::vector<uint8_t> l;
    huffman_lengths(std::vector<uint32_t>(freq, freq + n), max_bits, l);
    std::memcpy(len, l.data(), l.size());
}

}  // extern "C"
