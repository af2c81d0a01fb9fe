// Image input: PNG (zlib/DEFLATE, the five scanline filters, bit depths 1-16,
// grey/RGB/palette/grey+alpha/RGBA) and Netpbm (P2/P3/P5/P6) decoding on the
// host, then one GPU kernel that turns the samples into the reference's
// float64 gray and RGB planes (read_pnm / read_png / read_image,
// stereonet.py:236-324, :3190-3613).
//
// The decoder writes, per pixel, three integer channel values and a scale
// rule: value * (1 / maxval) (every sample type) or value / 255 (palette
// entries, which the reference divides).  The GPU kernel applies the rule
// and gray = 0.299 r + 0.587 g + 0.114 b in float64, without contraction,
// except for one-channel Netpbm files, whose gray is the scaled sample
// itself (read_pnm keeps the samples as gray), so gray and RGB equal the
// reference's bit for bit.
//
// DEFLATE (RFC 1951): stored, fixed-Huffman and dynamic-Huffman blocks;
// canonical codes decoded one bit at a time through per-length first-code
// counts (the count/symbol form of the RFC's code construction).
#include <algorithm>
#include <new>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

namespace {

// ---------------------------------------------------------------- inflate
// LSB-first bit reader over a 64-bit buffer; reads past the end supply
// zero bytes and are detected afterwards from the count of bits consumed.
struct Bits {
    const uint8_t* d;
    size_t n, pos = 0;
    uint64_t buf = 0;
    int cnt = 0;
    void refill() {
        if (pos + 8 <= n) {  // whole-word refill: top up to 56..63 bits
            uint64_t w;
            std::memcpy(&w, d + pos, 8);
            buf |= w << cnt;
            pos += (size_t)((63 - cnt) >> 3);
            cnt |= 56;
            return;
        }
        while (cnt <= 56) {
            buf |= (uint64_t)(pos < n ? d[pos] : 0) << cnt;
            ++pos;
            cnt += 8;
        }
    }
    int get(int k) {  // k <= 16
        if (cnt < k) refill();
        const int v = (int)(buf & ((1u << k) - 1u));
        buf >>= k;
        cnt -= k;
        return v;
    }
    bool bad() const { return 8 * pos - (size_t)cnt > 8 * n; }
    void align() {  // drop to the byte boundary; unread whole bytes go back
        buf = 0;
        pos -= (size_t)(cnt >> 3);
        cnt = 0;
    }
};

constexpr int kLut = 10;  // codes up to this length resolve in one lookup

struct Code {
    int count[16];          // codes of each length
    int symbol[320];        // symbols ordered by (length, value)
    uint16_t lut[1 << kLut];  // next kLut stream bits -> symbol << 4 | length
};

// returns false for an over-subscribed code
bool build(Code& c, const int* len, int n) {
    std::memset(c.count, 0, sizeof(c.count));
    for (int s = 0; s < n; ++s) c.count[len[s]]++;
    c.count[0] = 0;
    int left = 1;
    for (int l = 1; l < 16; ++l) {
        left = (left << 1) - c.count[l];
        if (left < 0) return false;
    }
    int offs[16];
    offs[1] = 0;
    for (int l = 1; l < 15; ++l) offs[l + 1] = offs[l] + c.count[l];
    for (int s = 0; s < n; ++s)
        if (len[s]) c.symbol[offs[len[s]]++] = s;
    // canonical code values (RFC 1951, 3.2.2), bit-reversed into the table
    std::memset(c.lut, 0, sizeof(c.lut));
    int next[16];
    int code = 0;
    next[0] = 0;
    for (int l = 1; l < 16; ++l) {
        code = (code + c.count[l - 1]) << 1;
        next[l] = code;
    }
    for (int s = 0; s < n; ++s) {
        const int l = len[s];
        if (!l) continue;
        const int v = next[l]++;
        if (l > kLut) continue;
        int r = 0;
        for (int k = 0; k < l; ++k) r |= ((v >> k) & 1) << (l - 1 - k);
        for (int j = r; j < (1 << kLut); j += 1 << l) c.lut[j] = (uint16_t)(s << 4 | l);
    }
    return true;
}

// codes arrive most-significant bit first within the LSB-first stream:
// one table lookup for short codes, else bit by bit over the length counts
int decode(Bits& b, const Code& c) {
    if (b.cnt < 15) b.refill();
    const int e = c.lut[b.buf & ((1u << kLut) - 1u)];
    if (e) {
        b.buf >>= (e & 15);
        b.cnt -= (e & 15);
        return e >> 4;
    }
    int code = 0, first = 0, index = 0;
    for (int l = 1; l < 16; ++l) {
        code |= b.get(1);
        const int cnt = c.count[l];
        if (code - first < cnt) return c.symbol[index + (code - first)];
        index += cnt;
        first = (first + cnt) << 1;
        code <<= 1;
    }
    return -1;
}

// RFC 1951, 3.2.5
const int kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                          35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const int kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2,
                           3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const int kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129,
                           193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097,
                           6145, 8193, 12289, 16385, 24577};
const int kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6,
                            6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
const int kClenOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

bool codes(Bits& b, const Code& lit, const Code& dist, std::vector<uint8_t>& out, size_t limit,
           std::string& err) {
    for (;;) {
        const int sym = decode(b, lit);
        if (sym < 0) {
            err = "invalid Huffman code in DEFLATE stream";
            return false;
        }
        if (sym < 256) {
            if (out.size() >= limit) {
                err = "DEFLATE output far beyond the image size";
                return false;
            }
            out.push_back((uint8_t)sym);
        } else if (sym == 256) {
            if (b.bad()) {
                err = "truncated DEFLATE stream";
                return false;
            }
            return true;
        } else {
            const int si = sym - 257;
            if (si >= 29) {
                err = "invalid length symbol";
                return false;
            }
            const int length = kLenBase[si] + b.get(kLenExtra[si]);
            const int ds = decode(b, dist);
            if (ds < 0 || ds >= 30) {
                err = "invalid distance symbol";
                return false;
            }
            const size_t distance = (size_t)(kDistBase[ds] + b.get(kDistExtra[ds]));
            if (b.bad()) {
                err = "truncated DEFLATE stream";
                return false;
            }
            const size_t o = out.size();
            if (o + length > limit) {
                err = "DEFLATE output far beyond the image size";
                return false;
            }
            if (distance > o) {
                err = "back-reference before start of output";
                return false;
            }
            out.resize(o + length);
            uint8_t* q = out.data() + o;
            const uint8_t* f = q - distance;
            for (int k = 0; k < length; ++k) q[k] = f[k];  // forward: overlap repeats
        }
    }
}

// limit: the most output accepted (a bound on memory for corrupt or hostile
// streams; a valid image needs far less)
bool inflate(const uint8_t* d, size_t n, std::vector<uint8_t>& out, size_t limit,
             std::string& err) {
    Bits b{d, n};
    int last = 0;
    while (!last) {
        last = b.get(1);
        const int type = b.get(2);
        if (b.bad()) {
            err = "truncated DEFLATE stream";
            return false;
        }
        if (type == 0) {
            b.align();
            if (b.pos + 4 > n) {
                err = "truncated stored block";
                return false;
            }
            const unsigned len = d[b.pos] | (d[b.pos + 1] << 8);
            const unsigned nlen = d[b.pos + 2] | (d[b.pos + 3] << 8);
            if (len != (~nlen & 0xFFFFu)) {
                err = "stored block length check failed";
                return false;
            }
            b.pos += 4;
            if (b.pos + len > n) {
                err = "truncated stored block";
                return false;
            }
            if (out.size() + len > limit) {
                err = "DEFLATE output far beyond the image size";
                return false;
            }
            out.insert(out.end(), d + b.pos, d + b.pos + len);
            b.pos += len;
        } else if (type == 1) {
            struct Fixed {
                Code lit, dist;
                Fixed() {
                    int l[288];
                    for (int s = 0; s < 288; ++s) l[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : 8;
                    build(lit, l, 288);
                    for (int s = 0; s < 30; ++s) l[s] = 5;
                    build(dist, l, 30);
                }
            };
            static const Fixed fixed;  // built once (thread-safe static initialisation)
            if (!codes(b, fixed.lit, fixed.dist, out, limit, err)) return false;
        } else if (type == 2) {
            const int hlit = b.get(5) + 257, hdist = b.get(5) + 1, hclen = b.get(4) + 4;
            int cl[19] = {0};
            for (int i = 0; i < hclen; ++i) cl[kClenOrder[i]] = b.get(3);
            Code clc;
            if (!build(clc, cl, 19)) {
                err = "bad code-length code";
                return false;
            }
            int lens[320];
            int k = 0;
            while (k < hlit + hdist) {
                const int sym = decode(b, clc);
                if (sym < 0 || b.bad()) {
                    err = "invalid code-length symbol";
                    return false;
                }
                if (sym < 16) {
                    lens[k++] = sym;
                } else {
                    int rep = 0, val = 0;
                    if (sym == 16) {
                        if (k == 0) {
                            err = "repeat with no previous length";
                            return false;
                        }
                        val = lens[k - 1];
                        rep = 3 + b.get(2);
                    } else if (sym == 17) {
                        rep = 3 + b.get(3);
                    } else {
                        rep = 11 + b.get(7);
                    }
                    if (k + rep > hlit + hdist) {
                        err = "too many code lengths";
                        return false;
                    }
                    while (rep--) lens[k++] = val;
                }
            }
            Code lit, dist;
            if (!build(lit, lens, hlit) || !build(dist, lens + hlit, hdist)) {
                err = "over-subscribed Huffman code";
                return false;
            }
            if (!codes(b, lit, dist, out, limit, err)) return false;
        } else {
            err = "reserved DEFLATE block type 3";
            return false;
        }
    }
    return true;
}

// ------------------------------------------------------------- PNG / PNM
constexpr long long kMaxPixels = 1LL << 28;  // 268 M pixels: 1.6 GB of samples
struct Decoded {
    int W = 0, H = 0;
    // 0: value * (1 / maxval), 1: value / 255 (palette), 2: as 0 with gray =
    // the first channel (one-channel Netpbm)
    int mode = 0;
    int maxval = 255;
    std::vector<uint16_t> s;  // 3 values per pixel
};

uint32_t be32(const uint8_t* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

// the Paeth predictor without branches (unpredictable on image data): with
// p = a + b - c, |p - a| = |b - c|, |p - b| = |a - c|, |p - c| = |a + b - 2c|;
// a wins ties with both, then b with c
inline int paeth(int a, int b, int c) {
    int pa = b - c, pb = a - c, pc = pa + pb;
    pa = pa < 0 ? -pa : pa;
    pb = pb < 0 ? -pb : pb;
    pc = pc < 0 ? -pc : pc;
    const int mbc = -(int)(pb <= pc);
    const int r = c ^ ((b ^ c) & mbc);
    const int ma = -(int)((pa <= pb) & (pa <= pc));
    return r ^ ((a ^ r) & ma);
}

// one Paeth-filtered row with bpp fixed: the left and upper-left bytes of
// each channel stay in registers (no reload of the byte just written)
template <int BPP>
void paeth_row(const uint8_t* row, const uint8_t* up, uint8_t* cur, size_t stride) {
    int a[BPP], c[BPP];
#pragma unroll
    for (int k = 0; k < BPP; ++k) a[k] = c[k] = 0;
    for (size_t i = 0; i + BPP <= stride; i += BPP) {
#pragma unroll
        for (int k = 0; k < BPP; ++k) {
            const int u = up[i + k];
            const int v = (row[i + k] + paeth(a[k], u, c[k])) & 0xFF;
            cur[i + k] = (uint8_t)v;
            a[k] = v;
            c[k] = u;
        }
    }
}

bool png(const uint8_t* d, size_t n, Decoded& o, std::string& err, bool header_only) {
    size_t pos = 8;
    int depth = 0, ctype = 0, interlace = 0;
    const uint8_t* pal = nullptr;
    size_t npal = 0;
    std::vector<uint8_t> idat;
    bool ihdr = false;
    while (pos + 8 <= n) {
        // a chunk running past the end of the file keeps the bytes that are
        // present (read_png slices the body and carries on, stereonet.py:
        // 3521-3524); missing image data fails later as too short
        const uint32_t len = (uint32_t)std::min<size_t>(be32(d + pos), n - pos - 8);
        const size_t step = 12 + (size_t)be32(d + pos);
        const uint8_t* type = d + pos + 4;
        const uint8_t* body = d + pos + 8;
        if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
            o.W = (int)be32(body);
            o.H = (int)be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
            ihdr = true;
            if (header_only) break;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            pal = body;
            npal = len;
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += step;
    }
    if (!ihdr || o.W <= 0 || o.H <= 0) {
        err = "PNG without a valid IHDR";
        return false;
    }
    if ((long long)o.W * o.H > kMaxPixels) {
        err = "image too large";
        return false;
    }
    if (interlace) {
        err = "Adam7 interlaced PNG is not supported; re-save without interlacing";
        return false;
    }
    if (depth != 1 && depth != 2 && depth != 4 && depth != 8 && depth != 16) {
        err = "unsupported PNG bit depth " + std::to_string(depth);
        return false;
    }
    const int channels = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (!channels) {
        err = "unsupported PNG colour type " + std::to_string(ctype);
        return false;
    }
    if (header_only) return true;
    // zlib wrapper, RFC 1950
    if (idat.size() < 2 || (idat[0] & 0x0F) != 8 || ((idat[0] << 8) | idat[1]) % 31 != 0 ||
        (idat[1] & 0x20)) {
        err = "bad or unsupported zlib header";
        return false;
    }
    std::vector<uint8_t> raw;
    const size_t need = ((size_t)o.W * channels * depth + 7) / 8 * o.H + o.H;
    raw.reserve(need);
    if (!inflate(idat.data() + 2, idat.size() - 2, raw, 4 * need + (1u << 20), err)) return false;
    const size_t stride = ((size_t)o.W * channels * depth + 7) / 8;
    const int bpp = std::max(1, channels * depth / 8);
    if (raw.size() < (stride + 1) * (size_t)o.H) {
        err = "PNG image data too short";
        return false;
    }
    std::vector<uint8_t> img((size_t)o.H * stride);
    const size_t B = (size_t)bpp;
    for (int y = 0; y < o.H; ++y) {
        const uint8_t f = raw[(size_t)y * (stride + 1)];
        const uint8_t* row = raw.data() + (size_t)y * (stride + 1) + 1;
        uint8_t* cur = img.data() + (size_t)y * stride;
        const uint8_t* up = y ? cur - stride : nullptr;
        switch (f) {
            case 0:
                std::memcpy(cur, row, stride);
                break;
            case 1:
                for (size_t i = 0; i < stride; ++i) cur[i] = (uint8_t)(row[i] + (i >= B ? cur[i - B] : 0));
                break;
            case 2:
                for (size_t i = 0; i < stride; ++i) cur[i] = (uint8_t)(row[i] + (up ? up[i] : 0));
                break;
            case 3:
                for (size_t i = 0; i < stride; ++i) {
                    const int a = i >= B ? cur[i - B] : 0, u = up ? up[i] : 0;
                    cur[i] = (uint8_t)(row[i] + ((a + u) >> 1));
                }
                break;
            case 4:
                if (!up) {  // first row: the predictor is the left byte
                    for (size_t i = 0; i < stride; ++i)
                        cur[i] = (uint8_t)(row[i] + paeth(i >= B ? cur[i - B] : 0, 0, 0));
                    break;
                }
                if (stride % B == 0 && (B == 3 || B == 4 || B == 6 || B == 8)) {
                    // whole pixels per row (8- and 16-bit colour)
                    if (B == 3) paeth_row<3>(row, up, cur, stride);
                    else if (B == 4) paeth_row<4>(row, up, cur, stride);
                    else if (B == 6) paeth_row<6>(row, up, cur, stride);
                    else paeth_row<8>(row, up, cur, stride);
                    break;
                }
                for (size_t i = 0; i < B && i < stride; ++i)
                    cur[i] = (uint8_t)(row[i] + paeth(0, up[i], 0));
                for (size_t i = B; i < stride; ++i)
                    cur[i] = (uint8_t)(row[i] + paeth(cur[i - B], up[i], up[i - B]));
                break;
            default:
                err = "unknown PNG filter type " + std::to_string(f) + " on row " + std::to_string(y);
                return false;
        }
    }
    o.maxval = depth == 16 ? 65535 : (1 << depth) - 1;
    o.mode = ctype == 3 ? 1 : 0;
    if (ctype == 3 && !pal) {
        err = "indexed PNG with no PLTE chunk";
        return false;
    }
    o.s.resize((size_t)3 * o.W * o.H);
    for (int y = 0; y < o.H; ++y) {
        const uint8_t* r = img.data() + (size_t)y * stride;
        uint16_t* o0 = o.s.data() + 3 * (size_t)y * o.W;
        if (depth == 8 && (ctype == 2 || ctype == 6)) {  // the common 8-bit colour files
            for (int x = 0; x < o.W; ++x) {
                const uint8_t* p = r + (size_t)x * channels;
                o0[3 * x] = p[0];
                o0[3 * x + 1] = p[1];
                o0[3 * x + 2] = p[2];
            }
            continue;
        }
        for (int x = 0; x < o.W; ++x) {
            unsigned s[4] = {0, 0, 0, 0};
            if (depth == 8) {
                for (int c = 0; c < channels; ++c) s[c] = r[(size_t)x * channels + c];
            } else if (depth == 16) {
                for (int c = 0; c < channels; ++c)
                    s[c] = (r[((size_t)x * channels + c) * 2] << 8) | r[((size_t)x * channels + c) * 2 + 1];
            } else {
                const size_t bit = (size_t)x * depth;
                const int shift = 8 - depth - (int)(bit & 7);
                s[0] = (r[bit >> 3] >> shift) & ((1u << depth) - 1u);
            }
            uint16_t* o3 = o.s.data() + 3 * ((size_t)y * o.W + x);
            if (ctype == 3) {
                const size_t p = 3 * (size_t)s[0];
                if (p + 3 > npal) {
                    err = "palette index out of range";
                    return false;
                }
                o3[0] = pal[p];
                o3[1] = pal[p + 1];
                o3[2] = pal[p + 2];
            } else if (ctype == 0 || ctype == 4) {
                o3[0] = o3[1] = o3[2] = (uint16_t)s[0];
            } else {
                o3[0] = (uint16_t)s[0];
                o3[1] = (uint16_t)s[1];
                o3[2] = (uint16_t)s[2];
            }
        }
    }
    return true;
}

bool pnm(const uint8_t* d, size_t n, Decoded& o, std::string& err, bool header_only) {
    if (n < 2 || d[0] != 'P' || (d[1] != '2' && d[1] != '3' && d[1] != '5' && d[1] != '6')) {
        err = "not a PGM/PPM file";
        return false;
    }
    const bool ascii = d[1] == '2' || d[1] == '3';
    const int ch = (d[1] == '3' || d[1] == '6') ? 3 : 1;
    long tok[3];
    int nt = 0;
    size_t i = 2;
    auto ws = [](uint8_t c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; };
    while (nt < 3 && i < n) {
        if (d[i] == '#') {
            while (i < n && d[i] != '\n' && d[i] != '\r') ++i;
        } else if (ws(d[i])) {
            ++i;
        } else {
            long v = 0;
            while (i < n && !ws(d[i])) {
                if (d[i] < '0' || d[i] > '9' || v > (1L << 31)) {
                    err = "bad Netpbm header";
                    return false;
                }
                v = v * 10 + (d[i] - '0');
                ++i;
            }
            tok[nt++] = v;
        }
    }
    if (nt < 3 || tok[0] <= 0 || tok[1] <= 0 || tok[2] <= 0 || tok[2] > 65535) {
        err = "bad Netpbm header";
        return false;
    }
    if (tok[0] > kMaxPixels || tok[1] > kMaxPixels || tok[0] * tok[1] > kMaxPixels) {
        err = "image too large";
        return false;
    }
    o.W = (int)tok[0];
    o.H = (int)tok[1];
    o.maxval = (int)tok[2];
    o.mode = ch == 1 ? 2 : 0;
    if (header_only) return true;
    const size_t count = (size_t)o.W * o.H * ch;
    std::vector<uint16_t> v(count);
    if (ascii) {
        size_t k = 0;
        while (k < count && i < n) {
            while (i < n && ws(d[i])) ++i;
            if (i >= n) break;
            long x = 0;
            while (i < n && !ws(d[i])) {
                if (d[i] < '0' || d[i] > '9' || x > 65535) {
                    err = "bad Netpbm sample";
                    return false;
                }
                x = x * 10 + (d[i++] - '0');
            }
            if (x > 65535) {
                err = "bad Netpbm sample";
                return false;
            }
            v[k++] = (uint16_t)x;
        }
        if (k < count) {
            err = "Netpbm data too short";
            return false;
        }
    } else {
        i += 1;  // the single whitespace byte after maxval
        const size_t bytes = count * (o.maxval < 256 ? 1 : 2);
        if (i + bytes > n) {
            err = "Netpbm data too short";
            return false;
        }
        for (size_t k = 0; k < count; ++k)
            v[k] = o.maxval < 256 ? d[i + k] : (uint16_t)((d[i + 2 * k] << 8) | d[i + 2 * k + 1]);
    }
    o.s.resize((size_t)3 * o.W * o.H);
    for (size_t p = 0; p < (size_t)o.W * o.H; ++p)
        for (int c = 0; c < 3; ++c) o.s[3 * p + c] = v[ch == 3 ? 3 * p + c : p];
    return true;
}

bool decode_any(const uint8_t* d, size_t n, Decoded& o, std::string& err, bool header_only) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (n >= 8 && !std::memcmp(d, sig, 8)) return png(d, n, o, err, header_only);
    return pnm(d, n, o, err, header_only);
}

// samples -> float64 RGB (interleaved) and gray, the reference's arithmetic
__device__ __forceinline__ double scale(unsigned v, int mode, double inv) {
    return mode == 1 ? (double)v / 255.0 : (double)v * inv;
}

__device__ __forceinline__ double gray_of(const double* c, int mode) {
    return mode == 2 ? c[0] : 0.299 * c[0] + 0.587 * c[1] + 0.114 * c[2];
}

// one pixel per thread; any alignment, either output may be absent
__global__ void k_image_expand(const uint16_t* __restrict__ s, long long npix, int mode,
                               int maxval, double* __restrict__ gray, double* __restrict__ rgb) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= npix) return;
    double c[3];
    const double inv = 1.0 / (double)maxval;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c[k] = scale(s[3 * i + k], mode, inv);
        if (rgb) rgb[3 * i + k] = c[k];
    }
    if (gray) gray[i] = gray_of(c, mode);
}

// two pixels per thread: three 4-byte sample loads, 16-byte stores (samples
// 4-byte, gray and rgb 16-byte aligned, both outputs present)
__global__ void k_image_expand2(const uint16_t* __restrict__ s, long long npix, int mode,
                                int maxval, double* __restrict__ gray, double* __restrict__ rgb) {
    const long long p = (blockIdx.x * 256LL + threadIdx.x) * 2;
    if (p >= npix) return;
    const double inv = 1.0 / (double)maxval;
    if (p + 1 < npix) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(s + 3 * p);
        const uint32_t w0 = w[0], w1 = w[1], w2 = w[2];
        double c[6];
        c[0] = scale(w0 & 0xFFFFu, mode, inv);
        c[1] = scale(w0 >> 16, mode, inv);
        c[2] = scale(w1 & 0xFFFFu, mode, inv);
        c[3] = scale(w1 >> 16, mode, inv);
        c[4] = scale(w2 & 0xFFFFu, mode, inv);
        c[5] = scale(w2 >> 16, mode, inv);
        double2* r2 = reinterpret_cast<double2*>(rgb + 3 * p);
        r2[0] = make_double2(c[0], c[1]);
        r2[1] = make_double2(c[2], c[3]);
        r2[2] = make_double2(c[4], c[5]);
        *reinterpret_cast<double2*>(gray + p) = make_double2(gray_of(c, mode), gray_of(c + 3, mode));
    } else {
        double c[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) rgb[3 * p + k] = c[k] = scale(s[3 * p + k], mode, inv);
        gray[p] = gray_of(c, mode);
    }
}

}  // namespace

extern "C" {

int sv_image_size(const uint8_t* data, size_t n, int* W, int* H) {
    SV_REQUIRE(data && W && H, "sv_image_size: bad arguments");
    Decoded o;
    std::string err;
    if (!decode_any(data, n, o, err, true)) return sv::set_error(SV_EINVAL, "%s", err.c_str());
    *W = o.W;
    *H = o.H;
    return SV_OK;
}

int sv_image_decode(const uint8_t* data, size_t n, uint16_t* samples, size_t capacity, int* mode,
                    int* maxval) {
    SV_REQUIRE(data && samples && mode && maxval, "sv_image_decode: bad arguments");
    Decoded o;
    std::string err;
    try {
        if (!decode_any(data, n, o, err, false)) return sv::set_error(SV_EINVAL, "%s", err.c_str());
    } catch (const std::bad_alloc&) {
        return sv::set_error(SV_EINVAL, "image too large for host memory");
    }
    if (o.s.size() > capacity)
        return sv::set_error(SV_EINVAL, "sv_image_decode: %zu samples do not fit in %zu", o.s.size(),
                             capacity);
    std::memcpy(samples, o.s.data(), o.s.size() * sizeof(uint16_t));
    *mode = o.mode;
    *maxval = o.maxval;
    return SV_OK;
}

int sv_image_expand(const uint16_t* samples, long long npix, int mode, int maxval, double* gray,
                    double* rgb, void* stream) {
    SV_REQUIRE(samples && npix > 0 && (gray || rgb) && mode >= 0 && mode <= 2 && maxval > 0,
               "sv_image_expand: bad arguments");
    const bool vec = gray && rgb && !((uintptr_t)samples & 3) && !((uintptr_t)gray & 15) &&
                     !((uintptr_t)rgb & 15);
    if (vec)
        hipLaunchKernelGGL(k_image_expand2, dim3((unsigned)((npix + 511) / 512)), dim3(256), 0,
                           (hipStream_t)stream, samples, npix, mode, maxval, gray, rgb);
    else
        hipLaunchKernelGGL(k_image_expand, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, samples, npix, mode, maxval, gray, rgb);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // extern "C"
