// PNG decoder and inflate for textures. Reads only what the file holds:
// every length is checked against the bytes left before it is used.
#include "png.h"

#include <array>
#include <cstring>
#include <stdexcept>

namespace raymotion {
namespace {

[[noreturn]] void bad(const std::string& why) { throw std::runtime_error("PNG: " + why); }

uint32_t be32(const uint8_t* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

// LSB-first bit reader over a byte range (RFC 1951 3.1.1).
class BitReader {
    const uint8_t* data;
    size_t size, pos = 0;
    uint64_t buf = 0;
    int have = 0;

public:
    BitReader(const uint8_t* d, size_t n) : data(d), size(n) {}
    uint32_t take(int k) {  // 0 <= k <= 24
        while (have < k) {
            if (pos >= size) bad("truncated deflate stream");
            buf |= uint64_t(data[pos++]) << have;
            have += 8;
        }
        const uint32_t v = uint32_t(buf & ((uint64_t(1) << k) - 1));
        buf >>= k;
        have -= k;
        return v;
    }
    void to_byte_boundary() {
        buf >>= have & 7;
        have &= ~7;
    }
};

// Canonical Huffman code decoded one bit at a time: count[l] codes of length
// l, symbols listed by (length, value).
struct Huffman {
    std::array<uint16_t, 16> count{};
    std::array<uint16_t, 320> symbol{};

    void build(const uint8_t* lengths, int n) {
        count.fill(0);
        for (int i = 0; i < n; ++i) ++count[lengths[i]];
        count[0] = 0;
        int left = 1;
        for (int l = 1; l < 16; ++l) {
            left = 2 * left - count[size_t(l)];
            if (left < 0) bad("over-subscribed Huffman code");
        }
        std::array<uint16_t, 16> next{};
        for (int l = 1; l < 15; ++l) next[size_t(l) + 1] = uint16_t(next[size_t(l)] + count[size_t(l)]);
        for (int i = 0; i < n; ++i)
            if (lengths[i]) symbol[next[lengths[i]]++] = uint16_t(i);
    }
    int decode(BitReader& in) const {
        int code = 0, first = 0, index = 0;
        for (int l = 1; l < 16; ++l) {
            code |= int(in.take(1));
            const int c = count[size_t(l)];
            if (code - first < c) return symbol[size_t(index + code - first)];
            index += c;
            first = (first + c) << 1;
            code <<= 1;
        }
        bad("invalid Huffman code");
    }
};

// RFC 1951 3.2.5: base values and extra bits of the length and distance codes.
constexpr uint16_t kLengthBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                      31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
constexpr uint8_t kLengthExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
constexpr uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                    193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
constexpr uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
// Order of the code-length code lengths in a dynamic block header.
constexpr uint8_t kClOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

void fixed_codes(Huffman& lit, Huffman& dist) {
    uint8_t l[288];
    for (int i = 0; i < 288; ++i) l[i] = i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8;
    lit.build(l, 288);
    uint8_t d[30];
    std::memset(d, 5, sizeof(d));
    dist.build(d, 30);
}

void dynamic_codes(BitReader& in, Huffman& lit, Huffman& dist) {
    const int nlit = int(in.take(5)) + 257, ndist = int(in.take(5)) + 1, ncl = int(in.take(4)) + 4;
    if (nlit > 286 || ndist > 30) bad("too many length or distance codes");
    uint8_t cl[19] = {};
    for (int i = 0; i < ncl; ++i) cl[kClOrder[i]] = uint8_t(in.take(3));
    Huffman clh;
    clh.build(cl, 19);
    uint8_t len[286 + 30] = {};
    int n = 0;
    while (n < nlit + ndist) {
        const int s = clh.decode(in);
        if (s < 16) {
            len[n++] = uint8_t(s);
            continue;
        }
        int rep = 0;
        uint8_t v = 0;
        if (s == 16) {
            if (n == 0) bad("repeat with no previous length");
            v = len[n - 1];
            rep = 3 + int(in.take(2));
        } else if (s == 17) {
            rep = 3 + int(in.take(3));
        } else {
            rep = 11 + int(in.take(7));
        }
        if (n + rep > nlit + ndist) bad("code lengths overrun");
        while (rep--) len[n++] = v;
    }
    if (len[256] == 0) bad("no end-of-block code");
    lit.build(len, nlit);
    dist.build(len + nlit, ndist);
}

}  // namespace

std::vector<uint8_t> zlib_inflate(const uint8_t* data, size_t size, size_t max_out) {
    if (size < 2) bad("zlib stream too short");
    const unsigned cmf = data[0], flg = data[1];
    if ((cmf & 15) != 8 || (cmf >> 4) > 7 || ((cmf << 8) | flg) % 31 != 0) bad("bad zlib header");
    if (flg & 32) bad("zlib preset dictionary");
    BitReader in(data + 2, size - 2);
    std::vector<uint8_t> out;
    out.reserve(max_out);
    bool last = false;
    // The Adler-32 trailer is not checked, and inflation stops once max_out
    // bytes are out (stb_image accepts the same streams).
    while (!last && out.size() < max_out) {
        last = in.take(1) != 0;
        const uint32_t type = in.take(2);
        if (type == 0) {
            in.to_byte_boundary();
            const uint32_t len = in.take(16), nlen = in.take(16);
            if ((len ^ 0xFFFFu) != nlen) bad("stored block length check");
            for (uint32_t i = 0; i < len && out.size() < max_out; ++i) out.push_back(uint8_t(in.take(8)));
        } else if (type == 1 || type == 2) {
            Huffman lit, dist;
            if (type == 1) fixed_codes(lit, dist);
            else dynamic_codes(in, lit, dist);
            for (;;) {
                int s = lit.decode(in);
                if (s < 256) {
                    if (out.size() >= max_out) break;
                    out.push_back(uint8_t(s));
                } else if (s == 256) {
                    break;
                } else {
                    s -= 257;
                    if (s >= 29) bad("invalid length code");
                    const size_t len = kLengthBase[s] + in.take(kLengthExtra[s]);
                    const int ds = dist.decode(in);
                    if (ds >= 30) bad("invalid distance code");
                    const size_t back = kDistBase[ds] + in.take(kDistExtra[ds]);
                    if (back > out.size()) bad("distance before the start of the data");
                    const size_t from = out.size() - back;
                    for (size_t i = 0; i < len && out.size() < max_out; ++i) {
                        const uint8_t b = out[from + i];
                        out.push_back(b);
                    }
                }
            }
        } else {
            bad("invalid deflate block type");
        }
    }
    return out;
}

bool is_png(const std::string& bytes) {
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    return bytes.size() >= 8 && std::memcmp(bytes.data(), sig, 8) == 0;
}

Rgb8Image decode_png(const std::string& bytes) {
    if (!is_png(bytes)) bad("not a PNG file");
    const uint8_t* d = reinterpret_cast<const uint8_t*>(bytes.data());
    const size_t n = bytes.size();
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    bool have_header = false, ended = false;
    std::vector<uint8_t> palette, idat;
    while (!ended) {
        if (n - pos < 12) bad("truncated chunk");
        const uint32_t len = be32(d + pos);
        const uint8_t* type = d + pos + 4;
        if (len > 0x7FFFFFFFu || n - pos - 12 < len) bad("truncated chunk");
        const uint8_t* c = d + pos + 8;
        auto is = [&](const char* t) { return std::memcmp(type, t, 4) == 0; };
        if (is("IHDR")) {
            if (have_header || len != 13) bad("bad IHDR");
            w = be32(c);
            h = be32(c + 4);
            depth = c[8];
            ctype = c[9];
            interlace = c[12];
            if (w == 0 || h == 0 || w > (1u << 24) || h > (1u << 24) || uint64_t(w) * h > (uint64_t(1) << 28))
                bad("unsupported image size");
            const bool ok_depth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                                  (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                                  ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
            if (!ok_depth) bad("bad colour type / bit depth");
            if (c[10] != 0 || c[11] != 0 || interlace > 1) bad("unsupported compression, filter or interlace method");
            have_header = true;
        } else if (!have_header) {
            bad("first chunk is not IHDR");
        } else if (is("PLTE")) {
            if (len == 0 || len % 3 != 0 || len > 768) bad("bad PLTE");
            palette.assign(c, c + len);
        } else if (is("IDAT")) {
            if (ctype == 3 && palette.empty()) bad("palette image without PLTE");
            idat.insert(idat.end(), c, c + len);
        } else if (is("IEND")) {
            ended = true;
        } else if (!(type[0] & 0x20)) {
            bad("unknown critical chunk");
        }
        pos += 12 + size_t(len);
    }
    if (!have_header || idat.empty()) bad("no image data");
    if (ctype == 3 && palette.empty()) bad("palette image without PLTE");

    const int channels = ctype == 2 ? 3 : ctype == 4 ? 2 : ctype == 6 ? 4 : 1;
    const int pixel_bits = channels * depth;
    const size_t filter_bpp = size_t(pixel_bits >= 8 ? pixel_bits / 8 : 1);
    // Adam7 passes (x0, y0, dx, dy); a plain image is one pass (0, 0, 1, 1).
    static const int kAdam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                     {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    static const int kPlain[1][4] = {{0, 0, 1, 1}};
    const int npass = interlace ? 7 : 1;
    const int (*passes)[4] = interlace ? kAdam7 : kPlain;
    size_t total = 0;
    for (int p = 0; p < npass; ++p) {
        const uint32_t pw = w > uint32_t(passes[p][0]) ? (w - passes[p][0] + passes[p][2] - 1) / passes[p][2] : 0;
        const uint32_t ph = h > uint32_t(passes[p][1]) ? (h - passes[p][1] + passes[p][3] - 1) / passes[p][3] : 0;
        if (pw && ph) total += size_t(ph) * (1 + (size_t(pw) * pixel_bits + 7) / 8);
    }
    std::vector<uint8_t> raw = zlib_inflate(idat.data(), idat.size(), total);
    if (raw.size() < total) bad("not enough image data");

    Rgb8Image img;
    img.width = int(w);
    img.height = int(h);
    img.rgb.assign(size_t(w) * h * 3, 0);
    static const uint8_t kGreyScale[9] = {0, 255, 85, 0, 17, 0, 0, 0, 1};
    size_t at = 0;
    for (int p = 0; p < npass; ++p) {
        const uint32_t pw = w > uint32_t(passes[p][0]) ? (w - passes[p][0] + passes[p][2] - 1) / passes[p][2] : 0;
        const uint32_t ph = h > uint32_t(passes[p][1]) ? (h - passes[p][1] + passes[p][3] - 1) / passes[p][3] : 0;
        if (!pw || !ph) continue;
        const size_t row_bytes = (size_t(pw) * pixel_bits + 7) / 8;
        std::vector<uint8_t> prev(row_bytes, 0), cur(row_bytes);
        for (uint32_t y = 0; y < ph; ++y) {
            const uint8_t filter = raw[at];
            std::memcpy(cur.data(), raw.data() + at + 1, row_bytes);
            at += 1 + row_bytes;
            for (size_t i = 0; i < row_bytes; ++i) {
                const int a = i >= filter_bpp ? cur[i - filter_bpp] : 0, b = prev[i],
                          c = i >= filter_bpp ? prev[i - filter_bpp] : 0;
                int pred = 0;
                switch (filter) {
                    case 0: pred = 0; break;
                    case 1: pred = a; break;
                    case 2: pred = b; break;
                    case 3: pred = (a + b) >> 1; break;
                    case 4: {
                        const int q = a + b - c, pa = q > a ? q - a : a - q, pb = q > b ? q - b : b - q,
                                  pc = q > c ? q - c : c - q;
                        pred = (pa <= pb && pa <= pc) ? a : pb <= pc ? b : c;
                        break;
                    }
                    default: bad("invalid row filter");
                }
                cur[i] = uint8_t(cur[i] + pred);
            }
            auto sample = [&](size_t x, int ch) -> int {
                if (depth == 16) return cur[(x * channels + size_t(ch)) * 2];  // high byte
                if (depth == 8) return cur[x * channels + size_t(ch)];
                const size_t bit = x * size_t(depth);
                return (cur[bit / 8] >> (8 - depth - int(bit % 8))) & ((1 << depth) - 1);
            };
            const size_t oy = size_t(passes[p][1]) + size_t(y) * passes[p][3];
            for (uint32_t x = 0; x < pw; ++x) {
                const size_t ox = size_t(passes[p][0]) + size_t(x) * passes[p][2];
                uint8_t* o = &img.rgb[(oy * w + ox) * 3];
                if (ctype == 2 || ctype == 6) {
                    o[0] = uint8_t(sample(x, 0));
                    o[1] = uint8_t(sample(x, 1));
                    o[2] = uint8_t(sample(x, 2));
                } else if (ctype == 3) {
                    const size_t k = size_t(sample(x, 0)) * 3;
                    if (k + 2 < palette.size()) {
                        o[0] = palette[k];
                        o[1] = palette[k + 1];
                        o[2] = palette[k + 2];
                    }
                } else {  // grey, grey + alpha
                    const int g = depth < 8 ? sample(x, 0) * kGreyScale[depth] : sample(x, 0);
                    o[0] = o[1] = o[2] = uint8_t(g);
                }
            }
            std::swap(prev, cur);
        }
    }
    return img;
}

}  // namespace raymotion
