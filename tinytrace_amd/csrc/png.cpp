// PNG decoding for glTF images, as the reference's load_png
// (src/asset/gltf.rs:122-137) uses it: a PNG to 8-bit RGBA (the png crate
// 0.17 with ALPHA | normalize_to_color8: palettes and low bit depths
// expanded, 16-bit samples stripped to their high byte, alpha 255 where the
// image has none, tRNS keys made transparent).  Interlaced (Adam7) images
// are deinterlaced.  Grey images are replicated into R, G and B: the png
// crate leaves them as grey + alpha, two bytes a pixel, which the
// reference's [u8; 4] texel cast then misreads.  Chunk CRCs are checked as
// the png crate does by default.  Inflate follows RFC
// 1950/1951 (Adler-32 checked); filtering follows the PNG specification
// (none, sub, up, average, Paeth).
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace tt {

namespace {

struct BitReader {
    const uint8_t* p;
    size_t n, pos = 0;
    uint32_t bitbuf = 0;
    int bitcount = 0;
    bool fail = false;
    uint32_t bits(int count) {
        while (bitcount < count) {
            if (pos >= n) {
                fail = true;
                return 0;
            }
            bitbuf |= uint32_t(p[pos++]) << bitcount;
            bitcount += 8;
        }
        uint32_t v = bitbuf & ((count == 32) ? 0xffffffffu : ((1u << count) - 1u));
        bitbuf >>= count;
        bitcount -= count;
        return v;
    }
    void align() {
        bitbuf = 0;
        bitcount = 0;
    }
};

// Canonical Huffman decoding table: counts per length and symbols in code
// order.
struct Huffman {
    uint16_t count[16] = {};
    std::vector<uint16_t> symbol;
    bool build(const uint8_t* lengths, int n) {
        std::memset(count, 0, sizeof(count));
        for (int i = 0; i < n; i++) count[lengths[i]]++;
        count[0] = 0;
        int left = 1;
        for (int len = 1; len < 16; len++) {
            left <<= 1;
            left -= count[len];
            if (left < 0) return false;  // over-subscribed
        }
        uint16_t offs[16] = {};
        for (int len = 1; len < 15; len++) offs[len + 1] = uint16_t(offs[len] + count[len]);
        symbol.assign(size_t(n), 0);
        for (int i = 0; i < n; i++)
            if (lengths[i]) symbol[offs[lengths[i]]++] = uint16_t(i);
        return true;
    }
    int decode(BitReader& br) const {
        int code = 0, first = 0, index = 0;
        for (int len = 1; len < 16; len++) {
            code |= int(br.bits(1));
            if (br.fail) return -1;
            int c = count[len];
            if (code - c < first) return symbol[size_t(index + (code - first))];
            index += c;
            first += c;
            first <<= 1;
            code <<= 1;
        }
        return -1;
    }
};

const uint16_t LENGTH_BASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                  31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t LENGTH_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t DIST_BASE[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

bool inflate_block(BitReader& br, const Huffman& lit, const Huffman& dist, std::vector<uint8_t>& out, size_t limit) {
    for (;;) {
        int sym = lit.decode(br);
        if (sym < 0) return false;
        if (sym < 256) {
            if (out.size() >= limit) return false;
            out.push_back(uint8_t(sym));
        } else if (sym == 256) {
            return true;
        } else {
            sym -= 257;
            if (sym >= 29) return false;
            size_t len = LENGTH_BASE[sym] + br.bits(LENGTH_EXTRA[sym]);
            int ds = dist.decode(br);
            if (ds < 0 || ds >= 30) return false;
            size_t d = DIST_BASE[ds] + br.bits(DIST_EXTRA[ds]);
            if (br.fail || d > out.size() || out.size() + len > limit) return false;
            size_t from = out.size() - d;
            for (size_t k = 0; k < len; k++) out.push_back(out[from + k]);
        }
    }
}

// zlib stream -> bytes (at most `limit`).
bool zlib_inflate(const uint8_t* data, size_t n, size_t limit, std::vector<uint8_t>& out) {
    if (n < 2) return false;
    uint8_t cmf = data[0], flg = data[1];
    if ((cmf & 0x0f) != 8 || ((uint32_t(cmf) << 8) | flg) % 31 != 0 || (flg & 0x20)) return false;
    BitReader br{data + 2, n - 2};
    for (;;) {
        uint32_t final = br.bits(1);
        uint32_t type = br.bits(2);
        if (br.fail) return false;
        if (type == 0) {
            br.align();
            if (br.pos + 4 > br.n) return false;
            uint32_t len = br.p[br.pos] | (uint32_t(br.p[br.pos + 1]) << 8);
            uint32_t nlen = br.p[br.pos + 2] | (uint32_t(br.p[br.pos + 3]) << 8);
            br.pos += 4;
            if ((len ^ 0xffffu) != nlen || br.pos + len > br.n || out.size() + len > limit) return false;
            out.insert(out.end(), br.p + br.pos, br.p + br.pos + len);
            br.pos += len;
        } else if (type == 1) {
            uint8_t lengths[288 + 30];
            for (int i = 0; i < 144; i++) lengths[i] = 8;
            for (int i = 144; i < 256; i++) lengths[i] = 9;
            for (int i = 256; i < 280; i++) lengths[i] = 7;
            for (int i = 280; i < 288; i++) lengths[i] = 8;
            for (int i = 0; i < 30; i++) lengths[288 + i] = 5;
            Huffman lit, dist;
            if (!lit.build(lengths, 288) || !dist.build(lengths + 288, 30)) return false;
            if (!inflate_block(br, lit, dist, out, limit)) return false;
        } else if (type == 2) {
            uint32_t hlit = br.bits(5) + 257, hdist = br.bits(5) + 1, hclen = br.bits(4) + 4;
            static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t cl[19] = {};
            for (uint32_t i = 0; i < hclen; i++) cl[order[i]] = uint8_t(br.bits(3));
            Huffman clh;
            if (br.fail || !clh.build(cl, 19)) return false;
            uint8_t lengths[288 + 32] = {};
            uint32_t i = 0;
            while (i < hlit + hdist) {
                int sym = clh.decode(br);
                if (sym < 0) return false;
                if (sym < 16) {
                    lengths[i++] = uint8_t(sym);
                } else {
                    uint8_t value = 0;
                    uint32_t repeat;
                    if (sym == 16) {
                        if (i == 0) return false;
                        value = lengths[i - 1];
                        repeat = 3 + br.bits(2);
                    } else if (sym == 17) {
                        repeat = 3 + br.bits(3);
                    } else {
                        repeat = 11 + br.bits(7);
                    }
                    if (br.fail || i + repeat > hlit + hdist) return false;
                    while (repeat--) lengths[i++] = value;
                }
            }
            Huffman lit, dist;
            if (!lit.build(lengths, int(hlit)) || !dist.build(lengths + hlit, int(hdist))) return false;
            if (!inflate_block(br, lit, dist, out, limit)) return false;
        } else {
            return false;
        }
        if (final) break;
    }
    // Adler-32 of the output.
    br.align();
    if (br.pos + 4 > br.n) return false;
    uint32_t want = (uint32_t(br.p[br.pos]) << 24) | (uint32_t(br.p[br.pos + 1]) << 16) |
                    (uint32_t(br.p[br.pos + 2]) << 8) | br.p[br.pos + 3];
    uint32_t a = 1, b = 0;
    for (uint8_t c : out) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    return ((b << 16) | a) == want;
}

uint32_t be32(const uint8_t* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

// CRC-32 (ISO 3309, reflected polynomial 0xedb88320) over a chunk's type
// and data.
uint32_t crc32(const uint8_t* p, size_t n) {
    struct Table {
        uint32_t v[256];
        Table() {
            for (uint32_t i = 0; i < 256; i++) {
                uint32_t c = i;
                for (int k = 0; k < 8; k++) c = (c & 1u) ? 0xedb88320u ^ (c >> 1) : c >> 1;
                v[i] = c;
            }
        }
    };
    static const Table table;  // initialised once, thread-safe
    uint32_t c = 0xffffffffu;
    for (size_t i = 0; i < n; i++) c = table.v[(c ^ p[i]) & 0xffu] ^ (c >> 8);
    return c ^ 0xffffffffu;
}

int paeth(int a, int b, int c) {
    int p = a + b - c;
    int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

}  // namespace

// PNG bytes -> width x height RGBA8 (rows from the top).  Empty error string
// on success; otherwise what was wrong.
std::string decode_png(const uint8_t* data, size_t n, uint32_t& width, uint32_t& height, std::vector<uint8_t>& rgba) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (n < 8 || std::memcmp(data, sig, 8) != 0) return "not a PNG";
    size_t pos = 8;
    uint32_t depth = 0, ctype = 0, interlace = 0;
    bool header = false;
    std::vector<uint8_t> idat, palette, trns;
    while (pos + 12 <= n) {
        uint32_t len = be32(data + pos);
        if (len > n - pos - 12) return "truncated chunk";
        const uint8_t* type = data + pos + 4;
        const uint8_t* body = data + pos + 8;
        if (crc32(type, size_t(len) + 4) != be32(body + len)) return "chunk CRC mismatch";
        if (std::memcmp(type, "IHDR", 4) == 0) {
            if (len != 13) return "bad IHDR";
            width = be32(body);
            height = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
            header = true;
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            palette.assign(body, body + len);
        } else if (std::memcmp(type, "tRNS", 4) == 0) {
            trns.assign(body, body + len);
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), body, body + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            break;
        }
        pos += 12 + size_t(len);
    }
    if (!header || width == 0 || height == 0 || width > 32768 || height > 32768) return "bad image size";
    if (interlace > 1) return "unknown interlace method";
    int channels = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                    (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                    ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!channels || !depth_ok) return "unsupported PNG format";
    if (ctype == 3 && (palette.empty() || palette.size() % 3 != 0)) return "missing palette";
    size_t bits_per_pixel = size_t(channels) * depth;
    size_t bpp = (bits_per_pixel + 7) / 8;  // filter unit

    // Passes: the whole image, or Adam7's seven (origin and step in x, y).
    struct Pass {
        uint32_t x0, y0, dx, dy;
    };
    static const Pass single[1] = {{0, 0, 1, 1}};
    static const Pass adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                  {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const Pass* passes = interlace ? adam7 : single;
    int pass_count = interlace ? 7 : 1;
    size_t total = 0;
    for (int k = 0; k < pass_count; k++) {
        const Pass& q = passes[k];
        size_t pw = width > q.x0 ? (width - q.x0 + q.dx - 1) / q.dx : 0;
        size_t ph = height > q.y0 ? (height - q.y0 + q.dy - 1) / q.dy : 0;
        if (pw && ph) total += ph * ((pw * bits_per_pixel + 7) / 8 + 1);
    }
    std::vector<uint8_t> raw;
    if (!zlib_inflate(idat.data(), idat.size(), total, raw) || raw.size() != total) return "corrupt image data";

    rgba.assign(size_t(width) * height * 4, 0);
    auto sample = [&](const uint8_t* row, size_t index) -> uint32_t {
        // index-th sample of the row at this bit depth (16-bit: high byte).
        if (depth == 8) return row[index];
        if (depth == 16) return row[2 * index];
        size_t bit = index * depth;
        return (row[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1u);
    };
    auto sample16 = [&](const uint8_t* row, size_t index) -> uint32_t {
        // the full sample, for tRNS keys
        return depth == 16 ? (uint32_t(row[2 * index]) << 8 | row[2 * index + 1]) : sample(row, index);
    };
    auto expand = [&](uint32_t v) { return depth < 8 ? v * 255u / ((1u << depth) - 1u) : v; };
    // One pixel of an unfiltered row -> RGBA8.
    auto pixel = [&](const uint8_t* row, size_t x, uint8_t* o) -> bool {
        size_t base = x * size_t(channels);
        if (ctype == 3) {
            uint32_t i = sample(row, x);
            if (3 * size_t(i) + 2 >= palette.size()) return false;
            o[0] = palette[3 * i];
            o[1] = palette[3 * i + 1];
            o[2] = palette[3 * i + 2];
            o[3] = i < trns.size() ? trns[i] : 255;
        } else if (ctype == 0 || ctype == 4) {
            uint8_t g = uint8_t(expand(sample(row, base)));
            o[0] = o[1] = o[2] = g;
            if (ctype == 4) {
                o[3] = uint8_t(sample(row, base + 1));
            } else {
                bool clear = trns.size() >= 2 && sample16(row, base) == ((uint32_t(trns[0]) << 8) | trns[1]);
                o[3] = clear ? 0 : 255;
            }
        } else {
            for (int c = 0; c < 3; c++) o[c] = uint8_t(sample(row, base + c));
            if (ctype == 6) {
                o[3] = uint8_t(sample(row, base + 3));
            } else {
                bool clear = trns.size() >= 6;
                for (int c = 0; c < 3 && clear; c++)
                    clear = sample16(row, base + c) == ((uint32_t(trns[2 * c]) << 8) | trns[2 * c + 1]);
                o[3] = clear ? 0 : 255;
            }
        }
        return true;
    };

    size_t offset = 0;
    for (int k = 0; k < pass_count; k++) {
        const Pass& q = passes[k];
        size_t pw = width > q.x0 ? (width - q.x0 + q.dx - 1) / q.dx : 0;
        size_t ph = height > q.y0 ? (height - q.y0 + q.dy - 1) / q.dy : 0;
        if (!pw || !ph) continue;
        size_t stride = (pw * bits_per_pixel + 7) / 8;
        std::vector<uint8_t> prev(stride, 0), row(stride);
        for (size_t y = 0; y < ph; y++) {
            uint8_t filter = raw[offset];
            const uint8_t* in = raw.data() + offset + 1;
            offset += stride + 1;
            for (size_t i = 0; i < stride; i++) {
                int a = i >= bpp ? row[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
                int v = in[i];
                switch (filter) {
                    case 0: break;
                    case 1: v += a; break;
                    case 2: v += b; break;
                    case 3: v += (a + b) / 2; break;
                    case 4: v += paeth(a, b, c); break;
                    default: return "bad filter";
                }
                row[i] = uint8_t(v);
            }
            size_t dst_y = q.y0 + y * q.dy;
            for (size_t x = 0; x < pw; x++) {
                size_t dst_x = q.x0 + x * q.dx;
                if (!pixel(row.data(), x, &rgba[(dst_y * width + dst_x) * 4])) return "palette index out of range";
            }
            prev.swap(row);
            row.assign(stride, 0);
        }
    }
    return std::string();
}

}  // namespace tt
