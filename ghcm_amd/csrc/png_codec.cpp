// png_codec.cpp -- PNG decoding and encoding for the field loaders and the result writers.
//
// The reference decodes images with stb_image (`stbi_loadf_from_memory`,
// src/field_loaders.cpp:36-37) and writes heat maps with stb_image_write
// (src/util_files.cpp:83-95).  Neither library is part of this build, so the codec is
// written here from the format specifications: zlib (RFC 1950), DEFLATE (RFC 1951) and
// PNG (ISO/IEC 15948): every colour type, bit depths 1-16, all five filters and Adam7
// interlacing.  The conversions applied after decoding follow stb's documented
// behaviour for `stbi_loadf` with a requested channel count:
//   * 16-bit samples are reduced to 8 bits by keeping the high byte;
//   * grey from RGB(A) is (77 r + 150 g + 29 b) >> 8 (alpha ignored);
//   * palette images are expanded to RGB(A) first, grey+alpha drops alpha;
//   * sub-byte grey samples are scaled to 0..255 (x255, x85, x17), palette indices are not.
// Chunk CRCs are not required to match on input (stb does not check them); they are
// written correctly on output.
#include "png_codec.h"

#include <algorithm>
#include <cstring>

namespace ghcm {
namespace {

// ---------------------------------------------------------------------------------------
// checksums
// ---------------------------------------------------------------------------------------
struct Crc32Table {
    uint32_t t[256];
    Crc32Table() {
        for (uint32_t n = 0; n < 256; ++n) {
            uint32_t c = n;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[n] = c;
        }
    }
};

uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
    static const Crc32Table tab;
    uint32_t c = crc ^ 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = tab.t[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

uint32_t adler32(const uint8_t* p, size_t n) {
    uint32_t a = 1, b = 0;
    while (n) {
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

uint32_t be32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

// ---------------------------------------------------------------------------------------
// inflate (RFC 1951)
// ---------------------------------------------------------------------------------------
struct BitReader {
    const uint8_t* p;
    const uint8_t* end;
    uint64_t buf = 0;
    int cnt = 0;
    size_t pad = 0;                        // zero bytes appended past the end of the input
    void refill() {
        while (cnt <= 56) {
            uint64_t byte = 0;
            if (p < end) byte = *p++;
            else ++pad;
            buf |= byte << cnt;
            cnt += 8;
        }
    }
    // True once a bit past the end of the input has been consumed.
    bool overrun() const { return (size_t)cnt < pad * 8; }
    uint32_t bits(int n) {                 // n <= 32
        if (n == 0) return 0;
        if (cnt < n) refill();
        const uint32_t v = (uint32_t)(buf & ((1ull << n) - 1));
        buf >>= n;
        cnt -= n;
        return v;
    }
    uint32_t peek(int n) {
        if (cnt < n) refill();
        return (uint32_t)(buf & ((1ull << n) - 1));
    }
    void drop(int n) {
        buf >>= n;
        cnt -= n;
    }
    void align_byte() { drop(cnt & 7); }
};

// Canonical Huffman table with a direct 15-bit lookup: entry = symbol << 4 | length.
struct Huffman {
    std::vector<uint32_t> lut;
    bool build(const uint8_t* lens, int n) {
        int count[16] = {0};
        for (int i = 0; i < n; ++i) count[lens[i]]++;
        count[0] = 0;
        int code = 0, next[16] = {0};
        for (int b = 1; b < 16; ++b) {
            code = (code + count[b - 1]) << 1;
            next[b] = code;
            if (count[b] > (1 << b)) return false;
        }
        lut.assign(1u << 15, 0u);
        for (int s = 0; s < n; ++s) {
            const int len = lens[s];
            if (!len) continue;
            const uint32_t c = (uint32_t)next[len]++;
            if (c >= (1u << len)) return false;           // over-subscribed
            // DEFLATE sends codes MSB first; the reader takes LSB first: reverse.
            uint32_t r = 0;
            for (int k = 0; k < len; ++k) r |= ((c >> k) & 1u) << (len - 1 - k);
            for (uint32_t f = r; f < (1u << 15); f += (1u << len)) lut[f] = (uint32_t)s << 4 | (uint32_t)len;
        }
        return true;
    }
    int decode(BitReader& br) const {
        const uint32_t e = lut[br.peek(15)];
        const int len = (int)(e & 15u);
        if (!len) return -1;
        br.drop(len);
        return (int)(e >> 4);
    }
};

const uint16_t kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                               35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2,
                               3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193,
                                257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145,
                                8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6,
                                7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

// Inflation stops once `limit` bytes are out (what follows is not needed), so a stream never
// grows the output past what the caller can use.
bool inflate_raw(BitReader& br, std::vector<uint8_t>& out, size_t limit, std::string& err) {
    Huffman lit, dist;
    for (;;) {
        const uint32_t final_block = br.bits(1);
        const uint32_t type = br.bits(2);
        if (type == 0) {
            br.align_byte();
            const uint32_t len = br.bits(16), nlen = br.bits(16);
            if ((len ^ 0xFFFFu) != nlen) { err = "stored block length mismatch"; return false; }
            const uint32_t take = (uint32_t)std::min<size_t>(len, limit - out.size());
            for (uint32_t i = 0; i < take; ++i) out.push_back((uint8_t)br.bits(8));
            if (take < len) return true;
        } else if (type == 1 || type == 2) {
            uint8_t lens[320];
            int hlit = 288, hdist = 32;
            if (type == 1) {
                for (int i = 0; i < 144; ++i) lens[i] = 8;
                for (int i = 144; i < 256; ++i) lens[i] = 9;
                for (int i = 256; i < 280; ++i) lens[i] = 7;
                for (int i = 280; i < 288; ++i) lens[i] = 8;
                for (int i = 0; i < 32; ++i) lens[288 + i] = 5;
            } else {
                hlit = (int)br.bits(5) + 257;
                hdist = (int)br.bits(5) + 1;
                const int hclen = (int)br.bits(4) + 4;
                static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
                uint8_t cl[19] = {0};
                for (int i = 0; i < hclen; ++i) cl[order[i]] = (uint8_t)br.bits(3);
                Huffman clh;
                if (!clh.build(cl, 19)) { err = "bad code-length code"; return false; }
                int i = 0;
                std::memset(lens, 0, sizeof lens);
                while (i < hlit + hdist) {
                    const int sym = clh.decode(br);
                    if (sym < 0) { err = "bad code length symbol"; return false; }
                    if (sym < 16) {
                        lens[i++] = (uint8_t)sym;
                    } else {
                        int rep, val = 0;
                        if (sym == 16) {
                            if (i == 0) { err = "repeat with no previous length"; return false; }
                            val = lens[i - 1];
                            rep = 3 + (int)br.bits(2);
                        } else if (sym == 17) {
                            rep = 3 + (int)br.bits(3);
                        } else {
                            rep = 11 + (int)br.bits(7);
                        }
                        if (i + rep > hlit + hdist) { err = "code lengths overflow"; return false; }
                        while (rep--) lens[i++] = (uint8_t)val;
                    }
                }
                // distance lengths follow the literal/length ones
                std::memmove(lens + 288, lens + hlit, (size_t)hdist);
                for (int k = hlit; k < 288; ++k) lens[k] = 0;
            }
            if (!lit.build(lens, 288) || !dist.build(lens + 288, 32)) { err = "bad Huffman table"; return false; }
            for (;;) {
                const int sym = lit.decode(br);
                if (sym < 0) { err = "bad literal/length code"; return false; }
                if (sym < 256) {
                    if (out.size() >= limit) return true;
                    out.push_back((uint8_t)sym);
                } else if (sym == 256) {
                    break;
                } else {
                    const int li = sym - 257;
                    if (li >= 29) { err = "bad length symbol"; return false; }
                    const uint32_t len = kLenBase[li] + br.bits(kLenExtra[li]);
                    const int ds = dist.decode(br);
                    if (ds < 0 || ds >= 30) { err = "bad distance code"; return false; }
                    const uint32_t d = kDistBase[ds] + br.bits(kDistExtra[ds]);
                    if (d > out.size()) { err = "distance beyond output"; return false; }
                    const size_t from = out.size() - d;
                    const uint32_t take = (uint32_t)std::min<size_t>(len, limit - out.size());
                    for (uint32_t k = 0; k < take; ++k) out.push_back(out[from + k]);
                    if (take < len) return true;
                }
                if (br.overrun()) { err = "truncated deflate stream"; return false; }
            }
        } else {
            err = "reserved block type";
            return false;
        }
        if (br.overrun()) { err = "truncated deflate stream"; return false; }
        if (final_block) return true;
    }
}

// ---------------------------------------------------------------------------------------
// PNG decoding
// ---------------------------------------------------------------------------------------
inline uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

// Reverse the per-scanline filters of one (sub)image in place; `rows` of `stride` bytes
// each preceded by a filter byte.  Returns the unfiltered rows packed back to back.
bool unfilter(const uint8_t* src, uint32_t rows, size_t stride, int bpp, std::vector<uint8_t>& dst,
              std::string& err) {
    dst.assign((size_t)rows * stride, 0);
    for (uint32_t y = 0; y < rows; ++y) {
        const uint8_t f = src[(size_t)y * (stride + 1)];
        const uint8_t* in = src + (size_t)y * (stride + 1) + 1;
        uint8_t* cur = dst.data() + (size_t)y * stride;
        const uint8_t* prev = y ? cur - stride : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)bpp ? cur[i - bpp] : 0;
            const int b = prev ? prev[i] : 0;
            const int c = (prev && i >= (size_t)bpp) ? prev[i - bpp] : 0;
            int v = in[i];
            switch (f) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth(a, b, c); break;
                default: err = "bad filter type"; return false;
            }
            cur[i] = (uint8_t)v;
        }
    }
    return true;
}

// Samples of one unfiltered row -> 16-bit values, one per channel.
void row_samples(const uint8_t* row, uint32_t w, int channels, int depth, uint16_t* out) {
    const size_t n = (size_t)w * channels;
    if (depth == 8) {
        for (size_t i = 0; i < n; ++i) out[i] = row[i];
    } else if (depth == 16) {
        for (size_t i = 0; i < n; ++i) out[i] = (uint16_t)(row[2 * i] << 8 | row[2 * i + 1]);
    } else {
        const int per = 8 / depth;
        const int mask = (1 << depth) - 1;
        for (size_t i = 0; i < n; ++i) {
            const int byte = row[i / per];
            const int shift = 8 - depth * (int)(i % per + 1);
            out[i] = (uint16_t)((byte >> shift) & mask);
        }
    }
}

}  // namespace

uint32_t png_crc32(const uint8_t* p, size_t n) { return crc32_update(0, p, n); }

bool zlib_inflate(const uint8_t* src, size_t n, std::vector<uint8_t>& out, std::string& err,
                  size_t limit) {
    if (n < 2) { err = "zlib stream too short"; return false; }
    const int cmf = src[0], flg = src[1];
    if ((cmf & 15) != 8 || ((cmf << 8) | flg) % 31 != 0 || (flg & 0x20)) {
        err = "bad zlib header";
        return false;
    }
    BitReader br{src + 2, src + n};
    if (!inflate_raw(br, out, limit, err)) return false;
    return true;
}

bool png_decode(const uint8_t* data, size_t n, PngImage& img, std::string& err) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || std::memcmp(data, sig, 8) != 0) { err = "not a PNG file"; return false; }
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    bool have_ihdr = false, have_iend = false;
    while (pos + 12 <= n) {
        const uint32_t len = be32(data + pos);
        const uint8_t* type = data + pos + 4;
        if (len > n - pos - 12) { err = "truncated chunk"; return false; }
        const uint8_t* body = data + pos + 8;
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) { err = "bad IHDR"; return false; }
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            if (body[10] != 0 || body[11] != 0) { err = "unknown compression or filter method"; return false; }
            interlace = body[12];
            have_ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            have_iend = true;
            break;
        } else if (!(type[0] & 0x20)) {
            err = std::string("unknown critical chunk ") + std::string((const char*)type, 4);
            return false;
        }
        pos += 12 + (size_t)len;
    }
    (void)have_iend;
    if (!have_ihdr) { err = "missing IHDR"; return false; }
    if (w == 0 || h == 0 || w > (1u << 24) || h > (1u << 24)) { err = "bad image size"; return false; }
    int channels;
    switch (ctype) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: err = "bad colour type"; return false;
    }
    const bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!depth_ok) { err = "bad bit depth for colour type"; return false; }
    if (interlace > 1) { err = "bad interlace method"; return false; }
    if (ctype == 3 && (plte.empty() || plte.size() % 3)) { err = "missing or bad palette"; return false; }

    const int bits_pp = channels * depth;
    static const int x0[7] = {0, 4, 0, 2, 0, 1, 0}, y0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const int dx[7] = {8, 8, 4, 4, 2, 2, 1}, dy[7] = {8, 8, 8, 4, 4, 2, 2};
    const int passes = interlace ? 7 : 1;
    // Filtered bytes the header implies (each row: a filter byte and its packed samples).
    // The inflated stream may not exceed it, and must reach it before anything is sized by
    // the header: a tiny IDAT under a huge IHDR allocates nothing large.
    size_t total = 0;
    for (int pass = 0; pass < passes; ++pass) {
        const int px0 = interlace ? x0[pass] : 0, py0 = interlace ? y0[pass] : 0;
        const int pdx = interlace ? dx[pass] : 1, pdy = interlace ? dy[pass] : 1;
        const size_t pw = w > (uint32_t)px0 ? (w - px0 + pdx - 1) / pdx : 0;
        const size_t ph = h > (uint32_t)py0 ? (h - py0 + pdy - 1) / pdy : 0;
        if (pw && ph) total += ((pw * bits_pp + 7) / 8 + 1) * ph;
    }
    // Decoded samples are held as 16-bit words: an image above 2^31 samples (4 GiB of
    // samples, whatever its bit depth) is refused before anything is inflated, so a small file
    // with a huge header and a highly compressible stream cannot ask for tens of GB.
    constexpr uint64_t kMaxSamples = uint64_t(1) << 31;
    if ((uint64_t)w * h * (uint64_t)channels > kMaxSamples) {
        err = "image too large (more than 2^31 samples)";
        return false;
    }
    std::vector<uint8_t> raw;
    if (!zlib_inflate(idat.data(), idat.size(), raw, err, total)) return false;
    if (raw.size() < total) { err = "image data too short"; return false; }

    const int bpp = std::max(1, bits_pp / 8);
    // Samples of the whole image, 16 bits each, channels interleaved.
    std::vector<uint16_t> smp((size_t)w * h * channels);
    size_t off = 0;
    std::vector<uint8_t> rows;
    std::vector<uint16_t> line;
    for (int pass = 0; pass < passes; ++pass) {
        uint32_t pw = w, ph = h;
        int px0 = 0, py0 = 0, pdx = 1, pdy = 1;
        if (interlace) {
            px0 = x0[pass]; py0 = y0[pass]; pdx = dx[pass]; pdy = dy[pass];
            pw = w > (uint32_t)px0 ? (w - px0 + pdx - 1) / pdx : 0;
            ph = h > (uint32_t)py0 ? (h - py0 + pdy - 1) / pdy : 0;
        }
        if (!pw || !ph) continue;
        const size_t stride = ((size_t)pw * bits_pp + 7) / 8;
        const size_t need = (stride + 1) * ph;
        if (off + need > raw.size()) { err = "image data too short"; return false; }
        if (!unfilter(raw.data() + off, ph, stride, bpp, rows, err)) return false;
        off += need;
        line.resize((size_t)pw * channels);
        for (uint32_t y = 0; y < ph; ++y) {
            row_samples(rows.data() + (size_t)y * stride, pw, channels, depth, line.data());
            const size_t oy = (size_t)(py0 + (int)y * pdy);
            for (uint32_t x = 0; x < pw; ++x) {
                const size_t ox = (size_t)(px0 + (int)x * pdx);
                for (int c = 0; c < channels; ++c)
                    smp[(oy * w + ox) * channels + c] = line[(size_t)x * channels + c];
            }
        }
    }

    img.w = w;
    img.h = h;
    img.depth16 = depth == 16;
    // Palette expansion (RGB, or RGBA when tRNS is present), and sub-byte grey scaling.
    if (ctype == 3) {
        const size_t np = plte.size() / 3;
        const bool alpha = !trns.empty();
        img.channels = alpha ? 4 : 3;
        img.samples.resize((size_t)w * h * img.channels);
        for (size_t i = 0; i < (size_t)w * h; ++i) {
            const size_t k = smp[i];
            uint16_t r = 0, g = 0, b = 0, a = 255;
            if (k < np) { r = plte[3 * k]; g = plte[3 * k + 1]; b = plte[3 * k + 2]; }
            if (alpha) a = k < trns.size() ? trns[k] : 255;
            uint16_t* o = &img.samples[i * img.channels];
            o[0] = r; o[1] = g; o[2] = b;
            if (alpha) o[3] = a;
        }
        img.depth16 = false;
    } else {
        img.channels = channels;
        if (depth < 8) {
            const int scale = depth == 1 ? 0xFF : depth == 2 ? 0x55 : 0x11;
            for (auto& v : smp) v = (uint16_t)(v * scale);
        }
        img.samples.swap(smp);
    }
    return true;
}

std::vector<uint8_t> png_to_8bit(const PngImage& img, int req_comp) {
    const size_t n = (size_t)img.w * img.h;
    const int in = img.channels;
    const int outc = req_comp ? req_comp : in;
    std::vector<uint8_t> out(n * outc);
    auto y_of = [](uint32_t r, uint32_t g, uint32_t b) { return (r * 77 + g * 150 + b * 29) >> 8; };
    for (size_t i = 0; i < n; ++i) {
        const uint16_t* s = &img.samples[i * in];
        uint32_t v[4] = {s[0], in > 1 ? s[1] : 0u, in > 2 ? s[2] : 0u, in > 3 ? s[3] : 0u};
        uint32_t o[4] = {0, 0, 0, 0};
        // channel conversion at the source precision, then 16 -> 8 bits
        if (outc == in) {
            for (int c = 0; c < in; ++c) o[c] = v[c];
        } else {
            const bool grey_src = in <= 2;
            const uint32_t top = img.depth16 ? 0xFFFFu : 0xFFu;
            const uint32_t alpha = (in == 2) ? v[1] : (in == 4) ? v[3] : top;
            const uint32_t grey = grey_src ? v[0] : y_of(v[0], v[1], v[2]);
            switch (outc) {
                case 1: o[0] = grey; break;
                case 2: o[0] = grey; o[1] = alpha; break;
                case 3:
                    if (grey_src) { o[0] = o[1] = o[2] = v[0]; }
                    else { o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; }
                    break;
                default:
                    if (grey_src) { o[0] = o[1] = o[2] = v[0]; }
                    else { o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; }
                    o[3] = alpha;
                    break;
            }
        }
        for (int c = 0; c < outc; ++c) out[i * outc + c] = (uint8_t)(img.depth16 ? (o[c] >> 8) : o[c]);
    }
    return out;
}

// ---------------------------------------------------------------------------------------
// encoding: zlib stream with fixed-Huffman DEFLATE blocks and greedy hash-chain matching
// ---------------------------------------------------------------------------------------
namespace {

struct BitWriter {
    std::vector<uint8_t>& out;
    uint64_t buf = 0;
    int cnt = 0;
    explicit BitWriter(std::vector<uint8_t>& o) : out(o) {}
    void put(uint32_t v, int n) {                  // LSB-first
        buf |= (uint64_t)v << cnt;
        cnt += n;
        while (cnt >= 8) {
            out.push_back((uint8_t)buf);
            buf >>= 8;
            cnt -= 8;
        }
    }
    void put_rev(uint32_t code, int len) {         // Huffman codes go MSB-first
        uint32_t r = 0;
        for (int k = 0; k < len; ++k) r |= ((code >> k) & 1u) << (len - 1 - k);
        put(r, len);
    }
    void flush() {
        if (cnt) out.push_back((uint8_t)buf);
        buf = 0;
        cnt = 0;
    }
};

void fixed_lit(BitWriter& bw, int sym) {
    if (sym < 144) bw.put_rev(0x30 + sym, 8);
    else if (sym < 256) bw.put_rev(0x190 + sym - 144, 9);
    else if (sym < 280) bw.put_rev(sym - 256, 7);
    else bw.put_rev(0xC0 + sym - 280, 8);
}

}  // namespace

std::vector<uint8_t> zlib_deflate(const uint8_t* src, size_t n) {
    std::vector<uint8_t> out;
    out.push_back(0x78);
    out.push_back(0x01);
    BitWriter bw(out);
    bw.put(1, 1);          // one final block
    bw.put(1, 2);          // fixed Huffman
    constexpr int kHashBits = 15, kWindow = 32768, kMaxChain = 16;
    std::vector<int32_t> head(1u << kHashBits, -1), prev(n ? n : 1, -1);
    auto hash3 = [&](size_t i) {
        return ((uint32_t)src[i] << 10 ^ (uint32_t)src[i + 1] << 5 ^ src[i + 2]) & ((1u << kHashBits) - 1);
    };
    size_t i = 0;
    while (i < n) {
        int best_len = 0, best_d = 0;
        if (i + 3 <= n) {
            const uint32_t hv = hash3(i);
            int32_t cand = head[hv];
            for (int chain = 0; cand >= 0 && chain < kMaxChain; ++chain) {
                const size_t d = i - (size_t)cand;
                if (d > kWindow) break;
                int l = 0;
                const int lim = (int)std::min<size_t>(258, n - i);
                while (l < lim && src[cand + l] == src[i + l]) ++l;
                if (l > best_len) { best_len = l; best_d = (int)d; if (l == 258) break; }
                cand = prev[cand];
            }
            prev[i] = head[hv];
            head[hv] = (int32_t)i;
        }
        if (best_len >= 3) {
            int li = 28;
            while (li > 0 && kLenBase[li] > best_len) --li;
            fixed_lit(bw, 257 + li);
            bw.put((uint32_t)(best_len - kLenBase[li]), kLenExtra[li]);
            int di = 29;
            while (di > 0 && kDistBase[di] > best_d) --di;
            bw.put_rev((uint32_t)di, 5);
            bw.put((uint32_t)(best_d - kDistBase[di]), kDistExtra[di]);
            for (int k = 1; k < best_len; ++k) {
                const size_t j = i + (size_t)k;
                if (j + 3 <= n) {
                    const uint32_t hv = hash3(j);
                    prev[j] = head[hv];
                    head[hv] = (int32_t)j;
                }
            }
            i += (size_t)best_len;
        } else {
            fixed_lit(bw, src[i]);
            ++i;
        }
    }
    fixed_lit(bw, 256);
    bw.flush();
    const uint32_t ad = adler32(src, n);
    for (int s = 24; s >= 0; s -= 8) out.push_back((uint8_t)(ad >> s));
    return out;
}

std::vector<uint8_t> png_encode(const uint8_t* pixels, uint32_t w, uint32_t h, int channels) {
    // Filter 0 per row keeps the encoder simple; the heat maps compress well as they are.
    const size_t stride = (size_t)w * channels;
    std::vector<uint8_t> raw((stride + 1) * h);
    for (uint32_t y = 0; y < h; ++y) {
        raw[(stride + 1) * y] = 0;
        std::memcpy(&raw[(stride + 1) * y + 1], pixels + stride * y, stride);
    }
    const std::vector<uint8_t> z = zlib_deflate(raw.data(), raw.size());
    std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
    auto chunk = [&](const char* type, const uint8_t* body, size_t len) {
        for (int s = 24; s >= 0; s -= 8) out.push_back((uint8_t)(len >> s));
        const size_t start = out.size();
        out.insert(out.end(), type, type + 4);
        out.insert(out.end(), body, body + len);
        const uint32_t c = crc32_update(0, out.data() + start, len + 4);
        for (int s = 24; s >= 0; s -= 8) out.push_back((uint8_t)(c >> s));
    };
    uint8_t ihdr[13];
    for (int s = 0; s < 4; ++s) {
        ihdr[s] = (uint8_t)(w >> (24 - 8 * s));
        ihdr[4 + s] = (uint8_t)(h >> (24 - 8 * s));
    }
    ihdr[8] = 8;
    ihdr[9] = (uint8_t)(channels == 1 ? 0 : channels == 2 ? 4 : channels == 3 ? 2 : 6);
    ihdr[10] = ihdr[11] = ihdr[12] = 0;
    chunk("IHDR", ihdr, 13);
    chunk("IDAT", z.data(), z.size());
    chunk("IEND", nullptr, 0);
    return out;
}

}  // namespace ghcm
