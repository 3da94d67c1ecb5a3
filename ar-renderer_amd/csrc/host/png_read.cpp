// PNG decoding for the scene loader's image textures, written from the PNG specification (W3C,
// second edition): chunk stream, zlib-compressed IDAT, the five per-row filters, Adam7
// interlacing and sub-byte sample packing. The reference reads PNG through stb_image asking for
// three 8-bit channels (src/core/image.cpp:90-101); the result here is that same 8-bit RGB:
// 16-bit samples keep their high byte, grey is replicated into R, G and B, palette indices are
// expanded, sub-byte grey is scaled to 0..255 by bit replication (1 bit x255, 2 bits x85,
// 4 bits x17) and alpha is dropped. Chunk CRCs are not checked and ancillary chunks (gAMA, tRNS,
// ...) are skipped, as stb_image does when asked for RGB.
#include "image_codecs.h"

#include <zlib.h>

#include <cstdint>
#include <cstring>

namespace {

uint32_t BE32(const unsigned char *p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}

int Paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Undoes one row's filter in place; prev is the previous unfiltered row of the same pass (or
// null for its first row), bpp the bytes of one whole pixel rounded up to at least one.
bool Unfilter(int type, unsigned char *row, const unsigned char *prev, size_t n, size_t bpp) {
    switch (type) {
    case 0: return true;
    case 1:
        for (size_t i = bpp; i < n; ++i) row[i] = (unsigned char)(row[i] + row[i - bpp]);
        return true;
    case 2:
        if (prev)
            for (size_t i = 0; i < n; ++i) row[i] = (unsigned char)(row[i] + prev[i]);
        return true;
    case 3:
        for (size_t i = 0; i < n; ++i) {
            const int a = i >= bpp ? row[i - bpp] : 0, b = prev ? prev[i] : 0;
            row[i] = (unsigned char)(row[i] + ((a + b) >> 1));
        }
        return true;
    case 4:
        for (size_t i = 0; i < n; ++i) {
            const int a = i >= bpp ? row[i - bpp] : 0, b = prev ? prev[i] : 0;
            const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            row[i] = (unsigned char)(row[i] + Paeth(a, b, c));
        }
        return true;
    default: return false;
    }
}

}  // namespace

bool DecodePNG(const std::vector<unsigned char> &file, int *w, int *h, std::vector<unsigned char> *rgb,
               std::string *err) {
    static const unsigned char kSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (file.size() < 8 || std::memcmp(file.data(), kSig, 8) != 0) {
        *err = "not a PNG file";
        return false;
    }
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> plte, idat;
    bool sawHeader = false, sawEnd = false;
    for (size_t p = 8; p + 12 <= file.size();) {
        const uint32_t len = BE32(&file[p]);
        if (len > file.size() - p - 12) {
            *err = "truncated PNG chunk";
            return false;
        }
        const unsigned char *type = &file[p + 4], *data = &file[p + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) {
                *err = "bad PNG header";
                return false;
            }
            W = BE32(data);
            H = BE32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
            if (data[10] != 0 || data[11] != 0 || interlace > 1) {
                *err = "bad PNG header (compression, filter or interlace method)";
                return false;
            }
            sawHeader = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            sawEnd = true;
            break;
        } else if (!(type[0] & 0x20)) {  // an unknown critical chunk
            *err = std::string("unsupported PNG chunk ") + std::string(reinterpret_cast<const char *>(type), 4);
            return false;
        }
        p += 12 + size_t(len);
    }
    if (!sawHeader || !sawEnd || idat.empty()) {
        *err = "truncated PNG (missing IHDR, IDAT or IEND)";
        return false;
    }
    int channels = 0;
    bool depthOK = false;
    switch (ctype) {
    case 0: channels = 1; depthOK = depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16; break;
    case 2: channels = 3; depthOK = depth == 8 || depth == 16; break;
    case 3: channels = 1; depthOK = depth == 1 || depth == 2 || depth == 4 || depth == 8; break;
    case 4: channels = 2; depthOK = depth == 8 || depth == 16; break;
    case 6: channels = 4; depthOK = depth == 8 || depth == 16; break;
    default: break;
    }
    if (!channels || !depthOK) {
        *err = "bad PNG colour type / bit depth";
        return false;
    }
    if (W == 0 || H == 0 || uint64_t(W) * H > (uint64_t(1) << 26)) {
        *err = "PNG size out of range";
        return false;
    }
    if (ctype == 3 && (plte.empty() || plte.size() % 3 != 0)) {
        *err = "PNG palette missing";
        return false;
    }
    const size_t bitsPerPixel = size_t(channels) * size_t(depth);
    const size_t bpp = bitsPerPixel >= 8 ? bitsPerPixel / 8 : 1;
    // Adam7 passes: (x0, y0, dx, dy); a non-interlaced image is one pass over every pixel.
    static const int kAdam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                     {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    static const int kWhole[1][4] = {{0, 0, 1, 1}};
    const int (*passes)[4] = interlace ? kAdam7 : kWhole;
    const int nPasses = interlace ? 7 : 1;
    size_t expected = 0;
    for (int k = 0; k < nPasses; ++k) {
        const uint64_t pw = (W + uint32_t(passes[k][2] - passes[k][0] - 1)) / uint32_t(passes[k][2]);
        const uint64_t ph = (H + uint32_t(passes[k][3] - passes[k][1] - 1)) / uint32_t(passes[k][3]);
        if (pw && ph) expected += size_t(ph) * (1 + (size_t(pw) * bitsPerPixel + 7) / 8);
    }
    std::vector<unsigned char> raw(expected);
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) {
        *err = "zlib initialisation failed";
        return false;
    }
    zs.next_in = idat.data();
    zs.avail_in = uInt(idat.size());
    zs.next_out = raw.data();
    zs.avail_out = uInt(raw.size());
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR && zr != Z_OK) || got != expected) {
        *err = "corrupt or truncated PNG image data";
        return false;
    }
    *w = int(W);
    *h = int(H);
    rgb->assign(size_t(W) * H * 3, 0);
    const int maxv = (1 << (depth < 8 ? depth : 8)) - 1;
    const int greyScale = ctype == 0 && depth < 8 ? 255 / maxv : 1;
    const size_t nPal = plte.size() / 3;
    size_t off = 0;
    for (int k = 0; k < nPasses; ++k) {
        const int x0 = passes[k][0], y0 = passes[k][1], dx = passes[k][2], dy = passes[k][3];
        const size_t pw = (W + uint32_t(dx - x0 - 1)) / uint32_t(dx);
        const size_t ph = (H + uint32_t(dy - y0 - 1)) / uint32_t(dy);
        if (!pw || !ph) continue;
        const size_t rowBytes = (pw * bitsPerPixel + 7) / 8;
        const unsigned char *prev = nullptr;
        for (size_t r = 0; r < ph; ++r, off += 1 + rowBytes) {
            unsigned char *row = &raw[off + 1];
            if (!Unfilter(raw[off], row, prev, rowBytes, bpp)) {
                *err = "bad PNG row filter";
                return false;
            }
            prev = row;
            const size_t y = size_t(y0) + r * size_t(dy);
            for (size_t c = 0; c < pw; ++c) {
                // the first sample's 8-bit value (high byte at 16 bits; bit-packed below 8)
                auto sample = [&](int s) -> int {
                    if (depth == 16) return row[(c * size_t(channels) + size_t(s)) * 2];
                    if (depth == 8) return row[c * size_t(channels) + size_t(s)];
                    const size_t bit = c * size_t(depth);
                    return (row[bit >> 3] >> (8 - depth - int(bit & 7))) & maxv;
                };
                unsigned char *o = &(*rgb)[3 * (y * W + size_t(x0) + c * size_t(dx))];
                if (ctype == 3) {
                    const size_t i = size_t(sample(0));
                    if (i < nPal) std::memcpy(o, &plte[3 * i], 3);
                } else if (channels >= 3) {
                    o[0] = (unsigned char)sample(0);
                    o[1] = (unsigned char)sample(1);
                    o[2] = (unsigned char)sample(2);
                } else {
                    o[0] = o[1] = o[2] = (unsigned char)(sample(0) * greyScale);
                }
            }
        }
    }
    return true;
}
