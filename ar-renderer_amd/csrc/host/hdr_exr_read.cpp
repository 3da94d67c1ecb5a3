// Radiance RGBE and OpenEXR decoding for the scene loader's environment maps and image textures,
// written from the formats' published descriptions (Ward's RGBE format in Graphics Gems II and the
// Radiance reference manual; the OpenEXR file layout document). The reference reads .hdr through
// stb_image's float loader and .exr through tinyexr's LoadEXR (src/core/image.cpp:79-89,
// src/core/exr_read.cpp:19-37); these restate what those produce for the files they accept:
// RGBE texels as c * 2^(e - 136), EXR channels R, G, B by name (a one-channel file is grey; a UINT
// channel's 32 bits are taken as a float, and a DECREASING_Y file is turned upside down, as
// LoadEXR returns them).
#include "image_codecs.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace {

// ---- Radiance RGBE ----------------------------------------------------------------------------

bool ReadLine(const std::vector<unsigned char> &f, size_t *p, std::string *line) {
    line->clear();
    while (*p < f.size() && f[*p] != '\n') line->push_back(char(f[(*p)++]));
    if (*p >= f.size()) return false;
    ++*p;
    return true;
}

void RgbeToFloat(const unsigned char *c, float *out) {
    if (c[3] == 0) {
        out[0] = out[1] = out[2] = 0;
        return;
    }
    const float f = std::ldexp(1.0f, int(c[3]) - (128 + 8));
    for (int k = 0; k < 3; ++k) out[k] = float(c[k]) * f;
}

// ---- OpenEXR ----------------------------------------------------------------------------------

uint32_t LE32(const unsigned char *p) {
    return uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16) | (uint32_t(p[3]) << 24);
}
uint64_t LE64(const unsigned char *p) { return uint64_t(LE32(p)) | (uint64_t(LE32(p + 4)) << 32); }

float HalfToFloat(uint16_t h) {
    const uint32_t s = uint32_t(h >> 15) << 31, e = (h >> 10) & 0x1f, m = h & 0x3ff;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = s;
        } else {  // subnormal: m * 2^-24
            const float v = std::ldexp(float(m), -24);
            std::memcpy(&bits, &v, 4);
            bits |= s;
        }
    } else if (e == 31) {
        bits = s | 0x7f800000u | (m << 13);
    } else {
        bits = s | ((e + 112) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

struct ExrChannel {
    std::string name;
    int type = 0;  // 0 UINT, 1 HALF, 2 FLOAT
    int xs = 1, ys = 1;
};

// The byte predictor and interleave that RLE and ZIP apply before compressing: undo the
// difference coding (each byte is the previous plus the stored value minus 128), then merge the
// two halves (even output bytes from the first, odd from the second).
void ExrUnpredict(std::vector<unsigned char> &t) {
    for (size_t i = 1; i < t.size(); ++i) t[i] = (unsigned char)(int(t[i - 1]) + int(t[i]) - 128);
    std::vector<unsigned char> out(t.size());
    const size_t half = (t.size() + 1) / 2;
    for (size_t i = 0; i < t.size(); ++i) out[i] = (i & 1) ? t[half + i / 2] : t[i / 2];
    t.swap(out);
}

bool ExrRle(const unsigned char *in, size_t n, std::vector<unsigned char> &out, size_t want) {
    out.clear();
    size_t p = 0;
    while (p < n && out.size() < want) {
        const int count = int(int8_t(in[p++]));
        if (count < 0) {
            if (p + size_t(-count) > n) return false;
            out.insert(out.end(), in + p, in + p - count);
            p += size_t(-count);
        } else {
            if (p >= n) return false;
            out.insert(out.end(), size_t(count) + 1, in[p++]);
        }
    }
    return out.size() == want;
}

}  // namespace

bool DecodeHDR(const std::vector<unsigned char> &f, int *w, int *h, std::vector<float> *rgb, std::string *err) {
    size_t p = 0;
    std::string line;
    if (!ReadLine(f, &p, &line) || (line != "#?RADIANCE" && line != "#?RGBE")) {
        *err = "not a Radiance HDR file";
        return false;
    }
    bool formatOK = true;
    while (ReadLine(f, &p, &line) && !line.empty())
        if (line.rfind("FORMAT=", 0) == 0 && line != "FORMAT=32-bit_rle_rgbe") formatOK = false;
    if (!formatOK) {
        *err = "unsupported Radiance HDR format (32-bit_rle_rgbe is read)";
        return false;
    }
    int H = 0, W = 0;
    char c1[3] = {0}, c2[3] = {0};
    if (!ReadLine(f, &p, &line) || std::sscanf(line.c_str(), "%2s %d %2s %d", c1, &H, c2, &W) != 4 ||
        std::strcmp(c1, "-Y") != 0 || std::strcmp(c2, "+X") != 0) {
        *err = "unsupported Radiance HDR layout (\"-Y h +X w\" is read)";
        return false;
    }
    if (W <= 0 || H <= 0 || int64_t(W) * H > (int64_t(1) << 26)) {
        *err = "Radiance HDR size out of range";
        return false;
    }
    rgb->assign(size_t(W) * size_t(H) * 3, 0.f);
    std::vector<unsigned char> scan(size_t(W) * 4);
    auto flat = [&](size_t from) {  // uncompressed RGBE texels for the rest of the image
        const size_t need = (size_t(W) * size_t(H) - from) * 4;
        if (f.size() - p < need) return false;
        for (size_t i = from; i < size_t(W) * size_t(H); ++i, p += 4) RgbeToFloat(&f[p], &(*rgb)[3 * i]);
        return true;
    };
    for (int y = 0; y < H; ++y) {
        const bool rle = W >= 8 && W < 32768 && p + 4 <= f.size() && f[p] == 2 && f[p + 1] == 2 && !(f[p + 2] & 0x80);
        if (!rle) {
            if (!flat(size_t(y) * size_t(W))) {
                *err = "truncated Radiance HDR data";
                return false;
            }
            break;
        }
        if ((int(f[p + 2]) << 8 | f[p + 3]) != W) {
            *err = "bad Radiance HDR scanline";
            return false;
        }
        p += 4;
        for (int c = 0; c < 4; ++c) {  // each component run-length coded over the scanline
            int x = 0;
            while (x < W) {
                if (p >= f.size()) {
                    *err = "truncated Radiance HDR data";
                    return false;
                }
                int count = f[p++];
                if (count > 128) {
                    count -= 128;
                    if (x + count > W || p >= f.size()) {
                        *err = "bad Radiance HDR run";
                        return false;
                    }
                    for (int k = 0; k < count; ++k) scan[size_t(x++) * 4 + size_t(c)] = f[p];
                    ++p;
                } else {
                    if (count == 0 || x + count > W || p + size_t(count) > f.size()) {
                        *err = "bad Radiance HDR run";
                        return false;
                    }
                    for (int k = 0; k < count; ++k) scan[size_t(x++) * 4 + size_t(c)] = f[p++];
                }
            }
        }
        for (int x = 0; x < W; ++x) RgbeToFloat(&scan[size_t(x) * 4], &(*rgb)[3 * (size_t(y) * size_t(W) + size_t(x))]);
    }
    *w = W;
    *h = H;
    return true;
}

bool DecodeEXR(const std::vector<unsigned char> &f, int *w, int *h, std::vector<float> *rgb, std::string *err) {
    if (f.size() < 8 || LE32(&f[0]) != 20000630u) {
        *err = "not an OpenEXR file";
        return false;
    }
    const uint32_t version = LE32(&f[4]);
    if ((version & 0xff) != 2 || (version & (0x200 | 0x800 | 0x1000))) {
        *err = "unsupported OpenEXR file (single-part scanline images are read, not tiled, deep or multi-part)";
        return false;
    }
    size_t p = 8;
    auto cstr = [&](std::string *s) {
        s->clear();
        while (p < f.size() && f[p]) s->push_back(char(f[p++]));
        if (p >= f.size()) return false;
        ++p;
        return true;
    };
    std::vector<ExrChannel> chans;
    int compression = -1, lineOrder = 0;
    int32_t box[4] = {0, 0, -1, -1};
    bool haveBox = false;
    for (;;) {
        std::string name, type;
        if (!cstr(&name)) {
            *err = "truncated OpenEXR header";
            return false;
        }
        if (name.empty()) break;
        if (!cstr(&type) || p + 4 > f.size()) {
            *err = "truncated OpenEXR header";
            return false;
        }
        const uint32_t size = LE32(&f[p]);
        p += 4;
        if (size > f.size() - p) {
            *err = "truncated OpenEXR header";
            return false;
        }
        const unsigned char *v = &f[p];
        if (name == "channels" && type == "chlist") {
            size_t q = 0;
            while (q < size && v[q]) {
                ExrChannel c;
                while (q < size && v[q]) c.name.push_back(char(v[q++]));
                if (q + 17 > size) {
                    *err = "bad OpenEXR channel list";
                    return false;
                }
                ++q;
                c.type = int(LE32(v + q));
                c.xs = int(LE32(v + q + 8));
                c.ys = int(LE32(v + q + 12));
                q += 16;
                chans.push_back(c);
            }
        } else if (name == "compression" && size >= 1) {
            compression = v[0];
        } else if (name == "lineOrder" && size >= 1) {
            lineOrder = v[0];
        } else if (name == "dataWindow" && type == "box2i" && size >= 16) {
            for (int k = 0; k < 4; ++k) box[k] = int32_t(LE32(v + 4 * k));
            haveBox = true;
        }
        p += size;
    }
    if (chans.empty() || !haveBox || compression < 0) {
        *err = "OpenEXR header lacks channels, compression or dataWindow";
        return false;
    }
    const int64_t W = int64_t(box[2]) - box[0] + 1, H = int64_t(box[3]) - box[1] + 1;
    if (W <= 0 || H <= 0 || W * H > (int64_t(1) << 26)) {
        *err = "OpenEXR data window out of range";
        return false;
    }
    size_t pixelBytes = 0;
    for (const ExrChannel &c : chans) {
        if (c.type < 0 || c.type > 2 || c.xs != 1 || c.ys != 1) {
            *err = "unsupported OpenEXR channel '" + c.name + "' (subsampled or of unknown type)";
            return false;
        }
        pixelBytes += c.type == 1 ? 2 : 4;
    }
    int lines;
    switch (compression) {
    case 0: case 1: case 2: lines = 1; break;  // NONE, RLE, ZIPS
    case 3: lines = 16; break;                 // ZIP
    default:
        *err = "unsupported OpenEXR compression " + std::to_string(compression) + " (NONE, RLE, ZIPS and ZIP are read)";
        return false;
    }
    // R, G, B by name, else the one channel as grey
    int src[3] = {-1, -1, -1};
    for (size_t i = 0; i < chans.size(); ++i) {
        if (chans[i].name == "R") src[0] = int(i);
        if (chans[i].name == "G") src[1] = int(i);
        if (chans[i].name == "B") src[2] = int(i);
    }
    if (chans.size() == 1) src[0] = src[1] = src[2] = 0;
    if (src[0] < 0 || src[1] < 0 || src[2] < 0) {
        *err = "OpenEXR file without R, G and B channels";
        return false;
    }
    const int64_t chunks = (H + lines - 1) / lines;
    if (p + size_t(chunks) * 8 > f.size()) {
        *err = "truncated OpenEXR offset table";
        return false;
    }
    rgb->assign(size_t(W) * size_t(H) * 3, 0.f);
    std::vector<unsigned char> block, tmp;
    for (int64_t k = 0; k < chunks; ++k) {
        const uint64_t off = LE64(&f[p + size_t(k) * 8]);
        if (off + 8 > f.size()) {
            *err = "bad OpenEXR chunk offset";
            return false;
        }
        const int32_t y0 = int32_t(LE32(&f[off]));
        const uint32_t n = LE32(&f[off + 4]);
        if (n > f.size() - off - 8 || y0 < box[1] || y0 > box[3]) {
            *err = "bad OpenEXR chunk";
            return false;
        }
        const int rows = int(std::min<int64_t>(lines, int64_t(box[3]) - y0 + 1));
        const size_t want = size_t(rows) * size_t(W) * pixelBytes;
        const unsigned char *data = &f[off + 8];
        if (n == want) {  // stored uncompressed when compression did not shrink the block
            block.assign(data, data + n);
        } else if (compression == 1) {
            if (!ExrRle(data, n, tmp, want)) {
                *err = "corrupt OpenEXR RLE block";
                return false;
            }
            block.swap(tmp);
            ExrUnpredict(block);
        } else if (compression == 2 || compression == 3) {
            block.assign(want, 0);
            uLongf got = uLongf(want);
            if (uncompress(block.data(), &got, data, uLong(n)) != Z_OK || got != want) {
                *err = "corrupt OpenEXR ZIP block";
                return false;
            }
            ExrUnpredict(block);
        } else {
            *err = "bad OpenEXR block size";
            return false;
        }
        // per scanline: each channel's W samples in turn
        size_t q = 0;
        for (int r = 0; r < rows; ++r) {
            // LoadEXR stores a DECREASING_Y file's rows bottom up, so the reference sees it flipped
            const size_t yr = size_t(y0 - box[1] + r), y = lineOrder == 1 ? size_t(H - 1) - yr : yr;
            for (size_t c = 0; c < chans.size(); ++c) {
                const int t = chans[c].type;
                for (int64_t x = 0; x < W; ++x) {
                    float v;
                    if (t == 1) {
                        v = HalfToFloat(uint16_t(block[q] | (block[q + 1] << 8)));
                        q += 2;
                    } else {  // FLOAT, and UINT as LoadEXR returns it: its 32 bits taken as a float
                        const uint32_t b = LE32(&block[q]);
                        std::memcpy(&v, &b, 4);
                        q += 4;
                    }
                    for (int o = 0; o < 3; ++o)
                        if (src[o] == int(c)) (*rgb)[3 * (y * size_t(W) + size_t(x)) + size_t(o)] = v;
                }
            }
        }
    }
    *w = int(W);
    *h = int(H);
    return true;
}
